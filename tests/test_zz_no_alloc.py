"""The Q bank entry point that allocates nothing (qpsdm_qbank_noalloc),
called through the C ABI, against qpsdm_qbank on the same traces: two calls
in a row, the second smaller (fewer, shorter traces, another parameter set,
so the device coefficient array is rewritten between them)."""
import ctypes as C

import numpy as np
import pytest
import torch

from q_psdm_amd import engine, synth
from q_psdm_amd import lib as L

pytestmark = pytest.mark.gpu


def _noalloc(sv, traces, dtype):
    out = torch.empty(sv.bank_shape(traces.shape[0], dtype), dtype=dtype, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(L.lib().qpsdm_qbank_noalloc(C.byref(sv.bp), C.c_void_p(traces.data_ptr()), traces.shape[0],
                                        C.c_void_p(out.data_ptr()), sv._BANK_CODES[dtype], stream))
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16])
def test_noalloc_bank_matches_qbank(gpu, dtype):
    for cfg, ntr in ((synth.config5(), 64), (synth.config1(), 9)):
        sv = engine.Survey(cfg, 0.05)
        rng = np.random.default_rng(ntr)
        x = torch.from_numpy(rng.standard_normal((ntr, cfg.nt)).astype(np.float32)).cuda()
        ref = sv.qbank(x, dtype=dtype)
        got = _noalloc(sv, x, dtype)
        torch.cuda.synchronize()
        a = sv.decode_bank(ref).cpu().numpy()
        b = sv.decode_bank(got).cpu().numpy()
        scale = np.abs(a).max()
        # the same FP64 coefficient formulas evaluated by the device's math
        # library instead of the host's: last-place differences of the
        # FP32-rounded coefficients, nothing more
        tol = 1e-3 if dtype == torch.float16 else 2e-6
        assert np.abs(a - b).max() <= tol * scale
        assert np.linalg.norm(a - b) <= (5e-4 if dtype == torch.float16 else 1e-6) * np.linalg.norm(a)


def test_noalloc_rejects_windows_beyond_its_array(gpu):
    cfg = synth.config1()
    sv = engine.Survey(cfg, 0.05)
    bp = L.BankParams()
    C.memmove(C.byref(bp), C.byref(sv.bp), C.sizeof(bp))
    bp.K = 1000   # K * nww beyond 131072
    x = torch.zeros((2, cfg.nt), device="cuda")
    out = torch.empty((2, bp.K, cfg.nt), device="cuda")
    rc = L.lib().qpsdm_qbank_noalloc(C.byref(bp), C.c_void_p(x.data_ptr()), 2, C.c_void_p(out.data_ptr()),
                                     L.QPSDM_F32, None)
    assert rc == -4   # QPSDM_ERANGE
    assert b"scratch" in L.lib().qpsdm_last_error()
