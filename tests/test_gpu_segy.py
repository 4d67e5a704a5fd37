"""SEG-Y rev0 on the GPU: the device decode against the host codec (itself
pinned to the reference's reader on the CPU) and the oracle, bit for bit,
for IBM and IEEE samples, moved header keys, special IBM words and a
full-size 2000-trace shot; the pipeline's SEG-Y intake against its SU intake
on the same traces (identical images)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from q_psdm_amd import engine, segy
from tests.helpers import segy_records
from tests.test_gpu_pipeline import _direct, _gp, _setup

pytestmark = pytest.mark.gpu

MOVES = {"sx": (181, 4), "sy": (185, 4), "gx": (189, 2), "gy": (191, 2), "cdp": (197, 4), "scalco": (201, 4),
         "dt": (205, 4), "offset": (29, 2)}


def _dev(rec):
    return torch.from_numpy(np.ascontiguousarray(rec).reshape(-1)).cuda()


def _special_words(rng, n):
    w = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    special = np.array([0, 0x80000000, 0x7FFFFFFF, 0xFFFFFFFF, 0x00000001, 0x7F000001, 0x00FFFFFF, 0x41100000,
                        0xC276A000, 0x20100000, 0x21100000, 0x61100000, 0x62100000, 0x1F000001], np.uint32)
    m = min(n, special.size)
    w[:m] = special[:m]
    return w


@pytest.mark.parametrize("fmt", [segy.IBM, segy.IEEE])
@pytest.mark.parametrize("moved", [False, True])
@pytest.mark.parametrize("ntr,ns", [(1, 1), (7, 33), (50, 1001)])
def test_device_decode_matches_host_and_oracle(gpu, fmt, moved, ntr, ns):
    rng = np.random.default_rng(ntr * 7 + ns + fmt + 2 * moved)
    keys = MOVES if moved else None
    k = segy.keyindex(**(keys or {}))
    ints = rng.integers(-2 ** 31, 2 ** 31, (ntr, 8))
    fields = {"ns": [ns] * ntr, "sx": ints[:, 0], "sy": ints[:, 1], "gx": ints[:, 2], "gy": ints[:, 3],
              "cdp": ints[:, 4], "offset": ints[:, 5], "scalco": ints[:, 6], "dt": ints[:, 7], "tracl": ints[:, 0] // 3,
              "fldr": ints[:, 1] // 5}
    rec = segy_records(np.zeros((ntr, ns), np.float32), fields, 5, keys=keys, rng=rng)
    w = _special_words(rng, ntr * ns)
    rec[:, 240:] = w.reshape(ntr, ns).astype(">u4").view(np.uint8).reshape(ntr, 4 * ns)
    host_s, host_h = segy.decode(rec, ns, fmt, k)
    print(f"decode fmt={fmt} moved={moved} ntr={ntr} ns={ns}", flush=True)
    dev_s, dev_h = segy.decode_dev(_dev(rec), ntr, ns, fmt, k)
    torch.cuda.synchronize()
    assert np.array_equal(dev_s.cpu().numpy().view(np.uint32), host_s.view(np.uint32))
    assert np.array_equal(segy.headers_view(dev_h), host_h)
    for i in range(min(ntr, 3)):  # the checker
        o = O.segy_samples(rec[i].tobytes(), ns, fmt)
        assert np.array_equal(dev_s[i].cpu().numpy().view(np.uint32), o.view(np.uint32))
    assert np.all(host_h["ns"] == ns)


def test_device_decode_full_shot(gpu):
    """One 2000-trace shot of 2001 samples (16.5 MB of IBM records): device
    decode equals the host decode; values of IBM-exact traces come back
    exactly."""
    rng = np.random.default_rng(11)
    ntr, ns = 2000, 2001
    x = (rng.standard_normal((ntr, ns)) * 1e-3).astype(np.float32)
    x = (x.view(np.uint32) & np.uint32(0xFFFFFFF8)).view(np.float32)
    rec = segy_records(x, {"ns": [ns] * ntr, "sx": np.arange(ntr), "gy": -np.arange(ntr)}, segy.IBM)
    dev_s, dev_h = segy.decode_dev(_dev(rec), ntr, ns, segy.IBM)
    host_s, host_h = segy.decode(rec, ns, segy.IBM)
    torch.cuda.synchronize()
    assert np.array_equal(dev_s.cpu().numpy(), x)
    assert np.array_equal(host_s, x)
    assert np.array_equal(segy.headers_view(dev_h), host_h)
    # no headers requested, empty batch
    s2, h2 = segy.decode_dev(_dev(rec), ntr, ns, segy.IBM, headers=False)
    assert h2 is None and torch.equal(s2, dev_s)
    s0, _ = segy.decode_dev(_dev(rec), 0, ns, segy.IBM)
    assert s0.shape == (0, ns)


@pytest.mark.parametrize("fmt,moved", [(segy.IBM, False), (segy.IEEE, True), (segy.IBM, True)])
def test_segy_pipeline_matches_su_pipeline(gpu, fmt, moved):
    """The same traces (IBM-exact samples) and integer coordinates pushed as
    SU records and as SEG-Y records (keys moved or not): identical images,
    and equal to the direct kernel path."""
    cfg, sv, tab, wxy, rel, traces = _setup()
    traces = (traces.view(np.uint32) & np.uint32(0xFFFFFFF8)).view(np.float32)
    ref = _direct(sv, tab, rel, traces)
    keys = {"sx": (181, 4), "sy": (185, 4), "gx": (189, 4), "gy": (193, 4)} if moved else None
    ntr = traces.shape[0]
    rec = segy_records(traces, {"ns": [cfg.nt] * ntr, "sx": wxy[:, 0], "sy": wxy[:, 1], "gx": wxy[:, 2],
                                "gy": wxy[:, 3]}, fmt, keys=keys)
    pl = engine.Pipeline(sv, _gp(cfg), batch_traces=6)
    pl.set_tables(tab)
    data = rec.tobytes()
    n = len(data) // 2 // rec.shape[1] * rec.shape[1]
    pl.push_segy(data[:n], fmt, segy.keyindex(**(keys or {})))
    pl.push_segy(data[n:], fmt, segy.keyindex(**(keys or {})))
    got = pl.image()
    pl.close()
    pl2 = engine.Pipeline(sv, _gp(cfg), batch_traces=6)
    pl2.set_tables(tab)
    pl2.push_su(engine.su_encode(traces, wxy))
    su_img = pl2.image()
    pl2.close()
    np.testing.assert_array_equal(got, su_img)
    np.testing.assert_array_equal(got, ref)


def test_segy_pipeline_rejects_bad_records(gpu):
    cfg, sv, tab, wxy, rel, traces = _setup()
    pl = engine.Pipeline(sv, _gp(cfg), batch_traces=6)
    pl.set_tables(tab)
    rec = segy_records(traces[:2], {"ns": [cfg.nt, cfg.nt + 1]}, segy.IBM)
    with pytest.raises(Exception):
        pl.push_segy(rec.tobytes(), segy.IBM)
    with pytest.raises(Exception):
        pl.push_segy(rec[:1].tobytes(), 3)
    with pytest.raises(Exception):
        pl.push_segy(rec[:1].tobytes()[:-4], segy.IBM)
    pl.close()
