"""Parity at the benchmark's full sizes: the 401 x 401 x 301 image at 10 m,
2000-receiver shots over 4 x 4 km, nt = 2001 at 2 ms, K = 32, 25 m tables on
5 x 5 nodes, the layered + Gaussian-anomaly model.  The oracle images the
whole volume for a subset of a shot's traces; size-independent properties
(linearity, bit equality of the batched and per-shot kernels over uneven
gathers with per-shot apertures) cover whole shots."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from q_psdm_amd import engine, synth
from tests.helpers import o_bparams, o_iparams, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full(gpu):
    cfg = synth.config5()
    probe = engine.Survey(cfg, 0.05)
    tab = probe.build_tables()
    torch.cuda.synchronize()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab))
    geom, ranges = synth.geometry(synth.shots_config2(nshots=1))
    return cfg, sv, tab, geom


def test_fullsize_qbank_matches_fp64_oracle(full):
    cfg, sv, tab, geom = full
    x = synth.reflectivity_traces(cfg, 6, seed=(cfg.seed, 7))
    tr = torch.from_numpy(x).cuda()
    ref = O.qbank(x, o_bparams(sv))
    b32 = sv.qbank(tr).cpu().numpy()
    b16 = sv.decode_bank(sv.qbank(tr, dtype=torch.float16)).cpu().numpy()
    for i in range(x.shape[0]):
        for j in range(0, cfg.K, 5):
            assert rel_l2(b32[i, j], ref[i, j]) < 2e-5, (i, j)
        assert rel_l2(b16[i], ref[i]) < 5e-4, i


def test_fullsize_fp32_image_matches_oracle(full):
    """Twelve traces spread over one 2000-receiver shot, imaged into the
    whole volume: the FP32 GPU form against the oracle's FP32 path (same
    operation order) and its FP64 path, and the update count."""
    cfg, sv, tab, geom = full
    sel = np.ascontiguousarray(geom[::167][:12])
    x = synth.reflectivity_traces(cfg, sel.shape[0], seed=(cfg.seed, 11))
    g_d = torch.from_numpy(sel).cuda()
    bank = sv.qbank(torch.from_numpy(x).cuda())
    img = sv.new_image()
    sv.image_shot(img, g_d, 0, sel.shape[0], bank, tab)
    cnt = sv.count_shot(g_d, 0, sel.shape[0])
    b = bank.cpu().numpy()
    t = tab.cpu().numpy()
    ref32, n32 = O.image(b, sel, t, o_iparams(sv), fp32=True)
    ref64, n64 = O.image(b.astype(np.float64), sel, t, o_iparams(sv))
    got = img.cpu().numpy()
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print(f"full volume, 12 traces: rel-L2 vs CPU FP32 {e32:.2e}, vs FP64 {e64:.2e}; updates {cnt} vs {n64}")
    assert e32 < 1e-5
    assert e64 < 1e-4
    assert abs(cnt - n64) <= 1e-5 * n64 + 2


def test_fullsize_linearity(full):
    """image(2a - 3b) = 2 image(a) - 3 image(b) over the whole volume for 64
    traces of a shot (the bank and the sum are linear; the bank level of a
    point does not depend on the data)."""
    cfg, sv, tab, geom = full
    sel = np.ascontiguousarray(geom[:64])
    g_d = torch.from_numpy(sel).cuda()
    a = torch.from_numpy(synth.reflectivity_traces(cfg, 64, seed=(cfg.seed, 21))).cuda()
    b = torch.from_numpy(synth.reflectivity_traces(cfg, 64, seed=(cfg.seed, 22))).cuda()

    def image(x):
        img = sv.new_image()
        sv.image_shot(img, g_d, 0, 64, sv.qbank(x), tab)
        return img.double()

    lhs = image(2.0 * a - 3.0 * b)
    rhs = 2.0 * image(a) - 3.0 * image(b)
    e = float((lhs - rhs).norm() / rhs.norm())
    print(f"linearity over the full volume: rel-L2 {e:.2e}")
    assert e < 1e-5


def test_fullsize_batched_equals_per_shot_uneven_gathers(full):
    """Three shot gathers of the batching variant (log-uniform trace counts,
    per-shot apertures 25-60 deg) over the full image, FP16 bank and U16X
    tables: the batched kernel (persistent and one group per tile) equals
    consecutive per-shot launches bit for bit."""
    cfg, sv, tab, _ = full
    shots = synth.shots_config4(nshots=3)
    geom, ranges = synth.geometry(shots)
    desc = [(tr0, n) + synth.dip_tans(a) for (tr0, n), (*_, a) in zip(ranges, shots)]
    x = torch.from_numpy(synth.reflectivity_traces(cfg, geom.shape[0], seed=(cfg.seed, 31))).cuda()
    g_d = torch.from_numpy(geom).cuda()
    bank = sv.qbank(x, dtype=torch.float16)
    q = sv.quantize_tables(tab, code="u16x")
    ref = sv.new_image()
    keep = (sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2)
    try:
        for tr0, n, *tans in desc:
            sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2 = tans
            sv.image_shot(ref, g_d, tr0, n, bank[tr0:tr0 + n], q, launch=(8, 8, 8, 0))
    finally:
        sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2 = keep
    sl = engine.ShotList(desc)
    print(f"gathers of {[n for _, n in ranges]} traces, apertures {[round(s[3], 1) for s in shots]} deg")
    for nwg in (0, -1):
        img = sv.new_image()
        sv.image_batch(img, g_d, bank, q, sl, launch=(8, 8, 8, 0), nwg=nwg)
        torch.cuda.synchronize()
        assert torch.equal(img, ref), nwg
    assert float(ref.abs().sum()) > 0


def test_fullsize_mixed_form_within_budget(full):
    """256 traces of a shot over the full volume: the mixed form (FP16 bank,
    U16X tables) against the FP32 form stays within 1e-3 rel-L2; a BF16 bank
    on the same data does not (its 8-bit significand)."""
    cfg, sv, tab, geom = full
    sel = np.ascontiguousarray(geom[:256])
    g_d = torch.from_numpy(sel).cuda()
    x = torch.from_numpy(synth.reflectivity_traces(cfg, 256, seed=(cfg.seed, 41))).cuda()
    q = sv.quantize_tables(tab, code="u16x")
    imgs = {}
    for name, dt, t in (("fp32", torch.float32, tab), ("f16_u16x", torch.float16, q),
                        ("bf16_u16x", torch.bfloat16, q)):
        img = sv.new_image()
        sv.image_shot(img, g_d, 0, 256, sv.qbank(x, dtype=dt), t)
        imgs[name] = img.double()
    e = {k: float((v - imgs["fp32"]).norm() / imgs["fp32"].norm()) for k, v in imgs.items() if k != "fp32"}
    print(f"full volume, 256 traces, rel-L2 vs FP32: {e}")
    assert e["f16_u16x"] < 1e-3
    assert e["bf16_u16x"] > e["f16_u16x"]
