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
from tests.helpers import o_bparams, o_iparams, o_model, rel_l2

pytestmark = pytest.mark.gpu

# the bench's span schedule: spans of 2, 4 and 8 table intervals from 0, 800
# and 1600 m down
BENCH_SPANS = ((0.0, 2), (800.0, 4), (1600.0, 8))


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


def _shot_list(ranges, tans):
    return engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges])


def test_fullsize_fast_form_against_fp64_and_surrogate_gain(full):
    """The surrogate variant at full size: a surrogate trained on the device
    (6000 Adam steps on 8M samples), judged (0.5 ms) and used for one
    2000-receiver shot's source table.
    * The fast form -- FP16 bank, U16Q tables, node form below 400 m, the
      shot's own source table -- against the FP64 oracle fed that same
      source table, on 32 traces of the shot over the whole volume: <= 1e-3.
    * Gain over the table methods without the network: the source leg's mean
      |dt| against FP64 rays at the true source below 0.6x that of the
      nearest node's table and of the four node tables blended, and the
      whole shot's image error against the exact-source image below 0.75x
      both."""
    cfg, sv, tab, geom = full
    shots = synth.shots_config2(nshots=1)
    g1, ranges = synth.geometry(shots)
    n = ranges[0][1]
    tans = cfg.dip_bounds()[:4]
    g_d = torch.from_numpy(g1).cuda()
    x = synth.reflectivity_traces_gpu(cfg, n, seed=77)
    sur = engine.Surrogate.for_survey(cfg, out_scale=(0.02, 0.004, 0.05))
    X, T = sur.samples(8_000_000, seed=0x5EED5)
    sur.train(X, T, steps=6000, batch=65536, lr0=4e-3, lr1=4e-5)
    del X, T
    mask, _ = sur.mask(sv.tg, tab, tol=0.5e-3)
    sl = _shot_list(ranges, tans)
    src = sur.src_tables(sv, g_d, sl, tab=tab, mask=mask)
    nus, nvs, nw = sur.src_grid(sv)
    hdr = src[:16].view(torch.float32).cpu().numpy()
    body = src[16:].view(torch.float32).view(nvs, nus, nw, 2)
    # fast form (in the span form, the bench's, and in the node form) vs the
    # FP64 oracle fed the same source table
    m = 32
    sub = np.ascontiguousarray(g1[:m])
    ref, _ = O.image(O.qbank(x[:m].cpu().numpy(), o_bparams(sv)), sub, tab.cpu().numpy(), o_iparams(sv),
                     src=(hdr, body.cpu().numpy()))
    for form, launch in (("spans", (8, 8, 4)), ("nodes", (8, 8, 4))):
        svm = engine.Survey(cfg, sv.tstar_max).set_form(form, 400.0, BENCH_SPANS)
        q = svm.quantize_tables(tab, code="u16q")
        img = svm.new_image()
        bank = svm.qbank(x[:m].contiguous(), dtype=torch.float16, pair=svm.pair_bank())
        svm.image_batch(img, torch.from_numpy(sub).cuda(), bank, q, _shot_list([(0, m)], tans), launch=launch,
                        nwg=-1, src=src)
        e = rel_l2(img.cpu().numpy(), ref)
        print(f"fast form ({form}) vs FP64 oracle with its own source table: {e:.2e}")
        assert e <= 1e-3
    # gain over the table methods
    exact = src.clone()
    engine.build_tables_into(cfg.model, dict(cfg.tgeom, nnx=1, nny=1, n0x=float(hdr[0]), n0y=float(hdr[1]), nu=nus,
                                             nv=nvs, u0=float(hdr[2]), v0=float(hdr[3])),
                             exact[16:].view(torch.float32))
    blend = sv.src_tables_nodes(g_d, sl, tab)
    tx = exact[16:].view(torch.float32).view(nvs, nus, nw, 2)[..., 0]
    g = cfg.tgeom
    iu0 = int(round((hdr[2] - g["u0"]) / g["du"]))
    iv0 = int(round((hdr[3] - g["v0"]) / g["dv"]))
    i = min(max(int(np.floor((hdr[0] - g["n0x"]) / g["ndx"] + 0.5)), 0), g["nnx"] - 1)
    j = min(max(int(np.floor((hdr[1] - g["n0y"]) / g["ndy"] + 0.5)), 0), g["nny"] - 1)
    near = tab[j * g["nnx"] + i, iv0:iv0 + nvs, iu0:iu0 + nus, :, 0]
    es = float((body[..., 0] - tx).abs().mean())
    eb = float((blend[16:].view(torch.float32).view(nvs, nus, nw, 2)[..., 0] - tx).abs().mean())
    en = float((near - tx).abs().mean())
    print(f"source leg mean |dt|: surrogate {es * 1e3:.3f} ms, node blend {eb * 1e3:.3f} ms, nearest node "
          f"{en * 1e3:.3f} ms; bricks falling back {100 * (1 - float(mask.float().mean())):.1f} %")
    assert es < 0.6 * min(eb, en)
    bank = sv.qbank(x)
    imgs = {}
    for name, s in (("exact", exact), ("surrogate", src), ("blend", blend), ("nearest", None)):
        im = sv.new_image()
        sv.image_batch(im, g_d, bank, tab, sl, src=s)
        imgs[name] = im.double()
    err = {k: float((v - imgs["exact"]).norm() / imgs["exact"].norm()) for k, v in imgs.items() if k != "exact"}
    print(f"whole shot, image rel-L2 vs exact source legs: {err}")
    assert err["surrogate"] < 0.75 * min(err["blend"], err["nearest"])


@pytest.mark.parametrize("form", ["exact", "nodes", "spans"])
def test_fullsize_config4_graph_replays_equal_eager(full, form):
    """Four shot gathers of the batching variant (uneven trace counts,
    per-shot apertures) over the full image through a captured graph of
    [Q bank -> batched imaging]: each of three replays, with eager work and
    synchronisations between them, equals the eager batched call bit for
    bit; in the node and span forms too."""
    cfg, sv0, tab, _ = full
    sv = engine.Survey(cfg, sv0.tstar_max).set_form(form, 400.0, BENCH_SPANS)
    shots = synth.shots_config4(nshots=4, seed=0x5EED0044)
    geom, ranges = synth.geometry(shots)
    desc = [(tr0, n) + synth.dip_tans(a) for (tr0, n), (*_, a) in zip(ranges, shots)]
    x = synth.reflectivity_traces_gpu(cfg, geom.shape[0], seed=91)
    g_d = torch.from_numpy(geom).cuda()
    q = sv.quantize_tables(tab, code="u16q")
    sl = engine.ShotList(desc)
    launch = (8, 8, 4) if form != "exact" else (8, 8, 8)
    eager = sv.new_image()
    sv.image_batch(eager, g_d, sv.qbank(x, dtype=torch.float16, pair=sv.pair_bank()), q, sl, launch=launch, nwg=-1)
    img = sv.new_image()
    bank = torch.empty(sv.bank_shape(x.shape[0], torch.float16, pair=sv.pair_bank()), dtype=torch.float16,
                       device="cuda")
    gr = engine.BatchGraph(sv, x, bank, img, g_d, q, sl, launch=launch, nwg=-1, stages=5)
    try:
        for _ in range(3):
            img.zero_()
            torch.cuda.synchronize()
            gr.launch()
            torch.cuda.synchronize()
            assert torch.equal(img, eager)
            _ = (eager * 2).sum().item()   # eager work between replays
    finally:
        gr.close()
    print(f"{form}: gathers of {[n for _, n in ranges]} traces, apertures {[round(s[3], 1) for s in shots]} deg")
    assert float(eager.abs().sum()) > 0


def _lossless(model):
    return dict(model, q0=1e12, ql=[1e12] * len(model["ql"]))


@pytest.mark.parametrize("mk", [synth.config1, synth.config2])
def test_config2_rays_match_oracle(gpu, mk):
    """The device ray integrals (qpsdm_rays) against the oracle's FP64 rays
    through config 1's v(z) (closed form) and config 2's layers and
    anomaly (quadrature)."""
    cfg = mk()
    rng = np.random.default_rng(12)
    pts = np.stack([rng.uniform(0, 4000, 64), rng.uniform(0, 4000, 64), rng.uniform(0, 4000, 64),
                    rng.uniform(0, 4000, 64), rng.uniform(10, 3000, 64)], 1)
    got = engine.rays(cfg.model, pts).cpu().numpy()
    m = o_model(cfg.model)
    for i, (sx, sy, x, y, z) in enumerate(pts):
        t, ts = O.ray_times(m, sx, sy, x, y, z)
        sg = O.ray_sigma(m, sx, sy, x, y, z)
        assert abs(got[i, 0] - t) <= 1e-11 * t and abs(got[i, 1] - ts) <= 1e-11 * ts, i
        assert abs(got[i, 2] - sg) <= 1e-11 * sg, i


def _exact_src(sv, model, g_d, sl, tab):
    """Per-shot source tables of the model at each shot's own source: the
    headers of qpsdm_src_tables_nodes, bodies rebuilt by the table builder
    for one node at (sx, sy) on the same relative sub-grid."""
    cfg, g = sv.cfg, sv.cfg.tgeom
    src = sv.src_tables_nodes(g_d, sl, tab)
    n = sl.n
    hdr = src[:16 * n].view(torch.float32).view(n, 4).cpu().numpy()
    body = src[16 * n:].view(torch.float32).view(n, -1)
    nus, nvs, _ = sv.src_grid()
    for k in range(n):
        engine.build_tables_into(model, dict(g, nnx=1, nny=1, n0x=float(hdr[k, 0]), n0y=float(hdr[k, 1]), nu=nus,
                                             nv=nvs, u0=float(hdr[k, 2]), v0=float(hdr[k, 3])), body[k])
    return src


def test_config2_q_compensation_and_focus(gpu):
    """Config 2 at full size (401 x 401 x 301 at 10 m, layers Q 40..200 with
    the Gaussian anomaly of Q x 0.5 and v x 0.8 at its centre, 25 m tables
    on 5 x 5 nodes, K = 32): three point scatterers -- one under the
    anomaly's centre -- recorded through the model by the FP64 ray
    integrals on the 65 shots (2000 receivers each) over the middle of the
    survey, imaged with each shot's own source leg.  Imaged through the
    bank, each focuses within one cell of its point and its peak is 0.85-1.02
    of the same scatterers recorded without attenuation and imaged at
    t* = 0 (tables of the lossless model); the gain cap (40 dB) bounds the
    recovery of the highest frequencies, more so under the anomaly.  Imaged
    without compensation, the peak is below 0.2 of the lossless one."""
    cfg = synth.config2()
    scat = np.array([[1500.0, 1500.0, 1000.0], [2000.0, 2000.0, 2000.0], [2500.0, 2300.0, 1500.0]])
    max_rel = (0.15, 0.3, 0.15)
    shots = [s for s in synth.shots_config2() if abs(s[0] - 2000.0) <= 1300 and abs(s[1] - 2000.0) <= 900]
    geom, ranges = synth.geometry(shots)
    tab = engine.Survey(cfg, 0.05).build_tables()
    torch.cuda.synchronize()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab))
    m0 = _lossless(cfg.model)
    tab0 = engine.build_tables_into(m0, cfg.tgeom, torch.empty_like(tab))
    att = synth.diffraction_traces_rays(cfg, geom, scat)
    clean_cfg = synth.config2()
    clean_cfg.model = m0
    clean = synth.diffraction_traces_rays(clean_cfg, geom, scat)
    g_d = torch.from_numpy(geom).cuda()
    sl = _shot_list(ranges, cfg.dip_bounds()[:4])
    src, src0 = _exact_src(sv, cfg.model, g_d, sl, tab), _exact_src(sv, m0, g_d, sl, tab0)

    def image(x, t, s):
        img = sv.new_image()
        sv.image_batch(img, g_d, sv.qbank(x), t, sl, src=s)
        torch.cuda.synchronize()
        return img.cpu().numpy().astype(np.float64)

    got, ideal, raw = image(att, tab, src), image(clean, tab0, src0), image(att, tab0, src0)
    for k, (x, y, z) in enumerate(scat):
        ix, iy, iz = int(round(x / cfg.dx)), int(round(y / cfg.dy)), int(round(z / cfg.dz))
        w = (slice(iy - 5, iy + 6), slice(ix - 5, ix + 6), slice(iz - 5, iz + 6))
        a, b, c = got[w], ideal[w], raw[w]
        pk = np.array(np.unravel_index(np.argmax(np.abs(a)), a.shape)) - 5
        pk0 = np.array(np.unravel_index(np.argmax(np.abs(b)), b.shape)) - 5
        ratio, e = np.abs(a).max() / np.abs(b).max(), rel_l2(a, b)
        raw_ratio = np.abs(c).max() / np.abs(b).max()
        print(f"scatterer {k} at {(x, y, z)}: peak offset {pk} (lossless {pk0}) (y, x, z); peak / lossless "
              f"{ratio:.3f}, 11^3 window rel-L2 {e:.3f}; uncompensated peak / lossless {raw_ratio:.3f}")
        assert np.all(np.abs(pk) <= 1) and np.all(np.abs(pk0) <= 1)
        assert 0.85 < ratio < 1.02 and e < max_rel[k]
        assert raw_ratio < 0.2
