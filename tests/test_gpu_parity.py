"""GPU parity of the HIP path against the CPU oracle (same seeded inputs)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from q_psdm_amd import engine, synth
from q_psdm_amd import lib as L
from tests.helpers import o_bparams, o_iparams, o_model, o_tgeom, o_window, rel_l2

pytestmark = pytest.mark.gpu


def _survey(cfg):
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    return engine.Survey(cfg, engine.tstar_max_from_tables(tab)), tab


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
def test_tables_match_oracle(gpu, mk):
    cfg = mk()
    sv, tab = _survey(cfg)
    ref = O.build_tables(o_model(cfg.model), o_tgeom(cfg.tgeom))
    got = tab.cpu().numpy()
    # FP64 quadrature on both sides, stored FP32: equal to a few ulp
    np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("nt,f1", [(1001, 1.0), (2001, 1.0), (1250, 1.0)])
def test_trace_spectra_match_reference_preprocessing(gpu, nt, f1):
    cfg = synth.config1()
    cfg.nt = nt
    cfg.dt = 0.004 if nt == 1250 else 0.002
    sv = engine.Survey(cfg, 0.02)
    rng = np.random.default_rng(7)
    traces = rng.standard_normal((5, nt)).astype(np.float32)
    re, im = sv.trace_spectra(torch.from_numpy(traces).cuda())
    for i in range(5):
        ore, oim = O.trace_spectra(traces[i], cfg.dt, o_window(sv.win))
        scale = np.abs(np.concatenate([ore, oim])).max()
        assert np.abs(re[i].cpu().numpy() - ore).max() < 2e-5 * scale
        assert np.abs(im[i].cpu().numpy() - oim).max() < 2e-5 * scale


@pytest.mark.parametrize("mk", [synth.tiny, synth.config1])
def test_qbank_matches_fp64_oracle(gpu, mk):
    cfg = mk()
    sv, tab = _survey(cfg)
    rng = np.random.default_rng(11)
    traces = rng.standard_normal((6, cfg.nt)).astype(np.float32)
    bank = sv.qbank(torch.from_numpy(traces).cuda()).cpu().numpy()
    ref = O.qbank(traces, o_bparams(sv))
    for i in range(traces.shape[0]):
        for j in range(cfg.K):
            assert rel_l2(bank[i, j], ref[i, j]) < 2e-5, (i, j)


def test_qbank_bf16_rounds_fp32(gpu):
    cfg = synth.config1()
    sv, tab = _survey(cfg)
    traces = torch.from_numpy(np.random.default_rng(3).standard_normal((4, cfg.nt)).astype(np.float32)).cuda()
    b32 = sv.qbank(traces).float()
    b16 = sv.qbank(traces, dtype=torch.bfloat16).float()
    assert rel_l2(b16.cpu().numpy(), b32.cpu().numpy()) < 4e-3


@pytest.mark.parametrize("mk", [synth.tiny, synth.config1])
def test_qbank_f16_scaled_rows(gpu, mk):
    """FP16 bank: each row's header holds a power-of-two scale and its
    inverse, every scaled sample is below 2^15 (the scale comes from a bound
    on the trace's largest sample), and the decoded bank matches the FP32
    bank and the FP64 oracle to FP16 rounding."""
    cfg = mk()
    sv, tab = _survey(cfg)
    rng = np.random.default_rng(17)
    x = rng.standard_normal((6, cfg.nt)).astype(np.float32)
    x[1] *= 1e-3   # traces of very different amplitudes get their own scales
    x[2] *= 1e4
    x[3] = 0.0     # a spike: its levels come close to the bound
    x[3, cfg.nt // 2] = 1.0
    tr = torch.from_numpy(x).cuda()
    b16 = sv.qbank(tr, dtype=torch.float16)
    assert tuple(b16.shape) == sv.bank_shape(6, torch.float16)
    hdr = b16[:, :8].contiguous().view(torch.float32).cpu().numpy()
    samples = b16[:, 8:8 + cfg.K * cfg.nt].float().cpu().numpy()
    for i in range(6):
        inv, sc = float(hdr[i, 0]), float(hdr[i, 1])
        e = np.log2(sc)
        assert e == np.round(e) and inv * sc == 1.0 and hdr[i, 2] == 0 and hdr[i, 3] == 0
        m = np.abs(samples[i]).max()
        assert 2.0 ** 5 <= m < 2.0 ** 15, (i, m)   # below 2^15, and the bound is not loose by over 2^10
    got = sv.decode_bank(b16).cpu().numpy()
    ref32 = sv.qbank(tr).cpu().numpy()
    ref64 = O.qbank(x, o_bparams(sv))
    for i in range(6):
        assert rel_l2(got[i], ref32[i]) < 5e-4, i
        assert rel_l2(got[i], ref64[i]) < 5e-4, i


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
def test_image_matches_oracle(gpu, mk):
    cfg = mk()
    sv, tab = _survey(cfg)
    geom, ranges = synth.geometry(synth.shots_tiny(3, 8, extent=((cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy)))
    rng = np.random.default_rng(5)
    traces = rng.standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    g_d = torch.from_numpy(geom).cuda()
    img = sv.new_image()
    cnt = 0
    for tr0, n in ranges:
        sv.image_shot(img, g_d, tr0, n, bank[tr0:tr0 + n], tab)
        cnt += sv.count_shot(g_d, tr0, n)
    b = bank.cpu().numpy()
    t = tab.cpu().numpy()
    ref32, n32 = O.image(b, geom, t, o_iparams(sv), fp32=True)
    ref64, n64 = O.image(b.astype(np.float64), geom, t, o_iparams(sv))
    got = img.cpu().numpy()
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print(f"rel-L2 GPU vs CPU FP32 {e32:.2e}, vs FP64 oracle {e64:.2e}")
    assert e32 < 1e-5          # plain FP32 GPU vs the FP32 CPU path
    assert e64 < 1e-4          # white-noise traces: FP32 time rounding vs FP64
    assert abs(cnt - n64) <= 1e-5 * n64 + 2


def test_count_matches_oracle_config1_shot(gpu):
    cfg = synth.config1()
    sv, tab = _survey(cfg)
    shots = synth.shots_config1()[:1]
    geom, ranges = synth.geometry(shots)
    geom = geom[::16]  # 64 of the 1024 receivers
    g_d = torch.from_numpy(np.ascontiguousarray(geom)).cuda()
    got = sv.count_shot(g_d, 0, geom.shape[0])
    ref = O.count(geom, o_iparams(sv))
    assert abs(got - ref) <= 1e-5 * ref + 2


def test_end_to_end_config1_subset_vs_fp64(gpu):
    """Traces -> GPU bank -> GPU image against FP64 bank -> FP64 image."""
    cfg = synth.config1()
    sv, tab = _survey(cfg)
    shots = synth.shots_config1()[5:6]
    geom, _ = synth.geometry(shots)
    geom = np.ascontiguousarray(geom[::8])
    traces = synth.diffraction_traces(cfg, geom, synth.scatterers(cfg))
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    img = sv.new_image()
    sv.image_shot(img, torch.from_numpy(geom).cuda(), 0, geom.shape[0], bank, tab)
    ob = O.qbank(traces, o_bparams(sv))
    ref, _ = O.image(ob, geom, tab.cpu().numpy(), o_iparams(sv))
    cpu32, _ = O.image(sv.qbank(torch.from_numpy(traces).cuda()).cpu().numpy(), geom,
                       tab.cpu().numpy(), o_iparams(sv), fp32=True)
    e64, e32 = rel_l2(img.cpu().numpy(), ref), rel_l2(img.cpu().numpy(), cpu32)
    print(f"config1 subset: GPU vs FP64 oracle {e64:.2e}, vs FP32 CPU path {e32:.2e}")
    assert e32 < 1e-5
    assert e64 < 1e-4


def _dequant(sv, q, code="f16"):
    """numpy decode of the 16-bit offset+scale bricks back to [node][v][u][w][2]."""
    g = sv.cfg.tgeom
    nbv, nbu, nbw = (g["nv"] + 7) // 8, (g["nu"] + 7) // 8, (g["nw"] + 7) // 8
    nn = g["nnx"] * g["nny"]
    raw = q.cpu().numpy().reshape(nn, nbv, nbu, nbw, 2064)
    hdr = raw[..., :16].copy().view(np.float32)            # [..., 4]
    if code == "f16":
        val = raw[..., 16:].copy().view(np.float16).astype(np.float32)
    else:
        val = raw[..., 16:].copy().view(np.uint16).astype(np.float32) / 65535.0
    val = val.reshape(nn, nbv, nbu, nbw, 8, 8, 8, 2)
    out = np.empty((nn, nbv * 8, nbu * 8, nbw * 8, 2), np.float32)
    t = hdr[..., 0, None, None, None] + hdr[..., 1, None, None, None] * val[..., 0]
    s = hdr[..., 2, None, None, None] + hdr[..., 3, None, None, None] * val[..., 1]
    out[..., 0] = t.transpose(0, 1, 4, 2, 5, 3, 6).reshape(nn, nbv * 8, nbu * 8, nbw * 8)
    out[..., 1] = s.transpose(0, 1, 4, 2, 5, 3, 6).reshape(nn, nbv * 8, nbu * 8, nbw * 8)
    return out[:, :g["nv"], :g["nu"], :g["nw"]], hdr


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
@pytest.mark.parametrize("code,tol", [("f32", 2e-7), ("u16", 1.5e-6), ("u16x", 1.5e-6), ("u16q", 1.5e-6),
                                      ("f16", 4e-5)])
def test_table_lookup_matches_oracle(gpu, mk, code, tol):
    """The device lookup of each table storage type against the oracle's
    FP64 trilinear lookup of the FP32 tables, over whole image volumes from
    surface positions on and between the nodes."""
    cfg = mk()
    sv, tab = _survey(cfg)
    t = tab if code == "f32" else sv.quantize_tables(tab, code=code)
    th = tab.cpu().numpy()
    og = o_tgeom(cfg.tgeom)
    rng = np.random.default_rng(17)
    ex, ey, ez = (cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy, (cfg.nz - 1) * cfg.dz
    for px, py in [(cfg.tgeom["n0x"], cfg.tgeom["n0y"])] + [tuple(rng.uniform(0, [ex, ey])) for _ in range(3)]:
        xyz = np.stack([rng.uniform(0, ex, 400), rng.uniform(0, ey, 400), rng.uniform(0, ez, 400)], 1)
        xyz = xyz.astype(np.float32)
        got = sv.lookup(t, float(px), float(py), torch.from_numpy(xyz).cuda()).cpu().numpy()
        ref = np.array([O.lookup(og, th, np.float32(px), np.float32(py), *p) for p in xyz.astype(np.float64)])
        err = np.abs(got - ref)
        assert err[:, 0].max() <= tol, (code, px, py, err[:, 0].max())
        assert err[:, 1].max() <= tol, (code, px, py, err[:, 1].max())


def _dequant_x(sv, q):
    """numpy decode of the edge-sharing 9x9x8 u16 bricks; every node is read
    from each brick that holds it, and all copies must agree to one step."""
    g = sv.cfg.tgeom
    nv, nu, nw = g["nv"], g["nu"], g["nw"]
    nbv, nbu, nbw = (nv - 2) // 8 + 1, (nu - 2) // 8 + 1, (nw + 7) // 8
    nn = g["nnx"] * g["nny"]
    raw = q.cpu().numpy().reshape(nn, nbv, nbu, nbw, 16 + 648 * 4)
    hdr = raw[..., :16].copy().view(np.float32)
    code = raw[..., 16:].copy().view(np.uint32).reshape(nn, nbv, nbu, nbw, 9, 9, 8)
    out = np.full((nn, nv, nu, nw, 2), np.nan, np.float32)
    step = np.zeros((nn, nv, nu, nw), np.float32)
    for bv in range(nbv):
        for bu in range(nbu):
            for bw in range(nbw):
                v = np.minimum(8 * bv + np.arange(9), nv - 1)
                u = np.minimum(8 * bu + np.arange(9), nu - 1)
                w = np.minimum(8 * bw + np.arange(8), nw - 1)
                h = hdr[:, bv, bu, bw]
                c = code[:, bv, bu, bw]
                t = h[:, 0, None, None, None] + h[:, 1, None, None, None] * (c & 0xFFFF).astype(np.float32)
                ts = h[:, 2, None, None, None] + h[:, 3, None, None, None] * (c >> 16).astype(np.float32)
                V, U, Wd = np.meshgrid(v, u, w, indexing="ij")
                prev = out[:, V, U, Wd, 0]
                ok = np.isnan(prev) | (np.abs(prev - t) <= 1.01 * np.maximum(step[:, V, U, Wd],
                                                                            h[:, 1, None, None, None]))
                assert ok.all(), (bv, bu, bw)
                out[:, V, U, Wd, 0] = t
                out[:, V, U, Wd, 1] = ts
                step[:, V, U, Wd] = h[:, 1, None, None, None]
    return out, step


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
def test_u16x_brick_tables_error_bound(gpu, mk):
    cfg = mk()
    sv, tab = _survey(cfg)
    q = sv.quantize_tables(tab, code="u16x")
    deq, step = _dequant_x(sv, q)
    ref = tab.cpu().numpy()
    err = np.abs(deq[..., 0] - ref[..., 0])
    # a code is the nearest of 65536 steps: half a step, plus FP32 rounding of off + step * code
    assert np.all(err <= 0.5 * step * 1.01 + 2e-7)
    assert err.max() < 0.5e-3


@pytest.mark.parametrize("mk", [synth.tiny])  # diffraction data are defined for v(z); the layered
def test_u16x_image_matches_f32_tables(gpu, mk):  # tables' reads are pinned by the lookup test
    """Images of diffraction data with both 16-bit layouts against FP32
    tables.  (On white-noise traces a t* perturbation of 1e-8 s flips the
    bank level of the few updates next to a rounding boundary, each an O(1)
    change, which makes such comparisons noisy; coherent data are not.)"""
    cfg = mk()
    sv, tab = _survey(cfg)
    geom, ranges = synth.geometry(synth.shots_tiny(3, 8, extent=((cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy)))
    traces = synth.diffraction_traces(cfg, geom, synth.scatterers(cfg, n=6))
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    g_d = torch.from_numpy(geom).cuda()
    qx = sv.quantize_tables(tab, code="u16x")
    qq = sv.quantize_tables(tab, code="u16q")
    qb = sv.quantize_tables(tab, code="u16")
    imgs = {}
    for name, t in (("f32", tab), ("u16x", qx), ("u16q", qq), ("u16", qb)):
        img = sv.new_image()
        for tr0, n in ranges:
            sv.image_shot(img, g_d, tr0, n, bank[tr0:tr0 + n], t)
        imgs[name] = img.cpu().numpy()
    ex, eb = rel_l2(imgs["u16x"], imgs["f32"]), rel_l2(imgs["u16"], imgs["f32"])
    eq, eqx = rel_l2(imgs["u16q"], imgs["f32"]), rel_l2(imgs["u16q"], imgs["u16x"])
    print(f"16-bit tables vs FP32 tables (diffractions): u16x {ex:.2e}, u16q {eq:.2e}, u16 {eb:.2e}; "
          f"u16q vs u16x {eqx:.2e}")
    assert ex < 1e-3 and eb < 1e-3 and eq < 1e-3
    assert eqx < 1e-3  # same codes; only the bilinear weights differ (16-bit fixed point)


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
def test_u16q_holds_the_u16x_codes(gpu, mk):
    """The corner-quad layout stores each U16X brick's header (its steps
    divided by 65535) and codes: every cell's word at every depth node holds
    the t codes of its four corners, then their t* codes, read back from the
    U16X brick that owns the cell."""
    cfg = mk()
    sv, tab = _survey(cfg)
    g = cfg.tgeom
    nv, nu, nw = g["nv"], g["nu"], g["nw"]
    nbv, nbu, nbw = (nv - 2) // 8 + 1, (nu - 2) // 8 + 1, (nw + 7) // 8
    nn = g["nnx"] * g["nny"]
    x = sv.quantize_tables(tab, code="u16x").cpu().numpy().reshape(nn, nbv, nbu, nbw, 16 + 648 * 4)
    q = sv.quantize_tables(tab, code="u16q").cpu().numpy().reshape(nn, nbv, nbu, nbw, 16 + 512 * 16)
    hx = x[..., :16].copy().view(np.float32)
    hq = q[..., :16].copy().view(np.float32)
    assert np.array_equal(hq[..., 0::2], hx[..., 0::2])  # offsets
    assert np.array_equal(hq[..., 1::2], hx[..., 1::2] * np.float32(1.0 / 65535.0))  # steps / 65535
    xc = x[..., 16:].copy().view(np.uint32).reshape(nn, nbv, nbu, nbw, 9, 9, 8)
    qc = q[..., 16:].copy().view(np.uint16).reshape(nn, nbv, nbu, nbw, 8, 8, 8, 8)
    # word halves: t00 t01 t10 t11 t*00 t*01 t*10 t*11
    for k, (dv, du) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        corner = xc[..., dv:dv + 8, du:du + 8, :]
        assert np.array_equal(qc[..., k], (corner & 0xffff).astype(np.uint16)), k
        assert np.array_equal(qc[..., 4 + k], (corner >> 16).astype(np.uint16)), k


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
@pytest.mark.parametrize("code,ulp", [("f16", 2.0 ** -11), ("u16", 2.0 ** -16)])
def test_16bit_brick_tables_error_bound(gpu, mk, code, ulp):
    cfg = mk()
    sv, tab = _survey(cfg)
    q = sv.quantize_tables(tab, code=code)
    deq, hdr = _dequant(sv, q, code)
    ref = tab.cpu().numpy()
    # a code in [0, 1] is within one half-spacing (2^-12 for FP16 in [0.5,1),
    # 2^-17 for 16-bit fixed point) of h, so the error per brick is at most
    # scale * ulp (plus the FP32 rounding of off + scale * h)
    err = np.abs(deq[..., 0] - ref[..., 0]).max()
    smax = max(hdr[..., 1].max(), 1e-9)
    assert err <= smax * ulp * 1.01 + 2e-7
    assert err < 0.5e-3  # inside the 0.5 ms budget


def test_mixed_precision_image_vs_fp32_and_fp64(gpu):
    """BF16 bank and FP16-brick tables with FP32 accumulation on a config-1
    shot subset; each reduced-precision store alone and both together, error
    stated against the FP64 oracle and the FP32 GPU form."""
    cfg = synth.config1()
    sv, tab = _survey(cfg)
    geom, _ = synth.geometry(synth.shots_config1()[5:7])
    geom = np.ascontiguousarray(geom[::16])
    traces = synth.diffraction_traces(cfg, geom, synth.scatterers(cfg))
    tr_d, g_d = torch.from_numpy(traces).cuda(), torch.from_numpy(geom).cuda()
    n = geom.shape[0]
    b32 = sv.qbank(tr_d)
    b16 = sv.qbank(tr_d, dtype=torch.bfloat16)
    bh = sv.qbank(tr_d, dtype=torch.float16)
    q = sv.quantize_tables(tab)
    qu = sv.quantize_tables(tab, code="u16")
    qx = sv.quantize_tables(tab, code="u16x")
    half = n // 2
    imgs = {}
    for name, bank, tb in (("fp32", b32, tab), ("bf16bank", b16, tab), ("f16bank", bh, tab), ("f16tab", b32, q),
                           ("u16tab", b32, qu), ("u16xtab", b32, qx), ("mixed", b16, q), ("mixed_u16", b16, qu),
                           ("mixed_u16x", b16, qx), ("mixed_f16_u16x", bh, qx)):
        img = sv.new_image()
        for tr0, m in ((0, half), (half, n - half)):
            sv.image_shot(img, g_d, tr0, m, bank[tr0:tr0 + m], tb)
        imgs[name] = img.cpu().numpy()
    ob = O.qbank(traces, o_bparams(sv))
    ref, _ = O.image(ob, geom, tab.cpu().numpy(), o_iparams(sv))
    err = {k: (rel_l2(v, ref), rel_l2(v, imgs["fp32"])) for k, v in imgs.items()}
    print("rel-L2 vs FP64 / vs FP32: " + ", ".join(f"{k} {a:.2e}/{b:.2e}" for k, (a, b) in err.items()))
    assert err["fp32"][0] < 1e-4
    assert err["mixed"][0] < 5e-3 and err["mixed"][1] < 5e-3
    assert err["u16tab"][0] < 2e-4
    assert err["mixed_u16"][0] < 1e-3      # the mixed form's stated bound vs FP64
    assert err["u16xtab"][0] < 2e-4
    assert err["mixed_u16x"][0] < 1e-3
    assert err["f16bank"][0] < 3e-4
    assert err["mixed_f16_u16x"][0] < 5e-4


def test_mixed_forms_on_incoherent_traces(gpu):
    """White-noise traces image to an incoherent sum, so the bank's rounding
    passes into the image at its own relative size: BF16 samples (8-bit
    significand) leave about 1e-3, FP16 samples under a per-trace scale
    (11-bit) about 1e-4.  Stated against the FP64 oracle."""
    cfg = synth.tiny_layers()
    sv, tab = _survey(cfg)
    geom, ranges = synth.geometry(synth.shots_tiny(3, 8, extent=((cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy)))
    x = np.random.default_rng(23).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    tr_d, g_d = torch.from_numpy(x).cuda(), torch.from_numpy(geom).cuda()
    qx = sv.quantize_tables(tab, code="u16x")
    imgs = {}
    for name, dt, tb in (("fp32", torch.float32, tab), ("bf16_u16x", torch.bfloat16, qx),
                         ("f16_u16x", torch.float16, qx), ("f16", torch.float16, tab)):
        bank = sv.qbank(tr_d, dtype=dt)
        img = sv.new_image()
        for tr0, n in ranges:
            sv.image_shot(img, g_d, tr0, n, bank[tr0:tr0 + n], tb)
        imgs[name] = img.cpu().numpy()
    ob = O.qbank(x, o_bparams(sv))
    ref, _ = O.image(ob, geom, tab.cpu().numpy(), o_iparams(sv))
    err = {k: rel_l2(v, ref) for k, v in imgs.items()}
    print("incoherent data, rel-L2 vs FP64: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()))
    assert err["fp32"] < 1e-4
    assert err["f16"] < 4e-4
    assert err["f16_u16x"] < 1e-3       # the mixed form's bound holds with FP16 samples
    assert err["bf16_u16x"] > err["f16_u16x"]


def test_qbank_long_window(gpu):
    """An 8 s trace at 1 ms: its window (nsb = 4576) needs about 120 KB of
    LDS per work-group, past the default 64 KB.  Where the device lets one
    work-group have that much, the bank matches the FP64 oracle; where it
    does not, the call reports the limit instead of failing the launch."""
    cfg = synth.tiny()
    cfg.nt, cfg.dt, cfg.K = 8001, 0.001, 3
    sv = engine.Survey(cfg, 0.02)
    x = np.random.default_rng(31).standard_normal((2, cfg.nt)).astype(np.float32)
    print(f"long window: nsb={sv.win.nsb} nww={sv.win.nww}", flush=True)
    try:
        got = sv.qbank(torch.from_numpy(x).cuda()).cpu().numpy()
    except L.QpsdmError as e:
        assert "LDS" in str(e)
        pytest.skip(f"device LDS limit: {e}")
    ref = O.qbank(x, o_bparams(sv))
    for i in range(2):
        for j in range(cfg.K):
            assert rel_l2(got[i, j], ref[i, j]) < 2e-5, (i, j)
