"""The learned source-leg surrogate on the GPU: training samples, forward
and gradients against the CPU oracle, training, the validity mask and its
table fallback, per-shot source tables in the imaging kernel, and what the
trained network gains over the nearest-node tables."""
import numpy as np
import pytest
import torch

from oracle import mlp as M
from oracle import oracle as O
from q_psdm_amd import engine, synth
from tests.helpers import o_model, rel_l2

pytestmark = pytest.mark.gpu

SCALE = (0.01, 0.002, 0.05)


def _sur(cfg, seed=7):
    return engine.Surrogate.for_survey(cfg, out_scale=SCALE, seed=seed)


def _survey(cfg):
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    return engine.Survey(cfg, engine.tstar_max_from_tables(tab)), tab


def _params(seed):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed) + rng.normal(0, 0.03, M.NPARAMS)
    return p.astype(np.float32)


def test_samples_match_oracle_ray_integrals(gpu):
    cfg = synth.tiny_layers()
    s = _sur(cfg)
    X, T = s.samples(300, seed=5)
    X, T = X.cpu().numpy().astype(np.float64), T.cpu().numpy().astype(np.float64)
    lo, hi = np.array(s.box_lo), np.array(s.box_hi)
    assert np.all(X >= lo - 1e-3) and np.all(X <= hi + 1e-3)
    assert np.all(X.std(0) > 0.2 * (hi - lo) / np.sqrt(12) * 0.8)
    om = o_model(cfg.model)
    for i in range(0, 300, 3):
        t, ts = O.ray_times(om, *X[i])
        sg = O.ray_sigma(om, *X[i])
        # samples are exact at the FP64 point; X is stored rounded to FP32
        assert abs(T[i, 0] - t) <= 2e-6 * t + 1e-7
        assert abs(T[i, 1] - ts) <= 2e-6 * ts + 1e-9
        assert abs(T[i, 2] + np.log(sg)) <= 2e-6 * abs(np.log(sg))


def test_forward_matches_oracle(gpu):
    cfg = synth.tiny_layers()
    s = _sur(cfg)
    p = _params(3)
    s.set_params(p)
    assert np.array_equal(s.params(), p)
    X, _ = s.samples(1000, seed=9)            # not a multiple of 64: ragged last group
    out = s.eval(X).cpu().numpy()
    ref = M.forward(p.astype(np.float64), X.cpu().numpy().astype(np.float64), s.box_lo, s.box_hi, cfg.model,
                    SCALE, bf16_round=True)
    err = np.abs(out - ref)
    print("forward |dev-oracle| max per output:", err.max(0), "mean:", err.mean(0))
    for o in range(3):
        assert err[:, o].max() <= 2e-3 * SCALE[o] and err[:, o].mean() <= 2e-4 * SCALE[o]


def test_gradient_matches_oracle(gpu):
    cfg = synth.tiny_layers()
    s = _sur(cfg)
    p = _params(4)
    s.set_params(p)
    X, T = s.samples(1024, seed=12)
    g, loss = s.grad(X, T)
    g = g.cpu().numpy().astype(np.float64)
    lref, gref = M.forward_backward(p.astype(np.float64), X.cpu().numpy().astype(np.float64),
                                    T.cpu().numpy().astype(np.float64), s.box_lo, s.box_hi, cfg.model, SCALE,
                                    bf16_round=True)
    print(f"loss dev {loss:.6e} oracle {lref:.6e}")
    assert abs(loss - lref) <= 1e-4 * lref
    k = 0
    for name, shp in M.SHAPES:
        n = int(np.prod(shp))
        e = rel_l2(g[k:k + n], gref[k:k + n])
        print(f"grad {name}: rel-L2 {e:.2e}")
        assert e < 2e-3, name
        k += n
    # a further call on the same data gives the same gradient (fixed-order reduction)
    g2, _ = s.grad(X, T)
    assert np.array_equal(g2.cpu().numpy().astype(np.float64), g)


@pytest.fixture(scope="module")
def trained():
    cfg = synth.tiny_layers()
    s = _sur(cfg, seed=11)
    X, T = s.samples(1 << 19, seed=21)
    hist = s.train(X, T, steps=4000, batch=16384, lr0=3e-3, lr1=3e-5)
    Xv, Tv = s.samples(16384, seed=99, offset=1 << 30)
    out = s.eval(Xv)
    dt = (out - Tv).abs().cpu().numpy()
    return cfg, s, hist, dt


def test_training_converges(gpu, trained):
    cfg, s, hist, dt = trained
    print(f"loss first {hist[:10].mean():.3e} last {hist[-50:].mean():.3e}; held-out |dt| "
          f"rms {np.sqrt((dt[:, 0] ** 2).mean()) * 1e3:.3f} ms max {dt[:, 0].max() * 1e3:.3f} ms, "
          f"|dt*| max {dt[:, 1].max() * 1e3:.3f} ms, |dlogA| max {dt[:, 2].max():.4f}")
    assert np.all(np.isfinite(hist))
    assert hist[-50:].mean() < hist[:10].mean() / 30
    assert np.sqrt((dt[:, 0] ** 2).mean()) < 0.25e-3


def _shots(cfg, n=3, seed=31):
    rng = np.random.default_rng(seed)
    ex, ey = (cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy
    shots = []
    for k in range(n):
        sx, sy = rng.uniform(0.1 * ex, 0.9 * ex), rng.uniform(0.1 * ey, 0.9 * ey)
        g = np.stack([rng.uniform(-40, ex + 40, 12), rng.uniform(-40, ey + 40, 12)], 1)
        shots.append((float(sx), float(sy), g, 40.0 + 5 * k))
    geom, ranges = synth.geometry(shots)
    desc = [(tr0, m) + synth.dip_tans(a) for (tr0, m), (*_, a) in zip(ranges, shots)]
    return geom, engine.ShotList(desc)


def _src_layout(s, sv, nshot, src):
    nus, nvs, nw = s.src_grid(sv)
    hdr = src[:16 * nshot].view(torch.float32).view(nshot, 4).cpu().numpy()
    tabs = src[16 * nshot:].view(torch.float32).view(nshot, nvs, nus, nw, 2)
    return hdr, tabs, (nus, nvs, nw)


def test_src_tables_fallback_and_surrogate(gpu):
    cfg = synth.tiny_layers()
    sv, tab = _survey(cfg)
    s = _sur(cfg)
    s.set_params(_params(6))
    geom, shots = _shots(cfg)
    g_d = torch.from_numpy(geom).cuda()
    nb = engine.L.lib().qpsdm_surrogate_mask_bricks(engine.C.byref(sv.tg))
    g = cfg.tgeom
    # every brick invalid: the table builder at each shot's own source (FP64
    # straight rays), against the oracle's ray integrals
    src = s.src_tables(sv, g_d, shots, tab=tab, mask=torch.zeros(nb, dtype=torch.uint8, device="cuda"))
    hdr, tabs, (nus, nvs, nw) = _src_layout(s, sv, shots.n, src)
    m = O.model_layers(cfg.model["vl"], cfg.model["ql"], cfg.model["layer_dz"],
                       (cfg.model["ax"], cfg.model["ay"], cfg.model["az"], cfg.model["asig"], cfg.model["av"],
                        cfg.model["aq"]))
    rng = np.random.default_rng(9)
    for k, (tr0, n, *_) in enumerate(shots.host):
        sx, sy = geom[tr0, 0], geom[tr0, 1]
        assert hdr[k, 0] == sx and hdr[k, 1] == sy
        got = tabs[k].cpu().numpy()
        for _ in range(30):
            iv, iu, iw = int(rng.integers(0, nvs)), int(rng.integers(0, nus)), int(rng.integers(0, nw))
            x = np.float32(sx + hdr[k, 2] + np.float32(iu * g["du"]))
            y = np.float32(sy + hdr[k, 3] + np.float32(iv * g["dv"]))
            t, ts = O.ray_times(m, float(sx), float(sy), float(x), float(y), iw * g["dw"])
            assert abs(got[iv, iu, iw, 0] - t) <= 2e-6 * t + 1e-7
            assert abs(got[iv, iu, iw, 1] - ts) <= 2e-6 * ts + 1e-8
    # every brick valid: the surrogate at those points
    src1 = s.src_tables(sv, g_d, shots, tab=tab, mask=torch.ones(nb, dtype=torch.uint8, device="cuda"))
    hdr1, tabs1, _ = _src_layout(s, sv, shots.n, src1)
    k = 1
    su0, sv0 = hdr1[k, 2], hdr1[k, 3]
    sx, sy = hdr1[k, 0], hdr1[k, 1]
    iv, iu, iw = np.meshgrid(np.arange(nvs), np.arange(nus), np.arange(nw), indexing="ij")
    X = np.stack([np.full(iv.size, sx), np.full(iv.size, sy), (sx + su0) + iu.ravel() * np.float32(g["du"]),
                  (sy + sv0) + iv.ravel() * np.float32(g["dv"]), iw.ravel() * np.float32(g["dw"])], 1)
    out = s.eval(torch.from_numpy(X.astype(np.float32)).cuda()).cpu().numpy()
    got = tabs1[k].cpu().numpy().reshape(-1, 2)
    assert np.abs(got - out[:, :2]).max() < 2e-6


def test_image_with_src_tables_matches_node_path_for_shots_on_nodes(gpu):
    """Shots whose sources sit on surface nodes: the table path's source
    tables (the surrogate fully fallen back) hold the node's own values, and
    the image equals the node-table path's."""
    cfg = synth.tiny_layers()
    sv, tab = _survey(cfg)
    s = _sur(cfg)
    geom, shots = _shots(cfg, n=4, seed=3)
    g = cfg.tgeom
    for k, (tr0, n, *_) in enumerate(shots.host):
        geom[tr0:tr0 + n, 0] = g["n0x"] + (k % g["nnx"]) * g["ndx"]
        geom[tr0:tr0 + n, 1] = g["n0y"] + ((k + 1) % g["nny"]) * g["ndy"]
    g_d = torch.from_numpy(geom).cuda()
    traces = torch.from_numpy(np.random.default_rng(1).standard_normal((geom.shape[0], cfg.nt))
                              .astype(np.float32)).cuda()
    bank = sv.qbank(traces)
    nb = engine.L.lib().qpsdm_surrogate_mask_bricks(engine.C.byref(sv.tg))
    src = s.src_tables(sv, g_d, shots, tab=tab, mask=torch.zeros(nb, dtype=torch.uint8, device="cuda"))
    a = sv.new_image()
    sv.image_batch(a, g_d, bank, tab, shots)
    b = sv.new_image()
    sv.image_batch(b, g_d, bank, tab, shots, src=src)
    e = rel_l2(b.cpu().numpy(), a.cpu().numpy())
    print(f"source tables (all fallback) vs node tables: rel-L2 {e:.2e}")
    assert e < 1e-5 and float(a.abs().sum()) > 0


def _exact_src(cfg, sv, s, g_d, shots, src_like):
    """Source tables built by the FP64 ray integrator at each shot's exact
    source position on the same relative sub-grids (the ideal the surrogate
    approximates)."""
    hdr, _, (nus, nvs, nw) = _src_layout(s, sv, shots.n, src_like)
    g = cfg.tgeom
    out = src_like.clone()
    body = out[16 * shots.n:].view(torch.float32)
    per = nvs * nus * nw * 2
    for k in range(shots.n):
        tg = dict(g, nnx=1, nny=1, n0x=float(hdr[k, 0]), n0y=float(hdr[k, 1]), nu=nus, nv=nvs,
                  u0=float(hdr[k, 2]), v0=float(hdr[k, 3]))
        engine.build_tables_into(cfg.model, tg, body[k * per:(k + 1) * per])
    return out


def test_surrogate_gain_over_nearest_node_tables(gpu, trained):
    """Off-node sources: the trained surrogate's source tables against the
    exact ones, compared with what the nearest-node table gives there; and
    the images built from each."""
    cfg, s, _, _ = trained
    sv, tab = _survey(cfg)
    m, err = s.mask(sv.tg, tab, tol=0.5e-3)
    frac_bad = 1.0 - float(m.float().mean())
    geom, shots = _shots(cfg, n=5, seed=44)
    g_d = torch.from_numpy(geom).cuda()
    src_s = s.src_tables(sv, g_d, shots, tab=tab, mask=m)
    src_x = _exact_src(cfg, sv, s, g_d, shots, src_s)
    _, ts, _ = _src_layout(s, sv, shots.n, src_s)
    _, tx, _ = _src_layout(s, sv, shots.n, src_x)
    src_b = sv.src_tables_nodes(g_d, shots, tab)   # the table path: four node tables blended at the source
    _, tb, _ = _src_layout(s, sv, shots.n, src_b)
    es = (ts - tx)[..., 0].abs()
    en = (tb - tx)[..., 0].abs()
    # the nearest node's own values at the same relative nodes
    hdr, _, (nus, nvs, nw) = _src_layout(s, sv, shots.n, src_s)
    g = cfg.tgeom
    t_all = tab.cpu().numpy()
    near = np.empty((shots.n, nvs, nus, nw))
    for k in range(shots.n):
        iu0 = int(round((hdr[k, 2] - g["u0"]) / g["du"]))
        iv0 = int(round((hdr[k, 3] - g["v0"]) / g["dv"]))
        i = min(max(int(np.floor((hdr[k, 0] - g["n0x"]) / g["ndx"] + 0.5)), 0), g["nnx"] - 1)
        j = min(max(int(np.floor((hdr[k, 1] - g["n0y"]) / g["ndy"] + 0.5)), 0), g["nny"] - 1)
        near[k] = t_all[j * g["nnx"] + i, iv0:iv0 + nvs, iu0:iu0 + nus, :, 0]
    eq = np.abs(near - tx[..., 0].cpu().numpy())
    print(f"mask: {frac_bad * 100:.1f}% bricks fall back; source-leg |dt| vs exact: surrogate mean "
          f"{float(es.mean()) * 1e3:.4f} ms max {float(es.max()) * 1e3:.3f} ms; node blend mean "
          f"{float(en.mean()) * 1e3:.4f} ms max {float(en.max()) * 1e3:.3f} ms; nearest node mean "
          f"{eq.mean() * 1e3:.4f} ms max {eq.max() * 1e3:.3f} ms")
    # the learned source leg against both table-interpolation methods (the
    # nearest node's table, and the four surrounding node tables blended)
    assert float(es.mean()) < 0.6 * min(float(en.mean()), eq.mean())
    traces = torch.from_numpy(np.random.default_rng(2).standard_normal((geom.shape[0], cfg.nt))
                              .astype(np.float32)).cuda()
    bank = sv.qbank(traces)
    imgs = {}
    for name, src in (("nodes", None), ("blend", src_b), ("surrogate", src_s), ("exact", src_x)):
        img = sv.new_image()
        sv.image_batch(img, g_d, bank, tab, shots, src=src)
        imgs[name] = img.cpu().numpy()
    e_s, e_n = rel_l2(imgs["surrogate"], imgs["exact"]), rel_l2(imgs["nodes"], imgs["exact"])
    e_b = rel_l2(imgs["blend"], imgs["exact"])
    print(f"image rel-L2 vs exact source legs: surrogate {e_s:.2e}, node blend {e_b:.2e}, "
          f"nearest-node tables {e_n:.2e}")
    assert e_s < 0.75 * min(e_n, e_b)
