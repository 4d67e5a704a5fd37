"""Batched imaging: the persistent kernel over uneven shot gathers with
per-shot apertures, against the per-shot operator and the FP64 oracle, and
the captured graph against eager launches."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from q_psdm_amd import engine, synth
from tests.helpers import o_bparams, o_iparams, rel_l2

pytestmark = pytest.mark.gpu


def _survey(cfg):
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    return engine.Survey(cfg, engine.tstar_max_from_tables(tab)), tab


def _uneven_batch(cfg, counts, apertures, seed=21):
    """Shots with the given receiver counts (0 allowed) and opening angles;
    one NaN-geometry (rejected) trace in the second shot."""
    rng = np.random.default_rng(seed)
    ext = ((cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy)
    shots = []
    for n, a in zip(counts, apertures):
        sx, sy = rng.uniform(0, ext[0]), rng.uniform(0, ext[1])
        g = np.stack([rng.uniform(-60, ext[0] + 60, n), rng.uniform(-60, ext[1] + 60, n)], 1)
        shots.append((float(sx), float(sy), g, float(a)))
    geom, ranges = synth.geometry(shots)
    if counts[1] > 2:
        geom[ranges[1][0] + 1, :] = np.nan
    desc = [(tr0, n) + synth.dip_tans(a) for (tr0, n), (*_, a) in zip(ranges, shots)]
    return geom, ranges, desc


def _per_shot(sv, img, g_d, bank, tab, desc, launch=None):
    keep = (sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2)
    try:
        for tr0, n, *tans in desc:
            sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2 = tans
            if n:
                sv.image_shot(img, g_d, tr0, n, bank[tr0:tr0 + n], tab, launch=launch)
    finally:
        sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2 = keep


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
@pytest.mark.parametrize("launch", [None] + [(tx, ty, zp, 0) for tx, ty in ((8, 8), (16, 4), (4, 16))
                                            for zp in (2, 3, 4, 6, 8)])
def test_batch_equals_per_shot(gpu, mk, launch):
    """Same tiles and per-tile summation order (shots, then traces): the
    persistent kernel reproduces the per-shot launches bit for bit."""
    cfg = mk()
    sv, tab = _survey(cfg)
    geom, ranges, desc = _uneven_batch(cfg, [5, 9, 0, 1, 14, 3], [25.0, 60.0, 40.0, 33.0, 47.0, 58.0])
    traces = np.random.default_rng(4).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    g_d = torch.from_numpy(geom).cuda()
    ref = sv.new_image()
    _per_shot(sv, ref, g_d, bank, tab, desc, launch)
    shots = engine.ShotList(desc)
    for nwg in (0, 3, -1):  # resident grid, a grid much smaller than the tile count, one group per tile
        img = sv.new_image()
        sv.image_batch(img, g_d, bank, tab, shots, launch=launch, nwg=nwg)
        torch.cuda.synchronize()
        if not torch.equal(img, ref):
            d = (img - ref).abs()
            bad = torch.nonzero(d > 0)
            print(f"launch {launch} nwg {nwg}: {bad.shape[0]} points differ, max {float(d.max()):.3e}, "
                  f"first {bad[:4].tolist()}, rel-L2 {rel_l2(img.cpu().numpy(), ref.cpu().numpy()):.2e}")
        assert torch.equal(img, ref), (launch, nwg)


@pytest.mark.parametrize("aperture", [20.0, 45.0, 60.0])
def test_aperture_cull_keeps_every_update(gpu, aperture):
    """The batched kernels skip a trace for a tile when a bound of the
    bisector's tangent over the tile puts it past a cut-off (tested 64 traces
    at a time); the per-shot kernel tests every point.  A dense fan of
    receivers around and beyond the image, with sources inside and outside
    it, puts many (tile, trace) pairs at the aperture's edge: the two images
    are equal bit for bit at every tile depth."""
    cfg = synth.tiny()
    sv, tab = _survey(cfg)
    ext = ((cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy)
    u = np.linspace(-1.5, 2.5, 23)
    gx, gy = np.meshgrid(u * ext[0], u * ext[1])
    rec = np.stack([gx.ravel(), gy.ravel()], 1)
    shots = [(0.5 * ext[0], 0.5 * ext[1], rec, aperture), (-0.8 * ext[0], 1.7 * ext[1], rec, aperture),
             (2.2 * ext[0], -0.4 * ext[1], rec[::3], aperture)]
    geom, ranges = synth.geometry(shots)
    desc = [(tr0, n) + synth.dip_tans(a) for (tr0, n), (*_, a) in zip(ranges, shots)]
    traces = np.random.default_rng(12).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    g_d = torch.from_numpy(geom).cuda()
    for zp in (2, 4, 8):
        launch = (8, 8, zp, 0)
        ref = sv.new_image()
        _per_shot(sv, ref, g_d, bank, tab, desc, launch=launch)
        img = sv.new_image()
        sv.image_batch(img, g_d, bank, tab, engine.ShotList(desc), launch=launch, nwg=-1)
        torch.cuda.synchronize()
        assert float(ref.abs().max()) > 0
        assert torch.equal(img, ref), zp


def test_batch_matches_fp64_oracle_and_count(gpu):
    cfg = synth.tiny_layers()
    sv, tab = _survey(cfg)
    geom, ranges, desc = _uneven_batch(cfg, [7, 4, 11], [28.0, 55.0, 42.0], seed=5)
    traces = np.random.default_rng(9).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    g_d = torch.from_numpy(geom).cuda()
    shots = engine.ShotList(desc)
    img = sv.new_image()
    sv.image_batch(img, g_d, bank, tab, shots)
    cnt = sv.count_batch(g_d, shots)
    ob = O.qbank(traces, o_bparams(sv))
    t = tab.cpu().numpy()
    ref = np.zeros((cfg.ny, cfg.nx, cfg.nz))
    nref = 0
    keep = (sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2)
    for tr0, n, *tans in desc:
        sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2 = tans
        r, c = O.image(ob[tr0:tr0 + n], np.ascontiguousarray(geom[tr0:tr0 + n]), t, o_iparams(sv))
        ref += r
        nref += c
    sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2 = keep
    e = rel_l2(img.cpu().numpy(), ref)
    print(f"batched image vs FP64 oracle: rel-L2 {e:.2e}; updates {cnt} vs {nref}")
    assert e < 1e-4
    assert abs(cnt - nref) <= 1e-5 * nref + 2


@pytest.mark.parametrize("bank_dtype,code", [(torch.bfloat16, "u16"), (torch.float16, "u16x"),
                                             (torch.bfloat16, "u16x"), (torch.float16, "u16q")])
def test_batch_mixed_precision_matches_per_shot(gpu, bank_dtype, code):
    cfg = synth.tiny()
    sv, tab = _survey(cfg)
    geom, ranges, desc = _uneven_batch(cfg, [6, 6, 6], [30.0, 45.0, 60.0], seed=8)
    traces = np.random.default_rng(2).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    traces[4] *= 50.0  # rows with different scales in the FP16 bank
    bank = sv.qbank(torch.from_numpy(traces).cuda(), dtype=bank_dtype)
    q = sv.quantize_tables(tab, code=code)
    g_d = torch.from_numpy(geom).cuda()
    ref = sv.new_image()
    _per_shot(sv, ref, g_d, bank, q, desc)
    for nwg in (0, -1):
        img = sv.new_image()
        sv.image_batch(img, g_d, bank, q, engine.ShotList(desc), nwg=nwg)
        torch.cuda.synchronize()
        assert torch.equal(img, ref), nwg


@pytest.mark.parametrize("bank_dtype,nwg", [(torch.float32, 0), (torch.float16, -1)])
def test_graph_replays_match_eager(gpu, bank_dtype, nwg):
    """The captured [Q bank -> batched imaging] graph, replayed three
    times with the program's own eager work (new traces, image reset,
    synchronisations) between replays; each replay equals the eager result."""
    cfg = synth.tiny_layers()
    sv, tab = _survey(cfg)
    geom, ranges, desc = _uneven_batch(cfg, [8, 5, 12, 2], [26.0, 59.0, 37.0, 44.0], seed=13)
    ntr = geom.shape[0]
    g_d = torch.from_numpy(geom).cuda()
    traces = torch.zeros((ntr, cfg.nt), dtype=torch.float32, device="cuda")
    bank = torch.empty(sv.bank_shape(ntr, bank_dtype), dtype=bank_dtype, device="cuda")
    img = sv.new_image()
    shots = engine.ShotList(desc)
    g = engine.BatchGraph(sv, traces, bank, img, g_d, tab, shots, nwg=nwg)
    try:
        for rep in range(3):
            x = torch.from_numpy(np.random.default_rng(100 + rep).standard_normal((ntr, cfg.nt))
                                 .astype(np.float32)).cuda()
            traces.copy_(x)
            img.zero_()
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            got = img.clone()
            eager = sv.new_image()
            b = sv.qbank(x, dtype=bank_dtype)
            sv.image_batch(eager, g_d, b, tab, shots, nwg=nwg)
            torch.cuda.synchronize()
            assert torch.equal(b, bank), rep
            assert torch.equal(got, eager), rep
            assert float(got.abs().sum()) > 0
    finally:
        g.close()


def test_autotuner_picks_a_valid_shape(gpu):
    from q_psdm_amd import autotune
    cfg = synth.tiny_layers()
    sv, tab = _survey(cfg)
    geom, ranges, desc = _uneven_batch(cfg, [9, 7], [40.0, 50.0], seed=3)
    traces = np.random.default_rng(1).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    g_d = torch.from_numpy(geom).cuda()
    shots = engine.ShotList(desc)
    launch, nwg, res = autotune.tune(sv, g_d, bank, tab, shots=shots, ntr=9, kind="batch")
    assert len(res) == len(autotune.TILES) * len(autotune.DEPTHS) * len(autotune.GRIDS)
    assert launch[0] * launch[1] == 64 and launch[2] in autotune.DEPTHS
    ref = sv.new_image()
    _per_shot(sv, ref, g_d, bank, tab, desc, launch)
    img = sv.new_image()
    sv.image_batch(img, g_d, bank, tab, shots, launch=launch, nwg=nwg)
    torch.cuda.synchronize()
    if not torch.equal(img, ref):
        d = (img - ref).abs()
        bad = torch.nonzero(d > 0)
        print(f"autotuned {launch} nwg {nwg}: {bad.shape[0]} points differ, max {float(d.max()):.3e}, "
              f"first {bad[:4].tolist()}")
        for nw in (0, 3, -1, nwg):
            im2 = sv.new_image()
            sv.image_batch(im2, g_d, bank, tab, shots, launch=launch, nwg=nw)
            torch.cuda.synchronize()
            print(f"  nwg {nw}: equal {torch.equal(im2, ref)}")
    assert torch.equal(img, ref)
