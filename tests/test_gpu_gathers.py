"""Dip-angle gathers on the GPU against the oracle's FP32 gathers (same
operation order: bit-identical), the GPU image's columns (summed over bins),
the mixed storage forms, accumulation across runs and the axis choice."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from q_psdm_amd import engine, synth
from tests.helpers import o_iparams, rel_l2

pytestmark = pytest.mark.gpu

COLS = np.array([[3, 4], [12, 10], [19, 15], [0, 0], [17, 2], [8, 14]], np.int32)


def _case(mk, nshot=2, ntr=5):
    cfg = mk()
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab))
    geom, _ = synth.geometry(synth.shots_tiny(nshot, ntr, extent=((cfg.nx - 1) * cfg.dx, (cfg.ny - 1) * cfg.dy)))
    traces = np.random.default_rng(21).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    tr_d = torch.from_numpy(traces).cuda()
    g_d = torch.from_numpy(geom).cuda()
    cols = COLS[(COLS[:, 0] < cfg.nx) & (COLS[:, 1] < cfg.ny)]
    ip = sv.ip
    tanc = max(abs(v) for v in (ip.tanx1, ip.tanx2, ip.tany1, ip.tany2))
    dmax = math.degrees(math.atan(tanc + ip.taperzone * (1 + tanc * tanc)))
    return cfg, sv, tab, geom, tr_d, g_d, np.ascontiguousarray(cols), dmax


@pytest.mark.parametrize("mk", [synth.tiny, synth.tiny_layers])
def test_gathers_match_oracle_and_image(gpu, mk):
    cfg, sv, tab, geom, tr_d, g_d, cols, dmax = _case(mk)
    bank = sv.qbank(tr_d)
    gp = sv.gather_params(dipx=math.ceil(dmax) + 2, dipy=math.ceil(dmax) + 2, ddipx=1.0, ddipy=1.0)
    _, nbx, nby = gp
    cols_d = torch.from_numpy(cols).cuda()
    gx, gy = sv.new_gathers(cols.shape[0], nbx), sv.new_gathers(cols.shape[0], nby)
    print(f"gathers {cfg.name}: {cols.shape[0]} columns, {nbx} x {nby} bins, {geom.shape[0]} traces", flush=True)
    sv.gathers_shot(gx, gy, cols_d, g_d, 0, geom.shape[0], bank, tab, gp)
    img = sv.new_image()
    sv.image_shot(img, g_d, 0, geom.shape[0], bank, tab)
    torch.cuda.synchronize()
    e = O.gather_edges(math.ceil(dmax) + 2, 1.0)
    rx, ry = O.gathers(bank.cpu().numpy(), geom, tab.cpu().numpy(), o_iparams(sv), cols, e, e)
    gxh, gyh = gx.cpu().numpy(), gy.cpu().numpy()
    same = (np.mean(gxh == rx), np.mean(gyh == ry))
    print(f"  vs oracle FP32 gathers: rel-L2 {rel_l2(gxh, rx):.2e} / {rel_l2(gyh, ry):.2e}, "
          f"bit-equal fraction {same[0]:.6f} / {same[1]:.6f}")
    # the kernel evaluates the oracle's FP32 expressions in the same order,
    # with IEEE division and square root and no contraction: bit-identical
    assert np.array_equal(gxh, rx) and np.array_equal(gyh, ry)
    im = img.cpu().numpy()
    for k, (ix, iy) in enumerate(cols):
        col = im[iy, ix].astype(np.float64)
        for g in (gxh, gyh):
            assert np.allclose(g[k].astype(np.float64).sum(axis=0), col, rtol=1e-5,
                               atol=1e-5 * np.abs(col).max() + 1e-30)
    assert np.abs(gxh).sum() > 0


def test_gathers_mixed_storage_and_accumulation(gpu):
    """FP16 bank + U16X tables against FP32 (< 1e-3); two runs accumulated
    into the same gathers equal one run over both; x-only and y-only equal
    the axes of "all"."""
    cfg, sv, tab, geom, tr_d, g_d, cols, dmax = _case(synth.tiny, nshot=3, ntr=6)
    n = geom.shape[0]
    gp = sv.gather_params(dipx=40.0, dipy=40.0, ddipx=2.0, ddipy=2.0)
    _, nbx, nby = gp
    cols_d = torch.from_numpy(cols).cuda()
    b32 = sv.qbank(tr_d)
    b16 = sv.qbank(tr_d, dtype=torch.float16)
    q = sv.quantize_tables(tab, code="u16x")
    gx, gy = sv.new_gathers(len(cols), nbx), sv.new_gathers(len(cols), nby)
    sv.gathers_shot(gx, gy, cols_d, g_d, 0, n, b32, tab, gp)
    mx, my = sv.new_gathers(len(cols), nbx), sv.new_gathers(len(cols), nby)
    sv.gathers_shot(mx, my, cols_d, g_d, 0, n, b16, q, gp)
    ax, ay = sv.new_gathers(len(cols), nbx), sv.new_gathers(len(cols), nby)
    h = n // 2
    sv.gathers_shot(ax, ay, cols_d, g_d, 0, h, b32[:h], tab, gp)
    sv.gathers_shot(ax, ay, cols_d, g_d, h, n - h, b32[h:], tab, gp)
    gpx = sv.gather_params(dipx=40.0, dipy=40.0, ddipx=2.0, ddipy=2.0, aspect="x")
    gpy = sv.gather_params(dipx=40.0, dipy=40.0, ddipx=2.0, ddipy=2.0, aspect="y")
    ox, oy = sv.new_gathers(len(cols), nbx), sv.new_gathers(len(cols), nby)
    sv.gathers_shot(ox, None, cols_d, g_d, 0, n, b32, tab, gpx)
    sv.gathers_shot(None, oy, cols_d, g_d, 0, n, b32, tab, gpy)
    torch.cuda.synchronize()
    qq = sv.quantize_tables(tab, code="u16q")
    qx_, qy_ = sv.new_gathers(len(cols), nbx), sv.new_gathers(len(cols), nby)
    sv.gathers_shot(qx_, qy_, cols_d, g_d, 0, n, b16, qq, gp)
    e = rel_l2(mx.cpu().numpy(), gx.cpu().numpy()), rel_l2(my.cpu().numpy(), gy.cpu().numpy())
    eq = rel_l2(qx_.cpu().numpy(), gx.cpu().numpy()), rel_l2(qy_.cpu().numpy(), gy.cpu().numpy())
    print(f"mixed storage vs FP32 gathers: {e[0]:.2e} / {e[1]:.2e} (U16Q tables {eq[0]:.2e} / {eq[1]:.2e})")
    assert max(e) < 1e-3 and max(eq) < 1e-3
    assert rel_l2(ax.cpu().numpy(), gx.cpu().numpy()) < 1e-6 and rel_l2(ay.cpu().numpy(), gy.cpu().numpy()) < 1e-6
    assert torch.equal(ox, gx) and torch.equal(oy, gy)
    none = sv.gather_params(aspect="none")
    sv.gathers_shot(None, None, cols_d, g_d, 0, n, b32, tab, none)   # nothing to do
    with pytest.raises(Exception):
        sv.gathers_shot(gx, gy, cols_d, g_d, 0, n, b32, sv.quantize_tables(tab, code="u16"), gp)


def test_pipeline_gathers_match_direct_calls(gpu):
    """Gathers set on the pipeline context and fed by SU records equal the
    operator called on the same batches (same kernel, same trace groups);
    normalised by nsb like the image; reset with it."""
    from tests.test_gpu_pipeline import _gp, _setup
    cfg, sv, tab, wxy, rel, traces = _setup()
    cols = COLS[(COLS[:, 0] < cfg.nx) & (COLS[:, 1] < cfg.ny)]
    batch = 6
    pl = engine.Pipeline(sv, _gp(cfg), batch_traces=batch)
    pl.set_tables(tab)
    pl.set_gathers(cols, dipx=40.0, dipy=40.0, ddipx=2.0, ddipy=2.0)
    pl.push_su(engine.su_encode(traces, wxy))
    gx, gy = pl.gathers()
    nx, ny = pl.gathers(normalize=True)
    gp = sv.gather_params(dipx=40.0, dipy=40.0, ddipx=2.0, ddipy=2.0)
    _, nbx, nby = gp
    cols_d = torch.from_numpy(np.ascontiguousarray(cols)).cuda()
    g_d = torch.from_numpy(rel).cuda()
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    dx, dy = sv.new_gathers(len(cols), nbx), sv.new_gathers(len(cols), nby)
    for tr0 in range(0, traces.shape[0], batch):
        n = min(batch, traces.shape[0] - tr0)
        sv.gathers_shot(dx, dy, cols_d, g_d, tr0, n, bank[tr0:tr0 + n], tab, gp)
    torch.cuda.synchronize()
    assert np.array_equal(gx, dx.cpu().numpy()) and np.array_equal(gy, dy.cpu().numpy())
    assert np.abs(gx).sum() > 0
    np.testing.assert_allclose(nx, gx / sv.win.nsb, rtol=1e-6)
    pl.reset_image()
    zx, zy = pl.gathers()
    assert not zx.any() and not zy.any()
    with pytest.raises(Exception):
        pl.set_gathers([[cfg.nx, 0]])
    pl.close()
