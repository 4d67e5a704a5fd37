"""Degenerate inputs through the operators: nothing to image is a no-op that
leaves the image as it was (empty runs, empty batches, shots without
traces, every trace rejected, empty gathers), and the counts agree."""
import numpy as np
import pytest
import torch

from q_psdm_amd import engine, synth

pytestmark = pytest.mark.gpu


def test_empty_and_rejected_inputs_leave_the_image(gpu):
    cfg = synth.tiny()
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab))
    geom, ranges = synth.geometry(synth.shots_tiny(2, 4))
    g_d = torch.from_numpy(geom).cuda()
    x = torch.from_numpy(np.random.default_rng(3).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)).cuda()
    bank = sv.qbank(x)
    img = sv.new_image().fill_(1.5)
    tans = tuple(cfg.dip_bounds()[:4])
    print("empty inputs", flush=True)
    # an empty run, an empty bank computation, an empty batch, a batch of
    # shots without traces (both grids)
    assert sv.qbank(x[:0]).shape[0] == 0
    sv.image_shot(img, g_d, 0, 0, bank, tab)
    assert sv.count_shot(g_d, 0, 0) == 0
    sv.image_batch(img, g_d, bank, tab, engine.ShotList([]))
    empty_shots = engine.ShotList([(0, 0) + tans, (3, 0) + tans])
    for nwg in (0, -1):
        sv.image_batch(img, g_d, bank, tab, empty_shots, nwg=nwg)
    assert sv.count_batch(g_d, empty_shots) == 0
    # every trace rejected upstream (NaN geometry)
    bad = torch.full_like(g_d, float("nan"))
    sv.image_shot(img, bad, 0, geom.shape[0], bank, tab)
    assert sv.count_shot(bad, 0, geom.shape[0]) == 0
    sv.image_batch(img, bad, bank, tab, engine.ShotList([(tr0, n) + tans for tr0, n in ranges]), nwg=-1)
    # gathers over an empty run and over rejected traces
    cols = torch.tensor([[3, 4], [10, 10]], dtype=torch.int32, device="cuda")
    gp = sv.gather_params(dipx=40.0, dipy=40.0, ddipx=2.0, ddipy=2.0)
    gx, gy = sv.new_gathers(2, gp[1]), sv.new_gathers(2, gp[2])
    sv.gathers_shot(gx, gy, cols, g_d, 0, 0, bank, tab, gp)
    sv.gathers_shot(gx, gy, cols, bad, 0, geom.shape[0], bank, tab, gp)
    torch.cuda.synchronize()
    assert torch.all(img == 1.5)
    assert not gx.any() and not gy.any()
    # and the same calls with the real geometry do change it
    sv.image_shot(img, g_d, 0, geom.shape[0], bank, tab)
    assert sv.count_shot(g_d, 0, geom.shape[0]) > 0
    torch.cuda.synchronize()
    assert not torch.all(img == 1.5)
