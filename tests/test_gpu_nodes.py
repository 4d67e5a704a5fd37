"""The node form of the imaging sum on the GPU against the oracle's FP32
restatement of the same form, and its batched launches against per-shot
launches."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from q_psdm_amd import engine, synth
from tests.helpers import o_iparams, rel_l2

pytestmark = pytest.mark.gpu


def _deep_survey():
    """A 1 km deep v(z) survey (8 x 6 columns x 100 depths at 10 m, 20 m
    tables): depth enough for the node range to matter."""
    cfg = synth.tiny()
    cfg.nx, cfg.ny, cfg.nz = 8, 6, 100
    cfg.nt = 1201
    cfg.tgeom = dict(cfg.tgeom, nw=52)
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    return cfg, engine.Survey(cfg, engine.tstar_max_from_tables(tab)), tab


@pytest.mark.parametrize("ni", [2, 3, 4])
@pytest.mark.parametrize("z_exact", [0.0, 300.0])
def test_node_form_matches_oracle(gpu, ni, z_exact):
    cfg, sv, tab = _deep_survey()
    geom, ranges = synth.geometry(synth.shots_tiny(2, 7, extent=(300.0, 300.0)))
    rng = np.random.default_rng(21)
    traces = torch.from_numpy(rng.standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)).cuda()
    bank = sv.qbank(traces)
    g = torch.from_numpy(geom).cuda()
    sv.set_form("nodes", z_exact)
    img = sv.new_image()
    for tr0, n in ranges:
        sv.image_shot(img, g, tr0, n, bank[tr0:tr0 + n], tab, launch=(8, 8, ni))
    torch.cuda.synchronize()
    ref, _ = O.image_nodes(bank.cpu().numpy(), geom, tab.cpu().numpy(), o_iparams(sv), z_exact)
    assert rel_l2(img.cpu().numpy(), ref) < 1e-5
    # batched (every shot in one call) = per-shot launches, bit for bit
    tans = cfg.dip_bounds()[:4]
    img2 = sv.new_image()
    sv.image_batch(img2, g, bank, tab, engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges]),
                   launch=(8, 8, ni), nwg=-1)
    torch.cuda.synchronize()
    assert torch.equal(img, img2)


def test_node_form_mixed_storage_against_fp64(gpu):
    """FP16 bank + U16Q tables in the node form, below 300 m of a 1 km
    image: within 1e-3 of the FP64 exact sum."""
    cfg, sv, tab = _deep_survey()
    geom, ranges = synth.geometry(synth.shots_tiny(2, 7, extent=(300.0, 300.0)))
    rng = np.random.default_rng(22)
    tr_h = rng.standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    bank16 = sv.qbank(torch.from_numpy(tr_h).cuda(), dtype=torch.float16)
    q = sv.quantize_tables(tab, code="u16q")
    g = torch.from_numpy(geom).cuda()
    sv.set_form("nodes", 300.0)
    tans = cfg.dip_bounds()[:4]
    img = sv.new_image()
    sv.image_batch(img, g, bank16, q, engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges]),
                   launch=(8, 8, 4), nwg=-1)
    torch.cuda.synchronize()
    from tests.helpers import o_bparams
    ref, _ = O.image(O.qbank(tr_h, o_bparams(sv)), geom, tab.cpu().numpy(), o_iparams(sv))
    assert rel_l2(img.cpu().numpy(), ref) < 1e-3


def test_node_form_rejects_unsupported_shapes(gpu):
    cfg, sv, tab = _deep_survey()
    geom, ranges = synth.geometry(synth.shots_tiny(1, 3))
    g = torch.from_numpy(geom).cuda()
    bank = sv.qbank(torch.zeros((geom.shape[0], cfg.nt), device="cuda"))
    sv.set_form("nodes", 100.0)
    from q_psdm_amd.lib import QpsdmError
    with pytest.raises(QpsdmError):
        sv.image_shot(sv.new_image(), g, 0, geom.shape[0], bank, tab, launch=(8, 8, 8))
