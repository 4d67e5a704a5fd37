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


def _deep_survey(nw=52):
    """A 1 km deep v(z) survey (8 x 6 columns x 100 depths at 10 m, 20 m
    tables, nw depth nodes): depth enough for the node range to matter."""
    cfg = synth.tiny()
    cfg.nx, cfg.ny, cfg.nz = 8, 6, 100
    cfg.nt = 1201
    cfg.tgeom = dict(cfg.tgeom, nw=nw)
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


@pytest.mark.parametrize("ni", [2, 4])
def test_node_form_with_tables_shallower_than_the_image(gpu, ni):
    """Tables ending at 780 m under a 1 km image: the points beyond the last
    depth node go through the exact body (depth range [pe, nz)), those above
    z_exact too, the rest through the node body; against the oracle's node
    form, and the batched call equal to per-shot launches."""
    cfg, sv, tab = _deep_survey(nw=40)
    sv.set_form("nodes", 200.0)
    ps, pe = O.node_range(o_iparams(sv), 200.0)
    assert 0 < ps < pe < cfg.nz
    geom, ranges = synth.geometry(synth.shots_tiny(2, 7, extent=(300.0, 300.0)))
    rng = np.random.default_rng(23)
    bank = sv.qbank(torch.from_numpy(rng.standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)).cuda())
    g = torch.from_numpy(geom).cuda()
    img = sv.new_image()
    for tr0, n in ranges:
        sv.image_shot(img, g, tr0, n, bank[tr0:tr0 + n], tab, launch=(8, 8, ni))
    torch.cuda.synchronize()
    ref, _ = O.image_nodes(bank.cpu().numpy(), geom, tab.cpu().numpy(), o_iparams(sv), 200.0)
    got = img.cpu().numpy()
    assert rel_l2(got, ref) < 1e-5
    assert float(np.abs(got[..., pe:]).sum()) > 0
    img2 = sv.new_image()
    tans = cfg.dip_bounds()[:4]
    sv.image_batch(img2, g, bank, tab, engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges]),
                   launch=(8, 8, ni), nwg=-1)
    torch.cuda.synchronize()
    assert torch.equal(img, img2)


def test_node_form_with_source_tables_matches_oracle(gpu):
    """Node form over the layered + anomaly model with each shot's source leg
    from its own table (the four node tables blended at the source): the
    batched kernels' source-table variants (node body and the exact body
    above z_exact) against the oracle's node form fed the same tables, shot
    by shot; a trace of another source in a shot is not imaged."""
    cfg = synth.tiny_layers()
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab)).set_form("nodes", 60.0)
    geom, ranges = synth.geometry(synth.shots_tiny(3, 9, extent=(190.0, 150.0)))
    geom[ranges[1][0] + 4, 0] += 7.0  # a foreign source inside shot 1
    g_d = torch.from_numpy(geom).cuda()
    rng = np.random.default_rng(24)
    bank = sv.qbank(torch.from_numpy(rng.standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)).cuda())
    tans = cfg.dip_bounds()[:4]
    shots = engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges])
    src = sv.src_tables_nodes(g_d, shots, tab)
    img = sv.new_image()
    sv.image_batch(img, g_d, bank, tab, shots, launch=(8, 8, 3), nwg=-1, src=src)
    torch.cuda.synchronize()
    nus, nvs, nw = sv.src_grid()
    ns = shots.n
    hdr = src[:16 * ns].view(torch.float32).view(ns, 4).cpu().numpy()
    body = src[16 * ns:].view(torch.float32).view(ns, nvs, nus, nw, 2).cpu().numpy()
    b = bank.cpu().numpy()
    t = tab.cpu().numpy()
    ref = np.zeros_like(img.cpu().numpy())
    for k, (tr0, n) in enumerate(ranges):
        r, _ = O.image_nodes(b[tr0:tr0 + n], geom[tr0:tr0 + n], t, o_iparams(sv), 60.0, src=(hdr[k], body[k]))
        ref += r
    got = img.cpu().numpy()
    assert rel_l2(got, ref) < 1e-5 and float(np.abs(got).sum()) > 0


@pytest.mark.parametrize("dw, nw", [(10.0, 31), (15.0, 21), (30.0, 11)])
def test_node_form_across_node_spacings(gpu, dw, nw):
    """Table depth nodes at 1, 1.5 and 3 image depths (1, 1-2 and 3 points
    per interval): the node form on the GPU against the oracle's, per shot
    and batched, with the node range from 60 m down."""
    cfg = synth.tiny()
    cfg.tgeom = dict(cfg.tgeom, dw=dw, nw=nw)
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab)).set_form("nodes", 60.0)
    geom, ranges = synth.geometry(synth.shots_tiny(2, 7))
    rng = np.random.default_rng(25)
    bank = sv.qbank(torch.from_numpy(rng.standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)).cuda())
    g = torch.from_numpy(geom).cuda()
    ref, _ = O.image_nodes(bank.cpu().numpy(), geom, tab.cpu().numpy(), o_iparams(sv), 60.0)
    tans = cfg.dip_bounds()[:4]
    for ni in (2, 3, 4):
        img = sv.new_image()
        for tr0, n in ranges:
            sv.image_shot(img, g, tr0, n, bank[tr0:tr0 + n], tab, launch=(8, 8, ni))
        img2 = sv.new_image()
        sv.image_batch(img2, g, bank, tab, engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges]),
                       launch=(8, 8, ni), nwg=-1)
        torch.cuda.synchronize()
        assert rel_l2(img.cpu().numpy(), ref) < 1e-5, (dw, ni)
        assert torch.equal(img, img2), (dw, ni)
