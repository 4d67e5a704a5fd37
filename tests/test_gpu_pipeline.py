"""The qpsdm_ctx pipeline on the GPU: SU bytes in pinned staging slots, GPU
decode, Q bank and imaging, against the direct kernel path on the same
traces (same kernels, same trace order: identical images)."""
import numpy as np
import pytest
import torch

from q_psdm_amd import engine, synth
from q_psdm_amd import lib as L
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

OX, OY = 670000.0, 5100000.0   # world origin removed in FP64 before the FP32 cast


def _setup():
    cfg = synth.tiny()
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab))
    geom, ranges = synth.geometry(synth.shots_tiny(3, 7))
    # integer world coordinates (scalco 0 -> scale 1) and the image-relative floats
    wxy = np.round(geom.astype(np.float64) + [OX, OY, OX, OY]).astype(np.int32)
    rel = (wxy.astype(np.float64) - [OX, OY, OX, OY]).astype(np.float32)
    traces = np.random.default_rng(9).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    return cfg, sv, tab, wxy, rel, traces


def _direct(sv, tab, rel, traces, keep=None):
    img = sv.new_image()
    g = rel.copy()
    if keep is not None:
        g[~keep] = np.nan
    g_d = torch.from_numpy(g).cuda()
    bank = sv.qbank(torch.from_numpy(traces).cuda())
    sv.image_shot(img, g_d, 0, g.shape[0], bank, tab)
    return img.cpu().numpy()


def _gp(cfg, **kw):
    return L.GeomParams(1.0, kw.get("fxv1", -1e30), kw.get("exv1", 1e30), kw.get("fyv1", -1e30),
                        kw.get("eyv1", 1e30), OX, OY, OX, OY, cfg.dx, cfg.dy)


@pytest.mark.parametrize("batch", [5, 64])
def test_su_pipeline_matches_direct_path(gpu, batch):
    cfg, sv, tab, wxy, rel, traces = _setup()
    ref = _direct(sv, tab, rel, traces)
    pl = engine.Pipeline(sv, _gp(cfg), batch_traces=batch)
    pl.set_tables(tab)
    su = engine.su_encode(traces, wxy)
    rec = 240 + 4 * cfg.nt
    # push in uneven pieces (whole records)
    cuts = [0, 3 * rec, 4 * rec, 13 * rec, len(su)]
    for a, b in zip(cuts[:-1], cuts[1:]):
        pl.push_su(su[a:b])
    got = pl.image()
    assert pl.traces_pushed() == traces.shape[0]
    np.testing.assert_array_equal(got, ref)
    nrm = pl.image(normalize=True)
    np.testing.assert_allclose(nrm, ref / sv.win.nsb, rtol=1e-6)
    pl.reset_image()
    assert not pl.image().any()
    pl.close()


def test_decoded_trace_paths_match(gpu):
    cfg, sv, tab, wxy, rel, traces = _setup()
    ref = _direct(sv, tab, rel, traces)
    pl = engine.Pipeline(sv, _gp(cfg), batch_traces=4, table_dtype="f32")
    pl.set_tables(tab.cpu().numpy())            # host tables
    pl.push_traces(traces[:10], rel[:10])
    su = engine.su_encode(traces[10:], wxy[10:])
    rec = 240 + 4 * cfg.nt
    for i in range(traces.shape[0] - 10):        # the per-trace psdm_kernel entry point
        pl.psdm_kernel(su[i * rec:(i + 1) * rec])
    np.testing.assert_array_equal(pl.image(), ref)
    pl.close()


def test_extent_rejection(gpu):
    cfg, sv, tab, wxy, rel, traces = _setup()
    # model extent ending at the median of max(sx, gx): about half the traces
    # have a source or receiver beyond it and are dropped (src/q-psdm.cu:2120-2123)
    edge = float(np.median(np.maximum(wxy[:, 0], wxy[:, 2])))
    keep = (wxy[:, 2] <= edge) & (wxy[:, 0] <= edge)
    assert 0 < keep.sum() < keep.size
    ref = _direct(sv, tab, rel, traces, keep)
    pl = engine.Pipeline(sv, _gp(cfg, exv1=edge), batch_traces=6)
    pl.set_tables(tab)
    pl.push_su(engine.su_encode(traces, wxy))
    np.testing.assert_array_equal(pl.image(), ref)
    pl.close()


def test_pipeline_mixed_dtypes_and_table_builder(gpu):
    cfg, sv, tab, wxy, rel, traces = _setup()
    ref = _direct(sv, tab, rel, traces)
    pl = engine.Pipeline(sv, _gp(cfg), bank_dtype="bf16", table_dtype="u16", batch_traces=8)
    pl.build_tables()                            # built on the device by the context
    pl.push_su(engine.su_encode(traces, wxy))
    assert rel_l2(pl.image(), ref) < 5e-3
    pl.close()
    pl = engine.Pipeline(sv, _gp(cfg), bank_dtype="f16", table_dtype="u16q", batch_traces=8)
    pl.build_tables()
    pl.push_su(engine.su_encode(traces, wxy))
    assert rel_l2(pl.image(), ref) < 5e-3
    pl.close()


@pytest.mark.parametrize("storage", [("f16p", "u16q"), ("f32p", "f32")])
def test_span_form_pipeline_matches_batched_span_image(gpu, storage):
    """A context in the span form (FP16 pair bank with U16Q tables, or FP32
    pair bank with FP32 tables) images each staging slot as one batched
    shot: SU records pushed through it give the image of one batched
    span-form call over the same traces, bit for bit (per point the traces
    are summed in the same order)."""
    bank_code, tab_code = storage
    cfg = synth.tiny()
    cfg.nx, cfg.ny, cfg.nz, cfg.nt = 8, 6, 100, 1201
    cfg.tgeom = dict(cfg.tgeom, nw=52, dw=20.0)
    tab = engine.Survey(cfg, 0.01).build_tables()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab)).set_form("spans", 200.0,
                                                                      ((0.0, 2), (500.0, 4), (700.0, 8)))
    geom, _ = synth.geometry(synth.shots_tiny(3, 7, extent=(300.0, 300.0)))
    wxy = np.round(geom.astype(np.float64) + [OX, OY, OX, OY]).astype(np.int32)
    rel = (wxy.astype(np.float64) - [OX, OY, OX, OY]).astype(np.float32)
    traces = np.random.default_rng(10).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    q = sv.quantize_tables(tab, code="u16q") if tab_code == "u16q" else tab
    ref = sv.new_image()
    bank = sv.qbank(torch.from_numpy(traces).cuda(), dtype=torch.float16 if bank_code == "f16p" else torch.float32,
                    pair=True)
    tans = (sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2)
    sv.image_batch(ref, torch.from_numpy(rel).cuda(), bank, q, engine.ShotList([(0, geom.shape[0]) + tans]),
                   launch=(8, 8, 4), nwg=-1)
    torch.cuda.synchronize()
    pl = engine.Pipeline(sv, _gp(cfg), bank_dtype=bank_code, table_dtype=tab_code, batch_traces=8,
                         launch=(8, 8, 4))
    pl.set_tables(tab)
    pl.push_su(engine.su_encode(traces, wxy))
    got = pl.image()
    assert np.abs(got).max() > 0
    np.testing.assert_array_equal(got, ref.cpu().numpy())
    pl.close()
    with pytest.raises(L.QpsdmError):  # the span form needs a pair bank
        engine.Pipeline(sv, _gp(cfg), bank_dtype="f16", table_dtype="u16q", batch_traces=8)


def test_span_form_pipeline_images_a_held_slot_under_its_tables(gpu):
    """The span-form context images its two slots' shots in one batched call,
    holding a banked slot until the other is banked: a table swap while a
    slot is held images that slot under the tables it was pushed under
    (here alone), and the later slots as a pair under the new ones."""
    cfg = synth.tiny()
    cfg.nx, cfg.ny, cfg.nz, cfg.nt = 8, 6, 100, 1201
    cfg.tgeom = dict(cfg.tgeom, nw=52, dw=20.0)
    tab = engine.Survey(cfg, 0.01).build_tables()
    tab2 = tab.clone()
    tab2[..., 0] *= 1.03  # the same tables with 3 % longer times
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab)).set_form("spans", 200.0,
                                                                      ((0.0, 2), (500.0, 4), (700.0, 8)))
    geom, _ = synth.geometry(synth.shots_tiny(3, 7, extent=(300.0, 300.0)))
    wxy = np.round(geom.astype(np.float64) + [OX, OY, OX, OY]).astype(np.int32)
    rel = torch.from_numpy((wxy.astype(np.float64) - [OX, OY, OX, OY]).astype(np.float32)).cuda()
    traces = np.random.default_rng(11).standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    bank = sv.qbank(torch.from_numpy(traces).cuda(), dtype=torch.float16, pair=True)
    tans = (sv.ip.tanx1, sv.ip.tanx2, sv.ip.tany1, sv.ip.tany2)
    ref = sv.new_image()
    sv.image_batch(ref, rel, bank, sv.quantize_tables(tab, code="u16q"), engine.ShotList([(0, 8) + tans]),
                   launch=(8, 8, 4), nwg=-1)
    sv.image_batch(ref, rel, bank, sv.quantize_tables(tab2, code="u16q"),
                   engine.ShotList([(8, 8) + tans, (16, 5) + tans]), launch=(8, 8, 4), nwg=-1)
    torch.cuda.synchronize()
    su = engine.su_encode(traces, wxy)
    rec = len(su) // geom.shape[0]
    pl = engine.Pipeline(sv, _gp(cfg), bank_dtype="f16p", table_dtype="u16q", batch_traces=8, launch=(8, 8, 4))
    pl.set_tables(tab)
    pl.push_su(su[:8 * rec])  # one full slot: banked and held
    pl.set_tables(tab2)
    pl.push_su(su[8 * rec:])  # a full slot and 5 staged traces, imaged as a pair at the flush
    got = pl.image()
    assert np.abs(got).max() > 0
    np.testing.assert_array_equal(got, ref.cpu().numpy())
    pl.close()

