"""The span form of the imaging sum and the pair banks on the GPU: the pair
banks against the plain banks, the span kernels against the oracle's FP32
restatement of the span form (every tile depth and span length, with and
without per-shot source tables), the FP16 pair bank with U16Q tables against
the FP64 exact sum, and captured graphs against eager launches."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from q_psdm_amd import engine, synth
from q_psdm_amd import lib as L
from tests.helpers import o_bparams, o_iparams, rel_l2

pytestmark = pytest.mark.gpu


def _deep_survey(nw=52, dw=20.0, nz=100):
    """A 1 km deep v(z) survey (8 x 6 columns x 100 depths at 10 m, 20 m
    tables): depth enough for every band and span length to matter.  With
    30 m tables every interval holds 3 points (not the 3, 2 slot pattern);
    with 25 m tables and 101 depths the pattern holds but for the last
    interval, which ends at the tables' last node and holds 3."""
    cfg = synth.tiny()
    cfg.nx, cfg.ny, cfg.nz = 8, 6, nz
    cfg.nt = 1201
    cfg.tgeom = dict(cfg.tgeom, nw=nw, dw=dw)
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    return cfg, engine.Survey(cfg, engine.tstar_max_from_tables(tab)), tab


def _data(cfg, nshot=2, nrec=7, seed=31):
    geom, ranges = synth.geometry(synth.shots_tiny(nshot, nrec, extent=(300.0, 300.0)))
    rng = np.random.default_rng(seed)
    tr = rng.standard_normal((geom.shape[0], cfg.nt)).astype(np.float32)
    return geom, ranges, tr


def test_pair_banks_hold_samples_and_differences(gpu):
    cfg, sv, _ = _deep_survey()
    _, _, tr = _data(cfg)
    x = torch.from_numpy(tr).cuda()
    b32 = sv.qbank(x)
    p32 = sv.qbank(x, pair=True)
    p16 = sv.qbank(x, dtype=torch.float16, pair=True)
    assert sv.bank_code(p32) == L.QPSDM_F32P and sv.bank_code(p16) == L.QPSDM_F16P
    torch.cuda.synchronize()
    b = b32.cpu().numpy()
    p = p32.cpu().numpy()
    nt = cfg.nt
    # interior entries: the sample and its forward difference, exactly
    assert np.array_equal(p[:, :, 1:nt - 1, 0], b[:, :, 1:nt - 1])
    assert np.array_equal(p[:, :, 1:nt - 1, 1], (b[:, :, 2:] - b[:, :, 1:nt - 1]).astype(np.float32))
    # entries 0 and nt - 1 of every row are zero
    assert not p[:, :, 0].any() and not p[:, :, nt - 1].any()
    # FP16 pairs: the same values under the row's power-of-two scale
    h = p16.cpu()
    hdr = h[:, :8].contiguous().view(torch.float32).numpy()
    assert np.all(hdr[:, 0] * hdr[:, 1] == 1.0)
    v = h[:, 8:8 + 2 * cfg.K * nt].float().numpy().reshape(-1, cfg.K, nt, 2) * hdr[:, 0, None, None, None]
    assert not v[:, :, 0].any() and not v[:, :, nt - 1].any()
    assert np.allclose(v[:, :, 1:nt - 1], p[:, :, 1:nt - 1], rtol=2e-3, atol=2e-3 * np.abs(b).max())
    # every stored half is finite and below 2^15
    assert np.all(np.abs(h[:, 8:].float().numpy()) < 2 ** 15)
    # decode_bank gives the samples back
    assert torch.equal(sv.decode_bank(p32)[:, :, 1:nt - 1], b32[:, :, 1:nt - 1])


def _span_image(sv, cfg, geom, ranges, bank, tab, ni, src=None):
    g = torch.from_numpy(geom).cuda()
    tans = cfg.dip_bounds()[:4]
    shots = engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges])
    img = sv.new_image()
    sv.image_batch(img, g, bank, tab, shots, launch=(8, 8, ni), nwg=-1, src=src)
    torch.cuda.synchronize()
    return img, shots, g


@pytest.mark.parametrize("ni", [4, 8])
@pytest.mark.parametrize("sched", [((0.0, 1),), ((0.0, 2), (500.0, 4), (700.0, 8)), ((0.0, 8),),
                                   ((450.0, 4),)])
@pytest.mark.parametrize("geo", [(20.0, 100, 52), (30.0, 100, 36), (25.0, 101, 41)])
def test_span_form_matches_oracle(gpu, ni, sched, geo):
    dw, nz, nw = geo
    cfg, sv, tab = _deep_survey(nw=nw, dw=dw, nz=nz)
    geom, ranges, tr = _data(cfg)
    x = torch.from_numpy(tr).cuda()
    b32 = sv.qbank(x)
    p32 = sv.qbank(x, pair=True)
    z_exact = 300.0
    sv.set_form("spans", z_exact, sched)
    img, _, _ = _span_image(sv, cfg, geom, ranges, p32, tab, ni)
    ip = o_iparams(sv)
    ref, _ = O.image_spans(b32.cpu().numpy(), geom, tab.cpu().numpy(), ip, z_exact,
                           O.span_table(ip, z_exact, sched))
    assert rel_l2(img.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("ni", [4, 8])
def test_span_form_with_source_tables_matches_oracle(gpu, ni):
    """Each shot's source leg from its own source table (the four-node
    blend), a trace of a foreign source in the gather skipped."""
    cfg, sv, tab = _deep_survey()
    geom, ranges, tr = _data(cfg, nshot=2, nrec=6)
    geom[1, 0] += 3.0  # a foreign source in shot 0
    x = torch.from_numpy(tr).cuda()
    b32 = sv.qbank(x)
    p32 = sv.qbank(x, pair=True)
    sched = ((0.0, 2), (600.0, 8))
    sv.set_form("spans", 200.0, sched)
    g = torch.from_numpy(geom).cuda()
    tans = cfg.dip_bounds()[:4]
    shots = engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges])
    src = sv.src_tables_nodes(g, shots, tab)
    img, _, _ = _span_image(sv, cfg, geom, ranges, p32, tab, ni, src=src)
    nus, nvs, nw = sv.src_grid()
    hdr = src[:16 * shots.n].view(torch.float32).view(shots.n, 4).cpu().numpy()
    body = src[16 * shots.n:].view(torch.float32).view(shots.n, nvs, nus, nw, 2).cpu().numpy()
    ip = o_iparams(sv)
    spans = O.span_table(ip, 200.0, sched)
    b = b32.cpu().numpy()
    ref = np.zeros(img.shape, np.float64)
    for k, (tr0, n) in enumerate(ranges):
        r, _ = O.image_spans(b[tr0:tr0 + n], geom[tr0:tr0 + n], tab.cpu().numpy(), ip, 200.0, spans,
                             src=(hdr[k], body[k]))
        ref += r
    assert rel_l2(img.cpu().numpy(), ref) < 1e-5


def test_span_form_fp16_pairs_u16q_against_fp64(gpu):
    """FP16 pair bank + U16Q tables in the span form below 300 m of a 1 km
    image: within 1e-3 of the FP64 exact sum."""
    cfg, sv, tab = _deep_survey()
    geom, ranges, tr = _data(cfg, seed=32)
    p16 = sv.qbank(torch.from_numpy(tr).cuda(), dtype=torch.float16, pair=True)
    q = sv.quantize_tables(tab, code="u16q")
    sv.set_form("spans", 300.0, ((0.0, 2), (500.0, 4), (700.0, 8)))
    img, _, _ = _span_image(sv, cfg, geom, ranges, p16, q, 4)
    ref, _ = O.image(O.qbank(tr, o_bparams(sv)), geom, tab.cpu().numpy(), o_iparams(sv))
    assert rel_l2(img.cpu().numpy(), ref) < 1e-3


def test_span_form_graph_replays_match_eager(gpu):
    """[Q bank -> span-form imaging] captured once and replayed three times,
    each replay bit-identical to the eager calls."""
    cfg, sv, tab = _deep_survey()
    geom, ranges, tr = _data(cfg, seed=33)
    x = torch.from_numpy(tr).cuda()
    q = sv.quantize_tables(tab, code="u16q")
    sv.set_form("spans", 300.0, ((0.0, 2), (600.0, 8)))
    tans = cfg.dip_bounds()[:4]
    shots = engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges])
    g = torch.from_numpy(geom).cuda()
    eager = sv.new_image()
    sv.image_batch(eager, g, sv.qbank(x, dtype=torch.float16, pair=True), q, shots, launch=(8, 8, 8), nwg=-1)
    bank = torch.empty(sv.bank_shape(x.shape[0], torch.float16, pair=True), dtype=torch.float16, device="cuda")
    img = sv.new_image()
    graph = engine.BatchGraph(sv, x, bank, img, g, q, shots, launch=(8, 8, 8), nwg=-1, stages=5)
    for _ in range(3):
        img.zero_()
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        assert torch.equal(img, eager)
    graph.close()


def test_span_form_band_split_and_tile_depth_leave_the_image_unchanged(gpu):
    """Band starts lie a multiple of 8 intervals from the node range's first,
    so a band of spans of 8 split by another band of spans of 8 (a schedule
    depth of 940 m starts it at 960 m, 32 intervals below the range's first
    at 320 m; four bands, as many as a schedule may hold) keeps every span,
    and tiles of 4 and 8 intervals image spans of at most 4 with the same
    per-point arithmetic: the images are bit-identical."""
    cfg, sv, tab = _deep_survey()
    geom, ranges, tr = _data(cfg, seed=34)
    p16 = sv.qbank(torch.from_numpy(tr).cuda(), dtype=torch.float16, pair=True)
    q = sv.quantize_tables(tab, code="u16q")
    imgs = {}
    for name, sched, ni in (("three", ((0.0, 2), (500.0, 4), (700.0, 8)), 8),
                            ("four", ((0.0, 2), (500.0, 4), (700.0, 8), (940.0, 8)), 8),
                            ("short4", ((0.0, 1), (500.0, 2), (700.0, 4)), 4),
                            ("short8", ((0.0, 1), (500.0, 2), (700.0, 4)), 8)):
        sv.set_form("spans", 300.0, sched)
        imgs[name], _, _ = _span_image(sv, cfg, geom, ranges, p16, q, ni)
    assert float(imgs["three"].abs().max()) > 0 and float(imgs["short4"].abs().max()) > 0
    assert torch.equal(imgs["three"], imgs["four"])
    assert torch.equal(imgs["short4"], imgs["short8"])


def test_span_form_rejects_other_storage_and_per_shot_calls(gpu):
    cfg, sv, tab = _deep_survey()
    geom, ranges, tr = _data(cfg)
    x = torch.from_numpy(tr).cuda()
    g = torch.from_numpy(geom).cuda()
    sv.set_form("spans", 300.0)
    tans = cfg.dip_bounds()[:4]
    shots = engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges])
    from q_psdm_amd.lib import QpsdmError
    with pytest.raises(QpsdmError):  # a plain bank
        sv.image_batch(sv.new_image(), g, sv.qbank(x), tab, shots, launch=(8, 8, 4), nwg=-1)
    for tz in (2, 3):  # tile_z outside 4, 8
        with pytest.raises(QpsdmError):
            sv.image_batch(sv.new_image(), g, sv.qbank(x, pair=True), tab, shots, launch=(8, 8, tz), nwg=-1)
    with pytest.raises(QpsdmError):  # the per-shot entry point
        sv.image_shot(sv.new_image(), g, 0, ranges[0][1], sv.qbank(x, pair=True), tab)


def test_span_form_calls_from_two_threads(gpu):
    """Two host threads, each on its own stream, image their own batches in
    the span form at the same time (the library's per-device band streams
    are forked and joined under a lock per call): each image equals the
    same calls made one after another."""
    import threading
    cfg, sv, tab = _deep_survey()
    sv.set_form("spans", 300.0, ((0.0, 2), (500.0, 4), (700.0, 8)))
    q = sv.quantize_tables(tab, code="u16q")
    tans = cfg.dip_bounds()[:4]
    jobs = []
    for seed in (41, 42):
        geom, ranges, tr = _data(cfg, nshot=3, nrec=9, seed=seed)
        g = torch.from_numpy(geom).cuda()
        bank = sv.qbank(torch.from_numpy(tr).cuda(), dtype=torch.float16, pair=True)
        jobs.append((g, bank, engine.ShotList([(tr0, n) + tuple(tans) for tr0, n in ranges])))
    torch.cuda.synchronize()

    def run(job, img, stream, reps):
        g, bank, shots = job
        with torch.cuda.stream(stream):
            for _ in range(reps):
                sv.image_batch(img, g, bank, q, shots, launch=(8, 8, 4), nwg=-1, stream=stream)
        stream.synchronize()

    ref = []
    for job in jobs:
        img = sv.new_image()
        run(job, img, torch.cuda.current_stream(), 3)
        ref.append(img)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    imgs = [sv.new_image(), sv.new_image()]
    torch.cuda.synchronize()
    ts = [threading.Thread(target=run, args=(jobs[i], imgs[i], streams[i], 3)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    torch.cuda.synchronize()
    for i in range(2):
        assert float(ref[i].abs().max()) > 0
        assert torch.equal(imgs[i], ref[i])

