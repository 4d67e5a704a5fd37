"""Launch-shape autotuner for the imaging kernels.

The imaging kernels are bound by vector-instruction issue, and the best
tile shape trades the per-trace column set-up (amortised over more depth
points per lane) against registers (fewer resident waves) and the
granularity of the per-trace tile skip.  The best shape depends on the
table and bank storage types, the image and table geometry and the trace
count, so it is measured: every candidate images a slice of a
representative shot, timed with HIP events, and the fastest is kept.

Candidates: column tiles 8x8, 16x4 and 4x16 (one wave of columns) times
depth points per lane ZP in {2, 3, 4, 6, 8} (in the node form: table depth
intervals per tile in {2, 3, 4}; in the span form {4, 8}); for the
batched kernel also the grid:
persistent at the resident capacity or twice it, or one work-group per tile
(nwg = -1; the node form always runs one work-group per tile).  Results are cached per (geometry, storage
types, kernel) in memory for the process.
"""
import itertools

import torch

TILES = ((8, 8), (16, 4), (4, 16))
DEPTHS = (2, 3, 4, 6, 8)
NODE_INTERVALS = (2, 3, 4)  # node form: table depth intervals per tile
SPAN_INTERVALS = (4, 8)  # span form: table depth intervals per tile
GRIDS = (0, -2, -1)  # batched kernel: resident, twice resident (-2), one group per tile (-1)
_cache = {}


def _key(sv, bank, tab, kind):
    c = sv.cfg
    return (c.nx, c.ny, c.nz, c.K, c.nt, tuple(sorted(c.tgeom.items())), sv.bank_code(bank),
            getattr(tab, "qpsdm_code", 0), kind, sv.form, float(sv.ip.z_exact),
            tuple(sv.spans()) if sv.form == "spans" else ())


def _time(fn, reps=2):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(sv, geom, bank, tab, shots=None, ntr=400, kind="batch", log=None, reps=1):
    """Best launch for imaging `bank`/`tab` on this survey.

    sv: engine.Survey; geom: device (sx, sy, gx, gy) rows; bank: its Q bank;
    tab: tables in the storage type to tune for; shots: engine.ShotList for
    kind="batch" (the first shot is sliced to ntr traces), or None for
    kind="shot" (rows [0, ntr) as one run).  Returns (launch tuple, nwg,
    table of (launch, nwg, ms))."""
    from . import engine
    key = _key(sv, bank, tab, kind)
    if key in _cache:
        return _cache[key]
    img = sv.new_image()
    n = min(ntr, geom.shape[0])
    results = []
    depths = DEPTHS
    if kind == "batch":
        tr0, n0, *tans = shots.host[0]
        one = engine.ShotList([(tr0, min(n, n0)) + tuple(tans)])
        grids = GRIDS
    else:
        grids = (0,)
    if sv.form == "nodes":  # node tiles: tile_z intervals, one work-group per tile
        depths, grids = NODE_INTERVALS, ((-1,) if kind == "batch" else (0,))
    if sv.form == "spans":  # span tiles: tile_z intervals, batched, one work-group per tile
        assert kind == "batch", "the span form runs through the batched kernel"
        depths, grids = SPAN_INTERVALS, (-1,)
    for (tx, ty), zp, g in itertools.product(TILES, depths, grids):
        launch = (tx, ty, zp, 0)
        if kind == "batch":
            nwg = g if g != -2 else 2 * _resident(sv, geom, bank, tab, one, launch)
            fn = lambda: sv.image_batch(img, geom, bank, tab, one, launch=launch, nwg=nwg)
        else:
            nwg = 0
            fn = lambda: sv.image_shot(img, geom, 0, n, bank, tab, launch=launch)
        ms = _time(fn, reps)
        results.append((launch, nwg, ms))
        if log:
            log(f"autotune {kind}: tile {tx}x{ty}x{zp} nwg {nwg or 'auto'}: {ms:.2f} ms")
    best = min(results, key=lambda r: r[2])
    _cache[key] = (best[0], best[1], results)
    return _cache[key]


def _resident(sv, geom, bank, tab, shots, launch):
    """Resident-capacity grid of the persistent kernel for this launch (the
    library's own choice when nwg = 0), read back by a zero-shot query."""
    from . import lib as L
    import ctypes as C
    out = C.c_int(0)
    bcode, tcode = sv._codes(bank, tab)
    lp = L.Launch(*launch[:3])
    L.check(L.lib().qpsdm_image_batch_grid(C.byref(sv.ip), C.byref(lp), bcode, tcode, C.byref(out)))
    return out.value
