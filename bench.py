"""Benchmark of the Q-compensated Kirchhoff PSDM imaging path on one MI355X.

Headline workload (the surrogate variant of the layered survey): the
401 x 401 x 301 image at 10 m, 2000 jittered receivers per shot over 4 x 4 km,
2001 samples at 2 ms, K = 32 t* levels, 25 m tables on 5 x 5 surface nodes,
the layered + Gaussian-anomaly model with each layer's Q drawn from
U[30, 200].  A step images `--shots-per-step` shots (Q bank of their traces
+ the Kirchhoff sum into the full image), inputs resident in HBM.  Three
forms image the same shots:

  fp32   FP32 bank, FP32 tables, exact form, one launch per shot
  mixed  FP16 pair bank (sample and difference per entry, a power-of-two
         scale per trace), 16-bit fixed-point tables stored as corner quads
         (U16Q), the batched kernel under captured hipGraphs, the span form
         of the sum (below z_exact the taper ramps and amplitude at the ends
         of spans of table intervals, interpolated in 1/z; --mixed-form picks
         another form, auto the autotuner's fastest); the launch shape is
         picked by the autotuner
  fast   mixed + the learned source leg: per-shot source tables from the
         BF16-MFMA surrogate (trained on the device first; table fallback
         where the network misses by more than 0.5 ms), in the same graphs

Every form's image of a subset of shot 0's traces over the whole volume is
compared with the FP64 oracle fed the same tables (the node tables, or the
fast form's own source table): the metric's <= 1e-3 bar.  `value` is the
fast form's image-point x trace updates/s when it meets that bar, else the
FP32 form's.  The line also carries the model error against exact FP64
source rays (separately labelled), the imaging kernels' VALU-issue
roofline from hardware counters (two rocprofv3 passes of this script), the
HBM traffic (FETCH_SIZE, WRITE_SIZE, scaled by a copy kernel of known size)
and a unique-bytes HBM figure, config 1 timed in full on the GPU, the
batching variant (uneven gathers with per-shot apertures through captured
graphs), the SEG-Y decode and the end-to-end pipeline rate, and the CPU
baseline (the oracle's FP32 path on bounded samples, OpenMP).
"""
import argparse
import csv
import glob
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NCU = 256
CLOCK_GHZ = 2.4          # MI355X peak engine clock
CALIB_MIB = 512


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def kernel_name(form, nwg, sum_form):
    """Main imaging kernel a form launches."""
    if form == "fp32":
        return "image_nodes_kernel" if sum_form == "nodes" else "image_kernel"
    if sum_form == "spans":
        return "image_batch_spans_kernel"
    if sum_form == "nodes":
        return "image_batch_nodes_kernel"
    return "image_batch_tiles_kernel" if nwg == -1 else "image_batch_kernel"


# ---- workload ----------------------------------------------------------------

class Workload:
    def __init__(self, args, torch, engine, synth):
        self.cfg = cfg = synth.config5()
        self.S = args.shots_per_step
        self.nstep = args.warmup + args.steps
        nshot = self.nstep * self.S
        t0 = time.time()
        probe = engine.Survey(cfg, 0.05)
        self.tab = probe.build_tables()
        torch.cuda.synchronize()
        tsm = engine.tstar_max_from_tables(self.tab)
        self.sv = engine.Survey(cfg, tsm)                                   # fp32: exact form
        self.svm = engine.Survey(cfg, tsm).set_form(args.mixed_form if args.mixed_form != "auto" else "exact",
                                                    args.z_exact, parse_spans(args.spans))  # mixed / fast
        log(f"tables {tuple(self.tab.shape)} {self.tab.numel() * 4 / 1e9:.2f} GB in {time.time() - t0:.1f}s; "
            f"dtstar={self.sv.dtstar * 1e3:.3f} ms nsb={self.sv.win.nsb} nww={self.sv.win.nww}")
        self.tab_u16 = self.sv.quantize_tables(self.tab, code="u16q")
        shots = synth.shots_config2(nshots=nshot)
        self.geom, self.ranges = synth.geometry(shots)
        self.g_d = torch.from_numpy(self.geom).cuda()
        t0 = time.time()
        self.traces = [synth.reflectivity_traces_gpu(cfg, n, seed=cfg.seed * 1000 + s)
                       for s, (_, n) in enumerate(self.ranges)]
        torch.cuda.synchronize()
        log(f"traces for {nshot} shots in {time.time() - t0:.1f}s")
        self.tans = cfg.dip_bounds()[:4]
        self.counts = [self.sv.count_shot(self.g_d, tr0, n) for tr0, n in self.ranges]
        torch.cuda.synchronize()
        log(f"updates per shot: {self.counts}")
        self.steps = []
        for k in range(self.nstep):
            idx = list(range(k * self.S, (k + 1) * self.S))
            tr = torch.cat([self.traces[i] for i in idx], 0).contiguous()
            rows = np.concatenate([self.geom[self.ranges[i][0]:self.ranges[i][0] + self.ranges[i][1]] for i in idx])
            desc, t = [], 0
            for i in idx:
                desc.append((t, self.ranges[i][1]) + tuple(self.tans))
                t += self.ranges[i][1]
            self.steps.append(dict(idx=idx, traces=tr, geom=torch.from_numpy(np.ascontiguousarray(rows)).cuda(),
                                   shots=engine.ShotList(desc), updates=sum(self.counts[i] for i in idx)))


def parse_spans(text):
    """--spans "z:S,z:S,...": from depth z (m) down, spans of S table intervals."""
    return [(float(a), int(b)) for a, b in (t.split(":") for t in text.split(",") if t)]


def tune_form(form, W, args, torch, engine):
    """Launch shape (and, for the mixed forms with --mixed-form auto, the form
    of the sum) from the autotuner on a 400-trace slice of the first shot
    (FP32) or on the whole first shot (mixed forms), unless --ztile pins it."""
    from q_psdm_amd import autotune
    if args.ztile:
        return (args.tile_x, args.tile_y, args.ztile), args.nwg
    st = W.steps[0]
    n = min(400, st["traces"].shape[0])
    if form == "fp32":
        bank = W.sv.qbank(st["traces"][:n])
        launch, nwg, res = autotune.tune(W.sv, st["geom"], bank, W.tab, ntr=n, kind="shot")
        best = min(r[2] for r in res)
    else:
        cands = []
        forms = ("exact", "nodes", "spans") if args.mixed_form == "auto" else (args.mixed_form,)
        # (the whole first shot, three timed calls per shape: the span form's
        # concurrent bands favour other shapes on a full shot than on a short
        # slice, and its best shapes lie within a few per cent of each other)
        n = W.ranges[0][1]
        for f in forms:
            W.svm.set_form(f, args.z_exact, parse_spans(args.spans))
            bank = W.svm.qbank(st["traces"][:n], dtype=torch.float16, pair=W.svm.pair_bank())
            launch, nwg, res = autotune.tune(W.svm, st["geom"], bank, W.tab_u16, shots=st["shots"], ntr=n,
                                             kind="batch", reps=3)
            cands.append((min(r[2] for r in res), f, launch, nwg))
            log(f"[{form}] autotune {f} form: {launch[:3]} nwg {nwg} {cands[-1][0]:.1f} ms on {n} traces")
        best, f, launch, nwg = min(cands)
        W.svm.set_form(f, args.z_exact, parse_spans(args.spans))
    log(f"[{form}] launch {launch[:3]} nwg {nwg or 'resident'} form {W.svm.form if form != 'fp32' else 'exact'} "
        f"({best:.1f} ms on the autotuner's traces)")
    return launch, nwg


def time_form(form, W, args, torch, engine, sur=None, mask=None):
    """Warmup + timed steps of one form; returns (image, stats)."""
    cfg = W.cfg
    sv = W.sv if form == "fp32" else W.svm
    img = sv.new_image()
    launch, nwg = tune_form(form, W, args, torch, engine)
    ev = []
    # the Q bank (and source tables) of each shot or step run on a stream of
    # their own into their own buffers, so nothing but their own imaging
    # waits for them (the FP32 form's next bank overlaps the current
    # imaging); the image stays on the main stream, in shot order
    main, side = torch.cuda.current_stream(), torch.cuda.Stream()
    if form == "fp32":
        nmax = max(n for _, n in W.ranges)
        banks = [torch.empty((nmax, cfg.K, cfg.nt), dtype=torch.float32, device="cuda") for _ in range(2)]
        freed = [None, None]  # per bank buffer: the image launch that last read it
        nshot_run = [0]

        def run(k):
            es = []
            for i in W.steps[k]["idx"]:
                tr0, n = W.ranges[i]
                b = nshot_run[0] % 2
                nshot_run[0] += 1
                e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
                with torch.cuda.stream(side):
                    if freed[b] is not None:
                        side.wait_event(freed[b])
                    e[0].record(side)
                    sv.qbank(W.traces[i], out=banks[b][:n], stream=side)
                    e[1].record(side)
                main.wait_event(e[1])
                e[2].record(main)
                sv.image_shot(img, W.g_d, tr0, n, banks[b], W.tab, launch=launch)
                e[3].record(main)
                freed[b] = e[3]
                es.append(e)
            ev.append(es)
    else:
        # per step two captured graphs on one stream: [Q bank -> source tables]
        # and [batched imaging], HIP events around the second
        graphs = []
        for k in range(W.nstep):
            st = W.steps[k]
            bank = torch.empty(sv.bank_shape(st["traces"].shape[0], torch.float16, pair=sv.pair_bank()),
                               dtype=torch.float16, device="cuda")
            su = sur if form == "fast" else None
            g_pre = engine.BatchGraph(sv, st["traces"], bank, img, st["geom"], W.tab_u16, st["shots"], launch=launch,
                                      nwg=nwg, surrogate=su, tab_f32=W.tab, mask=mask, stages=3)
            g_img = engine.BatchGraph(sv, st["traces"], bank, img, st["geom"], W.tab_u16, st["shots"], launch=launch,
                                      nwg=nwg, surrogate=su, tab_f32=W.tab, mask=mask, src=g_pre.src, stages=4)
            graphs.append((g_pre, g_img, bank))

        def run(k):
            e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            e[0].record(side)
            graphs[k][0].launch(side)
            e[1].record(side)
            main.wait_event(e[1])
            e[2].record(main)
            graphs[k][1].launch(main)
            e[3].record(main)
            ev.append([e])

    for k in range(args.warmup):
        run(k)
    torch.cuda.synchronize()
    ev.clear()
    t0 = time.perf_counter()
    for k in range(args.warmup, W.nstep):
        run(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    upd = sum(W.steps[k]["updates"] for k in range(args.warmup, W.nstep))
    if form == "fp32":
        img_ms = [e[2].elapsed_time(e[3]) for es in ev for e in es]
        bank_ms = [e[0].elapsed_time(e[1]) for es in ev for e in es]
        per = 1
    else:
        img_ms = [es[0][2].elapsed_time(es[0][3]) for es in ev]
        pre_ms = [es[0][0].elapsed_time(es[0][1]) for es in ev]
        bank_ms = None
        per = W.S
    # per shot (one imaging dispatch per shot in every form, plus the exact
    # body's shallow dispatch in the node form), as rocprofv3 averages it
    avg_shot_ms = float(np.mean(img_ms)) / per
    node_form = sv.form in ("nodes", "spans")
    stats = {"value": upd / elapsed, "ms_per_step": elapsed / args.steps * 1e3, "launch": list(launch[:3]),
             "nwg": nwg, "form": sv.form, "z_exact": sv.ip.z_exact if node_form else None,
             "spans": sv.spans() if sv.form == "spans" else None,
             "imaging_kernel": kernel_name(form, nwg, sv.form), "imaging_ms_per_shot": avg_shot_ms,
             "updates_per_shot": upd / (len(img_ms) * per),
             "imaging_share_of_step": sum(img_ms) / (elapsed * 1e3)}
    if form != "fp32":
        stats["bank_and_source_tables_ms_per_step"] = float(np.mean(pre_ms))
    if bank_ms:
        nb = W.ranges[0][1]
        stats["qbank_avg_launch_ms"] = float(np.mean(bank_ms))
        stats["qbank_GBs"] = nb * (4 * cfg.nt + cfg.K * cfg.nt * 4) / (float(np.mean(bank_ms)) / 1e3) / 1e9
    log(f"[{form}] {stats['value']:.3e} updates/s, {stats['ms_per_step']:.1f} ms/step, imaging "
        f"{avg_shot_ms:.1f} ms per shot ({stats['imaging_share_of_step'] * 100:.0f}% of the step)")
    if form != "fp32":
        for g_pre, g_img, _ in graphs:
            g_img.close()
            g_pre.close()
    return img, stats


def time_fp32_batched(W, args, torch, engine):
    """The FP32 form (FP32 bank and tables, exact form of the sum) through the
    batched imaging instead of one launch per shot: the same per-point
    arithmetic with the batched kernels' trace cull (64 traces tested per
    pass against the tile's aperture).  Reported beside the headline so that
    the fast form is also compared with the best FP32 imaging, not only the
    plain per-shot one."""
    from q_psdm_amd import autotune
    cfg = W.cfg
    sv = W.sv
    st = W.steps[0]
    n = min(400, st["traces"].shape[0])
    launch, nwg, res = autotune.tune(sv, st["geom"], sv.qbank(st["traces"][:n]), W.tab, shots=st["shots"], ntr=n,
                                     kind="batch")
    img = sv.new_image()
    nmax = max(int(W.steps[k]["traces"].shape[0]) for k in range(W.nstep))
    bank = torch.empty((nmax, cfg.K, cfg.nt), dtype=torch.float32, device="cuda")
    ev = []

    def run(k):
        stk = W.steps[k]
        ntr = int(stk["traces"].shape[0])
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        sv.qbank(stk["traces"], out=bank[:ntr])
        e[1].record()
        sv.image_batch(img, stk["geom"], bank[:ntr], W.tab, stk["shots"], launch=launch, nwg=nwg)
        e[2].record()
        ev.append(e)

    for k in range(args.warmup):
        run(k)
    torch.cuda.synchronize()
    ev.clear()
    t0 = time.perf_counter()
    for k in range(args.warmup, W.nstep):
        run(k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    upd = sum(W.steps[k]["updates"] for k in range(args.warmup, W.nstep))
    img_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev])) / W.S
    out = {"value": upd / elapsed, "ms_per_step": elapsed / args.steps * 1e3, "launch": list(launch[:3]), "nwg": nwg,
           "form": "exact", "imaging_kernel": kernel_name("mixed", nwg, "exact"), "imaging_ms_per_shot": img_ms,
           "note": "fp32 bank and tables, exact form, batched imaging with its aperture cull"}
    log(f"[fp32 batched] {out['value']:.3e} updates/s, {out['ms_per_step']:.1f} ms/step, imaging {img_ms:.1f} ms "
        f"per shot, launch {launch[:3]} nwg {nwg}")
    del bank
    return out


def train_surrogate(W, args, torch, engine):
    cfg = W.cfg
    sur = engine.Surrogate.for_survey(cfg, out_scale=(0.02, 0.004, 0.05))
    t0 = time.time()
    X, T = sur.samples(args.train_samples, seed=0x5EED5)
    hist = sur.train(X, T, steps=args.train_steps, batch=args.train_batch, lr0=4e-3, lr1=4e-5)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    del X, T
    Xv, Tv = sur.samples(1 << 18, seed=0x5EED6, offset=1 << 40)
    d = (sur.eval(Xv) - Tv).abs()
    mask, err = sur.mask(W.sv.tg, W.tab, tol=0.5e-3)
    torch.cuda.synchronize()
    info = {"train_seconds": t_train, "samples": args.train_samples, "steps": args.train_steps,
            "batch": args.train_batch, "loss_first": float(hist[:10].mean()), "loss_last": float(hist[-20:].mean()),
            "heldout_dt_rms_ms": float((d[:, 0] ** 2).mean().sqrt()) * 1e3,
            "heldout_dt_max_ms": float(d[:, 0].max()) * 1e3,
            "heldout_dtstar_max_ms": float(d[:, 1].max()) * 1e3,
            "fallback_brick_fraction": 1.0 - float(mask.float().mean())}
    del Xv, Tv, d, err
    log(f"surrogate: trained {args.train_steps} steps in {t_train:.1f}s, loss {info['loss_first']:.3e} -> "
        f"{info['loss_last']:.3e}; held-out |dt| rms {info['heldout_dt_rms_ms']:.3f} ms max "
        f"{info['heldout_dt_max_ms']:.2f} ms; {info['fallback_brick_fraction'] * 100:.1f}% bricks fall back")
    return sur, mask, info


# ---- accuracy ------------------------------------------------------------------

def src_layout(sur, sv, n, src, torch):
    nus, nvs, nw = sur.src_grid(sv)
    hdr = src[:16 * n].view(torch.float32).view(n, 4)
    body = src[16 * n:].view(torch.float32).view(n, nvs, nus, nw, 2)
    return hdr, body


def accuracy_vs_fp64(W, args, forms, shapes, sur, mask, torch, engine):
    """Each form's image of the first `oracle_traces` traces of shot 0 over
    the whole volume against the FP64 oracle (FP64 Q bank of the same
    traces, FP64 sum) fed the tables that form images with: the FP32 node
    tables for fp32 and mixed (whose own tables are the U16Q quantisation of
    them), the fast form's own source table for fast."""
    from oracle import oracle as O
    from tests.helpers import o_bparams, o_iparams
    n = min(args.oracle_traces, W.ranges[0][1])
    tr0 = W.ranges[0][0]
    geom = np.ascontiguousarray(W.geom[tr0:tr0 + n])
    g_d = torch.from_numpy(geom).cuda()
    x = W.traces[0][:n].contiguous()
    shots = engine.ShotList([(0, n) + tuple(W.tans)])
    tab_h = W.tab.cpu().numpy()
    bank64 = O.qbank(x.cpu().numpy(), o_bparams(W.sv))
    t0 = time.time()
    ref_nodes, _ = O.image(bank64, geom, tab_h, o_iparams(W.sv))
    log(f"accuracy: FP64 oracle image of {n} traces in {time.time() - t0:.1f}s")
    out, imgs = {}, {}
    if "fp32" in forms:
        img = W.sv.new_image()
        W.sv.image_shot(img, g_d, 0, n, W.sv.qbank(x), W.tab, launch=shapes["fp32"][0])
        imgs["fp32"] = img.cpu().numpy()
        out["fp32_vs_fp64"] = rel_l2(imgs["fp32"], ref_nodes)
    sv = W.svm
    b16 = sv.qbank(x, dtype=torch.float16, pair=sv.pair_bank())
    if "mixed" in forms:
        img = sv.new_image()
        sv.image_batch(img, g_d, b16, W.tab_u16, shots, launch=shapes["mixed"][0], nwg=shapes["mixed"][1])
        imgs["mixed"] = img.cpu().numpy()
        out["mixed_vs_fp64"] = rel_l2(imgs["mixed"], ref_nodes)
    if "fast" in forms:
        src = sur.src_tables(sv, g_d, shots, tab=W.tab, mask=mask)
        hdr, body = src_layout(sur, sv, 1, src, torch)
        ref_src, _ = O.image(bank64, geom, tab_h, o_iparams(W.sv),
                             src=(hdr[0].cpu().numpy(), body[0].cpu().numpy()))
        img = sv.new_image()
        sv.image_batch(img, g_d, b16, W.tab_u16, shots, launch=shapes["fast"][0], nwg=shapes["fast"][1], src=src)
        out["fast_vs_fp64"] = rel_l2(img.cpu().numpy(), ref_src)
    if "mixed" in forms and sv.form in ("nodes", "spans"):
        # the form of the sum alone: FP32 bank (pairs for the span form) and
        # FP32 tables
        img = sv.new_image()
        sv.image_batch(img, g_d, sv.qbank(x, pair=sv.pair_bank()), W.tab, shots, launch=shapes["mixed"][0], nwg=-1)
        out[f"{sv.form}_form_fp32_storage_vs_fp64"] = rel_l2(img.cpu().numpy(), ref_nodes)
    if "mixed" in forms:
        # config 3 as specified (BF16 bank, FP16 codes in 8^3 bricks under an
        # FP32 offset and scale), exact form, per shot: why the mixed form
        # stores the bank in FP16 and the tables as 16-bit fixed point
        f16b = W.sv.quantize_tables(W.tab, code="f16")
        img = W.sv.new_image()
        W.sv.image_shot(img, g_d, 0, n, W.sv.qbank(x, dtype=torch.bfloat16), f16b,
                        launch=shapes.get("fp32", ((8, 8, 8), -1))[0])
        out["config3_bf16_bank_f16b_tables_vs_fp64"] = rel_l2(img.cpu().numpy(), ref_nodes)
        del f16b
    out["subset"] = f"shot 0, first {n} traces, full 401x401x301 volume"
    log(f"accuracy vs FP64 oracle: {out}")
    return out


def source_leg_model_error(W, sur, mask, forms, shapes, torch, engine):
    """Model error, apart from precision: per-shot source tables of the first
    step's shots -- surrogate (with fallback), the table path (four node
    tables blended at the source), the nearest node -- against FP64 rays at
    the exact source positions, and the fast and fp32 forms' images against
    the image built with the exact source legs (all three on the same
    receiver tables)."""
    st = W.steps[0]
    sv, cfg = W.sv, W.cfg
    n = st["shots"].n
    src_s = sur.src_tables(sv, st["geom"], st["shots"], tab=W.tab, mask=mask)
    src_b = sv.src_tables_nodes(st["geom"], st["shots"], W.tab)
    nus, nvs, nw = sur.src_grid(sv)
    hdr, _ = src_layout(sur, sv, n, src_s, torch)
    hdr = hdr.cpu().numpy()
    exact = src_s.clone()
    body = exact[16 * n:].view(torch.float32)
    per = nvs * nus * nw * 2
    for k in range(n):
        tg = dict(cfg.tgeom, nnx=1, nny=1, n0x=float(hdr[k, 0]), n0y=float(hdr[k, 1]), nu=nus, nv=nvs,
                  u0=float(hdr[k, 2]), v0=float(hdr[k, 3]))
        engine.build_tables_into(cfg.model, tg, body[k * per:(k + 1) * per])
    _, ts = src_layout(sur, sv, n, src_s, torch)
    _, tb = src_layout(sur, sv, n, src_b, torch)
    _, tx = src_layout(sur, sv, n, exact, torch)
    g = cfg.tgeom
    t_all = W.tab
    near = torch.empty_like(tx[..., 0])
    for k in range(n):
        iu0 = int(round((hdr[k, 2] - g["u0"]) / g["du"]))
        iv0 = int(round((hdr[k, 3] - g["v0"]) / g["dv"]))
        i = min(max(int(math.floor((hdr[k, 0] - g["n0x"]) / g["ndx"] + 0.5)), 0), g["nnx"] - 1)
        j = min(max(int(math.floor((hdr[k, 1] - g["n0y"]) / g["ndy"] + 0.5)), 0), g["nny"] - 1)
        near[k] = t_all[j * g["nnx"] + i, iv0:iv0 + nvs, iu0:iu0 + nus, :, 0]
    e = {"src_dt_" + name + suf: float(fn(d)) * 1e3
         for name, d in (("surrogate", (ts - tx)[..., 0].abs()), ("node_blend", (tb - tx)[..., 0].abs()),
                         ("nearest_node", (near - tx[..., 0]).abs()))
         for suf, fn in (("_mean_ms", torch.mean), ("_max_ms", torch.max))}
    # images of the first step (FP32 storage) with each source leg
    bank = sv.qbank(st["traces"])
    imgs = {}
    for name, src in (("exact", exact), ("surrogate", src_s), ("node_blend", src_b), ("nearest_node", None)):
        img = sv.new_image()
        sv.image_batch(img, st["geom"], bank, W.tab, st["shots"], src=src)
        imgs[name] = img.cpu().numpy()
    for name in ("surrogate", "node_blend", "nearest_node"):
        e[f"image_{name}_vs_exact_source"] = rel_l2(imgs[name], imgs["exact"])
    e["note"] = ("model error of the source leg against FP64 rays at the true source positions (first step's "
                 "shots, FP32 bank and tables; the receiver legs are the node tables' in every image)")
    log(f"source-leg model error: {e}")
    return e


# ---- counters and rooflines -------------------------------------------------------

def unique_bytes_per_update(W, form_sv, torch, nsample=24, zp=8):
    """HBM bytes a tile must read at least once per trace, per update: for a
    sample of 8x8-column x zp-depth tiles and every trace of shot 0 that
    updates them, the distinct FP16 bank samples read (2 B each; a pair
    bank's sample-and-difference entries, 4 B each), the
    distinct U16Q corner words of the receiver table (16 B each per depth
    node, plus a 16 B brick header per cell brick), divided by the updates
    (points inside the dip aperture); plus the image read and write once per
    tile per dispatch.  Table reads use the library's lookup operator."""
    cfg = W.cfg
    sv = W.sv
    g = cfg.tgeom
    rng = np.random.default_rng(7)
    tr0, n = W.ranges[0]
    geom = W.geom[tr0:tr0 + n].astype(np.float64)
    tx1, tx2, ty1, ty2, tz = cfg.dip_bounds()
    lo = tx1 - tz * (1 + tx1 * tx1)
    hi = tx2 + tz * (1 + tx2 * tx2)
    tot_bytes = tot_upd = 0
    tiles = 0
    for _ in range(nsample):
        cx0 = int(rng.integers(0, cfg.nx // 8)) * 8
        cy0 = int(rng.integers(0, cfg.ny // 8)) * 8
        cz0 = int(rng.integers(1, (cfg.nz - zp) // zp)) * zp
        xs = (cx0 + np.arange(8)) * cfg.dx
        ys = (cy0 + np.arange(8)) * cfg.dy
        zs = (cz0 + np.arange(zp)) * cfg.dz
        X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
        pts = torch.from_numpy(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)).cuda()
        # source leg (one source per shot)
        sx, sy = geom[0, 0], geom[0, 1]
        tsrc = sv.lookup(W.tab, float(sx), float(sy), pts).cpu().numpy().astype(np.float64)
        tiles += 1
        tile_bytes = 8 * 64 * zp
        for k in range(n):
            gx, gy = geom[k, 2], geom[k, 3]
            dsx, dsy, dgx, dgy = sx - X, sy - Y, gx - X, gy - Y
            ds = np.sqrt(dsx ** 2 + dsy ** 2 + Z ** 2)
            dg = np.sqrt(dgx ** 2 + dgy ** 2 + Z ** 2)
            Xb, Yb, Zb = dsx / ds + dgx / dg, dsy / ds + dgy / dg, Z / ds + Z / dg
            inside = ((Xb > lo * Zb) & (Xb < hi * Zb) & (Yb > lo * Zb) & (Yb < hi * Zb)).ravel()
            m = int(inside.sum())
            if m == 0:
                continue
            trec = sv.lookup(W.tab, float(gx), float(gy), pts).cpu().numpy().astype(np.float64)
            t = tsrc[:, 0] + trec[:, 0]
            lev = np.clip(np.floor((tsrc[:, 1] + trec[:, 1]) / sv.dtstar + 0.5), 0, cfg.K - 1)
            i0 = np.floor(t / cfg.dt)
            ok = inside & (i0 > 0) & (i0 + 1 < cfg.nt)
            if form_sv.pair_bank():
                samp = np.unique(lev[ok] * cfg.nt + i0[ok])
                sbytes = 4
            else:
                samp = np.unique(np.concatenate([lev[ok] * cfg.nt + i0[ok], lev[ok] * cfg.nt + i0[ok] + 1]))
                sbytes = 2
            # receiver table: corner cells under the 64 columns, depth nodes spanned
            iu = np.floor((xs - gx - g["u0"]) / g["du"]).astype(int)
            iv = np.floor((ys - gy - g["v0"]) / g["dv"]).astype(int)
            cells = len(np.unique(iu)) * len(np.unique(iv))
            w0, w1 = int(zs[0] // g["dw"]), int(zs[-1] // g["dw"]) + 1
            nodes = w1 - w0 + 1
            bricks = len(np.unique(iu // 8)) * len(np.unique(iv // 8)) * len(np.unique(np.arange(w0, w1 + 1) // 8))
            tile_bytes += sbytes * samp.size + 16 * cells * nodes + 16 * bricks
            tot_upd += m
        tot_bytes += tile_bytes
    return tot_bytes / max(tot_upd, 1)


def probe(args):
    """Child of a rocprofv3 counter run: one step of the fp32 form and of the
    form named by --probe (an untrained surrogate for fast: the imaging
    kernel's work does not depend on the weights), with the launch shapes
    and form pinned, and a calibration copy of CALIB_MIB MiB."""
    import torch

    from q_psdm_amd import engine, synth
    args.warmup, args.steps = 0, 1
    args.mixed_form = args.probe_form
    W = Workload(args, torch, engine, synth)
    a = torch.ones(CALIB_MIB << 18, dtype=torch.float32, device="cuda")
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    engine.copy_f32(a, c)
    torch.cuda.synchronize()
    del a, c
    sur = engine.Surrogate.for_survey(W.cfg, out_scale=(0.02, 0.004, 0.05))
    mask = torch.zeros(engine.L.lib().qpsdm_surrogate_mask_bricks(engine.C.byref(W.sv.tg)), dtype=torch.uint8,
                       device="cuda")
    z = (args.ztile, args.tile_x, args.tile_y, args.nwg)
    args.ztile, args.tile_x, args.tile_y, args.nwg = args.fp32_shape
    time_form("fp32", W, args, torch, engine)
    args.ztile, args.tile_x, args.tile_y, args.nwg = z
    time_form(args.probe, W, args, torch, engine, sur=sur, mask=mask)


def counter_run(counters, args, tmp, tag):
    rp = shutil.which("rocprofv3")
    if not rp:
        return None, "rocprofv3 not found"
    d = os.path.join(tmp, tag)
    cmd = [rp, "--pmc"] + counters + ["--output-format", "csv", "-d", d, "-o", "probe", "--", sys.executable,
                                      os.path.join(ROOT, "bench.py"), "--probe", args.head,
                                      "--probe-form", args.head_form, "--z-exact", str(args.z_exact),
                                      "--spans", args.spans,
                                      "--shots-per-step", str(args.shots_per_step),
                                      "--tile-x", str(args.head_shape[0][0]), "--tile-y", str(args.head_shape[0][1]),
                                      "--ztile", str(args.head_shape[0][2]), "--nwg", str(args.head_shape[1]),
                                      "--fp32-shape"] + [str(v) for v in args.fp32_shape]
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=sys.stderr, stderr=sys.stderr,
                           timeout=args.probe_timeout)
    except subprocess.TimeoutExpired:
        return None, f"{tag} run timed out"
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        return None, f"{tag} run rc={r.returncode}, no counter file"
    rows = list(csv.DictReader(open(files[0])))
    out = {}
    for x in rows:
        k = (x["Kernel_Name"], x["Dispatch_Id"])
        out.setdefault(k, {})[x["Counter_Name"]] = out.get(k, {}).get(x["Counter_Name"], 0.0) + float(
            x["Counter_Value"])
        # the dispatch's own duration in the counter run (dispatches run one at a time there)
        out[k]["_dispatch_ns"] = float(int(x["End_Timestamp"]) - int(x["Start_Timestamp"]))
    return out, None


def per_shot(rows, names, nshot):
    """Counter totals of every dispatch of the kernels whose names contain
    one of `names`, per shot of the probe step (the node form dispatches the
    exact body's shallow part beside the node kernel)."""
    sel = [v for (k, _), v in rows.items() if any(n in k for n in names) and "count" not in k]
    if not sel:
        return None
    keys = set().union(*sel)
    return {c: float(np.sum([v.get(c, 0.0) for v in sel])) / nshot for c in keys}, len(sel)


def kernel_family(form):
    return ("image_kernel<", "image_nodes_kernel<") if form == "fp32" else ("image_batch",)


def measure_counters(args):
    """Three counter passes of a probe step: SQ (VALU issue) with the
    texture address / data units' busy cycles, FETCH_SIZE, WRITE_SIZE."""
    tmp = tempfile.mkdtemp(prefix="qpsdm_pmc_", dir="/tmp")
    try:
        res = {}
        for tag, ctrs in (("sq", ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_BUSY_CYCLES",
                                  "SQ_INSTS_VALU_TRANS_F32", "TA_TA_BUSY_sum", "TD_TD_BUSY_sum"]),
                          ("fetch", ["FETCH_SIZE"]), ("write", ["WRITE_SIZE"])):
            rows, err = counter_run(ctrs, args, tmp, tag)
            if err:
                log(f"counters: {err}")
                return None, err
            res[tag] = rows
            log(f"counters: {tag} pass done ({len(rows)} dispatches)")
        return res, None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def rooflines(res, results, head, upd_per_shot, ub_per_update, nshot):
    """VALU-issue roofline per imaging kernel (the binding resource) and the
    HBM figures of the headline kernel, per shot."""
    out = {}
    for form in ("fp32", head):
        R = results[form]
        kname = R["imaging_kernel"]
        got = per_shot(res["sq"], kernel_family(form), nshot)
        if not got:
            continue
        c, nd = got
        t = R["imaging_ms_per_shot"] * 1e-3
        valu = c["SQ_INSTS_VALU"]
        lane_per_upd = valu * 64 / upd_per_shot[form]
        out[form] = {"kernel": kname, "dispatches": nd, "valu_wave_instr": valu,
                     "valu_issue_frac": valu / (NCU * CLOCK_GHZ * 1e9 * t),
                     "lane_instr_per_update": lane_per_upd,
                     "transcendental_wave_instr": c.get("SQ_INSTS_VALU_TRANS_F32"),
                     "salu_wave_instr": c.get("SQ_INSTS_SALU"), "vmem_rd_wave_instr": c.get("SQ_INSTS_VMEM_RD"),
                     "clock_GHz_from_busy_cycles": c["SQ_BUSY_CYCLES"] / 32 / (c["_dispatch_ns"] * 1e-9) / 1e9
                     if c.get("_dispatch_ns") else None}
        # busy cycles of the per-CU texture addresser (TA) and texture data
        # unit (TD, which returns every vector load's data to the lanes) per
        # CU cycle of the counter run's own dispatch time (there the
        # dispatches run one at a time; the timed run may overlap them)
        tc = c.get("_dispatch_ns", 0.0) * 1e-9
        out[form]["counter_run_ms_per_shot"] = tc * 1e3
        for unit, ctr in (("ta", "TA_TA_BUSY_sum"), ("td", "TD_TD_BUSY_sum")):
            if c.get(ctr) is not None and tc > 0:
                out[form][f"{unit}_busy_frac"] = c[ctr] / (NCU * CLOCK_GHZ * 1e9 * tc)
    # HBM traffic of the headline kernel (calibrated against the copy kernel)
    traffic = None
    kname = results[head]["imaging_kernel"]
    det = {}
    for tag, ctr in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        got = per_shot(res[tag], kernel_family(head), nshot)
        copy = [v[ctr] for (k, _), v in res[tag].items() if k.startswith("(anonymous namespace)::copy_kernel")]
        if not got or not copy:
            det = None
            break
        factor = copy[0] / (CALIB_MIB * 1024.0)
        det[ctr] = {"kernel_KiB": got[0][ctr], "copy_KiB": copy[0], "factor": factor,
                    "bytes": got[0][ctr] * 1024.0 / factor}
    if det:
        traffic = det["FETCH_SIZE"]["bytes"] + det["WRITE_SIZE"]["bytes"]
    t = results[head]["imaging_ms_per_shot"] * 1e-3
    hbm = {"unique_bytes_per_update": ub_per_update,
           "achieved_GBs": ub_per_update * upd_per_shot[head] / t / 1e9 if ub_per_update else None,
           "peak": HBM_PEAK_GBS, "traffic_detail": det,
           "measured_GBs": traffic / t / 1e9 if traffic else None}
    if hbm["achieved_GBs"]:
        hbm["frac"] = hbm["achieved_GBs"] / HBM_PEAK_GBS
    if hbm["measured_GBs"]:
        hbm["measured_frac"] = hbm["measured_GBs"] / HBM_PEAK_GBS
    return out, traffic, hbm


# ---- other workloads -----------------------------------------------------------

def time_config1(args, torch, engine, synth):
    """Config 1 in full on the GPU: 101^3 image, 16 shots x 1024 receivers,
    K = 16, diffraction data from 20 scatterers; FP32 form per shot and the
    mixed form (FP16 bank, U16Q tables) batched."""
    cfg = synth.config1()
    probe = engine.Survey(cfg, 0.01)
    tab = probe.build_tables()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab))
    geom, ranges = synth.geometry(synth.shots_config1())
    traces = synth.diffraction_traces(cfg, geom, synth.scatterers(cfg, 20))
    g = torch.from_numpy(geom).cuda()
    x = torch.from_numpy(traces).cuda()
    upd = sum(sv.count_shot(g, tr0, n) for tr0, n in ranges)
    q = sv.quantize_tables(tab, code="u16q")
    shots = engine.ShotList([(tr0, n) + tuple(cfg.dip_bounds()[:4]) for tr0, n in ranges])
    out = {"workload": "config 1: 101x101x101 image @10 m, 16 shots x 1024 receivers, nt=1001 @2 ms, K=16, "
                       "20 m tables, v(z) = 1500 + 0.6 z, Q = 100", "updates": upd}
    for form in ("fp32", "mixed"):
        img = sv.new_image()
        if form == "fp32":
            bank = torch.empty((x.shape[0], cfg.K, cfg.nt), device="cuda")

            def run():
                sv.qbank(x, out=bank)
                for tr0, n in ranges:
                    sv.image_shot(img, g, tr0, n, bank[tr0:tr0 + n], tab, launch=(8, 8, 8))
        else:
            bank = torch.empty(sv.bank_shape(x.shape[0], torch.float16), dtype=torch.float16, device="cuda")

            def run():
                sv.qbank(x, out=bank)
                sv.image_batch(img, g, bank, q, shots, launch=(8, 8, 8), nwg=-1)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 3
        out[form] = {"seconds": dt, "updates_per_s": upd / dt}
    log(f"[config 1] {out}")
    return out, (cfg, sv, tab, geom, traces)


def cpu_config1(c1, args):
    """The oracle's FP32 path on config 1 (Q bank + sum), the first
    --cpu-config1-shots shots on all host threads."""
    from oracle import oracle as O
    from tests.helpers import o_bparams, o_iparams
    cfg, sv, tab, geom, traces = c1
    ns = args.cpu_config1_shots
    n = 1024 * ns
    t0 = time.perf_counter()
    b = O.qbank(traces[:n], o_bparams(sv), dtype=np.float32)
    _, upd = O.image(b, geom[:n], tab.cpu().numpy(), o_iparams(sv), fp32=True)
    dt = time.perf_counter() - t0
    log(f"[config 1 cpu] {ns} shots: {dt:.1f}s, {upd / dt:.3e} updates/s")
    return {"shots": ns, "seconds": dt, "updates": int(upd), "updates_per_s": upd / dt,
            "threads": O.num_threads(), "full": ns == 16}


def time_config4(args, torch, engine, synth, W):
    """The batching variant: --config4-shots uneven shot gathers (trace
    counts floor(exp(U[ln 200, ln 4000])), per-shot aperture U[25, 60] deg)
    of config 4 on the headline model and tables, imaged through captured
    hipGraphs of [Q bank -> batched imaging] in the mixed form, in batches of
    about 8000 traces; and the same shots in the FP32 form, one launch per
    shot."""
    cfg = W.cfg
    shots = synth.shots_config4(nshots=args.config4_shots)
    geom, ranges = synth.geometry(shots)
    g = torch.from_numpy(geom).cuda()
    x = synth.reflectivity_traces_gpu(cfg, geom.shape[0], seed=0x5EED0004)
    aps = [s[3] for s in shots]
    sv, svm = W.sv, W.svm
    desc = [(tr0, n) + synth.dip_tans(a) for (tr0, n), a in zip(ranges, aps)]
    upd = svm.count_batch(g, engine.ShotList(desc))
    # batches of consecutive shots, ~8000 traces each
    batches, cur, tot = [], [], 0
    for i, (tr0, n) in enumerate(ranges):
        cur.append(i)
        tot += n
        if tot >= 8000 or i == len(ranges) - 1:
            batches.append(cur)
            cur, tot = [], 0
    img = svm.new_image()
    graphs = []
    launch, nwg = args.mixed_shape
    for bidx in batches:
        r0 = ranges[bidx[0]][0]
        r1 = ranges[bidx[-1]][0] + ranges[bidx[-1]][1]
        xt = x[r0:r1]
        gg = g[r0:r1]
        sl = engine.ShotList([(ranges[i][0] - r0, ranges[i][1]) + desc[i][2:] for i in bidx])
        bank = torch.empty(svm.bank_shape(r1 - r0, torch.float16, pair=svm.pair_bank()), dtype=torch.float16,
                           device="cuda")
        graphs.append(engine.BatchGraph(svm, xt, bank, img, gg, W.tab_u16, sl, launch=launch, nwg=nwg, stages=5))
    for gr in graphs:
        gr.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for gr in graphs:
        gr.launch()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    for gr in graphs:
        gr.close()
    out = {"workload": f"config 4: {len(shots)} shots, {geom.shape[0]} traces (log-uniform 200-4000 per shot), "
                       f"per-shot aperture 25-60 deg, headline model and tables",
           "shots": len(shots), "traces": int(geom.shape[0]), "updates": upd, "graphs": len(graphs),
           "mixed": {"seconds": dt, "updates_per_s": upd / dt, "form": svm.form, "launch": list(launch[:3]),
                     "nwg": nwg}}
    # FP32 form, per-shot launches
    bank = torch.empty((max(n for _, n in ranges), cfg.K, cfg.nt), device="cuda")
    img32 = sv.new_image()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for (tr0, n), d in zip(ranges, desc):
        sv.qbank(x[tr0:tr0 + n], out=bank[:n])
        sv.image_batch(img32, g[tr0:tr0 + n], bank, W.tab, engine.ShotList([(0, n) + d[2:]]),
                       launch=args.fp32_launch, nwg=-1)
    torch.cuda.synchronize()
    dt32 = time.perf_counter() - t0
    out["fp32"] = {"seconds": dt32, "updates_per_s": upd / dt32}
    out["mixed_speedup_vs_fp32"] = dt32 / dt
    log(f"[config 4] {out}")
    return out


def cpu_baseline(W, ntr):
    """The oracle's FP32 path (FP32 Q bank + FP32 sum, OpenMP over image
    lines) on the first ntr traces of shot 0, on this host."""
    from oracle import oracle as O
    from tests.helpers import o_bparams, o_iparams

    sv = W.sv
    tr0, _ = W.ranges[0]
    g = np.ascontiguousarray(W.geom[tr0:tr0 + ntr])
    x = W.traces[0][:ntr].cpu().numpy()
    t_host = W.tab.cpu().numpy()
    cores = O.num_threads()
    t0 = time.perf_counter()
    b = O.qbank(x, o_bparams(sv), dtype=np.float32)
    t1 = time.perf_counter()
    _, n = O.image(b, g, t_host, o_iparams(sv), fp32=True)
    t2 = time.perf_counter()
    log(f"cpu baseline: bank {t1 - t0:.1f}s image {t2 - t1:.1f}s updates {n} ({cores} threads)")
    n1tr = max(1, ntr // 16)
    O.set_num_threads(1)
    try:
        s0 = time.perf_counter()
        b1 = O.qbank(x[:n1tr], o_bparams(sv), dtype=np.float32)
        _, n1 = O.image(b1, np.ascontiguousarray(g[:n1tr]), t_host, o_iparams(sv), fp32=True)
        s1 = time.perf_counter()
    finally:
        O.set_num_threads(cores)
    log(f"cpu baseline, 1 thread: {s1 - s0:.1f}s updates {n1}")
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n / (t2 - t0), "unit": "updates/s", "cores": cores, "kind": "port",
            "sample": f"shot 0, first {ntr} of its traces, full 401x401x301 image (FP32 Q bank + FP32 "
                      f"Kirchhoff sum, OpenMP over image lines)", "seconds": t2 - t0,
            "value_1thread": n1 / (s1 - s0), "sample_1thread": f"shot 0, first {n1tr} traces, one thread",
            "seconds_1thread": s1 - s0, "cpu_model": model, "nproc": os.cpu_count()}


def time_end_to_end(W, torch, engine, nshot=2, launch=None):
    """End-to-end rate through the pipeline context: big-endian SU records of
    nshot shots in host memory -> pinned staging slots -> H2D copy (its own
    stream) -> device decode -> FP16 Q bank -> imaging with U16Q tables ->
    image copied back to host memory.  The mixed form's survey: in the span
    form the context stores a pair bank and images each slot as one batched
    shot (the node tables' source legs: no source tables here)."""
    from q_psdm_amd import lib as L
    sv, cfg = W.svm, W.cfg
    spans = sv.form == "spans"
    idx = list(range(min(nshot, len(W.ranges))))
    rows = np.concatenate([W.geom[W.ranges[i][0]:W.ranges[i][0] + W.ranges[i][1]] for i in idx])
    traces = torch.cat([W.traces[i] for i in idx], 0).cpu().numpy()
    wxy = np.round(rows.astype(np.float64)).astype(np.int32)
    su = engine.su_encode(traces, wxy)
    gp = L.GeomParams(1.0, -1e30, 1e30, -1e30, 1e30, 0.0, 0.0, 0.0, 0.0, cfg.dx, cfg.dy)
    pl = engine.Pipeline(sv, gp, bank_dtype="f16p" if spans else "f16", table_dtype="u16q",
                         batch_traces=max(W.ranges[i][1] for i in idx), launch=launch)
    pl.set_tables(W.tab)
    pl.push_su(su[:240 + 4 * cfg.nt])
    pl.image()
    pl.reset_image()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pl.push_su(su)
    img = pl.image()
    wall = time.perf_counter() - t0
    pl.close()
    upd = sum(W.counts[i] for i in idx)
    log(f"[end-to-end] {len(idx)} shots of SU bytes to a host image: {wall * 1e3:.1f} ms, {upd / wall:.3e} updates/s")
    del img
    return {"shots": len(idx), "traces": int(rows.shape[0]), "su_bytes": len(su), "wall_ms": wall * 1e3,
            "updates_per_s": upd / wall, "image_bytes_to_host": cfg.nx * cfg.ny * cfg.nz * 4,
            "path": "SU bytes in host memory -> pinned slot -> H2D (copy stream) -> device decode -> f16 "
                    + ("pair Q bank -> batched span-form imaging" if spans else "Q bank -> per-shot imaging")
                    + f" (u16q tables, {sv.form} form) -> image to host"}


def time_segy_decode(torch, ntr=20000, ns=2001, reps=5):
    """Device decode of SEG-Y IBM records (random words, valid ns keys)."""
    from q_psdm_amd import segy
    rec = 240 + 4 * ns
    raw = torch.randint(0, 256, (ntr, rec), dtype=torch.uint8, device="cuda")
    raw[:, 114] = ns >> 8
    raw[:, 115] = ns & 0xff
    raw = raw.reshape(-1)
    segy.decode_dev(raw, ntr, ns, segy.IBM)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        segy.decode_dev(raw, ntr, ns, segy.IBM)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    per = rec + 4 * ns + 40
    gbs = ntr * per / (ms * 1e-3) / 1e9
    log(f"[segy] IBM decode of {ntr} traces x {ns} samples: {ms:.3f} ms, {gbs:.0f} GB/s")
    return {"kernel": "segy_decode_kernel<1> (IBM)", "traces_per_launch": ntr, "ns": ns, "avg_launch_ms": ms,
            "algorithmic_bytes_per_trace": per, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots-per-step", type=int, default=2)
    ap.add_argument("--forms", default="fp32,mixed,fast")
    ap.add_argument("--mixed-form", choices=["auto", "exact", "nodes", "spans"], default="spans",
                    help="form of the sum for the mixed and fast forms (auto: the autotuner's fastest)")
    ap.add_argument("--z-exact", type=float, default=400.0, help="node and span forms: depth imaged exactly (m)")
    ap.add_argument("--spans", default="0:2,800:4,1600:8",
                    help="span form: z:S,... -- from depth z (m) down, weights at the ends of spans of S table "
                         "intervals")
    ap.add_argument("--ztile", type=int, default=0, help="pin depth points (node form: intervals) per tile")
    ap.add_argument("--tile-x", type=int, default=8)
    ap.add_argument("--tile-y", type=int, default=8)
    ap.add_argument("--nwg", type=int, default=0,
                    help="batched grid with a pinned shape (0 = resident persistent, -1 = one group per tile)")
    ap.add_argument("--train-samples", type=int, default=20_000_000)
    ap.add_argument("--train-steps", type=int, default=15000)
    ap.add_argument("--train-batch", type=int, default=65536)
    ap.add_argument("--oracle-traces", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-traces", type=int, default=256)
    ap.add_argument("--cpu-config1-shots", type=int, default=16)
    ap.add_argument("--config4-shots", type=int, default=50)
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip config 1, config 4, SEG-Y and end-to-end")
    ap.add_argument("--probe", choices=["fp32", "mixed", "fast"], default=None, help=argparse.SUPPRESS)
    ap.add_argument("--probe-form", default="exact", help=argparse.SUPPRESS)
    ap.add_argument("--fp32-shape", type=int, nargs=4, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--probe-timeout", type=int, default=240)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only (--gpus 1): sharding the imaging across GPUs "
                          "(by image slab or by shot) is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.probe:
        probe(args)
        return

    import torch

    from q_psdm_amd import engine, synth
    forms = [f for f in args.forms.split(",") if f]
    W = Workload(args, torch, engine, synth)
    sur = mask = sinfo = None
    if "fast" in forms:
        sur, mask, sinfo = train_surrogate(W, args, torch, engine)
    results, shapes = {}, {}
    for f in forms:
        _, results[f] = time_form(f, W, args, torch, engine, sur=sur, mask=mask)
        shapes[f] = (tuple(results[f]["launch"]), results[f]["nwg"])
    if "fp32" in forms and not args.no_extras:
        results["fp32_batched"] = time_fp32_batched(W, args, torch, engine)
    torch.cuda.synchronize()
    acc = accuracy_vs_fp64(W, args, forms, shapes, sur, mask, torch, engine)
    model_err = source_leg_model_error(W, sur, mask, forms, shapes, torch, engine) if sur is not None else None
    head = "fast" if "fast" in results else forms[-1]
    head_ok = acc.get(f"{head}_vs_fp64", 1.0) <= 1e-3
    value_form = head if head_ok or "fp32" not in results else "fp32"
    R = results[value_form]
    # counters: VALU issue of the fp32 and headline imaging kernels; HBM traffic of the headline one
    roof, traffic, hbm = {}, None, None
    upd_per_shot = {f: results[f]["updates_per_shot"] for f in results if "updates_per_shot" in results[f]}
    ub = unique_bytes_per_update(W, W.svm, torch)
    log(f"unique bytes per update: {ub:.3f}")
    if not args.no_traffic and "fp32" in results:
        args.head = head
        args.head_form = W.svm.form
        args.head_shape = shapes[head]
        args.fp32_shape = [shapes["fp32"][0][2], shapes["fp32"][0][0], shapes["fp32"][0][1], 0]
        torch.cuda.empty_cache()
        res, err = measure_counters(args)
        if res:
            roof, traffic, hbm = rooflines(res, results, head, upd_per_shot, ub, args.shots_per_step)
        else:
            hbm = {"note": err}
    extras = {}
    if not args.no_extras:
        c1, c1data = time_config1(args, torch, engine, synth)
        extras["config1"] = c1
        args.mixed_shape = shapes.get("mixed", ((8, 8, 8), -1))
        args.fp32_launch = shapes["fp32"][0] if "fp32" in shapes else (8, 8, 8)
        extras["config4"] = time_config4(args, torch, engine, synth, W)
        extras["segy_decode"] = time_segy_decode(torch)
        extras["end_to_end"] = time_end_to_end(W, torch, engine, launch=shapes.get("mixed", ((8, 8, 4), -1))[0])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(W, args.cpu_traces)
        if not args.no_extras:
            cpu["config1"] = cpu_config1(c1data, args)
    hr = roof.get(value_form, {})
    bank_kind = "pair bank (sample, difference)" if W.svm.pair_bank() else "bank"
    out = {
        "metric": "image-point x trace updates/s @1 GPU (FP32 vs mixed+surrogate+batched, <=1e-3 rel-L2 vs FP64)",
        "value": R["value"],
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded layered + anomaly model (layer Q ~ U[30,200]), jittered geometry, band-limited "
                "reflectivity traces with 5% noise",
        "config": {"workload": "401x401x301 image @10 m; 2000 receivers/shot over 4x4 km; nt=2001 @2 ms; K=32; "
                               "25 m tables on 5x5 nodes; layered+anomaly model with randomized Q",
                   "form": value_form, "sum_form": R["form"], "z_exact": R["z_exact"],
                   "spans": R.get("spans"),
                   "storage": {"fast": f"f16 {bank_kind} (per-trace power-of-two scale), u16 offset+step tables "
                                       "as corner quads, f32 per-shot source tables from the bf16-mfma surrogate",
                               "mixed": f"f16 {bank_kind} (per-trace power-of-two scale), u16 offset+step tables "
                                        "as corner quads",
                               "fp32": "f32 bank, f32 tables"}[value_form],
                   "shots_per_step": args.shots_per_step, "tile": R["launch"],
                   "updates_per_step": R["updates_per_shot"] * args.shots_per_step,
                   "dense_updates_per_step": W.cfg.nx * W.cfg.ny * W.cfg.nz * int(W.steps[0]["traces"].shape[0])},
        "value_meets_1e-3_vs_fp64": acc.get(f"{value_form}_vs_fp64", 0.0) <= 1e-3,
        "roofline": {"bound": "valu", "kernel": R["imaging_kernel"],
                     "achieved": (hr.get("valu_wave_instr", 0) * 64 / (R["imaging_ms_per_shot"] * 1e-3) / 1e12)
                     if hr else None,
                     "peak": NCU * 64 * CLOCK_GHZ * 1e9 / 1e12, "unit": "T lane-instr/s",
                     "frac": hr.get("valu_issue_frac"), "traffic": traffic if value_form == head else None,
                     "lane_instr_per_update": hr.get("lane_instr_per_update"),
                     "hbm": hbm if value_form == head else None, "per_form": roof,
                     "imaging_ms_per_shot": R["imaging_ms_per_shot"],
                     "note": "the imaging kernels issue vector-ALU instructions at the fraction `frac` of the "
                             "chip's ceiling (one wave64 instruction per CU per cycle, 256 CUs at 2.4 GHz) over "
                             "the timed imaging (the span form's dispatches overlap there) while the per-CU "
                             "texture data unit, which hands every vector load's data to the lanes, is busy "
                             "beside it (per_form td_busy_frac, per cycle of the counter run, whose dispatches "
                             "run one at a time): the two bind together; `hbm` gives the unique bytes a tile "
                             "must read once per trace (FP16 bank samples, U16Q corner words, brick headers, "
                             "image) per update and the measured HBM traffic, both far from 8 TB/s"},
        "forms": results,
        "speedup_vs_fp32": (results[head]["value"] / results["fp32"]["value"]) if "fp32" in results else None,
        "speedup_vs_fp32_batched": (results[head]["value"] / results["fp32_batched"]["value"])
        if "fp32_batched" in results else None,
        "image_rel_l2_vs_fp64": acc,
        "source_leg_model_error": model_err,
        "surrogate": sinfo,
        "extras": extras,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
