"""Benchmark of the Q-compensated Kirchhoff PSDM imaging path on one MI355X.

Workload: the 401 x 401 x 301 (10 m) layered + anomaly survey, 200 shots x
2000 jittered receivers, 2001 samples at 2 ms, K = 32 t* levels, 25 m
tables.  One step = one shot: the Q filter bank of its 2000 traces plus
the Kirchhoff sum of those traces into the full image, with traces, tables
and image already resident in HBM.  value = image-point x trace updates/s
(updates = pairs inside the dip aperture, counted once per shot outside the
timed region by the count kernel).

The cpu_baseline leg times the CPU oracle's FP32 path (Q bank + imaging
sum, OpenMP) on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"fp32": 72.0, "mixed": 36.0}
BANK_DTYPE = {"fp32": "f32", "mixed": "bf16"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--form", choices=["fp32", "mixed"], default="fp32")
    ap.add_argument("--ztile", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-traces", type=int, default=256)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="device copy of this many MiB before the steps (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described "
                          "in DESIGN.md and not run", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from q_psdm_amd import engine, synth

    cfg = synth.config2()
    nshot = args.warmup + args.steps
    t0 = time.time()
    probe = engine.Survey(cfg, 0.05)
    tab = probe.build_tables()
    torch.cuda.synchronize()
    sv = engine.Survey(cfg, engine.tstar_max_from_tables(tab))
    log(f"tables built: {tuple(tab.shape)} {tab.numel() * 4 / 1e9:.2f} GB in {time.time() - t0:.1f}s; "
        f"dtstar={sv.dtstar * 1e3:.3f} ms nsb={sv.win.nsb} nww={sv.win.nww}")
    if args.form == "mixed":
        tab_use = sv.quantize_tables(tab)
        bank_dtype = torch.bfloat16
    else:
        tab_use = tab
        bank_dtype = torch.float32

    shots = synth.shots_config2(nshots=nshot)
    geom, ranges = synth.geometry(shots)
    g_d = torch.from_numpy(geom).cuda()
    t0 = time.time()
    traces = [torch.from_numpy(synth.reflectivity_traces(cfg, n, seed=(cfg.seed, s))).cuda()
              for s, (_, n) in enumerate(ranges)]
    log(f"traces generated for {nshot} shots in {time.time() - t0:.1f}s")
    counts = [sv.count_shot(g_d, tr0, n) for tr0, n in ranges]
    torch.cuda.synchronize()
    log(f"updates per shot: {counts}")

    if args.calib_copy_mib:
        a = torch.empty(args.calib_copy_mib << 18, dtype=torch.float32, device="cuda")
        a.fill_(1.0)
        c = torch.empty_like(a)
        torch.cuda.synchronize()
        c.copy_(a)
        torch.cuda.synchronize()
        del a, c
    img = sv.new_image()
    bank = torch.empty((max(n for _, n in ranges), cfg.K, cfg.nt), dtype=bank_dtype, device="cuda")
    launch = (8, 8, args.ztile, 1)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(nshot)]

    def step(s):
        tr0, n = ranges[s]
        e = ev[s]
        e[0].record()
        sv.qbank(traces[s], out=bank[:n])
        e[1].record()
        sv.image_shot(img, g_d, tr0, n, bank, tab_use, launch=launch)
        e[2].record()

    for s in range(args.warmup):
        step(s)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for s in range(args.warmup, nshot):
        step(s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    upd = sum(counts[args.warmup:])
    value = upd / elapsed
    bank_ms = [ev[s][0].elapsed_time(ev[s][1]) for s in range(args.warmup, nshot)]
    img_ms = [ev[s][1].elapsed_time(ev[s][2]) for s in range(args.warmup, nshot)]
    avg_img_s = float(np.mean(img_ms)) / 1e3
    upd_per_launch = upd / args.steps
    achieved = upd_per_launch * BYTES_PER_UPDATE[args.form] / avg_img_s / 1e9
    bank_bytes = [ranges[s][1] * (4 * cfg.nt + cfg.K * cfg.nt * (4 if args.form == "fp32" else 2))
                  for s in range(args.warmup, nshot)]
    bank_gbs = float(np.mean([b / (m / 1e3) / 1e9 for b, m in zip(bank_bytes, bank_ms)]))
    log(f"bank ms {['%.2f' % x for x in bank_ms]}  image ms {['%.1f' % x for x in img_ms]}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, sv, tab, geom, ranges, traces, args.cpu_traces)

    out = {
        "metric": "image-point x trace updates/s @1 GPU",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.form == "fp32" else "f32-accum/bf16-bank/f16-table",
        "data": "synthetic (seeded layered+anomaly model, jittered geometry, reflectivity traces)",
        "config": {"workload": "config2: image 401x401x301 @10 m, 200 shots x 2000 receivers, "
                               "nt=2001 @2 ms, K=32, tables 25 m; step = one shot",
                   "form": args.form, "tile": list(launch), "updates_per_step": upd_per_launch},
        "roofline": {"bound": "hbm", "kernel": "image_kernel", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None,
                     "algorithmic_bytes_per_update": BYTES_PER_UPDATE[args.form],
                     "avg_launch_ms": avg_img_s * 1e3},
        "qbank": {"avg_launch_ms": float(np.mean(bank_ms)), "achieved_GBs": bank_gbs,
                  "frac": bank_gbs / HBM_PEAK_GBS, "bytes_per_trace": bank_bytes[0] / ranges[-1][1]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(cfg, sv, tab, geom, ranges, traces, ntr):
    """Oracle FP32 path (bank + sum) on the first ntr traces of shot 0."""
    from oracle import oracle as O
    from tests.helpers import o_bparams, o_iparams

    tr0, _ = ranges[0]
    g = np.ascontiguousarray(geom[tr0:tr0 + ntr])
    x = traces[0][:ntr].cpu().numpy()
    t_host = tab.cpu().numpy()
    t0 = time.perf_counter()
    b = O.qbank(x, o_bparams(sv), dtype=np.float32)
    t1 = time.perf_counter()
    _, n = O.image(b, g, t_host, o_iparams(sv), fp32=True)
    t2 = time.perf_counter()
    log(f"cpu baseline: bank {t1 - t0:.1f}s image {t2 - t1:.1f}s updates {n}")
    return {"value": n / (t2 - t0), "unit": "updates/s", "cores": O.num_threads(),
            "kind": "port", "sample": f"shot 0, first {ntr} traces, full 401x401x301 image "
                                       f"(FP32 Q bank + FP32 sum, OpenMP)",
            "seconds": t2 - t0, "image_seconds": t2 - t1}


if __name__ == "__main__":
    main()
