#!/usr/bin/env python3
"""Benchmark of the plain GPU drop-in: one depth request = census 7x7 -> cost volume -> 8-path SGM with the
reference's schedule -> WTA + subpixel + L/R, on one 4160x3120 uint16 10-bit pair, D=256 (P1=8, P2=96).

Prints ONE JSON line. `value` = Gcells/s = W*H*D*8 paths / time per request, inputs resident in HBM.
`roofline` = the dominant kernel's algorithmic HBM bytes per launch / its average duration (HIP events on
the stream the kernels run on). `cpu_baseline` = the CPU oracle (the reference's algorithm restated in C:
row-parallel cost, serial aggregation) on a bounded band of the same workload, on this host.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H, D = 4160, 3120, 256
P1, P2, TILE, RADIUS, BITS = 8, 96, 64, 3, 10
HBM_PEAK_GBS = 8000.0
PATH_BYTES_PER_CELL = 7  # write raw 1 + pass +1 (raw 1, agg 2) + pass -1 (raw 1, agg 2), WTA fused
# algorithmic HBM bytes per cell of each stage (what the stage must move at minimum)
# algorithmic HBM bytes per cell of each timed stage:
#   rows+ (pass +1 path ->, census cost fused): writes raw (1) and agg (2)
#   vd+k / vd-k (one vertical or diagonal path sweep): reads raw (1), read-modify-writes agg (2 + 2);
#     with fused bands vd+1 / vd-1 is the one launch for the pass's three paths, still 5 B/cell
#   rows- (pass -1 path <-, winner-take-all fused): reads raw (1) and agg (2)
STAGE_BYTES = {"rows+": 3, "rows-": 3}
for _k in ("vd+1", "vd+2", "vd+3", "vd-1", "vd-2", "vd-3"):
    STAGE_BYTES[_k] = 5
KERNEL_NAMES = {"rows+": "sgm_row_fwd (pass +1 path ->, census cost fused)",
                "rows-": "sgm_row_bwd_wta (pass -1 path <-, winner-take-all fused)",
                "vd+1": "sgm_sweep<pass +1, path 1 (down-right)>", "vd+2": "sgm_sweep<pass +1, path 2 (down)>",
                "vd+3": "sgm_sweep<pass +1, path 3 (down-left)>", "vd-1": "sgm_sweep<pass -1, path 1 (up-left)>",
                "vd-2": "sgm_sweep<pass -1, path 2 (up)>", "vd-3": "sgm_sweep<pass -1, path 3 (up-right)>"}
BAND_NAMES = {"vd+1": "sgm_band<pass +1> (paths 1-3 fused, ticketed band tiles)",
              "vd-1": "sgm_band<pass -1> (paths 1-3 fused, ticketed band tiles)"}


def host_info():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(left, right, rows, threads):
    from oracle import oracle

    band_l = left[:rows].copy()
    band_r = right[:rows].copy()
    t0 = time.perf_counter()
    out = oracle.rectified(band_l, band_r, D, cost="census", radius=RADIUS, bits=BITS, p1=P1, p2=P2, T=TILE,
                           tol=1, threads=threads)
    t = time.perf_counter() - t0
    cells = W * rows * D
    return cells * 8 / t / 1e9, t, out["times_ms"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-rows", type=int, default=1560, help="rows of the CPU-baseline band (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bands", type=int, default=0, help="1: fused band tiles for the vertical/diagonal paths")
    ap.add_argument("--profile-run", action="store_true",
                    help="for counter runs: a few requests, then a 1 GiB device copy as the byte-count calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU layout is a design note.",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py is single-process, single-GPU", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from lux_amd import scene, sgm

    left, right, truth = scene.make_pair(W, H, D - 1, planes=6, sigma=4.0, seed=1, bits=BITS)
    digest = scene.sha256(left, right)
    dev = torch.device("cuda:0")
    dl = torch.from_numpy(left.view(np.int16)).to(dev)
    dr = torch.from_numpy(right.view(np.int16)).to(dev)
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    sub = torch.empty((H, W), dtype=torch.float32, device=dev)
    valid = torch.empty((H, W), dtype=torch.uint8, device=dev)
    ctx = sgm.Context(0)
    ctx.set_option(sgm.OPT_FUSED_BANDS, args.bands)
    p = sgm.rect_params(W, H, D, bits=BITS, cost="census", radius=RADIUS, p1=P1, p2=P2, tile=TILE, tol=1)
    # a stream of our own (default priority): the library enqueues on it, the events below are recorded on it
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    def request():
        sgm.rectified_device(ctx, p, dl.data_ptr(), dr.data_ptr(), disp.data_ptr(), sub.data_ptr(), valid.data_ptr(),
                             0, stream.cuda_stream)

    for _ in range(args.warmup):
        request()
    torch.cuda.synchronize()
    if args.profile_run:
        for _ in range(args.steps):
            request()
        torch.cuda.synchronize()
        src = torch.ones(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)  # calibration: reads and writes exactly 1 GiB
        torch.cuda.synchronize()
        print(json.dumps({"profile_run": True, "requests": args.steps, "copy_bytes": 1 << 30}))
        ctx.close()
        return
    # timed region: K requests back to back, HIP events on the stream the kernels run on
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        request()
    e1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1) / args.steps
    gcells = W * H * D * 8 / (ms / 1e3) / 1e9

    # per-request latency and per-kernel durations (profiled requests, one at a time)
    ctx.set_profiling(True)
    lat, stages = [], {}
    for _ in range(max(args.steps, 5)):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        request()
        b.record(stream)
        st = ctx.stage_times()
        torch.cuda.synchronize()
        lat.append(a.elapsed_time(b))
        for k, v in st.items():
            stages.setdefault(k, []).append(v)
    ctx.set_profiling(False)
    avg = {k: float(np.mean(v)) for k, v in stages.items()}
    dom = max((k for k in avg if k in STAGE_BYTES), key=lambda k: avg[k])
    names = dict(KERNEL_NAMES, **(BAND_NAMES if args.bands else {}))
    cells = W * H * D
    achieved = STAGE_BYTES[dom] * cells / (avg[dom] / 1e3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        try:
            with open(tf) as fh:
                traffic = json.load(fh).get("bands" if args.bands else "sweeps", {}).get(dom)
        except (OSError, ValueError):
            traffic = None

    d = disp.cpu().numpy().view(np.uint16)
    v = valid.cpu().numpy() > 0
    acc = float((np.abs(d.astype(np.int64) - truth)[v] <= 1).mean()) if v.any() else 0.0

    cpu = None
    if not args.no_cpu and args.cpu_rows > 0:
        model, ncpu = host_info()
        threads = min(16, ncpu)
        val, secs, tms = cpu_baseline(left, right, args.cpu_rows, threads)
        cpu = {"value": round(val, 4), "unit": "Gcells/s", "cores": threads, "kind": "port",
               "sample": (f"oracle/lux_oracle.c rectified pipeline on rows 0..{args.cpu_rows - 1} of the same "
                          f"4160x{H} pair (D=256, census 7x7): {secs:.1f} s (cost {tms[0] / 1e3:.1f} s on {threads} "
                          f"threads, aggregation {tms[1] / 1e3:.1f} s serial as DenseStereoSgm.cs:171, "
                          f"WTA {tms[2] / 1e3:.1f} s); host {model}, {ncpu} CPUs")}

    out = {
        "metric": "SGM Gcells/s (W·H·D·paths) + p50/p99 depth-request latency, served vs plain",
        "value": round(gcells, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": "plain GPU drop-in, one 4160x3120 uint16 10-bit pair, census 7x7, D=256, 8 paths, "
                               "P1=8, P2=96, WTA + subpixel + L/R check (tol 1), reference schedule T=64",
                   "vertical_diagonal": "fused bands" if args.bands else "per-path sweeps",
                   "width": W, "height": H, "disparities": D, "cells": cells, "inputs_sha256": digest},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                       "form": "plain"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": names.get(dom, dom), "avg_ms": round(avg[dom], 4),
                     "algorithmic_bytes": STAGE_BYTES[dom] * cells},
        "path_roofline": {"bytes_per_cell": PATH_BYTES_PER_CELL,
                          "achieved": round(PATH_BYTES_PER_CELL * cells / (ms / 1e3) / 1e9, 1),
                          "frac": round(PATH_BYTES_PER_CELL * cells / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "stage_ms": {k: round(v, 4) for k, v in avg.items()},
        "accuracy": {"valid_frac": round(float(v.mean()), 4), "within1_of_truth_on_valid": round(acc, 4)},
        "wall_s": round(t_wall, 3),
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
