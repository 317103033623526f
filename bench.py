#!/usr/bin/env python3
"""Benchmark of the plain GPU drop-in: one depth request = census 7x7 -> cost volume -> 8-path SGM with the
reference's schedule -> WTA + subpixel + L/R, on one 4160x3120 uint16 10-bit pair, D=256 (P1=8, P2=96).

Prints ONE JSON line. `value` = Gcells/s = W*H*D*8 paths / time per request, inputs resident in HBM.
`roofline` = the path's algorithmic HBM bytes (7 per cell: the reference's two-pass decomposition, write raw 1, pass +1
raw 1 + agg 2, pass -1 raw 1 + agg 2) / the request's time (HIP events on the stream the kernels run on), against the
8 TB/s peak; `roofline.traffic` = the HBM bytes one request moves, counted by rocprofv3 (FETCH_SIZE, WRITE_SIZE) and
read from profiles/traffic.json when that file was made from the same kernel sources. `kernel_roofline` = the
dominant kernel's own bytes over its time. `cpu_baseline` = the CPU oracle (the reference's algorithm restated in C:
row-parallel cost, serial aggregation) on a bounded band of the same workload, on this host: the median of five runs
after one warm-up; `cpu_baseline_c0` is the same for the 640x480 uint8 D=64 workload, whole.

More legs go into the same line:
  `batched`: one capture of 10 pairs (6 x 2080x1560 D=128, 4 x 3120x2340 D=192, independent seeds) as one graph
             launch with the autotuned schedule (graph branches, or the persistent executor in one launch or one
             launch per pass group), against the same pairs run one after another (plain drop-in); the executor's
             two forms are also timed on their own (`batched.executor`);
  `served`:  a mixed stream of depth requests (sizes uniform over 640x480 D64, 1040x780 D96, 2080x1560 D128,
             4160x3120 D256; host images in, host results out) through the request loop under a 96 GB budget:
             latency p50/p99 at Poisson arrivals of 40 req/s, served and plain (the plain drop-in serving the same
             arrivals in order, one at a time), and throughput at saturation against the plain drop-in in order;
             the same stream again through the device-resident loop (`served.device_loop`: the scheduler a
             persistent kernel, admission into a device pool of the budget);
  `coarse_to_fine`: 2000 3-level requests (1040x780 D=64 -> 2080x1560 +-16 -> 4160x3120 +-8) in bursts of 16
             through the request loop under a 192 GB budget, latency p50/p99 against plain sequential execution;
  `l16_pyramid*`: DenseStereoPyramid.Run at 2080x1560 (device form), its oracle on the host, and a plane span whose
             levels spill past their line-buffer segments (the real-capture case).
"""
import argparse
import glob
import hashlib
import json
import os
import platform
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, H, D = 4160, 3120, 256
P1, P2, TILE, RADIUS, BITS = 8, 96, 64, 3, 10
HBM_PEAK_GBS = 8000.0
PATH_BYTES_PER_CELL = 7  # write raw 1 + pass +1 (raw 1, agg 2) + pass -1 (raw 1, agg 2), WTA fused
# HBM bytes per cell each timed stage of this design moves (what it must move at minimum):
#   byte path volumes (census requests whose path costs fit a byte, lux_amd/csrc/engine.hip path_bytes):
#     rows+ (pass +1 path ->, census cost fused): writes raw (1) and its path volume (1)
#     vd+ / vd- (the three vertical and diagonal paths of a pass, one launch): each reads raw (1) and writes its path
#     volume (1)
#     rows- (pass -1 path <-, census cost recomputed from the codes, the sum of the eight paths and the winner-take-all
#     fused): seven volumes (7)
#   u16 aggregate (every other request): rows+ raw 1 + agg 2; a sweep raw 1 + agg read 2 + agg write 2 (three a pass);
#     rows- raw 1 + agg 2
STAGE_BYTES_PV = {"rows+": 2, "vd+": 6, "vd-": 6, "rows-": 7}
STAGE_BYTES_U16 = {"rows+": 3, "vd+": 15, "vd-": 15, "rows-": 3}
KERNEL_NAMES = {"rows+": "sgm_row_fwd (pass +1 path ->, census cost fused)",
                "rows-": "sgm_row_bwd_wta (pass -1 path <-, eight-path sum and winner-take-all fused)",
                "vd+": "the sweeps of pass +1 (down-right, down, down-left; sgm_sweeps_pv: one launch)",
                "vd-": "the sweeps of pass -1 (up-left, up, up-right; sgm_sweeps_pv: one launch)"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def path_bytes(d, radius, p2, align=8):
    """Whether a census request runs with byte path volumes: the rule of path_bytes in lux_amd/csrc/engine.hip (no
    padding planes, census maximum + P2 <= 255, at most 256 planes)."""
    cap = -(-d // align) * align
    return cap == d and d <= 256 and (2 * radius + 1) ** 2 - 1 + p2 <= 255


def source_stamp():
    """Digest of the kernel sources a counter profile was made from (the library's .hip/.h/.cpp and its header)."""
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(ROOT, "lux_amd", "csrc", "*.hip")) +
                   glob.glob(os.path.join(ROOT, "lux_amd", "csrc", "*.h")) +
                   glob.glob(os.path.join(ROOT, "lux_amd", "csrc", "*.cpp")) +
                   [os.path.join(ROOT, "include", "lux_sgm.h")])
    for f in files:
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


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


def cpu_baseline(left, right, rows, threads, d=D, radius=RADIUS, bits=BITS, p1=P1, p2=P2, reps=5):
    """The oracle on rows 0..rows-1 of the pair: one warm-up, then `reps` timed runs; the median."""
    from oracle import oracle

    band_l = left[:rows].copy()
    band_r = right[:rows].copy()
    kw = dict(cost="census", radius=radius, bits=bits, p1=p1, p2=p2, T=TILE, tol=1, threads=threads)
    oracle.rectified(band_l, band_r, d, **kw)
    ts, parts = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = oracle.rectified(band_l, band_r, d, **kw)
        ts.append(time.perf_counter() - t0)
        parts.append(out["times_ms"])
    import numpy as np

    k = int(np.argsort(ts)[len(ts) // 2])
    cells = left.shape[1] * rows * d
    return cells * 8 / ts[k] / 1e9, ts[k], parts[k], sorted(ts)


def hbm_ceilings(torch, gib=4, reps=5):
    """What plain streaming kernels reach on this device (torch's own copy, fill and max kernels over `gib` GiB), for
    reading the roofline fractions: the published 8 TB/s is not reached by any access mix, and a mix of reads and
    writes (the sweeps) streams slower than reads alone (the backward row path)."""
    n = gib << 30
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def rate(fn, moved):
        fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return round(moved * reps / (ev[0].elapsed_time(ev[1]) / 1e3) / 1e9, 1)

    out = {"copy_gbs": rate(lambda: b.copy_(a), 2 * n),
           "read_gbs": rate(lambda: a.view(torch.int64).max(), n),
           "write_gbs": rate(lambda: b.fill_(3), n),
           "how": f"torch copy_ (reads + writes), int64 max (reads) and fill_ (writes) over {gib} GiB, "
                  f"mean of {reps} after one"}
    del a, b
    torch.cuda.empty_cache()
    return out


def load_traffic():
    """The counted HBM bytes of one request (profiles/traffic.json), used only when the file was made from these kernel
    sources: (data or None, a note saying which)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, "no counter profile (profiles/traffic.json)"
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError) as e:
        return None, f"profiles/traffic.json unreadable: {e}"
    stamp = source_stamp()
    if t.get("source_sha16") != stamp:
        return None, (f"profiles/traffic.json counts kernels of sources {t.get('source_sha16')}, these are {stamp}: "
                      f"not used")
    return t, f"profiles/traffic.json, rocprofv3 FETCH_SIZE/WRITE_SIZE of these kernel sources ({stamp})"


def bench_c0(np, torch, sgm, scene, threads, model, ncpu, reps=20):
    """The CPU reference path's workload, whole: 640x480 uint8, census 5x5, D=64, P1=10, P2=120, L/R tolerance 1. The
    oracle on the host (median of five runs after one warm-up) and the device (inputs in HBM, HIP events)."""
    from oracle import oracle

    w, h, d = 640, 480, 64
    left, right, _ = scene.make_pair(w, h, d - 1, planes=4, sigma=2.0, seed=0, bits=8)
    kw = dict(cost="census", radius=2, bits=8, p1=10, p2=120, T=TILE, tol=1, threads=threads)
    oracle.rectified(left, right, d, **kw)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        want = oracle.rectified(left, right, d, **kw)
        ts.append(time.perf_counter() - t0)
    cpu_s = float(np.median(ts))
    dev = torch.device("cuda:0")
    dl, dr = torch.from_numpy(left).to(dev), torch.from_numpy(right).to(dev)
    od = torch.empty((h, w), dtype=torch.int16, device=dev)
    osub = torch.empty((h, w), dtype=torch.float32, device=dev)
    ov = torch.empty((h, w), dtype=torch.uint8, device=dev)
    p = sgm.rect_params(w, h, d, bits=8, cost="census", radius=2, p1=10, p2=120, tile=TILE, tol=1)
    st = torch.cuda.Stream()
    ctx = sgm.Context(0)
    call = lambda: sgm.rectified_device(ctx, p, dl.data_ptr(), dr.data_ptr(), od.data_ptr(), osub.data_ptr(),  # noqa: E731
                                        ov.data_ptr(), 0, st.cuda_stream)
    call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        call()
    e1.record(st)
    torch.cuda.synchronize()
    gpu_ms = e0.elapsed_time(e1) / reps
    same = bool(np.array_equal(od.cpu().numpy().view(np.uint16), want["disp"]) and
                np.array_equal(ov.cpu().numpy(), want["valid"]) and np.array_equal(osub.cpu().numpy(), want["sub"]))
    ctx.close()
    cells = w * h * d
    return {"workload": "640x480 uint8, census 5x5, D=64, P1=10, P2=120, L/R tolerance 1 (4 planes, sigma 2, seed 0)",
            "value": round(cells * 8 / cpu_s / 1e9, 4), "unit": "Gcells/s", "cores": threads, "kind": "port",
            "sample": f"the whole pair, median of 5 runs after 1 warm-up: {cpu_s * 1e3:.1f} ms; host {model}, {ncpu} CPUs",
            "runs_ms": [round(t * 1e3, 2) for t in sorted(ts)],
            "gpu_ms": round(gpu_ms, 4), "gpu_gcells_s": round(cells * 8 / gpu_ms / 1e6, 2),
            "gpu_vs_cpu": round(cpu_s * 1e3 / gpu_ms, 1), "identical_to_oracle": same}


def _stage_of(kernel):
    """The bench stage a kernel's dispatches belong to (rocprofv3's demangled names), or None."""
    m = re.search(r"sgm_sweeps?(?:_pv)?<(\d+), (-?1)", kernel)
    if m:
        return f"vd{'+' if m.group(2) == '1' else '-'}"
    if re.search(r"sgm_row_fwd(?:_pv)?<", kernel):
        return "rows+"
    if re.search(r"sgm_row_bwd_wta(?:_pv)?<", kernel):
        return "rows-"
    return None


def summarize_pmc(fetch_dir, write_dir, copy_bytes=1 << 30):
    """profiles/traffic.json from two rocprofv3 --pmc runs of `bench.py --profile-run` (FETCH_SIZE, then WRITE_SIZE, in
    KiB per dispatch): per kernel the mean bytes of a dispatch, scaled by what each counter reads for the 1 GiB device
    copy the run ends with, times the dispatches of one request (a request launches the census kernel twice)."""
    import csv

    def read(d, counter):
        per = {}
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        for f in files:
            with open(f, newline="") as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    per.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]) * 1024.0)
        return per

    fetch, write = read(fetch_dir, "FETCH_SIZE"), read(write_dir, "WRITE_SIZE")
    copy = [k for k in fetch if "copyBuffer" in k]
    if not copy:
        raise SystemExit("no copy kernel in the profile: run bench.py --profile-run")
    rf = copy_bytes / max(fetch[copy[0]][-1], 1.0)
    wf = copy_bytes / max(write[copy[0]][-1], 1.0)
    census = [k for k in fetch if "census_kernel" in k]
    nreq = sum(len(fetch[k]) for k in census) / 2
    kernels, stages, total = {}, {}, 0.0
    for k in fetch:
        if "lux::" not in k:  # the calibration copy and torch's own kernels (the copy's source fill)
            continue
        rd = sum(fetch[k]) / nreq * rf
        wr = sum(write.get(k, [0.0])) / nreq * wf
        kernels[k] = {"read": round(rd), "write": round(wr), "dispatches_per_request": len(fetch[k]) / nreq}
        st = _stage_of(k)
        if st:
            stages[st] = stages.get(st, 0) + round(rd + wr)
        total += rd + wr
    out = {"source_sha16": source_stamp(), "workload": f"{W}x{H} D={D} census 7x7 P1={P1} P2={P2}",
           "read_factor": round(rf, 4), "write_factor": round(wf, 4), "requests": nreq,
           "request_bytes": round(total), "stages": stages, "kernels": kernels}
    with open(TRAFFIC_FILE, "w") as fh:
        json.dump(out, fh, indent=1)
    cells = W * H * D
    print(f"read factor {rf:.4f}, write factor {wf:.4f}, {nreq:g} requests; per request {total / 1e9:.3f} GB = "
          f"{total / cells:.3f} B/cell")
    print("| kernel | read GB | write GB | total GB | B/cell |\n|---|---:|---:|---:|---:|")
    for k, v in sorted(kernels.items(), key=lambda kv: -(kv[1]["read"] + kv[1]["write"])):
        t = v["read"] + v["write"]
        name = re.sub(r"^void |lux::|\(anonymous namespace\)::|\(.*$", "", k)
        print(f"| `{name}` | {v['read'] / 1e9:.3f} | {v['write'] / 1e9:.3f} | {t / 1e9:.3f} | {t / cells:.3f} |")


BATCH = [(2080, 1560, 128)] * 6 + [(3120, 2340, 192)] * 4
SERVE_SIZES = [(640, 480, 64), (1040, 780, 96), (2080, 1560, 128), (4160, 3120, 256)]


def scene_params(sgm, w, h, d):
    return sgm.rect_params(w, h, d, bits=BITS, cost="census", radius=RADIUS, p1=P1, p2=P2, tile=TILE, tol=1)


def bench_batched(np, torch, sgm, scene, reps):
    """Ten pairs of one capture: plain one-by-one on one stream vs one tuned graph launch."""
    dev = torch.device("cuda:0")
    pairs, keep, cells = [], [], 0
    for i, (w, h, d) in enumerate(BATCH):
        left, right, _ = scene.make_pair(w, h, d - 1, planes=6, sigma=4.0, seed=100 + i, bits=BITS)
        tl = torch.from_numpy(left.view(np.int16)).to(dev)
        tr = torch.from_numpy(right.view(np.int16)).to(dev)
        od = torch.empty((h, w), dtype=torch.int16, device=dev)
        osub = torch.empty((h, w), dtype=torch.float32, device=dev)
        ov = torch.empty((h, w), dtype=torch.uint8, device=dev)
        pr = sgm.Pair()
        pr.params = scene_params(sgm, w, h, d)
        pr.left, pr.right, pr.disp, pr.sub, pr.valid = (tl.data_ptr(), tr.data_ptr(), od.data_ptr(), osub.data_ptr(),
                                                        ov.data_ptr())
        pairs.append(pr)
        keep += [tl, tr, od, osub, ov]
        cells += w * h * d
    stream = torch.cuda.Stream()
    ctx = sgm.Context(0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def plain_all():
        for pr in pairs:
            sgm.rectified_device(ctx, pr.params, pr.left, pr.right, pr.disp, pr.sub, pr.valid, 0, stream.cuda_stream)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_plain = timed(plain_all)
    ref = [t.clone() for t in keep[2::5]]  # plain disparities
    best, tried = sgm.batch_tune(ctx, pairs, stream=stream.cuda_stream, reps=2)
    ms_batch = timed(lambda: sgm.batch(ctx, pairs, stream=stream.cuda_stream))
    same = all(bool(torch.equal(a, b)) for a, b in zip(ref, keep[2::5]))
    # the persistent executor on its own, whatever the tuner picked: one launch per pass group, and one launch
    executor = {}
    for groups, key in ((7, "per_pass_group"), (1, "one_launch")):
        for t in keep[2::5]:
            t.fill_(-1)
        sgm.batch_set_schedule(ctx, pairs, executor=groups, order=1)
        ms = timed(lambda: sgm.batch(ctx, pairs, stream=stream.cuda_stream))
        executor[key] = {"launches": groups, "ms_per_batch": round(ms, 3), "gcells_s": round(cells * 8 / ms / 1e6, 3),
                         "speedup_vs_plain": round(ms_plain / ms, 3),
                         "identical_to_plain": all(bool(torch.equal(a, b)) for a, b in zip(ref, keep[2::5]))}
    ctx.close()
    del keep
    return {"workload": "6 x 2080x1560 D=128 + 4 x 3120x2340 D=192, uint16 10-bit, census 7x7, seeds 100..109",
            "cells": cells, "ms_per_batch": round(ms_batch, 3), "gcells_s": round(cells * 8 / ms_batch / 1e6, 3),
            "plain_ms": round(ms_plain, 3), "plain_gcells_s": round(cells * 8 / ms_plain / 1e6, 3),
            "speedup_vs_plain": round(ms_plain / ms_batch, 3), "schedule": best,
            "candidates": tried, "identical_to_plain": same, "executor": executor}


def bench_served(np, sgm, scene, n_req, rate, budget_gb, slots, window):
    """Mixed request stream: plain drop-in in order, served at saturation, served at Poisson arrivals."""
    import threading
    import queue

    rng = np.random.default_rng(7)
    sizes = rng.integers(0, len(SERVE_SIZES), n_req)
    gaps = rng.exponential(1.0 / rate, n_req)
    scenes = {}
    for k, (w, h, d) in enumerate(SERVE_SIZES):
        scenes[k] = [scene.make_pair(w, h, d - 1, planes=6, sigma=4.0, seed=200 + 10 * k + j, bits=BITS)[:2]
                     for j in range(2)]
    params = {k: scene_params(sgm, *SERVE_SIZES[k]) for k in scenes}
    cells = int(sum(SERVE_SIZES[k][0] * SERVE_SIZES[k][1] * SERVE_SIZES[k][2] for k in sizes))

    def outputs(k):
        w, h, _ = SERVE_SIZES[k]
        return (np.empty((h, w), np.uint16), np.empty((h, w), np.float32), np.empty((h, w), np.uint8))

    # plain drop-in: lux_sgm_rectified (host images in, host results out), one request after another
    ctx = sgm.Context(0)
    for k in scenes:  # warm every size's workspace once
        sgm.rectified(ctx, *scenes[k][0], SERVE_SIZES[k][2], bits=BITS, radius=RADIUS, p1=P1, p2=P2, tile=TILE)
    plain_lat = []
    t0 = time.perf_counter()
    for i, k in enumerate(sizes):
        a = time.perf_counter()
        sgm.rectified(ctx, *scenes[k][i % 2], SERVE_SIZES[k][2], bits=BITS, radius=RADIUS, p1=P1, p2=P2, tile=TILE)
        plain_lat.append((time.perf_counter() - a) * 1e3)
    plain_s = time.perf_counter() - t0
    # the plain drop-in under the same Poisson arrivals: one request at a time in arrival order (a FIFO), latency from
    # arrival to results in host memory, as for the served requests
    arrivals = np.cumsum(gaps)
    plain_poi = []
    t_start = time.perf_counter()
    for i, k in enumerate(sizes):
        arrive = t_start + arrivals[i]
        delay = arrive - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        sgm.rectified(ctx, *scenes[k][i % 2], SERVE_SIZES[k][2], bits=BITS, radius=RADIUS, p1=P1, p2=P2, tile=TILE)
        plain_poi.append((time.perf_counter() - arrive) * 1e3)
    ctx.close()

    def serve(srv):
        """Warm-up, then the stream at saturation and at the Poisson arrivals through server srv."""

        def run(arrivals):
            """Submit in order (at arrival times if given, else keeping `window` in flight); collect on a thread."""
            q = queue.Queue()
            stats, free = [], {k: [] for k in scenes}
            sem = threading.Semaphore(window if arrivals is None else n_req)

            def collector():
                while True:
                    item = q.get()
                    if item is None:
                        return
                    t, k, outs = item
                    stats.append(srv.wait(t))
                    free[k].append(outs)
                    sem.release()

            th = threading.Thread(target=collector)
            th.start()
            t_start = time.perf_counter()
            for i, k in enumerate(sizes):
                if arrivals is not None:
                    delay = t_start + arrivals[i] - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                sem.acquire()
                outs = free[k].pop() if free[k] else outputs(k)
                t = srv.submit(params[k], *scenes[k][i % 2], disp=outs[0], sub=outs[1], valid=outs[2])
                q.put((t, k, outs))
            q.put(None)
            th.join()
            return time.perf_counter() - t_start, stats

        for k in scenes:  # warm every slot's workspace and the pinned staging of a full window, outputs included
            keep = [outputs(k) for _ in range(window)]
            run_w = [srv.submit(params[k], *scenes[k][0], disp=o[0], sub=o[1], valid=o[2]) for o in keep]
            for t in run_w:
                srv.wait(t)
        sat_s, sat_stats = run(None)
        print(f"bench: served at saturation {sat_s:.2f} s", file=sys.stderr, flush=True)
        poi_s, poi_stats = run(arrivals)
        print(f"bench: served at Poisson arrivals {poi_s:.2f} s", file=sys.stderr, flush=True)
        info = srv.info()
        srv.close()
        return sat_s, sat_stats, poi_s, poi_stats, info

    sat_s, sat_stats, poi_s, poi_stats, info = serve(sgm.Server(0, int(budget_gb * 1e9), slots=slots))
    # the same stream through the device-resident loop (its scheduler a persistent kernel, the pool the budget)
    d_sat_s, d_sat_stats, d_poi_s, d_poi_stats, d_info = serve(
        sgm.Server(0, int(budget_gb * 1e9), slots=slots, device_loop=True))

    def pct(v, q):
        return round(float(np.percentile(np.array([st["total_ms"] for st in v]), q)), 3)

    device_loop = {"poisson": {"rate_per_s": rate, "p50_ms": pct(d_poi_stats, 50), "p99_ms": pct(d_poi_stats, 99),
                               "wall_s": round(d_poi_s, 3)},
                   "saturation": {"window": window, "wall_s": round(d_sat_s, 3),
                                  "gcells_s": round(cells * 8 / d_sat_s / 1e9, 3),
                                  "req_per_s": round(n_req / d_sat_s, 3), "p50_ms": pct(d_sat_stats, 50),
                                  "p99_ms": pct(d_sat_stats, 99)},
                   "vs_plain_throughput": round(plain_s / d_sat_s, 3),
                   "bytes_peak_gb": round(d_info["bytes_peak"] / 1e9, 3), "failed": d_info["failed"],
                   "latency": "submit -> results in host memory"}
    lat = np.array([st["total_ms"] for st in poi_stats])
    sat_lat = np.array([st["total_ms"] for st in sat_stats])
    return {"workload": f"{n_req} requests, sizes uniform over 640x480 D64 / 1040x780 D96 / 2080x1560 D128 / "
                        f"4160x3120 D256 (seed 7), uint16 10-bit, census 7x7; host images in, host results out",
            "cells": cells, "slots": slots, "budget_gb": budget_gb,
            "poisson": {"rate_per_s": rate, "p50_ms": round(float(np.percentile(lat, 50)), 3),
                        "p99_ms": round(float(np.percentile(lat, 99)), 3), "wall_s": round(poi_s, 3),
                        "latency": "submit -> device work and result download complete"},
            "saturation": {"window": window, "wall_s": round(sat_s, 3),
                           "gcells_s": round(cells * 8 / sat_s / 1e9, 3),
                           "req_per_s": round(n_req / sat_s, 3),
                           "p50_ms": round(float(np.percentile(sat_lat, 50)), 3),
                           "p99_ms": round(float(np.percentile(sat_lat, 99)), 3)},
            "plain": {"wall_s": round(plain_s, 3), "gcells_s": round(cells * 8 / plain_s / 1e9, 3),
                      "req_per_s": round(n_req / plain_s, 3),
                      "service_p50_ms": round(float(np.percentile(plain_lat, 50)), 3),
                      "service_p99_ms": round(float(np.percentile(plain_lat, 99)), 3),
                      "poisson": {"rate_per_s": rate, "p50_ms": round(float(np.percentile(plain_poi, 50)), 3),
                                  "p99_ms": round(float(np.percentile(plain_poi, 99)), 3),
                                  "latency": "arrival -> results in host memory, requests served in arrival order"}},
            "served_vs_plain_throughput": round(plain_s / sat_s, 3),
            "bytes_peak_gb": round(info["bytes_peak"] / 1e9, 3), "failed": info["failed"],
            "device_loop": device_loop}


PYR_W, PYR_H, PYR_D, PYR_RADII = 4160, 3120, 256, (16, 8)


def bench_pyramid(np, sgm, scene, n_req, burst, budget_gb, slots):
    """Coarse-to-fine queue: 3-level requests (1040x780 D=64 -> 2080x1560 +-16 -> 4160x3120 +-8) arriving in
    bursts; latency from a burst's arrival to each request's completion, served vs the same requests one after
    another (plain sequential execution)."""
    scenes = [scene.make_pair(PYR_W, PYR_H, PYR_D - 1, planes=6, sigma=4.0, seed=300 + j, bits=BITS, fg_frac=0.1)[:2]
              for j in range(2)]
    pp = sgm.pyramid_params(PYR_W, PYR_H, PYR_D, PYR_RADII, bits=BITS, radius=RADIUS, p1=P1, p2=P2, tile=TILE)
    cells_per = (PYR_W // 4) * (PYR_H // 4) * (PYR_D // 4) + (PYR_W // 2) * (PYR_H // 2) * 2 * PYR_RADII[0] + \
        PYR_W * PYR_H * 2 * PYR_RADII[1]
    outs = [(np.empty((PYR_H, PYR_W), np.uint16), np.empty((PYR_H, PYR_W), np.float32)) for _ in range(burst)]
    # plain sequential: each burst's requests one after another through lux_sgm_pyramid
    ctx = sgm.Context(0)
    sgm.pyramid(ctx, *scenes[0], PYR_D, PYR_RADII, bits=BITS, radius=RADIUS, p1=P1, p2=P2, tile=TILE)
    plain_lat, svc = [], []
    t0 = time.perf_counter()
    for b in range(n_req // burst):
        arrive = time.perf_counter()
        for i in range(burst):
            a = time.perf_counter()
            sgm.pyramid(ctx, *scenes[(b + i) % 2], PYR_D, PYR_RADII, bits=BITS, radius=RADIUS, p1=P1, p2=P2, tile=TILE)
            now = time.perf_counter()
            svc.append((now - a) * 1e3)
            plain_lat.append((now - arrive) * 1e3)
    plain_s = time.perf_counter() - t0
    ctx.close()
    srv = sgm.Server(0, int(budget_gb * 1e9), slots=slots)
    # warm every slot and the pinned staging of a whole burst, outputs included (staging is made at first use)
    for t in [srv.submit_pyramid(pp, *scenes[0], disp=outs[i][0], sub=outs[i][1]) for i in range(burst)]:
        srv.wait(t)
    lat = []
    t0 = time.perf_counter()
    for b in range(n_req // burst):
        arrive = time.perf_counter()
        ts = [srv.submit_pyramid(pp, *scenes[(b + i) % 2], disp=outs[i][0], sub=outs[i][1]) for i in range(burst)]
        for t in ts:
            srv.wait(t)
            lat.append((time.perf_counter() - arrive) * 1e3)
    served_s = time.perf_counter() - t0
    info = srv.info()
    srv.close()
    n = (n_req // burst) * burst
    pct = lambda v, q: round(float(np.percentile(v, q)), 3)  # noqa: E731
    return {"workload": f"{n} coarse-to-fine requests (1040x780 D=64 -> 2080x1560 +-16 -> 4160x3120 +-8), "
                        f"uint16 10-bit, census 7x7, 10% foreground, bursts of {burst}; host images in, host results out",
            "requests": n, "burst": burst, "slots": slots, "budget_gb": budget_gb, "cells_per_request": cells_per,
            "served": {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "wall_s": round(served_s, 3),
                       "req_per_s": round(n / served_s, 3), "gcells_s": round(n * cells_per * 8 / served_s / 1e9, 3)},
            "plain_sequential": {"p50_ms": pct(plain_lat, 50), "p99_ms": pct(plain_lat, 99), "wall_s": round(plain_s, 3),
                                 "req_per_s": round(n / plain_s, 3), "service_ms_p50": pct(svc, 50),
                                 "gcells_s": round(n * cells_per * 8 / plain_s / 1e9, 3)},
            "served_vs_plain_throughput": round(plain_s / served_s, 3),
            "latency": "burst arrival -> request complete (results in host memory)",
            "bytes_peak_gb": round(info["bytes_peak"] / 1e9, 3), "failed": info["failed"]}


L16_W, L16_H, L16_N8, L16_N2 = 2080, 1560, 304, 302
L16_SPAN = (500.0, 20000.0)        # plane depths: ranges up to 88 planes, every level on the path sweeps
L16_SPAN_WIDE = (1000.0, 2200.0)   # the same planes over the scene's own depth span: ranges up to 288 planes at the
                                   # depth edges, levels 1-5 on the sweeps with four packed pairs a lane


L16_SPILL = (1000.0, 2200.0, 64, 62)  # 64 planes over the scene's own span: MinLo 0 and ranges ending at the top
                                      # plane, so thousands of cells a level spill past their line-buffer segment


def bench_l16(np, torch, sgm, scene, reps, cpu_threads, span=L16_SPAN, upsample=True, n8=L16_N8, n2=L16_N2):
    """The L16 dense stereo end to end (DenseStereoPyramid.Run, six levels from 65x49 to 2080x1560, then the
    upsample layer to 4160x3120): one reference and four cameras of a rendered two-depth scene, 304 planes over the
    depth span `span`. The device call takes host images (box filter, levels, depth downloads included); the
    upsample is timed on device-resident buffers. The oracle runs the same pyramid on the host cores (cost
    row-parallel, aggregation serial as the reference) and its depths are compared bit for bit."""
    imgs, gray, warps, _ = scene.multiview(L16_W, L16_H, seed=7)
    p8, p2 = scene.plane_list(span[0], span[1], n8), scene.plane_list(span[0], span[1], n2)
    ctx = sgm.Context(0)
    depths, aligns = sgm.dense_pyramid(ctx, imgs, gray, warps, p8, p2)
    host_times = []
    for _ in range(reps):
        t = time.perf_counter()
        depths, aligns = sgm.dense_pyramid(ctx, imgs, gray, warps, p8, p2)
        host_times.append((time.perf_counter() - t) * 1e3)
    # images resident in HBM, depths left there: the call's wall time (it synchronises before returning)
    d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
    d_depths = [torch.empty(d.shape, dtype=torch.float32, device="cuda") for d in depths]
    shapes = [im.shape[:2] for im in imgs]
    args = ([t.data_ptr() for t in d_imgs], shapes, gray, warps, p8, p2, [t.data_ptr() for t in d_depths])
    sgm.dense_pyramid_device(ctx, *args)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        sgm.dense_pyramid_device(ctx, *args)
        times.append((time.perf_counter() - t) * 1e3)
    same = all(np.array_equal(d.cpu().numpy().view(np.uint32), h.view(np.uint32)) for d, h in zip(d_depths, depths))
    out = {"workload": f"DenseStereoPyramid.Run at {L16_W}x{L16_H} (6 levels, scales 32..1, 1 reference + 4 cameras, "
                       f"{n8} planes over z {span[0]:g}..{span[1]:g}, adaptive P2, skip masks on the last two)"
                       + (f", then the upsample layer to {2 * L16_W}x{2 * L16_H}" if upsample else "")
                       + "; rendered two-depth scene",
           "levels": [{"scale": s, "align": a, "shape": list(d.shape)} for s, a, d in zip(sgm.L16_SCALES, aligns, depths)],
           "gpu_ms_p50": round(float(np.percentile(times, 50)), 3), "gpu_ms_min": round(min(times), 3),
           "gpu_includes": "images resident in HBM: box filter, six levels (ranges, cost, aggregation, depth)",
           "host_buffers_ms_p50": round(float(np.percentile(host_times, 50)), 3),
           "device_equals_host_form": same}
    if cpu_threads:
        t = time.perf_counter()
        want = o_pyramid(imgs, gray, warps, p8, p2, cpu_threads)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value_s": round(cpu_s, 3), "cores": cpu_threads, "kind": "port",
                               "sample": "the whole pyramid at full size (the same scene and planes)"}
        out["gpu_vs_cpu"] = round(cpu_s * 1e3 / out["gpu_ms_p50"], 1)
        out["identical_to_oracle"] = all(np.array_equal(d.view(np.uint32), w["depth"].view(np.uint32))
                                         for d, w in zip(depths, want))
        for lv, w in zip(out["levels"], want):
            if w["count"] is None:   # layer 0: every pixel searches all planes
                continue
            a = w["align"]
            caps = (w["count"].astype(np.int64) + a - 1) // a * a
            maxhi = -(-int(w["maxhi"]) // a) * a
            lv["max_cap"] = int(caps.max())
            lv["spilling_cells"] = int((w["start"].astype(np.int64) - int(w["minlo"]) + caps > maxhi).sum())
    if not upsample:
        ctx.close()
        return out
    # the upsample layer on device buffers, timed with events on the stream it runs on
    guide = np.repeat(np.repeat(imgs[0], 2, 0), 2, 1)
    st = torch.cuda.Stream()
    d_prev = torch.from_numpy(depths[-1]).cuda()
    d_guide = torch.from_numpy(guide).cuda()
    d_out = torch.empty((2 * L16_H, 2 * L16_W), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(st):
        call = lambda: sgm.upsample_layer_device(ctx, d_prev.data_ptr(), L16_W, L16_H, d_guide.data_ptr(), 2 * L16_W,  # noqa: E731
                                                 2 * L16_H, L16_W, L16_H, d_out.data_ptr(), stream=st.cuda_stream)
        call()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(20):
            call()
        e1.record(st)
    torch.cuda.synchronize()
    ups_ms = e0.elapsed_time(e1) / 20
    ups_bytes = (2 * L16_W) * (2 * L16_H) * (4 + 4) + L16_W * L16_H * (4 + 4)
    ctx.close()
    out["upsample_ms"] = round(ups_ms, 4)
    out["upsample_gb_s"] = round(ups_bytes / (ups_ms / 1e3) / 1e9, 1)
    return out


def o_pyramid(imgs, gray, warps, p8, p2, threads):
    """The oracle's pyramid (CPU baseline and checker of the L16 leg)."""
    from oracle import oracle as o
    from lux_amd import sgm

    return o.dense_pyramid(imgs, gray, warps, p8, p2, gw=sgm.L16_GUIDANCE, threads=threads)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-rows", type=int, default=520, help="rows of the CPU-baseline band (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bands", type=int, default=0, help="1: fused band tiles for the vertical/diagonal paths")
    ap.add_argument("--forms", default="batch,serve,pyramid,l16,l16wide,l16spill",
                    help="extra legs: batch, serve, pyramid, l16, l16wide, l16spill (comma list; none for none)")
    ap.add_argument("--l16-span", default="default", choices=["default", "wide", "spill"],
                    help="--profile-l16: the plane span (wide: ranges up to 288 planes, four packed pairs a lane; spill: "
                         "64 planes whose ranges spill past their line-buffer segments)")
    ap.add_argument("--force-literal", action="store_true",
                    help="--profile-l16: every layer on the line-buffer-literal path (LUX_SGM_OPT_LITERAL)")
    ap.add_argument("--pyr-requests", type=int, default=2000)
    ap.add_argument("--pyr-burst", type=int, default=16)
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--serve-rate", type=float, default=40.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--profile-run", action="store_true",
                    help="for counter runs: a few requests, then a 1 GiB device copy as the byte-count calibration")
    ap.add_argument("--profile-l16", action="store_true",
                    help="for trace runs: --steps L16 pyramids (device form, images in HBM) and nothing else")
    ap.add_argument("--profile-c2f", action="store_true",
                    help="for trace runs: --steps coarse-to-fine requests (plain, host buffers) and nothing else")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic.json from two rocprofv3 --pmc runs of --profile-run (no GPU needed)")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
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

    if args.profile_c2f:
        left, right, _ = scene.make_pair(PYR_W, PYR_H, PYR_D - 1, planes=6, sigma=4.0, seed=300, bits=BITS, fg_frac=0.1)
        ctx = sgm.Context(0)
        for _ in range(args.warmup + args.steps):
            t = time.perf_counter()
            sgm.pyramid(ctx, left, right, PYR_D, PYR_RADII, bits=BITS, radius=RADIUS, p1=P1, p2=P2, tile=TILE)
            print(f"c2f request {(time.perf_counter() - t) * 1e3:.3f} ms", file=sys.stderr, flush=True)
        ctx.close()
        print(json.dumps({"profile_c2f": True, "requests": args.warmup + args.steps}))
        return

    if args.profile_l16:
        imgs, gray, warps, _ = scene.multiview(L16_W, L16_H, seed=7)
        span, n8, n2 = {"wide": (L16_SPAN_WIDE, L16_N8, L16_N2), "spill": (L16_SPILL[:2], L16_SPILL[2], L16_SPILL[3])}.get(
            args.l16_span, (L16_SPAN, L16_N8, L16_N2))
        p8, p2 = scene.plane_list(span[0], span[1], n8), scene.plane_list(span[0], span[1], n2)
        ctx = sgm.Context(0)
        if args.force_literal:
            ctx.set_option(sgm.OPT_LITERAL, 1)
        d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
        d_depths = [torch.empty((-(-L16_H // sc), -(-L16_W // sc)), dtype=torch.float32, device="cuda")
                    for sc in sgm.L16_SCALES]
        a = ([t.data_ptr() for t in d_imgs], [im.shape[:2] for im in imgs], gray, warps, p8, p2,
             [t.data_ptr() for t in d_depths])
        for _ in range(args.warmup + args.steps):
            t = time.perf_counter()
            sgm.dense_pyramid_device(ctx, *a)
            print(f"l16 pyramid {(time.perf_counter() - t) * 1e3:.3f} ms", file=sys.stderr, flush=True)
        ctx.close()
        print(json.dumps({"profile_l16": True, "pyramids": args.warmup + args.steps}))
        return

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
    for sign in "+-":  # the library marks each path of a pass; the byte-volume form runs the three in one launch
        parts = [avg.pop(f"vd{sign}{k}") for k in (1, 2, 3) if f"vd{sign}{k}" in avg]
        if parts:
            avg[f"vd{sign}"] = sum(parts)
    pv = path_bytes(D, RADIUS, P2)
    stage_bytes = STAGE_BYTES_PV if pv else STAGE_BYTES_U16
    dom = max((k for k in avg if k in stage_bytes), key=lambda k: avg[k])
    cells = W * H * D
    traffic, traffic_note = load_traffic()
    design_bpc = sum(stage_bytes.values())

    d = disp.cpu().numpy().view(np.uint16)
    v = valid.cpu().numpy() > 0
    acc = float((np.abs(d.astype(np.int64) - truth)[v] <= 1).mean()) if v.any() else 0.0

    cpu = cpu0 = None
    if not args.no_cpu and args.cpu_rows > 0:
        model, ncpu = host_info()
        threads = min(16, ncpu)
        print(f"bench: CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        val, secs, tms, all_s = cpu_baseline(left, right, args.cpu_rows, threads)
        cpu = {"value": round(val, 4), "unit": "Gcells/s", "cores": threads, "kind": "port",
               "runs_s": [round(t, 3) for t in all_s],
               "sample": (f"oracle/lux_oracle.c rectified pipeline on rows 0..{args.cpu_rows - 1} of the same "
                          f"4160x{H} pair (D=256, census 7x7): median of 5 runs after 1 warm-up, {secs:.2f} s (cost "
                          f"{tms[0] / 1e3:.2f} s on {threads} threads, aggregation {tms[1] / 1e3:.2f} s serial as "
                          f"DenseStereoSgm.cs:171, WTA {tms[2] / 1e3:.2f} s); host {model}, {ncpu} CPUs"),
               "gpu_vs_cpu": round(gcells / val, 1)}
        cpu0 = bench_c0(np, torch, sgm, scene, threads, model, ncpu)

    ctx.close()
    del dl, dr
    print("bench: HBM streaming ceilings", file=sys.stderr, flush=True)
    ceilings = hbm_ceilings(torch)
    forms = [f for f in args.forms.split(",") if f]
    batched = served = None
    if "batch" in forms:
        print("bench: batched leg", file=sys.stderr, flush=True)
        batched = bench_batched(np, torch, sgm, scene, max(2, args.steps // 4))
    if "serve" in forms:
        print("bench: served leg", file=sys.stderr, flush=True)
        served = bench_served(np, sgm, scene, args.serve_requests, args.serve_rate, 96, args.serve_slots, 16)

    pyramid = None
    if "pyramid" in forms:
        print("bench: coarse-to-fine leg", file=sys.stderr, flush=True)
        pyramid = bench_pyramid(np, sgm, scene, args.pyr_requests, args.pyr_burst, 192, args.serve_slots)

    l16 = None
    if "l16" in forms:
        print("bench: L16 dense pyramid leg", file=sys.stderr, flush=True)
        l16 = bench_l16(np, torch, sgm, scene, 5, 0 if args.no_cpu else min(16, os.cpu_count() or 1))
    l16s = None
    if "l16spill" in forms:
        print("bench: L16 dense pyramid leg, spilling padding", file=sys.stderr, flush=True)
        l16s = bench_l16(np, torch, sgm, scene, 3, 0 if args.no_cpu else min(16, os.cpu_count() or 1),
                         span=L16_SPILL[:2], upsample=False, n8=L16_SPILL[2], n2=L16_SPILL[3])
    l16w = None
    if "l16wide" in forms:
        print("bench: L16 dense pyramid leg, wide ranges", file=sys.stderr, flush=True)
        l16w = bench_l16(np, torch, sgm, scene, 3, 0 if args.no_cpu else min(16, os.cpu_count() or 1),
                         span=L16_SPAN_WIDE, upsample=False)

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
                   "paths": "byte path volumes" if pv else ("fused bands" if args.bands else "per-path sweeps"),
                   "width": W, "height": H, "disparities": D, "cells": cells, "inputs_sha256": digest},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4), "p99": round(float(np.percentile(lat, 99)), 4),
                       "form": "plain"},
        "roofline": {"bound": "hbm", "achieved": round(PATH_BYTES_PER_CELL * cells / (ms / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(PATH_BYTES_PER_CELL * cells / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic["request_bytes"] if traffic else None,
                     "kernel": "the whole request (census codes, the eight path kernels, border replays, finish)",
                     "avg_ms": round(ms, 4), "bytes_per_cell": PATH_BYTES_PER_CELL,
                     "algorithmic_bytes": PATH_BYTES_PER_CELL * cells, "traffic_source": traffic_note},
        "kernel_roofline": {"kernel": KERNEL_NAMES.get(dom, dom), "stage": dom, "avg_ms": round(avg[dom], 4),
                            "bytes_per_cell": stage_bytes[dom],
                            "achieved": round(stage_bytes[dom] * cells / (avg[dom] / 1e3) / 1e9, 1),
                            "frac": round(stage_bytes[dom] * cells / (avg[dom] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                            "traffic": (traffic["stages"].get(dom) if traffic else None),
                            "algorithmic_bytes": stage_bytes[dom] * cells},
        "design_bytes": {"form": "byte path volumes" if pv else "u16 aggregate", "bytes_per_cell": design_bpc,
                         "stages": stage_bytes,
                         "achieved": round(design_bpc * cells / (ms / 1e3) / 1e9, 1),
                         "frac": round(design_bpc * cells / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "hbm_measured": ceilings,
        "stage_ms": {k: round(v, 4) for k, v in avg.items()},
        "accuracy": {"valid_frac": round(float(v.mean()), 4), "within1_of_truth_on_valid": round(acc, 4)},
        "wall_s": round(t_wall, 3),
        "cpu_baseline": cpu,
        "cpu_baseline_c0": cpu0,
        "batched": batched,
        "served": served,
        "coarse_to_fine": pyramid,
        "l16_pyramid": l16,
        "l16_pyramid_wide_ranges": l16w,
        "l16_pyramid_spilling": l16s,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
