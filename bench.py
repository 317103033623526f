#!/usr/bin/env python3
"""Benchmark of the plain GPU drop-in: one depth request = census 7x7 -> cost volume -> 8-path SGM with the
reference's schedule -> WTA + subpixel + L/R, on one 4160x3120 uint16 10-bit pair, D=256 (P1=8, P2=96).

Prints ONE JSON line. `value` = Gcells/s = W*H*D*8 paths / time per request, inputs resident in HBM.
`roofline` = the dominant kernel's algorithmic HBM bytes per launch / its average duration (HIP events on
the stream the kernels run on). `cpu_baseline` = the CPU oracle (the reference's algorithm restated in C:
row-parallel cost, serial aggregation) on a bounded band of the same workload, on this host.

Two more legs go into the same line:
  `batched`: one capture of 10 pairs (6 x 2080x1560 D=128, 4 x 3120x2340 D=192, independent seeds) as one graph
             launch with the autotuned schedule, against the same pairs run one after another (plain drop-in);
  `served`:  a mixed stream of depth requests (sizes uniform over 640x480 D64, 1040x780 D96, 2080x1560 D128,
             4160x3120 D256; host images in, host results out) through the request loop under a 96 GB budget:
             latency p50/p99 at Poisson arrivals of 40 req/s, and throughput at saturation against the plain
             drop-in running the same requests in order;
  `coarse_to_fine`: 3-level requests (1040x780 D=64 -> 2080x1560 +-16 -> 4160x3120 +-8) in bursts of 16 through the
             request loop under a 192 GB budget, latency p50/p99 against plain sequential execution.
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
    ctx.close()
    del keep
    return {"workload": "6 x 2080x1560 D=128 + 4 x 3120x2340 D=192, uint16 10-bit, census 7x7, seeds 100..109",
            "cells": cells, "ms_per_batch": round(ms_batch, 3), "gcells_s": round(cells * 8 / ms_batch / 1e6, 3),
            "plain_ms": round(ms_plain, 3), "plain_gcells_s": round(cells * 8 / ms_plain / 1e6, 3),
            "speedup_vs_plain": round(ms_plain / ms_batch, 3), "schedule": best,
            "candidates": tried, "identical_to_plain": same}


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
    ctx.close()

    srv = sgm.Server(0, int(budget_gb * 1e9), slots=slots)

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

    for k in scenes:  # warm every slot's workspace for every size
        run_w = [srv.submit(params[k], *scenes[k][0]) for _ in range(slots)]
        for t in run_w:
            srv.wait(t)
    sat_s, sat_stats = run(None)
    poi_s, poi_stats = run(np.cumsum(gaps))
    info = srv.info()
    srv.close()
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
                      "p50_ms": round(float(np.percentile(plain_lat, 50)), 3),
                      "p99_ms": round(float(np.percentile(plain_lat, 99)), 3)},
            "served_vs_plain_throughput": round(plain_s / sat_s, 3),
            "bytes_peak_gb": round(info["bytes_peak"] / 1e9, 3), "failed": info["failed"]}


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
    for t in [srv.submit_pyramid(pp, *scenes[0]) for _ in range(slots)]:  # warm every slot
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


def bench_l16(np, torch, sgm, scene, reps, cpu_threads, span=L16_SPAN, upsample=True):
    """The L16 dense stereo end to end (DenseStereoPyramid.Run, six levels from 65x49 to 2080x1560, then the
    upsample layer to 4160x3120): one reference and four cameras of a rendered two-depth scene, 304 planes over the
    depth span `span`. The device call takes host images (box filter, levels, depth downloads included); the
    upsample is timed on device-resident buffers. The oracle runs the same pyramid on the host cores (cost
    row-parallel, aggregation serial as the reference) and its depths are compared bit for bit."""
    imgs, gray, warps, _ = scene.multiview(L16_W, L16_H, seed=7)
    p8, p2 = scene.plane_list(span[0], span[1], L16_N8), scene.plane_list(span[0], span[1], L16_N2)
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
                       f"{L16_N8} planes over z {span[0]:g}..{span[1]:g}, adaptive P2, skip masks on the last two)"
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
    ap.add_argument("--cpu-rows", type=int, default=1560, help="rows of the CPU-baseline band (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bands", type=int, default=0, help="1: fused band tiles for the vertical/diagonal paths")
    ap.add_argument("--forms", default="batch,serve,pyramid,l16,l16wide",
                    help="extra legs: batch, serve, pyramid, l16, l16wide (comma list; '' for none)")
    ap.add_argument("--l16-span", default="default", choices=["default", "wide"],
                    help="--profile-l16: the plane span (wide: ranges up to 288 planes, four packed pairs a lane)")
    ap.add_argument("--force-literal", action="store_true",
                    help="--profile-l16: every layer on the line-buffer-literal path (LUX_SGM_OPT_LITERAL)")
    ap.add_argument("--pyr-requests", type=int, default=128)
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
        span = L16_SPAN_WIDE if args.l16_span == "wide" else L16_SPAN
        p8, p2 = scene.plane_list(span[0], span[1], L16_N8), scene.plane_list(span[0], span[1], L16_N2)
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

    ctx.close()
    del dl, dr
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
        "batched": batched,
        "served": served,
        "coarse_to_fine": pyramid,
        "l16_pyramid": l16,
        "l16_pyramid_wide_ranges": l16w,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
