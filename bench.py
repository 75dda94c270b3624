"""Benchmark of the MI355X scan-alignment path (one JSON line on stdout).

Workload: CT-ICP with deskew, scan to map -- a 128-beam, 0.1 s sweep of
131,072 points (per-point stamps, 10 m/s forward, 30 deg/s yaw) against a
2,000,000-point outdoor map, 12-DoF pose + velocity, 15 Gauss-Newton
iterations, leaf policy defaults.

A step is one served batch: `--batch` concurrent alignments of that workload
(the same map resident in HBM, each request with its own scan tree and its
own initial pose / velocity) run to completion by kdh_batch_align -- one
persistent launch per Gauss-Newton iteration, replayed from a HIP graph.
value = scan points of completed alignments per second over the timed steps
(all iterations included; inputs resident in HBM before timing starts).

Beside it the line reports the plain one-alignment-per-call path (host
Gauss-Newton loop, one kernel build + host FP64 solve per iteration), the
device-resident single-alignment loop, per-alignment latency percentiles,
the roofline of the tile kernel (HIP events around each launch) and the CPU
oracle timed on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_iter(depth, ct, n_tiles):
    """Per Gauss-Newton iteration: per query 24 B moving mean+normal, 4 B dts
    (CT), 4 B leaf index, 32 B per internal node visited, 24 B fixed leaf;
    plus one partial row (27 or 90 floats) per tile."""
    per_q = 24 + (4 if ct else 0) + 4 + 24
    return float(per_q * len(depth) + 32 * int(np.sum(depth)) + n_tiles * 4 * (90 if ct else 27))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=32, help="concurrent alignments per served step")
    ap.add_argument("--plain-runs", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the alignment path shards by pair "
              "with no communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from kd_slam2_amd import (CTICPHIP3D, Context, DeviceTree, TreeGenerator3D, adapter_align, batch_align,
                              default_params, synth)

    t0 = time.time()
    w = synth.config2()
    g = TreeGenerator3D(parallel_level=3)
    fixed = g.makeTree(w["map"])
    moving = g.makeTree(w["scan"], w["stamps"])
    n_pts = len(w["scan"])
    log(f"workload ready in {time.time() - t0:.1f}s: fixed {fixed.num_nodes} nodes, moving {moving.num_leaves} leaves")

    ctx = Context(0)
    F = DeviceTree(ctx, fixed)
    M = DeviceTree(ctx, moving)
    p = default_params()
    p.max_error = 0.5
    p.max_iterations = 15
    p.velocity_damping = 10.0

    def aligner():
        al = CTICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(M)
        al.setFixedX(w["Xf"])
        al.setX(w["X0"])
        al.setVelocities(w["v0"])
        return al

    # descent depths at the initial cache (for algorithmic bytes)
    _, depth = aligner().correspondences()
    n_tiles = (moving.num_leaves + 255) // 256
    b_iter = algorithmic_bytes_per_iter(depth, 1, n_tiles)

    # ---- plain per-call path: CTICPHIP3D::compute() of the C++ plugin -----
    # (host Gauss-Newton loop, one kdh_cticp_build + host FP64 solve per
    # iteration -- the drop-in for CTICPCUDA3D)
    plain_lat = []
    for i in range(args.plain_runs + 1):
        t = time.perf_counter()
        Xp, vp, res_p, _, _ = adapter_align("CTICPHIP3D", ctx, F, M, w["X0"], w["Xf"], p, v0=w["v0"])
        dt = time.perf_counter() - t
        if i > 0:
            plain_lat.append(dt)
    plain_pps = n_pts / np.median(plain_lat)
    log(f"plain per-call: median {1e3 * np.median(plain_lat):.2f} ms/alignment, {plain_pps / 1e6:.1f} Mpts/s")

    # ---- device-resident loop, one alignment ------------------------------
    dev_lat = []
    kern_us = []
    for i in range(args.plain_runs + 1):
        al = aligner()
        t = time.perf_counter()
        al.computeOnDevice()
        dt = time.perf_counter() - t
        if i > 0:
            dev_lat.append(dt)
            kern_us.append(ctx.last_kernel_time()[0])
    dev_pps = n_pts / np.median(dev_lat)
    log(f"device loop: median {1e3 * np.median(dev_lat):.2f} ms/alignment, kernel {np.median(kern_us):.1f} us/iter")

    # ---- served batches ---------------------------------------------------
    rng = np.random.default_rng(np.random.PCG64(99))
    pairs = []
    for k in range(args.batch):
        X0 = synth.perturb(w["X_gt"], rng, 1.0, 0.3)
        pairs.append(dict(fixed=F, moving=M, ct=1, X0=X0, Xf=w["Xf"], v0=w["v0"], params=p))
    for _ in range(args.warmup):
        batch_align(ctx, pairs)
    step_t = []
    batch_kern = []
    for _ in range(args.steps):
        t = time.perf_counter()
        out = batch_align(ctx, pairs)
        step_t.append(time.perf_counter() - t)
        batch_kern.append(ctx.last_kernel_time()[0])
    total = float(np.sum(step_t))
    value = args.batch * n_pts * args.steps / total
    errs = [synth.pose_error(o["X"], w["X_gt"]) for o in out]
    log(f"served: {1e3 * np.mean(step_t):.2f} ms/step of {args.batch}, {value / 1e6:.1f} Mpts/s, "
        f"launch {np.median(batch_kern):.1f} us, max err {max(e[0] for e in errs):.4f} m")

    # roofline of the tile kernel in the served launch: one extra profiled
    # batch (no graph, HIP events around the tile kernel and the finish
    # kernel separately, on the context stream)
    ctx.set_profiling(True)
    prof_tile, prof_fin = [], []
    for _ in range(3):
        batch_align(ctx, pairs)
        tu, fu = ctx.kernel_breakdown()
        prof_tile.append(tu)
        prof_fin.append(fu)
    ctx.set_profiling(False)
    tile_us = float(np.median(prof_tile))
    achieved = args.batch * b_iter / (tile_us * 1e-6) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None)
    log(f"tile kernel {tile_us:.1f} us, finish kernel {np.median(prof_fin):.1f} us per iteration "
        f"-> {achieved:.0f} GB/s algorithmic")

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as O

        op = O.from_kdh_params(p, num_threads=8)
        t = time.perf_counter()
        n_done = 0
        while time.perf_counter() - t < args.cpu_seconds or n_done == 0:
            O.compute(fixed, moving, w["X0"], w["Xf"], op, ct=1, v0=w["v0"], fast=True)
            n_done += 1
        el = time.perf_counter() - t
        cpu = dict(value=round(n_done * n_pts / el, 1), unit="scan-pts/s", cores=8, kind="port",
                   sample=f"{n_done} full CT-ICP alignments (15 iterations) of the same workload on 8 threads "
                          f"(oracle/liboracle_fast.so, std::async chunks as ICP_CPU_), {el:.1f} s")
        log(f"cpu oracle: {cpu['value'] / 1e6:.2f} Mpts/s")

    line = dict(
        metric="scan-pts aligned/s + p99 align latency, served vs plain HIP, <=1mm/1e-3deg",
        value=round(value, 1),
        unit="scan-pts/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=round(1e3 * total / args.steps, 3),
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f32",
        data="synthetic",
        config=dict(workload="CT-ICP deskew scan-to-map, 131072-pt 128-beam sweep vs 2M-pt map, 15 iterations",
                    served_batch=args.batch, moving_leaves=int(moving.num_leaves),
                    fixed_nodes=int(fixed.num_nodes), mean_depth=round(float(np.mean(depth)), 2)),
        served_speedup_vs_plain=round(value / plain_pps, 2),
        plain_pts_per_s=round(plain_pps, 1),
        device_loop_pts_per_s=round(dev_pps, 1),
        latency_ms=dict(plain_p50=round(1e3 * float(np.percentile(plain_lat, 50)), 3),
                        plain_p99=round(1e3 * float(np.percentile(plain_lat, 99)), 3),
                        device_loop_p50=round(1e3 * float(np.percentile(dev_lat, 50)), 3),
                        device_loop_p99=round(1e3 * float(np.percentile(dev_lat, 99)), 3),
                        served_batch_p50=round(1e3 * float(np.percentile(step_t, 50)), 3),
                        served_batch_p99=round(1e3 * float(np.percentile(step_t, 99)), 3)),
        kernel_us_per_iter=dict(device_loop=round(float(np.median(kern_us)), 2),
                                served_launch=round(float(np.median(batch_kern)), 2),
                                served_tile_kernel=round(tile_us, 2),
                                served_finish_kernel=round(float(np.median(prof_fin)), 2)),
        algorithmic_bytes_per_iter=b_iter,
        roofline=roofline,
        cpu_baseline=cpu,
    )
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
