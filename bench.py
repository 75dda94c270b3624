"""Benchmark of the MI355X scan-alignment path (one JSON line on stdout).

Headline workload (`value`): CT-ICP with deskew, scan to map -- a 128-beam,
0.1 s sweep of 131,072 points (per-point stamps, 10 m/s forward, 30 deg/s
yaw) against a 2,000,000-point outdoor map, 12-DoF pose + velocity, 15
Gauss-Newton iterations, leaf policy defaults.

A step is one served batch: `--batch` concurrent alignments of that workload
(the map resident in HBM, each alignment with its own initial pose and
velocity) run to completion by kdh_batch_align -- one tile launch and one
finish launch per Gauss-Newton iteration, all iterations in one HIP graph.
value = scan points of completed alignments per second over the timed steps
(all iterations included; inputs resident in HBM before timing starts).

Beside it the line reports
  * the plain one-alignment-per-call path (CTICPHIP3D::compute from the C++
    plugin: host Gauss-Newton loop, one kernel build + host FP64 solve per
    iteration) and the device-resident single-alignment loop;
  * the serving mix: 4 odometry sequences at 10 Hz (64k-point CT sweeps
    against their own resident 4M-point maps) and loop-closure requests at
    20/s (rigid, 8k-200k points, map 3x), Poisson arrivals into the
    continuous-batching server with a 64 GB budget, p50/p99 latency from
    arrival to host-visible result;
  * the tile kernel's roofline (HIP events on the context stream) and its
    HBM traffic from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE)
    scaled by the same counters on a device copy of known size;
  * the precision stress: a 2M-point scan against an 8M-point map with the
    world frame 10 km away -- naive FP32, re-centred FP32 and the reference's
    FP32 form against the oracle's FP64 twin and FP64 loop;
  * the CPU oracle (restated ICP_CPU_ / CTICP_CPU_) on 8 host threads.
"""
import argparse
import csv
import ctypes
import gc
import glob
import json
import multiprocessing as mp
import os
import shutil
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_per_iter(depth, ct, n_tiles):
    """Per Gauss-Newton iteration: per query 24 B moving mean+normal, 4 B dts
    (CT), 4 B leaf index, 32 B per internal node visited, 24 B fixed leaf;
    plus one partial row (27 or 90 floats) per tile."""
    per_q = 24 + (4 if ct else 0) + 4 + 24
    return float(per_q * len(depth) + 32 * int(np.sum(depth)) + n_tiles * 4 * (90 if ct else 27))


# ---------------------------------------------------------------------------
# workload generation (host only; runs in worker processes before any GPU use)


def _gen_cfg2(_):
    from kd_slam2_amd import TreeGenerator3D, synth

    w = synth.config2()
    g = TreeGenerator3D(parallel_level=2)
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"], w["stamps"])
    return w


def _gen_sweep5(_):
    """One 64k-point odometry sweep of the serving mix (sequence 0's
    generator; its map is not needed here)."""
    from kd_slam2_amd import synth

    f = synth.config5_sequence(0, n_map=1000, n_frames=1)["frames"][0]
    return dict(scan=f["scan"], stamps=f["stamps"])


def _gen_seq(s):
    from kd_slam2_amd import TreeGenerator3D, synth

    q = synth.config5_sequence(s)
    g = TreeGenerator3D(parallel_level=2)
    q["fixed"] = g.makeTree(q["map"])
    del q["map"]
    for f in q["frames"]:
        f["moving"] = g.makeTree(f["scan"], f["stamps"])
        f["n_pts"] = len(f["scan"])
        del f["scan"], f["stamps"]
    return q


def _gen_pair3(i_n):
    """Config 3 pair i: config-1 generator, seed 3000+i, scan n, map 3n."""
    from kd_slam2_amd import TreeGenerator3D, synth

    i, n = i_n
    w = synth.config3_pair(i, n)
    g = TreeGenerator3D(leaf_max_points=16)
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"])
    w["n_pts"] = len(w["scan"])
    del w["map"], w["scan"], w["stamps"]
    return w


def _gen_loop(j):
    from kd_slam2_amd import TreeGenerator3D, synth

    w = synth.config5_loop_pair(j)
    g = TreeGenerator3D(leaf_max_points=16)
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"])
    w["n_pts"] = len(w["scan"])
    del w["map"], w["scan"]
    return w


def _gen_cfg4_block(kb):
    """Precision stress: block b of the 2M-point scan (kind 0) or of the
    8M-point map (kind 1)."""
    from kd_slam2_amd import synth

    kind, b = kb
    return synth.config4_scan_block(b) if kind == 0 else synth.config4_map_block(b)


# ---------------------------------------------------------------------------
# PMC traffic: rocprofv3 --pmc passes over a child run of this script


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def pmc_child(args):
    """One served batch of the headline workload plus a device copy of
    COPY_BYTES bytes, for the counter passes."""
    from kd_slam2_amd import Context, DeviceTree, batch_align, default_params, synth

    w = _gen_cfg2(0)
    ctx = Context(0)
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = _cfg2_params(default_params)
    rng = np.random.default_rng(np.random.PCG64(99))
    pairs = [dict(fixed=F, moving=M, ct=1, X0=synth.perturb(w["X_gt"], rng, 1.0, 0.3), Xf=w["Xf"], v0=w["v0"],
                  params=p) for _ in range(args.batch)]
    batch_align(ctx, pairs)
    hip = _hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipMalloc(ctypes.byref(a), COPY_BYTES)
    hip.hipMalloc(ctypes.byref(b), COPY_BYTES)
    hip.hipMemcpy(b, a, COPY_BYTES, 3)  # device to device
    hip.hipFree(a)
    hip.hipFree(b)
    print("pmc child done", flush=True)


def pmc_pass(counter, batch, tiles_per_alignment):
    """Per-dispatch counter (KiB) for the served tile kernel and for the copy
    kernel, from one rocprofv3 --pmc run.  Returns (tile, copy) or None."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    out = tempfile.mkdtemp(prefix="kdh_pmc_")
    cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", counter, "-d", out, "-o", "pmc", "--output-format", "csv",
           "--", sys.executable, os.path.abspath(__file__), "--pmc-child", "--batch", str(batch)]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=170, cwd=ROOT,
                           env=dict(os.environ, TMPDIR="/tmp"))
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            log(f"pmc {counter}: no counter file (rc {r.returncode})")
            return None
        tile, copy = [], []
        with open(files[0]) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                grid = int(row.get("Grid_Size", row.get("Grid_Size_X", "0")) or 0)
                v = float(row["Counter_Value"])
                if "tiles_kernel" in name and grid == 256 * batch * tiles_per_alignment:
                    tile.append(v)
                elif "copyBuffer" in name and grid > 1 << 16:
                    copy.append(v)
        if not tile or not copy:
            log(f"pmc {counter}: tile {len(tile)} copy {len(copy)} dispatches")
            return None
        return float(np.median(tile)), float(max(copy))
    except Exception as e:  # counters are reported as null when unavailable
        log(f"pmc {counter}: {e}")
        return None
    finally:
        shutil.rmtree(out, ignore_errors=True)


def _cfg2_params(default_params):
    p = default_params()
    p.max_error = 0.5
    p.max_iterations = 15
    p.velocity_damping = 10.0
    return p


# ---------------------------------------------------------------------------


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=32, help="concurrent alignments per served step")
    ap.add_argument("--plain-runs", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="length of the serving-mix arrival window")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--pairs3", type=int, default=256, help="loop-closure pairs of the batched config (0: skip)")
    ap.add_argument("--no-pmc", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the 10 km precision-stress workload")
    ap.add_argument("--no-tree-build", action="store_true", help="skip the tree-generation timings")
    ap.add_argument("--no-next-rows", action="store_true", help="skip tree maintenance, factor forms, KDTree3f")
    ap.add_argument("--tree-build-only", action="store_true",
                    help="only the tree-generation timings (for kernel traces of the builder)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); alignments shard by pair with no "
              "communication across GPUs: run one process per GPU, each on its own share of the pairs", file=sys.stderr)
        sys.exit(2)

    if args.tree_build_only:
        from kd_slam2_amd import Context, DeviceTree, TreeGenerator3D, default_params, synth

        w = synth.config2()
        sw5 = _gen_sweep5(0)
        ctx = Context(0)
        F = TreeGenerator3D().makeTreeOnDevice(ctx, w["map"], download=False)
        print(json.dumps(run_tree_build(ctx, w, sw5, DeviceTree, F, _cfg2_params(default_params))), flush=True)
        return

    t0 = time.time()
    n_loop_pool = 16
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 4)) as pool:
        j_cfg2 = pool.apply_async(_gen_cfg2, (0,))
        j_sw5 = None if args.no_tree_build else pool.apply_async(_gen_sweep5, (0,))
        j_seq = None if args.no_serve else pool.map_async(_gen_seq, range(4))
        j_loop = None if args.no_serve else pool.map_async(_gen_loop, range(n_loop_pool))
        j_p3 = None
        if args.pairs3:
            from kd_slam2_amd import synth as _s

            sizes = _s.config3_sizes(args.pairs3)
            j_p3 = pool.map_async(_gen_pair3, list(enumerate(sizes)))
        j_c4 = None if args.no_precision else pool.map_async(_gen_cfg4_block, [(k, b) for k in (0, 1)
                                                                                for b in range(8)])
        w = j_cfg2.get()
        sweep5 = j_sw5.get() if j_sw5 else None
        c4_blocks = j_c4.get() if j_c4 else None
        seqs = j_seq.get() if j_seq else []
        loops = j_loop.get() if j_loop else []
        pairs3 = j_p3.get() if j_p3 else []
    w4 = None
    if c4_blocks is not None:
        from kd_slam2_amd import TreeGenerator3D, synth as _s

        scan4 = np.concatenate(c4_blocks[:8])
        g4 = TreeGenerator3D(parallel_level=3)
        X_gt4, X04, Xf4 = _s.config4_poses()
        w4 = dict(n_pts=len(scan4), moving=g4.makeTree(scan4), fixed=g4.makeTree(np.concatenate(c4_blocks[8:])),
                  X_gt=X_gt4, X0=X04, Xf=Xf4)
        del scan4, c4_blocks
    fixed, moving = w["fixed"], w["moving"]
    n_pts = 131_072
    log(f"workloads ready in {time.time() - t0:.1f}s: fixed {fixed.num_nodes} nodes, moving {moving.num_leaves} "
        f"leaves")

    from kd_slam2_amd import (CTICPHIP3D, Context, DeviceTree, Server, adapter_align, batch_align,
                              default_params, synth)

    ctx = Context(0)
    F = DeviceTree(ctx, fixed)
    M = DeviceTree(ctx, moving)
    p = _cfg2_params(default_params)

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
    plain_lat = []
    gc.collect()
    gc.disable()
    for i in range(args.plain_runs + 1):
        t = time.perf_counter()
        adapter_align("CTICPHIP3D", ctx, F, M, w["X0"], w["Xf"], p, v0=w["v0"])
        dt = time.perf_counter() - t
        if i > 0:
            plain_lat.append(dt)
    gc.enable()
    plain_pps = n_pts / np.median(plain_lat)
    log(f"plain per-call: median {1e3 * np.median(plain_lat):.2f} ms/alignment, {plain_pps / 1e6:.1f} Mpts/s")

    # ---- device-resident loop, one alignment ------------------------------
    dev_lat, kern_us = [], []
    gc.disable()
    for i in range(args.plain_runs + 1):
        al = aligner()
        t = time.perf_counter()
        al.computeOnDevice()
        dt = time.perf_counter() - t
        if i > 0:
            dev_lat.append(dt)
            kern_us.append(ctx.last_kernel_time()[0])
    gc.enable()
    dev_pps = n_pts / np.median(dev_lat)
    log(f"device loop: median {1e3 * np.median(dev_lat):.2f} ms/alignment, {np.median(kern_us):.1f} us/iter")

    # ---- served batches (headline) ----------------------------------------
    rng = np.random.default_rng(np.random.PCG64(99))
    pairs = [dict(fixed=F, moving=M, ct=1, X0=synth.perturb(w["X_gt"], rng, 1.0, 0.3), Xf=w["Xf"], v0=w["v0"],
                  params=p) for _ in range(args.batch)]
    for _ in range(args.warmup):
        batch_align(ctx, pairs)
    # The Python collector stays off inside timed regions: a full collection
    # of the workload's large object graph would be timed as kernel work.
    gc.collect()
    gc.disable()
    step_t, batch_kern = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        out = batch_align(ctx, pairs)
        step_t.append(time.perf_counter() - t)
        batch_kern.append(ctx.last_kernel_time()[0])
    gc.enable()
    total = float(np.sum(step_t))
    value = args.batch * n_pts * args.steps / total
    errs = [synth.pose_error(o["X"], w["X_gt"]) for o in out]
    slow = [f"{i}:{1e3 * x:.2f}" for i, x in enumerate(step_t) if x > 3 * np.median(step_t)]
    log(f"served steps: median {1e3 * np.median(step_t):.3f} ms, slow (>3x median) {slow}")
    log(f"served: {1e3 * np.mean(step_t):.2f} ms/step of {args.batch}, {value / 1e6:.1f} Mpts/s, "
        f"{np.median(batch_kern):.1f} us/iter, max err {max(e[0] for e in errs):.4f} m")

    # the same served step at other batch sizes (not the headline)
    sweep = {}
    for nb in (1, 8, 16, 64, 128):
        pb = [dict(pairs[i % len(pairs)]) for i in range(nb)]
        batch_align(ctx, pb)
        gc.disable()
        ts = []
        for _ in range(5):
            t = time.perf_counter()
            batch_align(ctx, pb)
            ts.append(time.perf_counter() - t)
        gc.enable()
        sweep[str(nb)] = dict(ms_per_step=round(1e3 * float(np.median(ts)), 3),
                              pts_per_s=round(nb * n_pts / float(np.median(ts)), 1))
    sweep[str(args.batch)] = dict(ms_per_step=round(1e3 * float(np.median(step_t)), 3),
                                  pts_per_s=round(args.batch * n_pts / float(np.median(step_t)), 1))
    log("batch sweep: " + ", ".join(f"{k}: {v['pts_per_s'] / 1e6:.0f} Mpts/s" for k, v in sweep.items()))

    # roofline of the tile kernel in the served launch (profiled batches: no
    # graph, HIP events around 4 back-to-back tile launches and the finish)
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
    log(f"tile kernel {tile_us:.1f} us, finish kernel {np.median(prof_fin):.1f} us per iteration "
        f"-> {achieved:.0f} GB/s algorithmic")

    # ---- serving mix --------------------------------------------------------
    serving = None
    if not args.no_serve:
        serving = run_serving(ctx, seqs, loops, args.serve_seconds, DeviceTree, Server, default_params)

    # ---- batched loop-closure verification (config 3) ----------------------
    batch3 = None
    if pairs3:
        batch3 = run_batch3(ctx, pairs3, DeviceTree, batch_align, adapter_align, default_params, synth)

    # ---- precision stress (config 4) ---------------------------------------
    precision = None
    if w4 is not None:
        precision = run_precision(ctx, w4, DeviceTree, default_params, synth)

    # ---- tree generation on the device (makeTree + upload) ------------------
    tree_build = None
    if not args.no_tree_build:
        tree_build = run_tree_build(ctx, w, sweep5, DeviceTree, F, p)
    w.pop("map", None)

    # ---- the other rows around the path: tree maintenance, batched factor
    # forms, the KDTree3f wire and bag format -----------------------------
    next_rows = None
    if not args.no_next_rows:
        next_rows = run_next_rows(ctx, w, DeviceTree, F, M, p)

    # ---- CPU baseline -------------------------------------------------------
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
                   sample=f"{n_done} full CT-ICP alignments (15 iterations) of the headline workload on 8 threads "
                          f"(oracle/liboracle_fast.so, std::async chunks as ICP_CPU_), {el:.1f} s")
        log(f"cpu oracle: {cpu['value'] / 1e6:.2f} Mpts/s")

    # ---- HBM traffic of the tile kernel -------------------------------------
    traffic = None
    pmc = {}
    if not args.no_pmc:
        f = pmc_pass("FETCH_SIZE", args.batch, n_tiles)
        wr = pmc_pass("WRITE_SIZE", args.batch, n_tiles)
        if f and wr:
            kib = COPY_BYTES / 1024.0
            f_scale = kib / f[1]
            w_scale = kib / wr[1]
            traffic = round((f[0] * f_scale + wr[0] * w_scale) * 1024.0, 1)
            pmc = dict(fetch_kib=round(f[0], 1), write_kib=round(wr[0], 1), copy_fetch_kib=round(f[1], 1),
                       copy_write_kib=round(wr[1], 1), fetch_scale=round(f_scale, 4), write_scale=round(w_scale, 4))
            log(f"pmc: tile fetch {f[0]:.0f} KiB, write {wr[0]:.0f} KiB per launch (copy {f[1]:.0f}/{wr[1]:.0f} "
                f"KiB for {kib:.0f} KiB)")

    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic)
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
        served_batch_sweep=sweep,
        plain_pts_per_s=round(plain_pps, 1),
        device_loop_pts_per_s=round(dev_pps, 1),
        latency_ms=dict(plain_p50=round(1e3 * float(np.percentile(plain_lat, 50)), 3),
                        plain_p99=round(1e3 * float(np.percentile(plain_lat, 99)), 3),
                        device_loop_p50=round(1e3 * float(np.percentile(dev_lat, 50)), 3),
                        device_loop_p99=round(1e3 * float(np.percentile(dev_lat, 99)), 3),
                        served_batch_p50=round(1e3 * float(np.percentile(step_t, 50)), 3),
                        served_batch_p99=round(1e3 * float(np.percentile(step_t, 99)), 3)),
        kernel_us_per_iter=dict(device_loop=round(float(np.median(kern_us)), 2),
                                served_iteration=round(float(np.median(batch_kern)), 2),
                                served_tile_kernel=round(tile_us, 2),
                                served_finish_kernel=round(float(np.median(prof_fin)), 2)),
        algorithmic_bytes_per_iter=b_iter,
        pmc=pmc or None,
        serving_mix=serving,
        loop_closure_batch=batch3,
        precision_stress=precision,
        tree_build=tree_build,
        next_rows=next_rows,
        roofline=roofline,
        cpu_baseline=cpu,
    )
    print(json.dumps(line), flush=True)


def run_tree_build(ctx, w, sweep5, DeviceTree, fixed_tree=None, params=None):
    """makeTree + upload on the device (kdh_tree_build, bit-identical to the
    host builder) against the host builder single-threaded and with 16
    threads plus kdh_tree_upload: wall time from host points to a resident
    tree.  Generator parameters as the workloads (leaf_max_points 30)."""
    from kd_slam2_amd import TreeGenerator3D

    g = TreeGenerator3D()
    g16 = TreeGenerator3D(parallel_level=4)
    cases = []
    if sweep5 is not None:
        cases.append(("sweep_64k", sweep5["scan"], sweep5["stamps"], 30, 3))
    cases.append(("sweep_131k", w["scan"], w["stamps"], 30, 3))
    cases.append(("map_2M", w["map"], None, 10, 1))
    out = {}
    for name, pts, st, reps, host_reps in cases:
        for _ in range(3):
            g.makeTreeOnDevice(ctx, pts, st, download=False).close()
        gc.disable()
        dev = []
        for _ in range(reps):
            t = time.perf_counter()
            tr = g.makeTreeOnDevice(ctx, pts, st, download=False)
            dev.append(time.perf_counter() - t)
            tr.close()
        devin = _build_from_hbm(ctx, g, pts, st, reps)
        host1, host16 = [], []
        for _ in range(host_reps):
            t = time.perf_counter()
            DeviceTree(ctx, g.makeTree(pts, st)).close()
            host1.append(time.perf_counter() - t)
            t = time.perf_counter()
            DeviceTree(ctx, g16.makeTree(pts, st)).close()
            host16.append(time.perf_counter() - t)
        gc.enable()
        out[name] = dict(points=len(pts), device_ms_p50=round(1e3 * float(np.median(dev)), 3),
                         device_ms_p99=round(1e3 * float(np.percentile(dev, 99)), 3),
                         device_from_hbm_ms_p50=round(1e3 * float(np.median(devin)), 3),
                         host_1thread_ms=round(1e3 * float(np.median(host1)), 2),
                         host_16thread_ms=round(1e3 * float(np.median(host16)), 2))
        log(f"tree build {name}: device {out[name]['device_ms_p50']:.2f} ms ({out[name]['device_from_hbm_ms_p50']:.2f} "
            f"ms from points in HBM), host 1 thread "
            f"{out[name]['host_1thread_ms']:.1f} ms, 16 threads {out[name]['host_16thread_ms']:.1f} ms")
    if fixed_tree is not None:
        out["odometry_from_points"] = run_points_to_pose(ctx, w, fixed_tree, params, g, g16, DeviceTree)
    return out


def _build_from_hbm(ctx, g, pts, st, reps):
    """Wall times of kdh_tree_build over points (and stamps) already in HBM
    (inputs_on_device = 1): the device build without the host-to-device copy
    of the cloud."""
    import ctypes as C

    from kd_slam2_amd import _lib as L

    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    x = np.ascontiguousarray(pts, dtype=np.float32)
    t = None if st is None else np.ascontiguousarray(st, dtype=np.float64)
    dx, dt = C.c_void_p(), C.c_void_p()
    out = []
    try:
        if hip.hipMalloc(C.byref(dx), x.nbytes) != 0 or hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, 1) != 0:
            raise RuntimeError("device copy of the points failed")
        if t is not None and (hip.hipMalloc(C.byref(dt), t.nbytes) != 0 or
                              hip.hipMemcpy(dt, t.ctypes.data, t.nbytes, 1) != 0):
            raise RuntimeError("device copy of the stamps failed")
        tp = None if t is None else C.cast(dt, L.F64P)
        for k in range(reps + 3):
            h = C.c_void_p()
            t0 = time.perf_counter()
            ctx.check(L.lib().kdh_tree_build(ctx.h, C.cast(dx, L.F32P), tp, len(x), C.byref(g._p), 1, C.byref(h)),
                      "kdh_tree_build")
            if k >= 3:
                out.append(time.perf_counter() - t0)
            if h.value:
                L.lib().kdh_tree_free(ctx.h, h)
    finally:
        if dx.value:
            hip.hipFree(dx)
        if dt.value:
            hip.hipFree(dt)
    return out


def _median_us(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return 1e6 * float(np.median(ts))


def run_next_rows(ctx, w, DeviceTree, F, M, params):
    """Each row beside the alignment path, measured where it runs:
    - tree maintenance (CT deskew of a whole tree, isometry, compressed-leaf
      extraction) on copies of the 2M-point map and the 131k sweep trees:
      wall time per call (launch, kernel, wait; extraction includes the
      download) and the node-record bytes it moves against HBM peak;
    - quadratic forms of bundle-adjustment factors: 32 CT factors (sweep vs
      map at perturbed poses) and 32 CT dual factors (sweep vs sweep) in one
      batched launch against one plain per-call build each;
    - the KDTree3f codec and MCAP file I/O on the map tree's message."""
    from kd_slam2_amd import batch_build, batch_build_dual, kdtree_msg, synth
    from kd_slam2_amd.api import adapter_build_dual

    out = {}
    v = np.array([2.0, 0.1, 0.0, 0.01, 0.02, 0.5])
    X = synth.expmap(np.array([0.5, -0.2, 0.1, 0.01, 0.02, 0.3]))
    maint = {}
    for name, host in (("map_2M", w["fixed"]), ("sweep_131k", w["moving"])):
        T = DeviceTree(ctx, host)
        nn, nl = len(host.nodes), len(host.leaves)
        rec = 32 * (nn + nl)  # node and leaf records, read and written
        r = dict(nodes=nn, leaves=nl)
        for op, fn, bytes_moved in (
                ("apply_velocities", lambda: T.applyVelocitiesInPlace(v), 2 * rec + 4 * nn),
                ("apply_isometry", lambda: T.applyIsometryInPlace(X), 2 * rec),
                ("extract_leaves", lambda: T.extractCompressedLeaves(v), 32 * nl + 24 * nl)):
            us = _median_us(fn, 30)
            r[op + "_us"] = round(us, 1)
            r[op + "_GBps"] = round(bytes_moved / (us * 1e-6) / 1e9, 1)
        T.close()
        maint[name] = r
        log(f"tree maintenance {name}: " + ", ".join(f"{k} {r[k]}" for k in r if k.endswith("_us")))
    out["tree_maintenance"] = maint

    rng = np.random.default_rng(np.random.PCG64(5))
    n_f = 32
    pairs = [dict(fixed=F, moving=M, ct=1, X0=synth.perturb(w["X_gt"], rng, 0.5, 0.1), Xf=w["Xf"], v0=w["v0"],
                  params=params) for _ in range(n_f)]
    batch_us = _median_us(lambda: batch_build(ctx, pairs), 10)

    def one_by_one():
        for p_ in pairs:
            batch_build(ctx, [p_])
    plain_us = _median_us(one_by_one, 3)
    # factors between the stamped sweep and itself at a perturbed pose: the
    # shape of a bundle-adjustment factor between consecutive sweeps
    facs = [dict(fixed=M, moving=M, X_moving=synth.perturb(w["X_gt"], rng, 0.3, 0.05), v_moving=w["v0"],
                 X_fixed=w["X_gt"], v_fixed=w["v0"], params=params) for _ in range(n_f)]
    dual_us = _median_us(lambda: batch_build_dual(ctx, facs), 10)

    def dual_one_by_one():
        for f_ in facs:
            adapter_build_dual(ctx, f_["fixed"], f_["moving"], f_["X_moving"], f_["v_moving"], f_["X_fixed"],
                               f_["v_fixed"], params)
    dual_plain_us = _median_us(dual_one_by_one, 3)
    out["factor_forms"] = dict(
        factors=n_f, ct_batched_ms=round(batch_us / 1e3, 3), ct_one_by_one_ms=round(plain_us / 1e3, 3),
        ct_speedup=round(plain_us / batch_us, 2), dual_batched_ms=round(dual_us / 1e3, 3),
        dual_one_by_one_ms=round(dual_plain_us / 1e3, 3), dual_speedup=round(dual_plain_us / dual_us, 2))
    log(f"factor forms: {n_f} CT {batch_us / 1e3:.2f} ms batched vs {plain_us / 1e3:.2f} ms one by one; dual "
        f"{dual_us / 1e3:.2f} ms vs {dual_plain_us / 1e3:.2f} ms")

    msg = kdtree_msg.message_from_tree(w["fixed"], frame_id="map", src_topic="/tree")
    data = kdtree_msg.encode(msg)
    enc_us = _median_us(lambda: kdtree_msg.encode(msg), 10)
    dec_us = _median_us(lambda: kdtree_msg.decode(data), 10)
    tmp = tempfile.mkdtemp()
    try:
        path = os.path.join(tmp, "trees.mcap")
        recs = [("/tree", 1_000_000_000 + k, data) for k in range(8)]
        wr_us = _median_us(lambda: kdtree_msg.write_mcap(path, recs), 3)
        rd_us = _median_us(lambda: kdtree_msg.read_mcap(path), 3)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    mb = len(data) / 1e6
    out["kdtree3f"] = dict(message_MB=round(mb, 2), nodes=len(msg.nodes), encode_ms=round(enc_us / 1e3, 2),
                           decode_ms=round(dec_us / 1e3, 2), encode_MBps=round(mb / (enc_us * 1e-6), 1),
                           decode_MBps=round(mb / (dec_us * 1e-6), 1), mcap_write_8_ms=round(wr_us / 1e3, 2),
                           mcap_read_8_ms=round(rd_us / 1e3, 2))
    log(f"kdtree3f: {mb:.1f} MB message, encode {enc_us / 1e3:.2f} ms, decode {dec_us / 1e3:.2f} ms, MCAP 8 "
        f"messages write {wr_us / 1e3:.1f} ms, read {rd_us / 1e3:.1f} ms")
    return out


def run_points_to_pose(ctx, w, F, params, g, g16, DeviceTree):
    """One CT odometry step from the raw 131k-point sweep to the aligned pose
    against the resident 2M-point map: moving tree built on the device and
    the device-resident Gauss-Newton loop, against the host builder (1 and 16
    threads) + upload + the same loop."""
    from kd_slam2_amd import CTICPHIP3D

    def align(M):
        al = CTICPHIP3D(ctx)
        al.params = params
        al.setFixed(F)
        al.setMoving(M)
        al.setFixedX(w["Xf"])
        al.setX(w["X0"])
        al.setVelocities(w["v0"])
        al.computeOnDevice()
        X = np.array(al.getX(), dtype=np.float64)
        return X

    res, poses = {}, {}
    for name, make in (("device_build", lambda: g.makeTreeOnDevice(ctx, w["scan"], w["stamps"], download=False)),
                       ("host_1thread_build", lambda: DeviceTree(ctx, g.makeTree(w["scan"], w["stamps"]))),
                       ("host_16thread_build", lambda: DeviceTree(ctx, g16.makeTree(w["scan"], w["stamps"])))):
        reps = 20 if name == "device_build" else 5
        ts, X = [], None
        for i in range(reps + 2):
            t = time.perf_counter()
            M = make()
            X = align(M)
            if i >= 2:
                ts.append(time.perf_counter() - t)
            M.close()
        res[name + "_ms_p50"] = round(1e3 * float(np.median(ts)), 3)
        poses[name] = X
    # the trees are bit-identical, so the three routes end on the same pose
    res["pose_max_abs_diff"] = float(max(np.abs(poses[k] - poses["device_build"]).max() for k in poses))
    log(f"odometry from points: device build {res['device_build_ms_p50']:.2f} ms, host build "
        f"{res['host_1thread_build_ms_p50']:.2f} ms (1 thread) / {res['host_16thread_build_ms_p50']:.2f} ms (16), "
        f"pose max diff {res['pose_max_abs_diff']:.3g}")
    return res


def run_precision(ctx, w4, DeviceTree, default_params, synth):
    """Precision stress: rigid ICP of a 2M-point, 128-beam scan (ranges up to
    200 m) against an 8M-point map, the world frame 10 km away.  At the
    initial FP64 state: H/b relative error and Gauss-Newton step error of
    naive FP32 (the reference's world-coordinate formulas), re-centred FP32
    and the reference's FP32 CPU form, each against the oracle's FP64 twin.
    After 10 iterations: pose error of each loop against the oracle's FP64
    loop, and against the generating pose."""
    from kd_slam2_amd import ICPHIP3D
    from oracle import pyoracle as O

    F, M = DeviceTree(ctx, w4["fixed"]), DeviceTree(ctx, w4["moving"])
    p = default_params()
    p.max_error = 0.5
    p.max_iterations = 10
    op = O.from_kdh_params(p, num_threads=8)

    def step(H, b):
        return np.linalg.solve(np.asarray(H, dtype=np.float64) + p.damping * np.eye(6), -np.asarray(b, np.float64))

    def sdiff(dx, ref):
        return [float(np.linalg.norm(dx[:3] - ref[:3])), float(np.degrees(np.linalg.norm(dx[3:] - ref[3:])))]

    def rel(A, B):
        return float(np.abs(np.asarray(A, np.float64) - B).max() / np.abs(B).max())

    ref = O.build_form_at(w4["fixed"], w4["moving"], w4["X0"], w4["Xf"], op, fast=True)
    dx64 = step(ref["H64"], ref["b64"])
    forms, poses, lat = {}, {}, {}
    for rc, name in ((0, "naive_fp32"), (1, "recentred_fp32")):
        pr = default_params()
        pr.max_error, pr.max_iterations, pr.recenter = 0.5, 10, rc

        def mk():
            al = ICPHIP3D(ctx)
            al.params = pr
            al.setFixed(F)
            al.setMoving(M)
            al.setX(w4["X0"])
            al.setFixedX(w4["Xf"])
            return al

        al = mk()
        al.buildQuadraticFormAt()
        forms[name] = dict(H_rel=rel(al.H, ref["H64"]), b_rel=rel(al.b, ref["b64"]),
                           step_err_m_deg=sdiff(step(al.H, al.b), dx64))
        ts = []
        for _ in range(4):
            al = mk()
            t = time.perf_counter()
            al.computeOnDevice()
            ts.append(time.perf_counter() - t)
        lat[name] = 1e3 * float(np.median(ts[1:]))
        poses[name] = al.getX()
    forms["reference_fp32_cpu"] = dict(H_rel=rel(ref["H"], ref["H64"]), b_rel=rel(ref["b"], ref["b64"]),
                                       step_err_m_deg=sdiff(step(ref["H"], ref["b"]), dx64))
    X64, _, _ = O.compute64(w4["fixed"], w4["moving"], w4["X0"], w4["Xf"], op, fast=True)
    X32, _, _, _ = O.compute(w4["fixed"], w4["moving"], w4["X0"], w4["Xf"], op, fast=True)
    poses["reference_fp32_cpu"] = X32
    out = dict(workload=f"rigid ICP, {w4['n_pts']}-pt 128-beam scan (ranges <= 200 m) vs 8M-pt map, world frame "
                        f"offset (10 km, 0, 0), 10 iterations",
               moving_leaves=int(w4["moving"].num_leaves), fixed_nodes=int(w4["fixed"].num_nodes),
               form_at_initial_state=forms,
               pose_vs_fp64_loop_m_deg={k: list(synth.pose_error(X, X64)) for k, X in poses.items()},
               fp64_loop_vs_ground_truth_m_deg=list(synth.pose_error(X64, w4["X_gt"])),
               device_loop_ms={k: round(v, 3) for k, v in lat.items()})
    out["recentred_pts_per_s"] = round(w4["n_pts"] / (lat["recentred_fp32"] * 1e-3), 1)
    log(f"precision: forms {forms}; poses vs FP64 {out['pose_vs_fp64_loop_m_deg']}; device loop {lat}")
    return out


def run_batch3(ctx, pairs3, DeviceTree, batch_align, adapter_align, default_params, synth):
    """Config 3: 256 independent rigid pairs (scan 4k-120k points, map 3x) in
    one batched call vs the plain per-call path; the launch shape is picked
    by the autotuner on this batch first."""
    from kd_slam2_amd import autotune

    p = default_params()
    p.max_error = 0.5
    p.max_iterations = 10
    trees = [(DeviceTree(ctx, q["fixed"]), DeviceTree(ctx, q["moving"])) for q in pairs3]
    pairs = [dict(fixed=F, moving=M, ct=0, X0=q["X0"], Xf=q["Xf"], params=p) for q, (F, M) in zip(pairs3, trees)]
    n_pts = sum(q["n_pts"] for q in pairs3)
    best, _ = autotune.tune(ctx, pairs, batch_align, ct=0)
    batch_align(ctx, pairs)  # warm (graph capture)
    tb = []
    for _ in range(3):
        t = time.perf_counter()
        out = batch_align(ctx, pairs)
        tb.append(time.perf_counter() - t)
    t_batch = float(np.median(tb))
    t = time.perf_counter()
    for q, (F, M) in zip(pairs3, trees):
        adapter_align("ICPHIP3D", ctx, F, M, q["X0"], q["Xf"], p)
    t_plain = time.perf_counter() - t
    ok = sum(1 for o, q in zip(out, pairs3) if synth.pose_error(o["X"], q["X_gt"])[0] < 0.05)
    ctx.set_launch_config()  # back to the defaults
    for F, M in trees:
        F.close()
        M.close()
    res = dict(pairs=len(pairs3), scan_points=int(n_pts), batch_ms=round(1e3 * t_batch, 3),
               batch_pts_per_s=round(n_pts / t_batch, 1), plain_ms=round(1e3 * t_plain, 3),
               plain_pts_per_s=round(n_pts / t_plain, 1), speedup=round(t_plain / t_batch, 2),
               autotuned=best, converged_within_5cm=ok)
    log(f"config 3: {len(pairs3)} pairs, batch {res['batch_ms']} ms vs plain {res['plain_ms']} ms "
        f"({res['speedup']}x), tuned {best}")
    return res


def run_serving(ctx, seqs, loops, seconds, DeviceTree, Server, default_params):
    """Config 5: Poisson arrivals into the continuous-batching server."""
    srv = Server(0, budget_bytes=64 << 30, max_inflight=64)
    keep = []
    seq_maps = []
    for q in seqs:
        T = DeviceTree(ctx, q["fixed"])
        srv.add_resident(T)
        seq_maps.append(T)
    loop_maps = []
    for lp in loops:
        T = DeviceTree(ctx, lp["fixed"])
        srv.add_resident(T)
        loop_maps.append(T)
    keep += seq_maps + loop_maps
    p_ct = default_params()
    p_ct.max_error = 0.5
    p_ct.max_iterations = 15
    p_ct.velocity_damping = 10.0
    p_icp = default_params()
    p_icp.max_error = 0.5
    p_icp.max_iterations = 10

    def submit(kind, idx, frame):
        if kind == "odom":
            q, f = seqs[idx], seqs[idx]["frames"][frame]
            return srv.submit(seq_maps[idx], f["moving"], f["X0"], q["Xf"], p_ct, ct=1, v0=f["v0"]), f["n_pts"]
        lp = loops[idx]
        return srv.submit(loop_maps[idx], lp["moving"], lp["X0"], lp["Xf"], p_icp, ct=0), lp["n_pts"]

    # warm the engine with one request of each kind (not timed)
    for kind in ("odom", "loop"):
        rid, _ = submit(kind, 0, 0)
        srv.wait(rid, 60.0)
    rng = np.random.default_rng(np.random.PCG64(5999))
    from kd_slam2_amd import synth

    events = []
    for s in range(len(seqs)):
        for k, t in enumerate(synth.poisson_arrivals(10.0, seconds, rng)):
            events.append((t, "odom", s, k % len(seqs[s]["frames"])))
    for t in synth.poisson_arrivals(20.0, seconds, rng):
        events.append((t, "loop", int(rng.integers(len(loops))), 0))
    events.sort()
    submitted = []
    lock = threading.Lock()

    def submitter():
        t0 = time.perf_counter()
        for t, kind, idx, frame in events:
            now = time.perf_counter() - t0
            if t > now:
                time.sleep(t - now)
            rid, npts = submit(kind, idx, frame)
            with lock:
                submitted.append((rid, kind, npts))

    th = threading.Thread(target=submitter)
    gc.collect()
    gc.disable()
    t_start = time.perf_counter()
    th.start()
    results = []
    k = 0
    while True:
        with lock:
            pending = submitted[k:] if k < len(submitted) else []
        if not pending:
            if not th.is_alive() and k >= len(submitted):
                break
            time.sleep(0.0005)
            continue
        for rid, kind, npts in pending:
            r = srv.wait(rid, 60.0)
            results.append((kind, npts, r))
            k += 1
    th.join()
    wall = time.perf_counter() - t_start
    gc.enable()
    srv_info = srv.info()
    srv.close()
    for T in keep:
        T.close()

    def pct(kind=None):
        lat = [r["latency_us"] / 1e3 for kk, _, r in results if kind is None or kk == kind]
        if not lat:
            return None
        return dict(n=len(lat), p50_ms=round(float(np.percentile(lat, 50)), 3),
                    p99_ms=round(float(np.percentile(lat, 99)), 3), max_ms=round(float(np.max(lat)), 3))

    ok = all(r["status"] == 0 for _, _, r in results)
    pts = sum(n for _, n, _ in results)
    out = dict(window_s=seconds, requests=len(results), all_ok=ok, odometry=pct("odom"), loop_closure=pct("loop"),
               all=pct(), pts_per_s=round(pts / wall, 1), engine_steps=srv_info["steps"],
               budget_gb=64, resident_maps=len(keep))
    log(f"serving: {len(results)} requests in {wall:.2f}s, odometry p50/p99 {out['odometry']['p50_ms']}/"
        f"{out['odometry']['p99_ms']} ms, loop p50/p99 {out['loop_closure']['p50_ms']}/"
        f"{out['loop_closure']['p99_ms']} ms")
    return out


if __name__ == "__main__":
    main()
