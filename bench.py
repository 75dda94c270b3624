"""Benchmark of the MI355X scan-alignment path (one JSON line on stdout).

Headline workload (`value`): batched loop-closure verification -- 256
independent rigid point-to-plane ICP pairs, each scan and map from the
room-scene generator with its own seed (3000 + i), scans log-uniform in
4k-120k points (8.5M in all) and maps 3x their scan, every map with its own
KD-tree (leaf size 16), initial errors up to 5 deg / 1 m, 10 Gauss-Newton
iterations, max_error 0.5.  Trees are resident in HBM before timing starts.

A step is one run of the 256 pairs prepared as a batch (PairBatch:
kdh_batch_prepare once, then kdh_batch_run with a new set of initial poses
per step): one graph holding the state initialisation, one tile launch and
one finish launch per iteration, and the result write-back to the host.  value = scan points of completed
alignments per second over the timed steps.  The launch shape (leaves per
lane, LDS staging depth, grid) is picked first by the project autotuner on
the same batch.

Beside it the line reports
  * the plain one-alignment-per-call path on the same 256 pairs
    (ICPHIP3D::compute from the C++ plugin, one kernel build + host FP64
    solve per iteration) and the batched form's speedup over it;
  * the batched poses against the CPU oracle (max over the 256 pairs) and the
    H/b relative error against the oracle's FP64 twin on a sample of pairs;
  * the tile kernel's roofline: HIP events around it on the context stream,
    and its traffic to the memory fabric from two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE) scaled by the same counters on a device copy of
    known size; beside it the cache-inclusive demand bytes (every node visit
    at 32 B);
  * CT-ICP served batches (a 131k-point 0.1 s sweep against a 2M-point map,
    15 iterations), the single-alignment device loop and the plain path;
  * the serving mix: 4 odometry sequences at 10 Hz (64k-point CT sweeps
    against their own resident 4M-point maps) and loop-closure requests at
    20/s (8k-200k points), Poisson arrivals into the continuous-batching
    server with a 64 GB budget, p50/p99 latency from arrival to host-visible
    result, and the highest arrival-rate multiple it sustains with odometry
    p99 under 10 ms;
  * the precision stress (2M-point scan, 8M-point map, 10 km from the
    origin), device tree generation, tree maintenance, batched factor forms
    and the KDTree3f codec;
  * the CPU oracle (restated ICP_CPU_) on the headline workload, 8 threads
    (the reference's cap) and 1 thread, with the host CPU model.
"""
import argparse
import contextlib
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
    """Cache-inclusive demand bytes of one Gauss-Newton iteration: per query
    24 B moving mean+normal, 4 B dts (CT), 4 B leaf index, 32 B per internal
    node visited, 24 B fixed leaf; plus one partial row (27 or 90 floats) per
    tile.  Every visit is counted at 32 B whether LDS, L2 or HBM serves it."""
    per_q = 24 + (4 if ct else 0) + 4 + 24
    return float(per_q * len(depth) + 32 * int(np.sum(depth)) + n_tiles * 4 * (90 if ct else 27))


def host_cpu():
    """(model name with its family / model / stepping numbers, logical CPUs)
    of this host: a virtual machine may report only the vendor's brand name,
    and the numbers still tell the processor generation."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if not line.strip():
                    break  # the first processor's block is enough
                k, _, v = line.partition(":")
                fields.setdefault(k.strip(), v.strip())
    except OSError:
        pass
    model = fields.get("model name", "unknown")
    ids = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
    if ids:
        model = f"{model} ({', '.join(ids)})"
    return model, os.cpu_count() or 0


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
    """Loop-closure pair i of the headline batch: room-scene generator, seed
    3000+i, scan n, map 3n; plus three more initial poses (seeds 7000+i) for
    the timed steps."""
    from kd_slam2_amd import TreeGenerator3D, synth

    i, n = i_n
    w = synth.config3_pair(i, n)
    g = TreeGenerator3D(leaf_max_points=16)
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"])
    w["n_pts"] = len(w["scan"])
    rng = np.random.default_rng(np.random.PCG64(7000 + i))
    w["X0_steps"] = [w["X0"]] + [synth.perturb(w["X_gt"], rng, 5.0, 1.0, uniform=True) for _ in range(3)]
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


@contextlib.contextmanager
def _worker_pool():
    """A fork pool of workload generators (at most 16, the CPU share of one
    GPU) that ends with close() and join(): the workers exit by themselves.
    Pool's own context exit terminates them with SIGTERM, which a process
    run under rocprofv3 catches in the profiler's preloaded handler, where it
    can stall, and the pool then waits on it for ever."""
    pool = mp.get_context("fork").Pool(min(16, os.cpu_count() or 4))
    try:
        yield pool
    finally:
        pool.close()
        pool.join()


def _gen_pairs3(n_pairs, pool):
    from kd_slam2_amd import synth

    sizes = synth.config3_sizes(256)[:n_pairs]
    return pool.map_async(_gen_pair3, list(enumerate(sizes)))


def _p3_params(default_params):
    p = default_params()
    p.max_error = 0.5
    p.max_iterations = 10
    return p


# ---------------------------------------------------------------------------
# PMC traffic: rocprofv3 --pmc passes over a child run of this script


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def pmc_child(args):
    """One batch of the headline workload at the tuned launch shape plus a
    device copy of COPY_BYTES bytes, for the counter passes.  With
    --pmc-copies k every pair's trees are uploaded k times (separate device
    copies, so separate traffic) and the batch holds k x the pairs."""
    from kd_slam2_amd import Context, DeviceTree, batch_align, default_params

    with _worker_pool() as pool:
        pairs3 = _gen_pairs3(args.pairs3, pool).get()
    ctx = Context(0)
    ctx.set_launch_config(grid=args.pmc_grid, lds_nodes=args.pmc_lds, leaves_per_lane=args.pmc_lpl,
                          tile_leaves=args.pmc_tile)
    p = _p3_params(default_params)
    job = []
    for _ in range(max(args.pmc_copies, 1)):
        trees = [(DeviceTree(ctx, q["fixed"]), DeviceTree(ctx, q["moving"])) for q in pairs3]
        job += [dict(fixed=F, moving=M, ct=0, X0=q["X0"], Xf=q["Xf"], params=p) for q, (F, M) in zip(pairs3, trees)]
    batch_align(ctx, job)
    print(f"pmc child: batch of {len(job)} pairs done", flush=True)
    _copy_probe()


def _copy_probe():
    """One device-to-device copy of COPY_BYTES bytes: the known traffic that
    scales the counters."""
    hip = _hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipMalloc(ctypes.byref(a), COPY_BYTES)
    hip.hipMalloc(ctypes.byref(b), COPY_BYTES)
    hip.hipMemcpy(b, a, COPY_BYTES, 3)  # device to device
    hip.hipFree(a)
    hip.hipFree(b)
    print("pmc child done", flush=True)


def pmc_child_ct(args):
    """One CT served batch (--batch alignments of the 131k-point sweep against
    the 2M-point map, 15 iterations, the initial states of the CT row) through
    kdh_batch_align, then the copy probe: the workload of the CT counter
    passes."""
    from kd_slam2_amd import Context, DeviceTree, batch_align, default_params, synth

    w = _gen_cfg2(0)
    ctx = Context(0)
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = _cfg2_params(default_params)
    rng = np.random.default_rng(np.random.PCG64(99))
    batch_align(ctx, [dict(fixed=F, moving=M, ct=1, X0=synth.perturb(w["X_gt"], rng, 1.0, 0.3), Xf=w["Xf"],
                           v0=w["v0"], params=p) for _ in range(args.batch)])
    print("pmc child: CT batch done", flush=True)
    _copy_probe()


def pmc_child_precision(args):
    """The precision-stress alignment (2M-point scan, 8M-point map, 10 km
    from the origin, re-centred FP32) run three times as device loops, plus a
    device copy of COPY_BYTES bytes: the workload of the traces and counter
    passes of the 8M-point map."""
    from kd_slam2_amd import Context, DeviceTree, ICPHIP3D, TreeGenerator3D, default_params, synth

    with _worker_pool() as pool:
        blocks = pool.map(_gen_cfg4_block, [(k, b) for k in (0, 1) for b in range(8)])
    g4 = TreeGenerator3D(parallel_level=3)
    X_gt4, X04, Xf4 = synth.config4_poses()
    moving, fixed = g4.makeTree(np.concatenate(blocks[:8])), g4.makeTree(np.concatenate(blocks[8:]))
    ctx = Context(0)
    F, M = DeviceTree(ctx, fixed), DeviceTree(ctx, moving)
    p = _p3_params(default_params)
    for _ in range(3):
        al = ICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(M)
        al.setX(X04)
        al.setFixedX(Xf4)
        al.computeOnDevice()
    print(f"precision device loop: {ctx.last_kernel_time()[0]:.2f} us/iter, {M.num_leaves} moving leaves, "
          f"{len(fixed.nodes)} fixed nodes", flush=True)
    _copy_probe()


def pmc_pass(counter, args, tuned):
    """Per-dispatch counter (KiB) of the headline tile kernel (median over its
    dispatches) and of the copy kernel, from one rocprofv3 --pmc run.
    Returns (tile, copy) or None."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    out = tempfile.mkdtemp(prefix="kdh_pmc_")
    cmd = ["timeout", "-s", "KILL", "150", exe, "--pmc", counter, "-d", out, "-o", "pmc", "--output-format", "csv",
           "--", sys.executable, os.path.abspath(__file__), "--pmc-child", "--pairs3", str(args.pairs3),
           "--pmc-grid", str(tuned["grid"]), "--pmc-lds", str(tuned["lds_nodes"]),
           "--pmc-lpl", str(tuned["leaves_per_lane"]), "--pmc-tile", str(tuned.get("tile_leaves", 256))]
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=170, cwd=ROOT,
                           env=dict(os.environ, TMPDIR="/tmp"))
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            tail = r.stdout.decode(errors="replace").strip().splitlines()[-3:] if r.stdout else []
            log(f"pmc {counter}: no counter file (rc {r.returncode}); child output ends: {tail}")
            return None
        tile, copy = [], []
        with open(files[0]) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                grid = int(row.get("Grid_Size", row.get("Grid_Size_X", "0")) or 0)
                v = float(row["Counter_Value"])
                if "tiles_kernel" in name:
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
    ap.add_argument("--pairs3", type=int, default=256, help="loop-closure pairs of the headline batch")
    ap.add_argument("--batch", type=int, default=32, help="concurrent CT alignments per served CT step")
    ap.add_argument("--plain-runs", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="length of the serving-mix arrival window")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-pmc", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the 10 km precision-stress workload")
    ap.add_argument("--no-tree-build", action="store_true", help="skip the tree-generation timings")
    ap.add_argument("--no-next-rows", action="store_true", help="skip tree maintenance, factor forms, KDTree3f")
    ap.add_argument("--no-scaling", action="store_true", help="skip the batch-size scaling rows")
    ap.add_argument("--scaling-steps", type=int, default=20, help="timed batches per batch size")
    ap.add_argument("--headline-only", action="store_true", help="only the headline batch and its roofline")
    ap.add_argument("--tree-build-only", action="store_true",
                    help="only the tree-generation timings (for kernel traces of the builder)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-child-ct", action="store_true",
                    help="only one CT served batch and the copy probe (for the CT counter passes)")
    ap.add_argument("--pmc-child-precision", action="store_true",
                    help="only the precision-stress device loop (for traces and counter passes of the 8M map)")
    ap.add_argument("--pmc-grid", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-lds", type=int, default=256, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-lpl", type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-tile", type=int, default=256, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-copies", type=int, default=1, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.pmc_child_precision:
        pmc_child_precision(args)
        return
    if args.pmc_child_ct:
        pmc_child_ct(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); alignments shard by pair with no "
              "communication across GPUs: run one process per GPU, each on its own share of the pairs", file=sys.stderr)
        sys.exit(2)
    if args.headline_only:
        args.no_serve = args.no_precision = args.no_tree_build = args.no_next_rows = args.no_scaling = True

    if args.tree_build_only:
        from kd_slam2_amd import Context, DeviceTree, TreeGenerator3D, default_params, synth

        w = synth.config2()
        sw5 = _gen_sweep5(0)
        ctx = Context(0)
        F = TreeGenerator3D().makeTreeOnDevice(ctx, w["map"], download=False)
        print(json.dumps(run_tree_build(ctx, w, sw5, DeviceTree, F, _cfg2_params(default_params))), flush=True)
        return

    t0 = time.time()
    with _worker_pool() as pool:
        j_p3 = _gen_pairs3(args.pairs3, pool)
        j_cfg2 = None if args.headline_only else pool.apply_async(_gen_cfg2, (0,))
        j_sw5 = None if args.no_tree_build else pool.apply_async(_gen_sweep5, (0,))
        j_seq = None if args.no_serve else pool.map_async(_gen_seq, range(4))
        j_loop = None if args.no_serve else pool.map_async(_gen_loop, range(16))
        j_c4 = None if args.no_precision else pool.map_async(_gen_cfg4_block, [(k, b) for k in (0, 1)
                                                                                for b in range(8)])
        pairs3 = j_p3.get()
        w = j_cfg2.get() if j_cfg2 else None
        sweep5 = j_sw5.get() if j_sw5 else None
        c4_blocks = j_c4.get() if j_c4 else None
        seqs = j_seq.get() if j_seq else []
        loops = j_loop.get() if j_loop else []
    w4 = None
    if c4_blocks is not None:
        from kd_slam2_amd import TreeGenerator3D, synth as _s

        scan4 = np.concatenate(c4_blocks[:8])
        g4 = TreeGenerator3D(parallel_level=3)
        X_gt4, X04, Xf4 = _s.config4_poses()
        w4 = dict(n_pts=len(scan4), moving=g4.makeTree(scan4), fixed=g4.makeTree(np.concatenate(c4_blocks[8:])),
                  X_gt=X_gt4, X0=X04, Xf=Xf4)
        del scan4, c4_blocks
    log(f"workloads ready in {time.time() - t0:.1f}s")

    from kd_slam2_amd import (CTICPHIP3D, Context, DeviceTree, Server, adapter_align, batch_align,
                              default_params, synth)

    ctx = Context(0)

    # ---- headline: 256-pair batched loop-closure verification --------------
    head = run_headline(ctx, pairs3, args)
    ctx.set_launch_config(leaves_per_lane=1, batch_streams=1, tile_leaves=256)  # the defaults for the other rows

    # ---- fabric traffic of the headline tile kernel: two counter passes over
    # a child run of this script, made while this process holds only the
    # headline workload ------------------------------------------------------
    traffic = None
    pmc = {}
    if not args.no_pmc:
        f = pmc_pass("FETCH_SIZE", args, head["tuned"])
        wr = pmc_pass("WRITE_SIZE", args, head["tuned"])
        if f and wr:
            kib = COPY_BYTES / 1024.0
            f_scale = kib / f[1]
            w_scale = kib / wr[1]
            traffic = round((f[0] * f_scale + wr[0] * w_scale) * 1024.0, 1)
            pmc = dict(fetch_kib=round(f[0], 1), write_kib=round(wr[0], 1), copy_fetch_kib=round(f[1], 1),
                       copy_write_kib=round(wr[1], 1), fetch_scale=round(f_scale, 4), write_scale=round(w_scale, 4))
            log(f"pmc: tile fetch {f[0]:.0f} KiB, write {wr[0]:.0f} KiB per launch (copy {f[1]:.0f}/{wr[1]:.0f} "
                f"KiB for {kib:.0f} KiB)")

    # ---- the same path at other batch sizes --------------------------------
    scaling = None if args.no_scaling else run_batch_scaling(ctx, pairs3, args)

    # ---- CT-ICP: served batches, device loop, plain path ---------------------
    ct = None
    F = M = None
    if w is not None:
        F = DeviceTree(ctx, w["fixed"])
        M = DeviceTree(ctx, w["moving"])
        ct = run_ct(ctx, w, F, M, args, CTICPHIP3D, adapter_align, batch_align, default_params, synth)

    # ---- serving mix -------------------------------------------------------
    serving = None
    if not args.no_serve:
        serving = run_serving(ctx, seqs, loops, args.serve_seconds, DeviceTree, Server, default_params)
        serving["capacity"] = run_capacity(ctx, seqs, loops, DeviceTree, Server, default_params)

    # ---- precision stress (config 4) ---------------------------------------
    precision = None
    if w4 is not None:
        precision = run_precision(ctx, w4, DeviceTree, default_params, synth)

    # ---- tree generation on the device (makeTree + upload) ------------------
    tree_build = None
    if not args.no_tree_build and w is not None:
        tree_build = run_tree_build(ctx, w, sweep5, DeviceTree, F, _cfg2_params(default_params))
    if w is not None:
        w.pop("map", None)

    # ---- the other rows around the path: tree maintenance, batched factor
    # forms, the KDTree3f wire and bag format -----------------------------
    next_rows = None
    if not args.no_next_rows and w is not None:
        next_rows = run_next_rows(ctx, w, DeviceTree, F, M, _cfg2_params(default_params))

    # ---- CPU baseline: the oracle on the headline workload --------------------
    cpu = None
    if not args.no_cpu:
        cpu = run_cpu_baseline(pairs3, args, default_params)

    tile_us = head["tile_us"]
    demand = head["demand_bytes_per_launch"]
    achieved = traffic / (tile_us * 1e-6) / 1e9 if traffic else None
    roofline = dict(bound="hbm", achieved=round(achieved, 1) if achieved else None, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4) if achieved else None, traffic=traffic,
                    kernel="tiles_kernel (rigid, 256 pairs)", kernel_us=round(tile_us, 2),
                    cache_inclusive=dict(demand_bytes=demand,
                                         achieved=round(demand / (tile_us * 1e-6) / 1e9, 1),
                                         frac=round(demand / (tile_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)))
    if achieved is not None:
        assert roofline["frac"] <= 1.0, roofline
    line = dict(
        metric="scan-pts aligned/s + p99 align latency, served vs plain HIP, <=1mm/1e-3deg",
        value=head["value"],
        unit="scan-pts/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=head["ms_per_step"],
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f32",
        data="synthetic",
        config=dict(workload=f"batched loop-closure verification: {head['pairs']} rigid ICP pairs, scans "
                             f"log-uniform 4k-120k pts ({head['scan_points']} in all), maps 3x, 10 iterations",
                    pairs=head["pairs"], scan_points=head["scan_points"], moving_leaves=head["moving_leaves"],
                    fixed_nodes=head["fixed_nodes"], mean_depth=head["mean_depth"], tuned=head["tuned"]),
        latency_ms=head["latency_ms"],
        one_at_a_time=head["one_at_a_time"],
        leaf_iterations_per_s=head["leaf_iterations_per_s"],
        served_speedup_vs_plain=head["speedup_vs_plain"],
        plain_pts_per_s=head["plain_pts_per_s"],
        pcie_inclusive=head["pcie_inclusive"],
        accuracy=head["accuracy"],
        kernel_us_per_iter=head["kernel_us_per_iter"],
        batch_scaling=scaling,
        pmc=pmc or None,
        ct_icp=ct,
        serving_mix=serving,
        precision_stress=precision,
        tree_build=tree_build,
        next_rows=next_rows,
        roofline=roofline,
        cpu_baseline=cpu,
    )
    print(json.dumps(line), flush=True)


def run_headline(ctx, pairs3, args):
    """The 256-pair batch: autotuned launch shape, W + K PairBatch steps,
    the plain per-call path on the same pairs, accuracy against the oracle,
    tile / finish kernel times and the cache-inclusive demand bytes."""
    from kd_slam2_amd import DeviceTree, ICPHIP3D, PairBatch, adapter_align, autotune, batch_align, default_params
    from kd_slam2_amd import synth
    from oracle import pyoracle as O

    p = _p3_params(default_params)
    trees = [(DeviceTree(ctx, q["fixed"]), DeviceTree(ctx, q["moving"])) for q in pairs3]
    pairs = [dict(fixed=F, moving=M, ct=0, X0=q["X0"], Xf=q["Xf"], params=p) for q, (F, M) in zip(pairs3, trees)]
    n_pts = int(sum(q["n_pts"] for q in pairs3))
    tuned, _ = autotune.tune(ctx, pairs, batch_align, ct=0)
    log(f"headline: {len(pairs)} pairs, {n_pts} scan points, tuned {tuned}")
    B = PairBatch(pairs)
    X0_sets = [np.stack([q["X0_steps"][k] for q in pairs3]) for k in range(len(pairs3[0]["X0_steps"]))]
    for k in range(args.warmup):
        B.set_initial(X0_sets[k % len(X0_sets)])
        B.run(ctx)
    gc.collect()
    gc.disable()
    step_t = []
    kern = []
    steps_out = []
    for k in range(args.steps):
        t = time.perf_counter()
        B.set_initial(X0_sets[k % len(X0_sets)])
        steps_out.append(B.run(ctx))
        step_t.append(time.perf_counter() - t)
        kern.append(ctx.last_kernel_time()[0])
    # the same K steps pipelined: step k + 1 is submitted before step k's
    # results are collected (kdh_batch_submit / kdh_batch_collect, two runs
    # in flight), so the device runs the batches back to back while the host
    # prepares inputs and reads results; `value` is this whole-job rate
    piped = []
    t = time.perf_counter()
    for k in range(args.steps):
        B.set_initial(X0_sets[k % len(X0_sets)])
        B.submit(ctx)
        if k > 0:
            piped.append(B.collect(ctx))
    piped.append(B.collect(ctx))
    total = time.perf_counter() - t
    gc.enable()
    same_piped = all(np.array_equal(a.X, b.X) and np.array_equal(a.iterations, b.iterations)
                     for a, b in zip(steps_out, piped))
    seq_total = float(np.sum(step_t))
    value = n_pts * args.steps / total
    # moving-leaf iterations: every pair's leaves times the rounds it ran in each step
    leaves_per_pair = np.array([len(q["moving"].leaves) for q in pairs3], dtype=np.float64)
    leaf_iters = float(sum(np.dot(leaves_per_pair, o.iterations.astype(np.float64)) for o in steps_out))
    # the first pose set once more, for the accuracy check
    B.set_initial(X0_sets[0])
    out = B.run(ctx)
    log(f"headline: {1e3 * total / args.steps:.3f} ms/step pipelined, {value / 1e9:.2f} G pts/s; one at a time "
        f"{1e3 * seq_total / args.steps:.3f} ms/step; pipelined results equal: {same_piped}")
    # kernel times (profiled: plain launches with events, tile kernel repeated)
    ctx.set_profiling(True)
    prof = []
    for _ in range(3):
        B.run(ctx)
        prof.append(ctx.kernel_breakdown())
    ctx.set_profiling(False)
    tile_us = float(np.median([x[0] for x in prof]))
    fin_us = float(np.median([x[1] for x in prof]))
    # plain per-call path on the same pairs (C++ plugin, host loop)
    t = time.perf_counter()
    for q, (F, M) in zip(pairs3, trees):
        adapter_align("ICPHIP3D", ctx, F, M, q["X0"], q["Xf"], p)
    t_plain = time.perf_counter() - t
    # demand bytes per launch: descent depth of every query at the initial cache
    depth_sum, n_q, n_tiles = 0, 0, 0
    for q, (F, M) in zip(pairs3, trees):
        al = ICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(M)
        al.setX(q["X0"])
        al.setFixedX(q["Xf"])
        _, d = al.correspondences()
        depth_sum += int(np.sum(d))
        n_q += len(d)
        n_tiles += (len(d) + tuned["tile_leaves"] - 1) // tuned["tile_leaves"]
    demand = float((24 + 4 + 24) * n_q + 32 * depth_sum + n_tiles * 4 * 27)
    # accuracy: every pair against the oracle; H/b against the FP64 twin on a sample
    op = O.from_kdh_params(p, 8)
    worst_t = worst_r = 0.0
    same_code = 0
    for i, q in enumerate(pairs3):
        Xo, _, _, res_o = O.compute(q["fixed"], q["moving"], q["X0"], q["Xf"], op, fast=False)
        dt, dr = synth.pose_error(out.X[i], Xo)
        worst_t, worst_r = max(worst_t, dt), max(worst_r, dr)
        same_code += int(out.result[i] == res_o)
    hb = []
    for i in range(0, len(pairs3), 32):
        q = pairs3[i]
        F, M = trees[i]
        al = ICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(M)
        al.setX(q["X0"])
        al.setFixedX(q["Xf"])
        al.buildQuadraticFormAt()
        ref = O.build_form_at(q["fixed"], q["moving"], q["X0"], q["Xf"], op, fast=True)
        hb.append((float(np.abs(al.H - ref["H64"]).max() / np.abs(ref["H64"]).max()),
                   float(np.abs(al.b - ref["b64"]).max() / np.abs(ref["b64"]).max())))
    n_gt = sum(1 for i, q in enumerate(pairs3) if synth.pose_error(out.X[i], q["X_gt"])[0] < 0.05)
    # host-buffer boundary: every tree of the batch uploaded from host memory
    # (record staging, allocation, PCIe copy) before the same batch step
    t = time.perf_counter()
    fresh = [(DeviceTree(ctx, q["fixed"]), DeviceTree(ctx, q["moving"])) for q in pairs3]
    t_upload = time.perf_counter() - t
    tree_bytes = int(sum(32 * (len(T.nodes) + len(T.leaves)) for q in pairs3 for T in (q["fixed"], q["moving"])))
    for F, M in fresh:
        F.close()
        M.close()
    res = dict(
        value=round(value, 1), ms_per_step=round(1e3 * total / args.steps, 3), pairs=len(pairs3),
        leaf_iterations_per_s=round(leaf_iters / total, 1),
        scan_points=n_pts, moving_leaves=int(n_q), fixed_nodes=int(sum(len(q["fixed"].nodes) for q in pairs3)),
        mean_depth=round(depth_sum / max(n_q, 1), 2), tuned=tuned,
        latency_ms=dict(step_p50=round(1e3 * float(np.percentile(step_t, 50)), 3),
                        step_p99=round(1e3 * float(np.percentile(step_t, 99)), 3),
                        plain_per_alignment_mean=round(1e3 * t_plain / len(pairs3), 3)),
        one_at_a_time=dict(ms_per_step=round(1e3 * seq_total / args.steps, 3),
                           pts_per_s=round(n_pts * args.steps / seq_total, 1),
                           pipelined_results_equal=bool(same_piped)),
        plain_pts_per_s=round(n_pts / t_plain, 1), speedup_vs_plain=round(value / (n_pts / t_plain), 2),
        accuracy=dict(pose_vs_oracle_max_m_deg=[worst_t, worst_r], result_code_equal=same_code,
                      hb_rel_err_vs_fp64_max=[max(x[0] for x in hb), max(x[1] for x in hb)],
                      hb_sample_pairs=len(hb), within_5cm_of_ground_truth=n_gt),
        kernel_us_per_iter=dict(iteration=round(float(np.median(kern)), 2), tile_kernel=round(tile_us, 2),
                                finish_kernel=round(fin_us, 2)),
        pcie_inclusive=dict(upload_ms=round(1e3 * t_upload, 2), tree_MB=round(tree_bytes / 1e6, 1),
                            upload_GBps=round(tree_bytes / t_upload / 1e9, 2),
                            pts_per_s=round(n_pts / (t_upload + total / args.steps), 1)),
        tile_us=tile_us, demand_bytes_per_launch=demand)
    log(f"headline: tile {tile_us:.1f} us, finish {fin_us:.1f} us, plain {1e3 * t_plain:.1f} ms, accuracy "
        f"{res['accuracy']}")
    for F, M in trees:
        F.close()
        M.close()
    return res


def run_cpu_baseline(pairs3, args, default_params):
    """The oracle (restated ICP_CPU_, timing build) on pairs of the headline
    batch, in pair order, for about --cpu-seconds on 8 threads (the
    reference's cap) and a third of that on 1 thread."""
    from oracle import pyoracle as O

    p = _p3_params(default_params)
    model, ncpu = host_cpu()
    out = {}
    for threads, secs in ((8, args.cpu_seconds), (1, args.cpu_seconds / 3)):
        op = O.from_kdh_params(p, num_threads=threads)
        t = time.perf_counter()
        n_done, pts = 0, 0
        while time.perf_counter() - t < secs or n_done == 0:
            q = pairs3[n_done % len(pairs3)]
            O.compute(q["fixed"], q["moving"], q["X0"], q["Xf"], op, fast=True)
            n_done += 1
            pts += q["n_pts"]
        el = time.perf_counter() - t
        out[threads] = (pts / el, n_done, el)
    v8, n8, e8 = out[8]
    v1, n1, e1 = out[1]
    cpu = dict(value=round(v8, 1), unit="scan-pts/s", cores=8, kind="port",
               sample=f"{n8} whole rigid alignments (10 iterations) of the headline pairs in pair order on 8 threads "
                      f"(oracle/liboracle_fast.so, std::async chunks as ICP_CPU_), {e8:.1f} s",
               one_thread=dict(value=round(v1, 1), alignments=n1, seconds=round(e1, 1)),
               host=dict(cpu_model=model, logical_cpus=ncpu))
    log(f"cpu oracle: {v8 / 1e6:.2f} Mpts/s on 8 threads, {v1 / 1e6:.2f} on 1 ({model}, {ncpu} CPUs)")
    return cpu


def run_batch_scaling(ctx, pairs3, args):
    """Throughput of prepared batches of the first n headline pairs, and of
    all 256 twice (every tree uploaded a second time: separate device
    copies, so the 512 alignments share no tree), each at its own tuned
    launch shape: where the GPU fills and the fixed cost of a launch (one
    dependent gather chain, the launch itself) stops dominating."""
    from kd_slam2_amd import DeviceTree, PairBatch, autotune, batch_align, default_params

    p = _p3_params(default_params)
    copies = [[(DeviceTree(ctx, q["fixed"]), DeviceTree(ctx, q["moving"])) for q in pairs3] for _ in range(2)]
    every = [dict(fixed=F, moving=M, ct=0, X0=q["X0"], Xf=q["Xf"], params=p)
             for trees in copies for q, (F, M) in zip(pairs3, trees)]
    pts = [q["n_pts"] for q in pairs3] * 2
    tiles = [(len(q["moving"].leaves) + 255) // 256 for q in pairs3] * 2
    out = []
    for n in (32, 64, 128, 256, 512):
        if n > len(every):
            continue
        job = every[:n]
        tuned, _ = autotune.tune(ctx, job, batch_align, ct=0)
        B = PairBatch(job)
        for _ in range(3):
            B.run(ctx)
        ts = []
        for _ in range(args.scaling_steps):
            t = time.perf_counter()
            B.run(ctx)
            ts.append(time.perf_counter() - t)
        # pipelined, as the headline: the next run submitted before the
        # previous one is collected
        t = time.perf_counter()
        for k in range(args.scaling_steps):
            B.submit(ctx)
            if k > 0:
                B.collect(ctx)
        B.collect(ctx)
        piped_ms = 1e3 * (time.perf_counter() - t) / args.scaling_steps
        ctx.set_profiling(True)
        B.run(ctx)
        tu, fu = ctx.kernel_breakdown()
        ctx.set_profiling(False)
        ms = 1e3 * float(np.median(ts))
        n_pts = int(sum(pts[:n]))
        out.append(dict(pairs=n, scan_points=n_pts, tiles=int(sum(tiles[:n])), ms_per_batch=round(piped_ms, 3),
                        pts_per_s=round(n_pts / (piped_ms * 1e-3), 1), one_at_a_time_ms=round(ms, 3),
                        tile_kernel_us=round(tu, 2),
                        finish_kernel_us=round(fu, 2),
                        tuned=dict(lds_nodes=tuned["lds_nodes"], leaves_per_lane=tuned["leaves_per_lane"],
                                   tile_leaves=tuned["tile_leaves"], grid=tuned["grid"],
                                   batch_streams=tuned["batch_streams"])))
        log(f"batch scaling: {out[-1]}")
    for trees in copies:
        for F, M in trees:
            F.close()
            M.close()
    ctx.set_launch_config(leaves_per_lane=1, batch_streams=1, tile_leaves=256)
    return out


def run_ct(ctx, w, F, M, args, CTICPHIP3D, adapter_align, batch_align, default_params, synth):
    """CT-ICP with deskew (131k-point sweep vs 2M-point map, 15 iterations):
    plain per-call path, device-resident single-alignment loop, served
    batches of --batch alignments (the map shared, own initial states)."""
    from kd_slam2_amd import PairBatch

    p = _cfg2_params(default_params)
    n_pts = len(w["scan"]) if "scan" in w else 131_072

    def aligner():
        al = CTICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(M)
        al.setFixedX(w["Xf"])
        al.setX(w["X0"])
        al.setVelocities(w["v0"])
        return al

    plain = []
    gc.disable()
    for i in range(args.plain_runs + 1):
        t = time.perf_counter()
        adapter_align("CTICPHIP3D", ctx, F, M, w["X0"], w["Xf"], p, v0=w["v0"])
        if i:
            plain.append(time.perf_counter() - t)
    dev, kern = [], []
    for i in range(args.plain_runs + 1):
        al = aligner()
        t = time.perf_counter()
        al.computeOnDevice()
        if i:
            dev.append(time.perf_counter() - t)
            kern.append(ctx.last_kernel_time()[0])
    gc.enable()
    rng = np.random.default_rng(np.random.PCG64(99))
    pairs = [dict(fixed=F, moving=M, ct=1, X0=synth.perturb(w["X_gt"], rng, 1.0, 0.3), Xf=w["Xf"], v0=w["v0"],
                  params=p) for _ in range(args.batch)]
    B = PairBatch(pairs)
    for _ in range(5):
        B.run(ctx)
    gc.disable()
    st = []
    for _ in range(50):
        t = time.perf_counter()
        B.run(ctx)
        st.append(time.perf_counter() - t)
    # the same 50 batches pipelined (two in flight), as the headline steps
    t = time.perf_counter()
    for k in range(50):
        B.submit(ctx)
        if k > 0:
            B.collect(ctx)
    B.collect(ctx)
    piped = (time.perf_counter() - t) / 50
    gc.enable()
    ctx.set_profiling(True)
    B.run(ctx)
    tu, fu = ctx.kernel_breakdown()
    ctx.set_profiling(False)
    med = float(np.median(st))
    out = dict(workload="CT-ICP deskew scan-to-map, 131072-pt 128-beam sweep vs 2M-pt map, 15 iterations",
               plain_ms_p50=round(1e3 * float(np.median(plain)), 3),
               device_loop_ms_p50=round(1e3 * float(np.median(dev)), 3),
               device_loop_ms_p99=round(1e3 * float(np.percentile(dev, 99)), 3),
               device_loop_us_per_iter=round(float(np.median(kern)), 2),
               served_batch=args.batch, served_ms_p50=round(1e3 * med, 3),
               served_ms_p99=round(1e3 * float(np.percentile(st, 99)), 3),
               served_pts_per_s=round(args.batch * n_pts / med, 1),
               served_pipelined_ms=round(1e3 * piped, 3),
               served_pipelined_pts_per_s=round(args.batch * n_pts / piped, 1),
               plain_pts_per_s=round(n_pts / float(np.median(plain)), 1),
               served_speedup_vs_plain=round(args.batch * float(np.median(plain)) / med, 2),
               served_tile_us=round(tu, 2), served_finish_us=round(fu, 2))
    log(f"ct: {out}")
    return out


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



def _serve_window(srv, seqs, loops, seq_maps, loop_maps, seconds, rate_mult, seed, default_params):
    """Poisson arrivals for `seconds` at rate_mult x (10 Hz per odometry
    sequence, 20/s loop closure); returns [(kind, n_pts, result)] and the
    wall time."""
    from kd_slam2_amd import synth

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

    rng = np.random.default_rng(np.random.PCG64(seed))
    events = []
    for s in range(len(seqs)):
        for k, t in enumerate(synth.poisson_arrivals(10.0 * rate_mult, seconds, rng)):
            events.append((t, "odom", s, k % len(seqs[s]["frames"])))
    for t in synth.poisson_arrivals(20.0 * rate_mult, seconds, rng):
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
    return results, wall


def _pct(results, kind=None):
    lat = [r["latency_us"] / 1e3 for kk, _, r in results if kind is None or kk == kind]
    if not lat:
        return None
    return dict(n=len(lat), p50_ms=round(float(np.percentile(lat, 50)), 3),
                p99_ms=round(float(np.percentile(lat, 99)), 3), max_ms=round(float(np.max(lat)), 3))


def _serving_setup(ctx, seqs, loops, DeviceTree, Server):
    srv = Server(0, budget_bytes=64 << 30, max_inflight=64)
    seq_maps = [DeviceTree(ctx, q["fixed"]) for q in seqs]
    loop_maps = [DeviceTree(ctx, lp["fixed"]) for lp in loops]
    for T in seq_maps + loop_maps:
        srv.add_resident(T)
    return srv, seq_maps, loop_maps


def run_serving(ctx, seqs, loops, seconds, DeviceTree, Server, default_params):
    """Config 5: Poisson arrivals into the continuous-batching server."""
    srv, seq_maps, loop_maps = _serving_setup(ctx, seqs, loops, DeviceTree, Server)
    # warm the engine (not timed)
    _serve_window(srv, seqs, loops, seq_maps, loop_maps, 0.3, 1.0, 5998, default_params)
    results, wall = _serve_window(srv, seqs, loops, seq_maps, loop_maps, seconds, 1.0, 5999, default_params)
    info = srv.info()
    srv.close()
    for T in seq_maps + loop_maps:
        T.close()
    ok = all(r["status"] == 0 for _, _, r in results)
    pts = sum(n for _, n, _ in results)
    out = dict(window_s=seconds, requests=len(results), all_ok=ok, odometry=_pct(results, "odom"),
               loop_closure=_pct(results, "loop"), all=_pct(results), pts_per_s=round(pts / wall, 1),
               engine_steps=info["steps"], budget_gb=64, resident_maps=len(seq_maps) + len(loop_maps))
    log(f"serving: {len(results)} requests in {wall:.2f}s, odometry {out['odometry']}, loop {out['loop_closure']}")
    return out


def run_capacity(ctx, seqs, loops, DeviceTree, Server, default_params, seconds=1.5,
                 mults=(4, 16, 64, 128, 256, 512)):
    """Highest arrival-rate multiple of the serving mix (x 4 sequences at
    10 Hz and loop closure at 20/s) at which every request completes and the
    odometry p99 latency stays under 10 ms; points/s served at that rate."""
    srv, seq_maps, loop_maps = _serving_setup(ctx, seqs, loops, DeviceTree, Server)
    runs = []
    best = None
    for m in mults:
        results, wall = _serve_window(srv, seqs, loops, seq_maps, loop_maps, seconds, float(m), 6000 + m,
                                      default_params)
        odo = _pct(results, "odom")
        ok = all(r["status"] == 0 for _, _, r in results)
        pts = sum(n for _, n, _ in results)
        run = dict(rate_mult=m, requests=len(results), requests_per_s=round(len(results) / wall, 1),
                   pts_per_s=round(pts / wall, 1), odometry_p99_ms=odo["p99_ms"] if odo else None,
                   loop_p99_ms=(_pct(results, "loop") or {}).get("p99_ms"), all_ok=ok)
        runs.append(run)
        log(f"capacity x{m}: {run}")
        if ok and odo and odo["p99_ms"] < 10.0:
            best = run
        else:
            break
    srv.close()
    for T in seq_maps + loop_maps:
        T.close()
    return dict(window_s=seconds, sustained=best, runs=runs)


if __name__ == "__main__":
    main()
