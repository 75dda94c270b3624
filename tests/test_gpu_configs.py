"""Parity on the full benchmark workloads (GPU):

* batched loop-closure verification: 256 rigid pairs, scans log-uniform in
  4k-120k points, maps 3x, initial errors up to 5 deg / 1 m, in one
  kdh_batch_align launch; every pair's pose and result code against the
  oracle's compute(), iteration-0 correspondences bit-exact on a sample;
* the serving mix: 4 resident 4M-point maps with 64k-point CT sweeps at
  10 Hz each and loop-closure requests of 8k-200k points at 20/s, Poisson
  arrivals into the continuous-batching server; every served result
  bit-identical to the same request served alone, a sample against the
  oracle, odometry p99 latency under 10 ms.

The precision-stress workload (2M-point scan, 8M-point map at 10 km) is in
test_gpu_precision.py."""
import functools
import multiprocessing as mp
import threading
import time

import numpy as np
import pytest

from kd_slam2_amd import DeviceTree, ICPHIP3D, Server, TreeGenerator3D, batch_align, synth, update_cache_f32
from oracle import pyoracle as O

from .workloads import icp_params

pytestmark = pytest.mark.gpu

POSE_T_TOL = 1e-3  # metres
POSE_R_TOL = 1e-3  # degrees
COUNTERS = ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn")


def _gen_pair3(i_n):
    i, n = i_n
    w = synth.config3_pair(i, int(n))
    g = TreeGenerator3D(leaf_max_points=16)
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"])
    w["n_pts"] = len(w["scan"])
    del w["map"], w["scan"], w["stamps"]
    return w


def _gen_seq(s):
    q = synth.config5_sequence(s)
    g = TreeGenerator3D(parallel_level=2)
    q["fixed"] = g.makeTree(q["map"])
    del q["map"]
    for f in q["frames"]:
        f["moving"] = g.makeTree(f["scan"], f["stamps"])
        f["n_pts"] = len(f["scan"])
        del f["scan"], f["stamps"]
    return q


def _gen_loop(j):
    w = synth.config5_loop_pair(j)
    g = TreeGenerator3D(leaf_max_points=16)
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"])
    w["n_pts"] = len(w["scan"])
    del w["map"], w["scan"]
    return w


@functools.lru_cache(maxsize=None)
def config3():
    sizes = synth.config3_sizes(256)
    with mp.get_context("fork").Pool(16) as pool:
        return pool.map(_gen_pair3, list(enumerate(sizes)))


@functools.lru_cache(maxsize=None)
def config5(n_loops=16):
    with mp.get_context("fork").Pool(16) as pool:
        j_seq = pool.map_async(_gen_seq, range(4))
        j_loop = pool.map_async(_gen_loop, range(n_loops))
        return j_seq.get(), j_loop.get()


def test_config3_256_pairs_batched_against_oracle(ctx):
    ws = config3()
    assert len(ws) == 256 and 8_000_000 < sum(w["n_pts"] for w in ws) < 9_000_000
    p = icp_params()
    trees = [(DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])) for w in ws]
    pairs = [dict(fixed=F, moving=M, ct=0, X0=w["X0"], Xf=w["Xf"], v0=np.zeros(6), params=p)
             for w, (F, M) in zip(ws, trees)]
    res = batch_align(ctx, pairs)
    op = O.from_kdh_params(p, 8)
    worst = (0.0, 0.0)
    n_gt = 0
    for i, (w, r) in enumerate(zip(ws, res)):
        Xo, _, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], op)
        dt, dr = synth.pose_error(r["X"], Xo)
        worst = (max(worst[0], dt), max(worst[1], dr))
        assert r["result"] == res_o, i
        assert dt < POSE_T_TOL and dr < POSE_R_TOL, (i, dt, dr)
        gt = synth.pose_error(Xo, w["X_gt"])
        n_gt += int(gt[0] < 0.05)
    print(f"256 pairs: worst pose vs oracle {worst[0]:.2e} m {worst[1]:.2e} deg; oracle within 5 cm of ground "
          f"truth on {n_gt}")
    # iteration-0 correspondences and counters, bit for bit, on a sample
    for i in list(range(0, 256, 16)) + [int(np.argmax([w["n_pts"] for w in ws]))]:
        w = ws[i]
        F, M = trees[i]
        al = ICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(M)
        Xm, Xf, Xmf = update_cache_f32(w["X0"], w["Xf"])
        leaf, depth = al.correspondences(Xm, Xf, Xmf)
        ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, op)
        assert np.array_equal(leaf, ref["corr"]), i
        assert np.array_equal(depth, ref["depth"]), i
    for F, M in trees:
        F.close()
        M.close()


def test_config5_serving_mix_alone_equals_mixed(ctx):
    seqs, loops = config5()
    srv = Server(0, budget_bytes=64 << 30, max_inflight=64)
    keep = []
    try:
        seq_maps = [DeviceTree(ctx, q["fixed"]) for q in seqs]
        loop_maps = [DeviceTree(ctx, lp["fixed"]) for lp in loops]
        keep = seq_maps + loop_maps
        for T in keep:
            srv.add_resident(T)
        p_ct = icp_params(max_iterations=15, velocity_damping=10.0)
        p_icp = icp_params()

        def submit(kind, idx, frame):
            if kind == "odom":
                q, f = seqs[idx], seqs[idx]["frames"][frame]
                return srv.submit(seq_maps[idx], f["moving"], f["X0"], q["Xf"], p_ct, ct=1, v0=f["v0"])
            lp = loops[idx]
            return srv.submit(loop_maps[idx], lp["moving"], lp["X0"], lp["Xf"], p_icp, ct=0)

        rng = np.random.default_rng(np.random.PCG64(5999))
        seconds = 1.5
        events = []
        for s in range(len(seqs)):
            for k, t in enumerate(synth.poisson_arrivals(10.0, seconds, rng)):
                events.append((t, "odom", s, k % len(seqs[s]["frames"])))
        for t in synth.poisson_arrivals(20.0, seconds, rng):
            events.append((t, "loop", int(rng.integers(len(loops))), 0))
        events.sort()
        assert len(events) > 30
        submitted = []
        lock = threading.Lock()

        def submitter():
            t0 = time.perf_counter()
            for t, kind, idx, frame in events:
                now = time.perf_counter() - t0
                if t > now:
                    time.sleep(t - now)
                rid = submit(kind, idx, frame)
                with lock:
                    submitted.append((rid, kind, idx, frame))

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        mixed = [(kind, idx, frame, srv.wait(rid, 60.0)) for rid, kind, idx, frame in submitted]
        lat = [r["latency_us"] / 1e3 for kind, _, _, r in mixed if kind == "odom"]
        p99 = float(np.percentile(lat, 99))
        print(f"{len(mixed)} requests, odometry p50/p99 {np.percentile(lat, 50):.3f}/{p99:.3f} ms")
        assert all(r["status"] == 0 for _, _, _, r in mixed)
        assert p99 < 10.0
        # every distinct request served alone, one at a time
        alone = {}
        for kind, idx, frame, _ in mixed:
            key = (kind, idx, frame)
            if key not in alone:
                alone[key] = srv.wait(submit(*key), 60.0)
        for kind, idx, frame, r in mixed:
            a = alone[(kind, idx, frame)]
            assert np.array_equal(a["X"], r["X"]) and np.array_equal(a["v"], r["v"]), (kind, idx, frame)
            assert a["result"] == r["result"] and a["last"] == r["last"]
        # a sample against the oracle
        checked = 0
        for (kind, idx, frame), a in sorted(alone.items())[::3]:
            if kind == "odom":
                q, f = seqs[idx], seqs[idx]["frames"][frame]
                Xo, vo, _, res_o = O.compute(q["fixed"], f["moving"], f["X0"], q["Xf"], O.from_kdh_params(p_ct, 8),
                                             ct=1, v0=f["v0"])
                assert np.abs(a["v"] - vo).max() < 1e-2
            else:
                lp = loops[idx]
                Xo, _, _, res_o = O.compute(lp["fixed"], lp["moving"], lp["X0"], lp["Xf"],
                                            O.from_kdh_params(p_icp, 8))
            dt, dr = synth.pose_error(a["X"], Xo)
            assert a["result"] == res_o
            assert dt < POSE_T_TOL and dr < POSE_R_TOL, (kind, idx, frame, dt, dr)
            checked += 1
        assert checked >= 4
    finally:
        srv.close()
        for T in keep:
            T.close()
