"""Tree generation on the device (kdh_tree_build) against the oracle's
restatement of makeTree: the same Node_ records, leaf list and permuted
points bit for bit, on scans, maps and degenerate clouds, from host or
device inputs, and at the full sizes of the CT workload.  Alignments over
device-built trees are the alignments over host-built ones.  GPU."""
import ctypes as C

import numpy as np
import pytest

from kd_slam2_amd import CTICPHIP3D, DeviceTree, TreeGenerator3D, synth
from kd_slam2_amd import _lib as L
from oracle import pyoracle as O

from .workloads import icp_params, tree_clouds

pytestmark = pytest.mark.gpu


def _same(dev, ref, name):
    assert (dev is None) == (ref is None), name
    if ref is None:
        return
    h = dev.host
    assert len(h.nodes) == len(ref["nodes"]), name
    assert h.nodes.tobytes() == ref["nodes"].tobytes(), name
    assert np.array_equal(h.leaves, ref["leaves"]), name
    assert np.array_equal(h.points, ref["points"]), name
    assert np.array_equal(dev.stamps, ref["stamps"]), name


def test_device_build_equals_oracle_on_every_cloud(ctx):
    for name, pts, st, kw in tree_clouds():
        print(f"cloud: {name} ({len(pts)} points)", flush=True)
        ref = O.tree_build(pts, st, **kw)
        dev = TreeGenerator3D(**kw).makeTreeOnDevice(ctx, pts, st)
        _same(dev, ref, name)


def test_device_build_full_size_ct_workload(ctx):
    """The 131k-point sweep and the 2M-point map of the CT workload."""
    w = synth.config2()
    g = TreeGenerator3D(leaf_max_points=16)
    for name, pts, st in (("sweep", w["scan"], w["stamps"]), ("map", w["map"], None)):
        print(f"cloud: {name} ({len(pts)} points)", flush=True)
        _same(g.makeTreeOnDevice(ctx, pts, st), O.tree_build(pts, st, leaf_max_points=16), name)


def test_workspace_reuse_small_large_small(ctx):
    clouds = tree_clouds()
    order = [clouds[4], clouds[1], clouds[0], clouds[4], clouds[7]]
    for name, pts, st, kw in order:
        _same(TreeGenerator3D(**kw).makeTreeOnDevice(ctx, pts, st), O.tree_build(pts, st, **kw), name)


def test_device_inputs(ctx):
    """Points and stamps already in HBM (inputs_on_device = 1).  The device
    buffers come from the HIP runtime the library itself links."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    name, pts, st, kw = tree_clouds()[0]
    x = np.ascontiguousarray(pts, dtype=np.float32)
    t = np.ascontiguousarray(st, dtype=np.float64)
    dx, dt = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(dx), x.nbytes) == 0 and hip.hipMalloc(C.byref(dt), t.nbytes) == 0
    try:
        assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, 1) == 0  # host to device
        assert hip.hipMemcpy(dt, t.ctypes.data, t.nbytes, 1) == 0
        gen = TreeGenerator3D(**kw)
        h = C.c_void_p()
        ctx.check(L.lib().kdh_tree_build(ctx.h, C.cast(dx, L.F32P), C.cast(dt, L.F64P), len(x), C.byref(gen._p), 1,
                                         C.byref(h)), "kdh_tree_build")
    finally:
        hip.hipFree(dx)
        hip.hipFree(dt)
    nn, nl = C.c_int64(), C.c_int64()
    L.lib().kdh_tree_info(h, C.byref(nn), C.byref(nl), None)
    nodes = np.zeros(nn.value, dtype=L.NODE_DTYPE)
    leaves = np.zeros(nl.value, dtype=np.uint64)
    ctx.check(L.lib().kdh_tree_download(ctx.h, h, nodes.ctypes.data_as(C.c_void_p),
                                        leaves.ctypes.data_as(C.c_void_p), None, None), "kdh_tree_download")
    L.lib().kdh_tree_free(ctx.h, h)
    ref = O.tree_build(pts, st, **kw)
    assert nodes.tobytes() == ref["nodes"].tobytes() and np.array_equal(leaves, ref["leaves"])


def test_alignment_over_device_built_trees(ctx):
    """A CT alignment whose map and sweep trees were built on the device ends
    at the pose of the same alignment over host-built trees, bit for bit."""
    w = synth.config2(n_map=400_000, n_scan=65_536)
    g = TreeGenerator3D(leaf_max_points=16)
    p = icp_params(max_error=1.0, velocity_damping=10.0, max_iterations=8)
    out = []
    for trees in (
        (DeviceTree(ctx, g.makeTree(w["map"])), DeviceTree(ctx, g.makeTree(w["scan"], w["stamps"]))),
        (g.makeTreeOnDevice(ctx, w["map"], download=False), g.makeTreeOnDevice(ctx, w["scan"], w["stamps"],
                                                                                download=False)),
    ):
        al = CTICPHIP3D(ctx)
        al.params = p
        al.setFixed(trees[0])
        al.setMoving(trees[1])
        al.setX(w["X0"])
        al.setVelocities(w["v0"])
        al.setFixedX(w["Xf"])
        al.computeOnDevice()
        out.append((al.getX(), al.velocities(), dict(al.stats)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]
    t, r = synth.pose_error(out[1][0], w["X_gt"])
    t0, r0 = synth.pose_error(w["X0"], w["X_gt"])
    assert t < t0 / 3 and r < r0 / 3


def test_no_tree_cases(ctx):
    g = TreeGenerator3D()
    assert g.makeTreeOnDevice(ctx, np.zeros((0, 3), dtype=np.float32)) is None
    assert g.makeTreeOnDevice(ctx, np.ones((4, 3), dtype=np.float32)) is None
    rng = np.random.default_rng(5)
    few = rng.normal(size=(40, 3)).astype(np.float32)
    assert (g.makeTreeOnDevice(ctx, few) is None) == (O.tree_build(few) is None)


def test_tree_from_a_kdtree3f_message_aligns_identically(ctx):
    """A device-built sweep tree read back, sent through the KDTree3f wire
    format and uploaded again gives the same alignment bit for bit."""
    from kd_slam2_amd import kdtree_msg as KM

    w = synth.config2(n_map=200_000, n_scan=32_768)
    g = TreeGenerator3D(leaf_max_points=16)
    F = g.makeTreeOnDevice(ctx, w["map"])
    M = g.makeTreeOnDevice(ctx, w["scan"], w["stamps"])
    M2 = DeviceTree(ctx, KM.tree_from_message(KM.decode(KM.encode(KM.message_from_tree(M.host, frame_id="base")))))
    p = icp_params(max_error=1.0, velocity_damping=10.0, max_iterations=6)
    out = []
    for moving in (M, M2):
        al = CTICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(moving)
        al.setX(w["X0"])
        al.setVelocities(w["v0"])
        al.setFixedX(w["Xf"])
        al.computeOnDevice()
        out.append((al.getX(), al.velocities()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_reference_arithmetic_trees_on_the_device(ctx):
    """Trees built with the reference's own arithmetic (label_mode 1, host
    builder) align on the device like any other: a rigid alignment against
    the FP32 oracle on the same trees, pose within 1 mm / 1e-3 deg and the
    same result code and per-round counters.  The device builder refuses the
    mode with an error naming the host builder."""
    from kd_slam2_amd import ICPHIP3D

    w = synth.config1(seed=4, n_map=120_000, n_scan=12_000)
    g = TreeGenerator3D(leaf_max_points=16, label_mode=1)
    Fh, Mh = g.makeTree(w["map"]), g.makeTree(w["scan"])
    p = icp_params(max_iterations=10)
    F, M = DeviceTree(ctx, Fh), DeviceTree(ctx, Mh)
    al = ICPHIP3D(ctx)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    res = al.computeOnDevice()
    Xo, _, log, res_o = O.compute(Fh, Mh, w["X0"], w["Xf"], O.from_kdh_params(p, 8))
    dt, dr = synth.pose_error(al.getX(), Xo)
    assert dt < 1e-3 and dr < 1e-3 and res == res_o
    for a, b in zip(al.stats_log, log):
        assert all(a[k] == b[k] for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"))
    with pytest.raises(RuntimeError, match="host"):
        g.makeTreeOnDevice(ctx, w["scan"])
    F.close()
    M.close()
