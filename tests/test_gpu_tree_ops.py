"""Tree maintenance on the device (tree/tree_cuda_impl_.h:43-173): in-place
deskew, rigid transform and compressed-leaf extraction, bit for bit against
the oracle's restatement, and the descent over a moved tree.  GPU."""
import numpy as np
import pytest

from kd_slam2_amd import DeviceTree, HostTree, ICPHIP3D, synth, update_cache_f32
from oracle import pyoracle as O

from .workloads import cfg1, cfg2, icp_params

pytestmark = pytest.mark.gpu

V = np.array([8.0, 0.3, -0.1, 0.01, -0.02, 0.5], dtype=np.float32)


def _nodes6(nodes):
    return np.stack([nodes["mean"], nodes["direction"]], axis=1).astype(np.float32)


@pytest.mark.parametrize("skip_leaves", [False, True])
def test_apply_velocities_bit_exact(ctx, skip_leaves):
    w = cfg2()
    t = w["moving"]
    D = DeviceTree(ctx, t)
    D.applyVelocitiesInPlace(V, skip_leaves=skip_leaves)
    ref = O.tree_apply_velocities(t.nodes, V, skip_leaves)
    assert np.array_equal(D.readNodes(), _nodes6(ref))
    got = D.extractCompressedLeaves()
    assert np.array_equal(got[:, 0], ref["mean"][t.leaves]) and np.array_equal(got[:, 1], ref["direction"][t.leaves])
    D.close()


def test_apply_isometry_and_extract_bit_exact(ctx):
    w = cfg2()
    t = w["moving"]
    D = DeviceTree(ctx, t)
    X = synth.expmap(np.array([12.0, -3.0, 1.5, 0.05, -0.1, 1.2]))
    D.applyIsometryInPlace(X)
    ref = O.tree_apply_isometry(t.nodes, X)
    assert np.array_equal(D.readNodes(), _nodes6(ref))
    assert np.array_equal(D.extractCompressedLeaves(V), O.extract_leaves(ref, t.leaves, V))
    D.close()


def test_extract_with_velocities_bit_exact(ctx):
    t = cfg2()["moving"]
    D = DeviceTree(ctx, t)
    assert np.array_equal(D.extractCompressedLeaves(), O.extract_leaves(t.nodes, t.leaves))
    assert np.array_equal(D.extractCompressedLeaves(V), O.extract_leaves(t.nodes, t.leaves, V))
    D.close()


def test_descent_over_a_moved_fixed_tree_matches_the_oracle(ctx):
    """The fixed tree moved on the device descends exactly like the oracle's
    moved copy: node records and leaf records stay consistent."""
    w = cfg1()
    F = DeviceTree(ctx, w["fixed"])
    M = DeviceTree(ctx, w["moving"])
    T = synth.expmap(np.array([0.7, -0.4, 0.2, 0.02, 0.01, 0.3]))
    F.applyIsometryInPlace(T)
    moved = HostTree(O.tree_apply_isometry(w["fixed"].nodes, T), w["fixed"].leaves)
    Xf = w["Xf"] @ np.linalg.inv(T)  # same world geometry
    Xm, Xfc, Xmf = update_cache_f32(w["X0"], Xf)
    al = ICPHIP3D(ctx)
    al.params = icp_params()
    al.setFixed(F)
    al.setMoving(M)
    leaf, depth = al.correspondences(Xm, Xfc, Xmf)
    ref = O.build_form(moved, w["moving"], Xm, Xfc, Xmf, O.from_kdh_params(icp_params(), 8))
    assert np.array_equal(leaf, ref["corr"]) and np.array_equal(depth, ref["depth"])


def test_moved_moving_tree_aligns_to_the_same_world_pose(ctx):
    """A moving tree moved by T, started at X0 T^-1, lands on the original
    result: the re-centring anchor follows the moved root."""
    w = cfg1()
    F = DeviceTree(ctx, w["fixed"])
    out = []
    for T in (np.eye(4), synth.expmap(np.array([2.0, 1.0, -0.5, 0.0, 0.0, 0.8]))):
        M = DeviceTree(ctx, w["moving"])
        M.applyIsometryInPlace(T)
        al = ICPHIP3D(ctx)
        al.params = icp_params()
        al.setFixed(F)
        al.setMoving(M)
        al.setX(w["X0"] @ np.linalg.inv(T))
        al.setFixedX(w["Xf"])
        al.computeOnDevice()
        out.append(al.getX() @ T)
    t, r = synth.pose_error(out[0], out[1])
    assert t < 1e-3 and r < 1e-2
