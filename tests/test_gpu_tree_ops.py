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


def test_clone_is_a_deep_copy(ctx):
    """TreeCUDA_::clone: the same records, and independent of the source."""
    t = cfg2()["moving"]
    D = DeviceTree(ctx, t)
    B = D.clone()
    assert B.device_bytes() == D.device_bytes() and B.num_leaves == D.num_leaves
    assert np.array_equal(B.readNodes(), D.readNodes())
    assert np.array_equal(B.extractCompressedLeaves(), D.extractCompressedLeaves())
    D.applyVelocitiesInPlace(V)
    assert np.array_equal(B.readNodes(), _nodes6(t.nodes))
    assert not np.array_equal(D.readNodes(), B.readNodes())
    B.close()
    D.close()


def test_bundle_adjustment_backup_deskew_restore_cycle(ctx):
    """optimizer_proc_solver_action_.h:39-58 with map/map_.h:88-102: save()
    clones the keyframe tree; every solver iteration restores the nodes from
    the backup (copyNodes) and re-deskews them with the current velocities,
    skipping the leaves.  Each deskew is bit-identical to the oracle's on the
    original nodes, each restore gives the original back, and an alignment on
    the restored tree is bit-identical to one on a fresh upload."""
    w = cfg2()
    t = w["moving"]
    D = DeviceTree(ctx, t)
    backup = D.clone()
    rng = np.random.default_rng(5)
    for k in range(3):
        backup.copyNodes(D)
        assert np.array_equal(D.readNodes(), _nodes6(t.nodes))
        v = (V * rng.uniform(0.5, 1.5, 6)).astype(np.float32)
        D.applyVelocitiesInPlace(v, skip_leaves=True)
        ref = O.tree_apply_velocities(t.nodes, v, True)
        assert np.array_equal(D.readNodes(), _nodes6(ref)), k
    backup.copyNodes(D)
    F = DeviceTree(ctx, w["fixed"])
    fresh = DeviceTree(ctx, t)
    p = icp_params(max_iterations=5, velocity_damping=10.0)
    out = []
    for M in (D, fresh):
        from kd_slam2_amd import CTICPHIP3D

        al = CTICPHIP3D(ctx)
        al.params = p
        al.setFixed(F)
        al.setMoving(M)
        al.setX(w["X0"])
        al.setFixedX(w["Xf"])
        al.setVelocities(w["v0"])
        al.computeOnDevice()
        out.append((al.getX(), al.velocities()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    # copyNodes between trees of different shape is refused ("size mismatch")
    other = DeviceTree(ctx, cfg1()["moving"])
    with pytest.raises(RuntimeError):
        backup.copyNodes(other)
    for T in (D, backup, F, fresh, other):
        T.close()


def test_device_built_full_records_follow_in_place_ops(ctx):
    """A device-built tree keeps its full Node_ records in step with the
    compact records through an in-place deskew and transform, and a clone
    carries them."""
    from kd_slam2_amd import TreeGenerator3D

    w = cfg2()
    g = TreeGenerator3D()
    D = g.makeTreeOnDevice(ctx, w["scan"], w["stamps"])
    D.applyVelocitiesInPlace(V)
    D.applyIsometryInPlace(synth.expmap(np.array([3.0, -1.0, 0.5, 0.02, -0.01, 0.3])))
    full = D.downloadNodes()
    assert np.array_equal(_nodes6(full), D.readNodes())
    C2 = D.clone()
    assert np.array_equal(C2.downloadNodes(), full)
    C2.close()
    D.close()
