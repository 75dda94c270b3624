"""The tile kernel's other shapes: two moving leaves per lane (two descents
in flight per lane, kdh_set_leaves_per_lane(ctx, 2)) and 128-leaf tiles
(kdh_set_tile_leaves(ctx, 128)): correspondences bit-exact against the
oracle, whole alignments within the pose bar, batched results the same alone
and in a mix, prepared batches equal to one-shot ones; the setters' limits.
GPU."""
import numpy as np
import pytest

from kd_slam2_amd import CTICPHIP3D, Context, DeviceTree, ICPHIP3D, batch_align, synth, update_cache_f32
from oracle import pyoracle as O

from .workloads import cfg1, cfg2, icp_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[(2, 256), (1, 128)], ids=["2_leaves_per_lane", "128_leaf_tiles"])
def ctx2(request):
    lpl, tile = request.param
    c = Context(0)
    c.set_launch_config(grid=0, lds_nodes=256, leaves_per_lane=lpl, tile_leaves=tile)
    yield c
    c.close()


@pytest.mark.parametrize("ct", [0, 1])
def test_correspondences_and_align_two_leaves_per_lane(ctx2, ct):
    w = cfg2() if ct else cfg1()
    F, M = DeviceTree(ctx2, w["fixed"]), DeviceTree(ctx2, w["moving"])
    p = icp_params(max_iterations=15 if ct else 10, velocity_damping=10.0)
    al = (CTICPHIP3D if ct else ICPHIP3D)(ctx2)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    if ct:
        al.setVelocities(w["v0"])
    Xm, Xf, Xmf = update_cache_f32(w["X0"], w["Xf"])
    v = np.asarray(w["v0"], dtype=np.float32) if ct else None
    leaf, depth = al.correspondences(Xm, Xf, Xmf, v) if ct else al.correspondences(Xm, Xf, Xmf)
    ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, O.from_kdh_params(p, 8), ct=ct, v=v)
    assert np.array_equal(leaf, ref["corr"]) and np.array_equal(depth, ref["depth"])
    al.buildQuadraticForm()
    for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
        assert al.stats[k] == ref["stats"][k], k
    res = al.computeOnDevice()
    Xo, vo, _, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 8), ct=ct,
                                 v0=w["v0"] if ct else None)
    dt, dr = synth.pose_error(al.getX(), Xo)
    assert res == res_o and dt < 1e-3 and dr < 1e-3, (dt, dr)
    F.close()
    M.close()


def test_batch_alone_equals_mixed_two_leaves_per_lane(ctx2):
    ws = []
    for i, n in enumerate([3000, 20000, 7000]):
        w = synth.config1(seed=3100 + i, n_map=3 * n, n_scan=n, rot_deg=5.0, trans_m=1.0, uniform_init=True)
        from kd_slam2_amd import TreeGenerator3D

        g = TreeGenerator3D(leaf_max_points=16)
        w["fixed"], w["moving"] = g.makeTree(w["map"]), g.makeTree(w["scan"])
        ws.append(w)
    trees = [(DeviceTree(ctx2, w["fixed"]), DeviceTree(ctx2, w["moving"])) for w in ws]
    p = icp_params()
    pairs = [dict(fixed=F, moving=M, ct=0, X0=w["X0"], Xf=w["Xf"], params=p) for w, (F, M) in zip(ws, trees)]
    mixed = batch_align(ctx2, pairs)
    from kd_slam2_amd import PairBatch

    B = PairBatch(pairs)
    prepared = list(B.run(ctx2))
    for a, b in zip(mixed, prepared):
        assert np.array_equal(a["X"], b["X"]) and a["last"] == b["last"] and a["result"] == b["result"]
    for i in range(len(pairs)):
        alone = batch_align(ctx2, [pairs[i]])[0]
        assert np.array_equal(alone["X"], mixed[i]["X"]) and alone["last"] == mixed[i]["last"]
        Xo, _, _, _ = O.compute(ws[i]["fixed"], ws[i]["moving"], ws[i]["X0"], ws[i]["Xf"], O.from_kdh_params(p))
        dt, dr = synth.pose_error(mixed[i]["X"], Xo)
        assert dt < 1e-3 and dr < 1e-3
    for F, M in trees:
        F.close()
        M.close()


def test_tile_shape_limits():
    """Two leaves per lane take 256-leaf tiles and 128-leaf tiles one leaf per
    lane; other sizes are refused; a batch prepared at one tile size is
    refused at the other."""
    from kd_slam2_amd import PairBatch, TreeGenerator3D

    c = Context(0)
    try:
        with pytest.raises(RuntimeError):
            c.set_launch_config(tile_leaves=64)
        c.set_launch_config(leaves_per_lane=2)
        with pytest.raises(RuntimeError):
            c.set_launch_config(tile_leaves=128)
        c.set_launch_config(leaves_per_lane=1, tile_leaves=128)
        with pytest.raises(RuntimeError):
            c.set_launch_config(leaves_per_lane=2)
        w = synth.config1(seed=3200, n_map=12000, n_scan=4000, rot_deg=5.0, trans_m=1.0, uniform_init=True)
        g = TreeGenerator3D(leaf_max_points=16)
        F, M = DeviceTree(c, g.makeTree(w["map"])), DeviceTree(c, g.makeTree(w["scan"]))
        B = PairBatch([dict(fixed=F, moving=M, ct=0, X0=w["X0"], Xf=w["Xf"], params=icp_params())])
        B.run(c)
        c.set_launch_config(tile_leaves=256)
        with pytest.raises(RuntimeError, match="tile size"):
            B.run(c)
        B.close()
        F.close()
        M.close()
    finally:
        c.close()
