"""Fixed forest (ICP_::_fixed_forest + makeFixedEntry, icp_impl_.h:53-71):
one moving scan against several fixed trees with their own poses, H, b and
stats summed each round.  Device forest (kdh_batch_align groups), per-call
forest (host loop over kdh_icp_build) and the FP32 oracle forest.  GPU."""
import numpy as np
import pytest

from kd_slam2_amd import DeviceTree, ICPHIP3D, batch_align, synth
from oracle import pyoracle as O

from .workloads import forest_cfg1, icp_params

pytestmark = pytest.mark.gpu

POSE_TOL_M = 1e-3
POSE_TOL_DEG = 1e-3


def _aligner(ctx, f, M, params):
    al = ICPHIP3D(ctx)
    al.params = params
    al.setMoving(M)
    al.setX(f["w"]["X0"])
    for ref, (T, Xf) in enumerate(zip(f["dev"], f["poses"])):
        al._fixed_forest[ref] = al.makeFixedEntry(T, Xf)
    return al


@pytest.fixture(scope="module")
def forest(ctx):
    f = dict(forest_cfg1())
    f["dev"] = [DeviceTree(ctx, t) for t in f["trees"]]
    f["M"] = DeviceTree(ctx, f["w"]["moving"])
    return f


def test_forest_device_and_per_call_match_oracle(ctx, forest):
    f, w = forest, forest["w"]
    p = icp_params()
    Xo, _, log, res, ent = O.compute_forest(f["trees"], f["poses"], w["moving"], w["X0"], O.from_kdh_params(p, 8))
    dev = _aligner(ctx, f, f["M"], icp_params())
    r_dev = dev.computeOnDevice()
    host = _aligner(ctx, f, f["M"], icp_params())
    r_host = host.compute()
    assert r_dev == res == r_host == p.max_iterations
    for X in (dev.getX(), host.getX()):
        t, r = synth.pose_error(X, Xo)
        assert t < POSE_TOL_M and r < POSE_TOL_DEG
    t, r = synth.pose_error(dev.getX(), w["X_gt"])
    assert t < 0.06 and r < 0.4
    # the entries' shares add up to the aligner's stats
    for al in (dev, host):
        for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
            assert sum(e.stats[k] for e in al._fixed_forest.values()) == al.stats[k]
        assert all(e.stats["num_inliers"] > 0 for e in al._fixed_forest.values())


def test_forest_entry_stats_are_the_per_entry_decisions(ctx, forest):
    """One round: each entry's share equals the oracle's decisions for that
    entry at the cache the device forms from the same FP64 state."""
    f, w = forest, forest["w"]
    p = icp_params(max_iterations=1)
    dev = _aligner(ctx, f, f["M"], p)
    dev.computeOnDevice()
    for ref, (T, D, Xf) in enumerate(zip(f["trees"], f["dev"], f["poses"])):
        one = ICPHIP3D(ctx)
        one.params = icp_params(max_iterations=1)
        one.setFixed(D)
        one.setMoving(f["M"])
        one.setX(w["X0"])
        one.setFixedX(Xf)
        one.computeOnDevice(cache_log=True)
        Xm, Xfc, Xmf = (one.cache_log[0][12 * i:12 * (i + 1)] for i in range(3))
        ref_form = O.build_form(T, w["moving"], Xm, Xfc, Xmf, O.from_kdh_params(p, 8))
        got = dev._fixed_forest[ref].stats
        for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
            assert got[k] == ref_form["stats"][k], (ref, k)


def test_single_entry_forest_is_the_plain_alignment(ctx, forest):
    f, w = forest, forest["w"]
    p = icp_params()
    base = dict(fixed=f["dev"][0], moving=f["M"], ct=0, X0=w["X0"], Xf=f["poses"][0], params=p)
    plain = batch_align(ctx, [base])[0]
    one = batch_align(ctx, [dict(base, forest=0)])[0]
    assert np.array_equal(plain["X"], one["X"]) and plain["last"] == one["entry"]


def test_forest_identical_alone_and_in_a_mix(ctx, forest):
    f, w = forest, forest["w"]
    p = icp_params()
    group = [dict(fixed=T, moving=f["M"], ct=0, X0=w["X0"], Xf=Xf, params=p, forest=len(f["dev"]) - 1 if i == 0 else 0)
             for i, (T, Xf) in enumerate(zip(f["dev"], f["poses"]))]
    other = dict(fixed=f["dev"][1], moving=f["M"], ct=0, X0=w["X0"], Xf=f["poses"][1], params=p)
    alone = batch_align(ctx, group)
    mix = batch_align(ctx, [other] + group + [other])
    for a, b in zip(alone, mix[1:-1]):
        assert np.array_equal(a["X"], b["X"]) and a["last"] == b["last"] and a["entry"] == b["entry"]
    assert np.array_equal(mix[0]["X"], mix[-1]["X"])


def test_forest_errors(ctx, forest):
    f, w = forest, forest["w"]
    p = icp_params()
    bad = [dict(fixed=f["dev"][0], moving=f["M"], ct=0, X0=w["X0"], Xf=f["poses"][0], params=p, forest=2),
           dict(fixed=f["dev"][1], moving=f["M"], ct=0, X0=w["X0"], Xf=f["poses"][1], params=p)]
    with pytest.raises(RuntimeError):
        batch_align(ctx, bad)  # runs past the last pair
    bad[0]["forest"] = 1
    bad[1]["ct"] = 1
    with pytest.raises(RuntimeError):
        batch_align(ctx, bad)  # entry of another kind
