"""Batched builds (kdh_batch_build): the per-factor quadratic forms of a
bundle adjustment in one launch.  Each form is bit-identical to the same
build alone (kdh_build_at), in any mix of rigid and CT pairs, and within
the FP64-twin tolerance of the oracle.  GPU."""
import numpy as np
import pytest

from kd_slam2_amd import CTICPHIP3D, DeviceTree, ICPHIP3D, batch_build, synth
from oracle import pyoracle as O

from .workloads import cfg1, cfg2, icp_params, rel_err

pytestmark = pytest.mark.gpu

HB_TOL = 2e-5


def _alone(ctx, cls, F, M, X, Xf, v, params, stats_mode=False):
    al = cls(ctx)
    al.params = params
    al.setFixed(F)
    al.setMoving(M)
    al.setX(X)
    al.setFixedX(Xf)
    if cls is CTICPHIP3D:
        al.setVelocities(v)
    al.buildQuadraticFormAt(stats_mode)
    return al.H.copy(), al.b.copy(), dict(al.stats)


@pytest.mark.parametrize("stats_mode", [False, True])
def test_batch_build_matches_builds_alone(ctx, stats_mode):
    w1, w2 = cfg1(), cfg2()
    F1, M1 = DeviceTree(ctx, w1["fixed"]), DeviceTree(ctx, w1["moving"])
    F2, M2 = DeviceTree(ctx, w2["fixed"]), DeviceTree(ctx, w2["moving"])
    rng = np.random.default_rng(np.random.PCG64(77))
    p1 = icp_params()
    p2 = icp_params(max_error=1.0, velocity_damping=10.0)
    pairs, alone = [], []
    for k in range(6):
        if k % 2 == 0:
            X = synth.perturb(w1["X_gt"], rng, 1.5, 0.2)
            pairs.append(dict(fixed=F1, moving=M1, ct=0, X0=X, Xf=w1["Xf"], params=p1))
            alone.append(_alone(ctx, ICPHIP3D, F1, M1, X, w1["Xf"], None, p1, stats_mode))
        else:
            X = synth.perturb(w2["X_gt"], rng, 0.5, 0.2)
            pairs.append(dict(fixed=F2, moving=M2, ct=1, X0=X, Xf=w2["Xf"], v0=w2["v0"], params=p2))
            alone.append(_alone(ctx, CTICPHIP3D, F2, M2, X, w2["Xf"], w2["v0"], p2, stats_mode))
    got = batch_build(ctx, pairs, stats_mode=stats_mode)
    for (H, b, st), (Ha, ba, sa) in zip(got, alone):
        assert np.array_equal(H, Ha) and np.array_equal(b, ba)
        for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn", "chi2", "chi2_kernelized"):
            assert st[k] == sa[k], k
    if stats_mode:
        return
    # and against the oracle's FP64 twin at the same state (rigid pairs)
    H, b, st = got[0]
    ref = O.build_form_at(w1["fixed"], w1["moving"], pairs[0]["X0"], w1["Xf"], O.from_kdh_params(p1, 8))
    assert rel_err(H, ref["H64"]) < HB_TOL and rel_err(b, ref["b64"]) < HB_TOL


def test_batch_built_normal_matrix_is_symmetric_psd(ctx):
    """H = sum g J^T J: symmetric and positive semi-definite after the FP64
    re-centring map-back."""
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params()
    H, b, st = batch_build(ctx, [dict(fixed=F, moving=M, ct=0, X0=w["X_gt"], Xf=w["Xf"], params=p)])[0]
    assert st["num_inliers"] > 100
    assert np.allclose(H, H.T, rtol=1e-12, atol=1e-9)
    ev = np.linalg.eigvalsh(H)
    assert ev.min() > -1e-6 * ev.max()
