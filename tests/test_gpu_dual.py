"""CT dual form (kdh_batch_build_dual): the bundle-adjustment factor between
two CT sweeps, with the fixed side's velocity blocks.  Against the oracle's
buildQuadraticFormDual at the same FP64 states: the outcome counts exactly,
every block within the FP64-twin tolerance (re-centred mode) or the FP32
reference's summation-order tolerance (reference mode).  Batched factors are
bit-identical to each factor alone, and the moving-side blocks are the CT
build's.  GPU."""
import numpy as np
import pytest

from kd_slam2_amd import CTICPHIP3D, DeviceTree, batch_build_dual
from kd_slam2_amd.api import adapter_build_dual
from oracle import pyoracle as O

from .workloads import dual_factors, icp_params, rel_err

pytestmark = pytest.mark.gpu

# FP32 terms summed per tile, tiles summed in FP64, against the FP64 twin
BLOCK_TOL = 2e-5
BLOCKS = (("H", "H"), ("b", "b"), ("H_dual_vel_B", "H_vel_B"), ("b_dual_vel_B", "b_vel_B"),
          ("H_dual_cross_pose_vel_B", "H_cross_pose_vel_B"), ("H_dual_cross_vel_AB", "H_cross_vel_AB"))
COUNTS = ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn")


def _params(**kw):
    return icp_params(max_error=1.0, velocity_damping=10.0, **kw)


def _device_factors(ctx, params):
    trees = {}
    out = []
    for f in dual_factors():
        for k in ("fixed", "moving"):
            if id(f[k]) not in trees:
                trees[id(f[k])] = DeviceTree(ctx, f[k])
        out.append(dict(f, fixed=trees[id(f["fixed"])], moving=trees[id(f["moving"])], params=params))
    return out, trees


def _oracle(f, params, stats_mode=False):
    return O.build_form_dual(f["fixed"], f["moving"], f["X_moving"], f["v_moving"], f["X_fixed"], f["v_fixed"],
                             O.from_kdh_params(params, 8), stats_mode=stats_mode)


@pytest.mark.parametrize("stats_mode", [False, True])
def test_dual_forms_match_oracle_fp64_twin(ctx, stats_mode):
    p = _params()
    fac, _keep = _device_factors(ctx, p)
    got = batch_build_dual(ctx, fac, stats_mode=stats_mode)
    for k, (g, f) in enumerate(zip(got, dual_factors())):
        ref = _oracle(f, p, stats_mode)
        for c in COUNTS:
            assert g["stats"][c] == ref["stats"][c], (k, c)
        assert g["stats"]["num_inliers"] > 300
        for mine, theirs in BLOCKS:
            e = rel_err(g[mine], ref[theirs + "64"])
            assert e < BLOCK_TOL, (k, mine, e)
        assert abs(g["stats"]["chi2"] - ref["stats"]["chi2"]) <= 1e-5 * ref["stats"]["chi2"]
        # J_vel_B is taken at the fixed leaves' stamps: the blocks are live
        assert np.abs(g["H_dual_vel_B"]).max() > 1.0
        assert np.allclose(g["H_dual_vel_B"], g["H_dual_vel_B"].T, rtol=1e-12, atol=1e-9)


def test_reference_mode_matches_fp32_oracle(ctx):
    """recenter=0 evaluates the reference's world-coordinate formulas: every
    block matches the FP32 oracle to summation-order error."""
    p = _params(recenter=0)
    fac, _keep = _device_factors(ctx, p)
    got = batch_build_dual(ctx, fac)
    for k, (g, f) in enumerate(zip(got, dual_factors())):
        ref = _oracle(f, p)
        for c in COUNTS:
            assert g["stats"][c] == ref["stats"][c], (k, c)
        for mine, theirs in BLOCKS:
            e = rel_err(g[mine], ref[theirs])
            assert e < 2e-5, (k, mine, e)


def test_batched_factors_equal_factors_alone(ctx):
    p = _params()
    fac, _keep = _device_factors(ctx, p)
    together = batch_build_dual(ctx, fac)
    for k, f in enumerate(fac):
        alone = batch_build_dual(ctx, [f])[0]
        for mine, _ in BLOCKS:
            assert np.array_equal(together[k][mine], alone[mine]), (k, mine)
        assert together[k]["stats"] == alone["stats"]


def test_dual_moving_side_is_the_ct_build(ctx):
    """Outside stats mode the dual weight rule is quadraticTerm's: H and b of
    the dual build are the CT build's at the same state, bit for bit, through
    the aligner's buildQuadraticFormDual."""
    p = _params()
    fac, _keep = _device_factors(ctx, p)
    f = fac[0]
    al = CTICPHIP3D(ctx)
    al.params = p
    al.setFixed(f["fixed"])
    al.setMoving(f["moving"])
    al.setX(f["X_moving"])
    al.setVelocities(f["v_moving"])
    al.setFixedX(f["X_fixed"])
    al.buildQuadraticFormAt()
    H, b, st = al.H.copy(), al.b.copy(), dict(al.stats)
    al.setFixedVelocities(f["v_fixed"])
    al.buildQuadraticFormDual()
    assert np.array_equal(al.H, H) and np.array_equal(al.b, b)
    for c in COUNTS:
        assert al.stats[c] == st[c]
    assert al.H_dual_cross_vel_AB.shape == (6, 6) and al.b_dual_vel_B.shape == (6,)


def test_plugin_build_dual_is_the_c_abi_build(ctx):
    """CTICPHIP3D::buildQuadraticFormDual of the C++ plugin gives the
    kdh_batch_build_dual blocks bit for bit."""
    p = _params()
    fac, _keep = _device_factors(ctx, p)
    f = fac[1]
    ref = batch_build_dual(ctx, [f])[0]
    got = adapter_build_dual(ctx, f["fixed"], f["moving"], f["X_moving"], f["v_moving"], f["X_fixed"], f["v_fixed"], p)
    for mine, _ in BLOCKS:
        assert np.array_equal(got[mine].reshape(ref[mine].shape), ref[mine]), mine
    for c in COUNTS:
        assert got["stats"][c] == ref["stats"][c]
