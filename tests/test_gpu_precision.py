"""Precision far from the origin (the precision-stress workload at full
size: an outdoor scene whose world frame sits 10 km away, an 8M-point map and
a ~1.9M-point 128-beam scan with ranges up to 200 m).  Iteration-0
correspondences and counters are bit-exact against the oracle at the
reference's FP32 cache.  At one FP64 state the test compares

- naive FP32 (recenter=0: the reference's world-coordinate e and J),
- re-centred FP32 (recenter=1),
- the FP64 yardstick (the oracle's FP64 twin evaluated at that state),

by the relative error of H and b and by the Gauss-Newton step each form
gives, and runs whole alignments against the oracle's FP64 loop.  Also: the
reference mode reproduces the FP32 oracle's form near the origin.  GPU."""
import functools
import multiprocessing as mp

import numpy as np
import pytest

from kd_slam2_amd import DeviceTree, ICPHIP3D, TreeGenerator3D, synth, update_cache_f32
from oracle import pyoracle as O

from .workloads import cfg1, icp_params, rel_err

pytestmark = pytest.mark.gpu

N_COLUMNS = 22000
N_MAP = 8_000_000


def _scan_block(b):
    return synth.config4_scan_block(b, n_blocks=8, n_columns=N_COLUMNS)


def _map_block(b):
    return synth.config4_map_block(b, n_blocks=8, n_map=N_MAP)


@functools.lru_cache(maxsize=None)
def far_workload():
    with mp.get_context("fork").Pool(16) as pool:
        scan = np.concatenate(pool.map(_scan_block, range(8)))
        M = np.concatenate(pool.map(_map_block, range(8)))
    g = TreeGenerator3D(parallel_level=3)
    X_gt, X0, Xf = synth.config4_poses()
    return dict(fixed=g.makeTree(M), moving=g.makeTree(scan), X_gt=X_gt, X0=X0, Xf=Xf)


def test_correspondences_far_from_origin_bit_exact(ctx):
    """At the reference's FP32 cache (updateCache, icp_impl_.h:16-21) every
    moving leaf's fixed leaf, descent depth and outcome counter equals the
    oracle's, 2M-point scan against the 8M-point map at 10 km."""
    w = far_workload()
    assert len(w["moving"].leaves) > 15_000 and len(w["fixed"].nodes) > 700_000
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params()
    Xm, Xf, Xmf = update_cache_f32(w["X0"], w["Xf"])
    ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, O.from_kdh_params(p, 8))
    for rc in (0, 1):
        al = ICPHIP3D(ctx)
        al.params = icp_params(recenter=rc)
        al.setFixed(F)
        al.setMoving(M)
        al.setX(w["X0"])
        al.setFixedX(w["Xf"])
        leaf, depth = al.correspondences(Xm, Xf, Xmf)
        assert np.array_equal(leaf, ref["corr"]) and np.array_equal(depth, ref["depth"])
        al.buildQuadraticForm()
        for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
            assert al.stats[k] == ref["stats"][k], (rc, k)
    print("leaves", len(leaf), "mean depth", depth.mean(), ref["stats"])
    F.close()
    M.close()


def gn_step(H, b, damping):
    """Damped Gauss-Newton step in FP64 (the solve of oneRound)."""
    A = np.asarray(H, dtype=np.float64) + damping * np.eye(6)
    return np.linalg.solve(A, -np.asarray(b, dtype=np.float64))


def step_diff(dx, ref):
    return float(np.linalg.norm(dx[:3] - ref[:3])), float(np.degrees(np.linalg.norm(dx[3:] - ref[3:])))


def test_forms_far_from_origin_against_fp64_yardstick(ctx):
    w = far_workload()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params()
    ref = O.build_form_at(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 8))
    dx64 = gn_step(ref["H64"], ref["b64"], p.damping)
    out = {}
    for rc in (0, 1):
        al = ICPHIP3D(ctx)
        al.params = icp_params(recenter=rc)
        al.setFixed(F)
        al.setMoving(M)
        al.setX(w["X0"])
        al.setFixedX(w["Xf"])
        al.buildQuadraticFormAt()
        out[rc] = dict(H=al.H.copy(), b=al.b.copy(), stats=dict(al.stats))
    # the discrete decisions do not depend on the mode
    for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
        assert out[0]["stats"][k] == out[1]["stats"][k]
    errs = {}
    for name, H, b in (("naive", out[0]["H"], out[0]["b"]), ("recentred", out[1]["H"], out[1]["b"]),
                       ("reference FP32", ref["H"], ref["b"])):
        errs[name] = (rel_err(H, ref["H64"]), rel_err(b, ref["b64"]), step_diff(gn_step(H, b, p.damping), dx64))
        print(f"{name:15s} H rel {errs[name][0]:.2e}  b rel {errs[name][1]:.2e}  "
              f"GN step vs FP64: {errs[name][2][0]:.2e} m {errs[name][2][1]:.2e} deg")
    # b keeps the FP32 rounding of the per-tile partial sums, amplified by the
    # cancellation between terms of both signs near convergence; the step is
    # what the pose sees.
    H_e, b_e, (st, sr) = errs["recentred"]
    assert H_e < 1e-6 and b_e < 5e-6
    assert st < 1e-5 and sr < 1e-5
    assert 5 * b_e < errs["naive"][1] and 50 * st < errs["naive"][2][0]


def test_alignment_far_from_origin_against_fp64_loop(ctx):
    w = far_workload()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params()
    X64, _, _ = O.compute64(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 8))
    X32, _, _, _ = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 8))
    out = {}
    for rc in (0, 1):
        al = ICPHIP3D(ctx)
        al.params = icp_params(recenter=rc)
        al.setFixed(F)
        al.setMoving(M)
        al.setX(w["X0"])
        al.setFixedX(w["Xf"])
        al.computeOnDevice()
        out[rc] = al.getX()
    d = {name: synth.pose_error(X, X64) for name, X in (("recentred", out[1]), ("naive", out[0]),
                                                          ("reference FP32", X32))}
    for name, (t, r) in d.items():
        print(f"{name:15s} pose vs FP64 loop: {t:.2e} m {r:.2e} deg")
    print(f"FP64 loop vs ground truth: {synth.pose_error(X64, w['X_gt'])}")
    assert d["recentred"][0] < 1e-5 and d["recentred"][1] < 1e-5
    # and within the pose bar of the FP32 oracle
    assert d["reference FP32"][0] < 1e-3 and d["reference FP32"][1] < 1e-3
    t32, r32 = synth.pose_error(out[1], X32)
    assert t32 < 1e-3 and r32 < 1e-3


def test_reference_mode_matches_fp32_oracle_near_origin(ctx):
    """recenter=0 evaluates the reference's own formulas: its form matches the
    FP32 oracle's to summation-order error."""
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    Xm, Xf, Xmf = update_cache_f32(w["X0"], w["Xf"])
    ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, O.from_kdh_params(icp_params(), 8))
    al = ICPHIP3D(ctx)
    al.params = icp_params(recenter=0)
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    al.buildQuadraticForm()
    assert rel_err(al.H, ref["H"]) < 2e-6 and rel_err(al.b, ref["b"]) < 2e-5
    assert al.stats["num_inliers"] == ref["stats"]["num_inliers"]
