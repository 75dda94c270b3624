"""The drop-in itself: ICPHIP3D / CTICPHIP3D created by type name through the
C++ plugin's registry (the class INTEGRATION.md binds into the reference's
tree, the ICP_CUDA_ pattern of icp/icp_cuda_impl_.h:95-132) and driven
through compute() -- the plain per-call path, host Gauss-Newton loop
(tree_aligner_base_.h:100-115) -- against the oracle's ICP_CPU_ /
CTICP_CPU_ on the same trees.  GPU."""
import numpy as np
import pytest

from kd_slam2_amd import DeviceTree, adapter_align, synth
from oracle import pyoracle as O

from .workloads import cfg1, cfg2, icp_params

pytestmark = pytest.mark.gpu

POSE_T_TOL = 1e-3  # metres
POSE_R_TOL = 1e-3  # degrees
COUNTERS = ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn")


def _check(log, log_o, res, res_o, n_leaves):
    assert res == res_o
    assert len(log) == len(log_o)
    # round 0 runs at the same FP32 cache on both sides: every counter equal
    for k in COUNTERS:
        assert log[0][k] == log_o[0][k], k
    # later rounds start from the FP64 solve's state on the device side and
    # the FP32 solve's in the oracle (states 1e-6 apart): a leaf may sit on a
    # decision boundary between them
    for it, (a, b) in enumerate(zip(log, log_o)):
        for k in COUNTERS:
            assert abs(a[k] - b[k]) <= max(2, n_leaves // 2000), (it, k, a[k], b[k])
        assert abs(a["chi2"] - b["chi2"]) <= 1e-3 * max(1.0, b["chi2"]), it


@pytest.mark.parametrize("on_device", [0, 1])
def test_icp_plugin_compute_cfg1(ctx, on_device):
    """Rigid ICP, 16k-point scan vs 200k-point map, 10 iterations."""
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params()
    log = []
    X, _, res, nit, last = adapter_align("ICPHIP3D", ctx, F, M, w["X0"], w["Xf"], p, on_device=bool(on_device),
                                         log=log)
    Xo, _, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p))
    dt, dr = synth.pose_error(X, Xo)
    print("plugin ICPHIP3D vs oracle", dt, dr, res, res_o, nit)
    assert nit == len(log_o) == p.max_iterations
    _check(log, log_o, res, res_o, len(w["moving"].leaves))
    assert dt < POSE_T_TOL and dr < POSE_R_TOL
    F.close()
    M.close()


@pytest.mark.parametrize("on_device", [0, 1])
def test_cticp_plugin_compute_cfg2(ctx, on_device):
    """CT-ICP with deskew, 131k-point sweep vs 2M-point map, 15 iterations."""
    w = cfg2()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params(max_iterations=15, velocity_damping=10.0)
    log = []
    X, v, res, nit, last = adapter_align("CTICPHIP3D", ctx, F, M, w["X0"], w["Xf"], p, v0=w["v0"],
                                         on_device=bool(on_device), log=log)
    Xo, vo, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 8), ct=1,
                                     v0=w["v0"])
    dt, dr = synth.pose_error(X, Xo)
    print("plugin CTICPHIP3D vs oracle", dt, dr, np.abs(v - vo).max(), res, res_o, nit)
    assert nit == len(log_o) == p.max_iterations
    _check(log, log_o, res, res_o, len(w["moving"].leaves))
    assert dt < POSE_T_TOL and dr < POSE_R_TOL
    assert np.abs(v - vo).max() < 1e-2
    F.close()
    M.close()


def test_plugin_score_criterion_stop(ctx):
    """compute() through the plugin with the score criterion configured
    (ICPTerminationCriteriaScore3D) returns the oracle's round index."""
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params(max_iterations=30, use_term_crit=1, term_d_pert=1e-3, term_d_score=1e-3)
    log = []
    X, _, res, nit, _ = adapter_align("ICPHIP3D", ctx, F, M, w["X0"], w["Xf"], p, log=log)
    Xo, _, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p))
    assert 0 < res_o < 29
    _check(log, log_o, res, res_o, len(w["moving"].leaves))
    assert nit == res_o + 1
    F.close()
    M.close()
