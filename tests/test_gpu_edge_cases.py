"""Edge cases and failure handling of the GPU path, against the oracle and
the reference's conventions (tree_aligner_base_.h:100-115, icp_impl_.h:76-80).
GPU."""
import ctypes

import numpy as np
import pytest

from kd_slam2_amd import CTICPHIP3D, DeviceTree, HostTree, ICPHIP3D, _lib, synth, update_cache_f32
from kd_slam2_amd._lib import NODE_DTYPE
from oracle import pyoracle as O

from .test_tree_and_descent import _brute_region, _node, _tiny_axis_tree
from .workloads import cfg1, icp_params

pytestmark = pytest.mark.gpu
I34 = np.eye(4)[:3, :4].reshape(12).astype(np.float32)


def _aligner(ctx, cls, F, M, w, p):
    al = cls(ctx)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    return al


def test_descent_brute_force_on_device(ctx):
    """Axis-aligned tree with a pruned child and on-plane queries: device
    correspondences equal the brute-force half-space regions."""
    fixed = HostTree(_tiny_axis_tree(), np.array([3, 4, 5], dtype=np.uint64))
    rng = np.random.default_rng(0)
    qs = rng.uniform(-2, 2, size=(1000, 3)).astype(np.float32)
    qs[:40, 0] = 0.0
    qs[40:80, 1] = 0.0
    moving = HostTree(np.concatenate([_node(q, [0, 0, 1]) for q in qs]), np.arange(len(qs), dtype=np.uint64))
    F, M = DeviceTree(ctx, fixed), DeviceTree(ctx, moving)
    al = ICPHIP3D(ctx)
    al.params = icp_params(max_error=1e9, min_normal_cos=-2.0)
    al.setFixed(F)
    al.setMoving(M)
    leaf, depth = al.correspondences(I34, I34, I34)
    expect = np.array([_brute_region(fixed.nodes, q) for q in qs])
    assert np.array_equal(leaf, expect)
    al.buildQuadraticForm()
    assert al.stats["num_bad"] == int((expect == -1).sum())


def test_empty_moving_tree_fails_round_zero(ctx):
    w = cfg1()
    F = DeviceTree(ctx, w["fixed"])
    empty = HostTree(np.concatenate([_node([0, 0, 1], [0, 0, -1])]), np.zeros(0, dtype=np.uint64))
    M = DeviceTree(ctx, empty)
    al = _aligner(ctx, ICPHIP3D, F, M, w, icp_params())
    res = al.computeOnDevice()
    assert res == 0 and len(al.stats_log) == 1  # compute() returns -0
    assert al.stats_log[0]["num_inliers"] == 0
    assert np.array_equal(al.getX(), w["X0"])  # no update
    # the per-call path agrees
    al2 = _aligner(ctx, ICPHIP3D, F, M, w, icp_params())
    assert al2.compute() == 0


def test_no_overlap_fails_min_good(ctx):
    """Moving cloud placed 1 km away: every leaf is rejected, round 0 fails."""
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    far = synth.pose(None, [1000.0, 0, 0]) @ w["X0"]
    al = _aligner(ctx, ICPHIP3D, F, M, dict(w, X0=far), icp_params())
    res = al.computeOnDevice()
    Xo, _, log_o, res_o = O.compute(w["fixed"], w["moving"], far, w["Xf"], O.from_kdh_params(al.params))
    assert res == res_o == 0
    for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
        assert al.stats_log[0][k] == log_o[0][k]


def test_stats_mode_one_round_no_update(ctx):
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    al = _aligner(ctx, ICPHIP3D, F, M, w, icp_params())
    res = al.computeOnDevice(stats_mode=True)
    Xo, _, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(al.params),
                                    stats_mode=True)
    assert res == res_o == al.params.max_iterations and len(al.stats_log) == 1
    assert np.array_equal(al.getX(), w["X0"])
    s, so = al.stats_log[0], log_o[0]
    for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
        assert s[k] == so[k]
    assert abs(s["chi2_kernelized"] - so["chi2_kernelized"]) <= 1e-4 * max(1.0, so["chi2_kernelized"])
    assert abs(s["pert_pose_norm"] - so["pert_pose_norm"]) <= 1e-3 * max(1e-3, so["pert_pose_norm"])


def _criterion_margins(log, d_pert, d_score):
    """Per round i > 0 of an oracle log: (pert_pose_norm / d_pert, relative
    score change / d_score) as ICPTerminationCriteriaScore_ evaluates them
    (icp_stats_.h:85-102), to tell a decision near its threshold."""
    f = np.float32
    out, prev = [], f(0.0)
    for i, s in enumerate(log):
        nl = s["num_inliers"] + s["num_outliers"] + s["num_bad"] + s["num_err_normal"] + s["num_err_fn"]
        score = f(f(f(s["num_inliers"]) / f(nl)) / (f(s["chi2_kernelized"]) / f(s["num_inliers"]) + f(1e-6)))
        if i > 0:
            d = f(score - prev)
            out.append((s["pert_pose_norm"] / d_pert, float(d) / (float(prev) + 1e-6) / d_score))
        prev = score
    return out


@pytest.mark.parametrize("stop_on_last", [False, True])
def test_termination_criterion_stops_like_oracle(ctx, stop_on_last):
    """The score criterion ends the device loop at the oracle's round, with
    compute()'s return value (the round index), also when the stop falls on
    the last allowed round (tree_aligner_base_.h:100-115 asks the criterion
    before the loop ends)."""
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params(max_iterations=30, use_term_crit=1, term_d_pert=1e-3, term_d_score=1e-3)
    Xo, _, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p))
    assert 0 < res_o < 29
    # the oracle's decisions are clear of their thresholds in every round
    # up to the stop, so the FP64 device loop must take the same ones
    for pm, sm in _criterion_margins(log_o, 1e-3, 1e-3):
        assert abs(pm - 1.0) > 0.02 and (sm <= 0 or abs(sm - 1.0) > 0.02)
    if stop_on_last:
        p = icp_params(max_iterations=res_o + 1, use_term_crit=1, term_d_pert=1e-3, term_d_score=1e-3)
        Xo, _, log_o, res_o2 = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p))
        assert res_o2 == res_o == p.max_iterations - 1
    al = _aligner(ctx, ICPHIP3D, F, M, w, p)
    res = al.computeOnDevice()
    print("criterion stop", res, res_o)
    assert res == res_o
    assert len(al.stats_log) == res + 1
    for a, b in zip(al.stats_log, log_o):
        for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
            assert a[k] == b[k]
    dt, dr = synth.pose_error(al.getX(), Xo)
    assert dt < 1e-3 and dr < 1e-3
    # the host per-call loop (the drop-in) returns the same code
    al2 = _aligner(ctx, ICPHIP3D, F, M, w, p)
    assert al2.compute() == res_o


def test_pose_prior_device_host_and_oracle(ctx):
    """A strong prior pulls the solution toward Z * X_ref; the device loop,
    the host per-call loop and the FP32 oracle agree."""
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    p = icp_params()
    info = np.diag([1e4] * 3 + [1e5] * 3)
    X_ref = w["X0"]
    al = _aligner(ctx, ICPHIP3D, F, M, w, p)
    al.setPosePrior(X_ref, np.eye(4), info)
    al.computeOnDevice()
    al2 = _aligner(ctx, ICPHIP3D, F, M, w, p)
    al2.setPosePrior(X_ref, np.eye(4), info)
    al2.compute()
    ZXi = np.linalg.inv(X_ref)[:3, :4].reshape(12)
    Xo, _, _, _ = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p), prior_ZXi=ZXi,
                            prior_info=info.reshape(36))
    d_dev = synth.pose_error(al.getX(), Xo)
    d_host = synth.pose_error(al2.getX(), Xo)
    d_prior = synth.pose_error(al.getX(), X_ref)
    d_free = synth.pose_error(w["X_gt"], X_ref)
    print("prior", d_dev, d_host, d_prior, d_free)
    assert d_dev[0] < 1e-3 and d_dev[1] < 1e-3 and d_host[0] < 1e-3 and d_host[1] < 1e-3
    assert d_prior[0] < 0.5 * d_free[0]  # the prior dominates


def test_nan_pose_propagates_without_fault(ctx):
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    X0 = w["X0"].copy()
    X0[0, 3] = np.nan
    al = _aligner(ctx, ICPHIP3D, F, M, dict(w, X0=X0), icp_params(max_iterations=3))
    Xm, Xf, Xmf = update_cache_f32(X0, w["Xf"])
    leaf, _ = al.correspondences(Xm, Xf, Xmf)
    ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, O.from_kdh_params(al.params))
    assert np.array_equal(leaf, ref["corr"])  # NaN comparisons route identically
    al.computeOnDevice()
    assert not np.all(np.isfinite(al.getX()))  # the caller's finiteness check fires


def test_invalid_inputs_are_errors(ctx):
    L = _lib.lib()
    bad = np.concatenate([_node([0, 0, 0], [1, 0, 0], 0, 1), _node([1, 0, 0], [0, 0, 1])])  # self link
    h = ctypes.c_void_p()
    leaves = np.array([1], dtype=np.uint64)
    rc = L.kdh_tree_upload(ctx.h, bad.ctypes.data, len(bad), leaves.ctypes.data, 1, ctypes.byref(h))
    assert rc == -1 and b"forward" in L.kdh_last_error(ctx.h)
    oob = np.concatenate([_node([0, 0, 0], [1, 0, 0], 1, 7), _node([1, 0, 0], [0, 0, 1])])
    rc = L.kdh_tree_upload(ctx.h, oob.ctypes.data, len(oob), leaves.ctypes.data, 1, ctypes.byref(h))
    assert rc == -1
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    al = _aligner(ctx, ICPHIP3D, F, M, w, icp_params(max_iterations=65))
    with pytest.raises(RuntimeError, match="max_iterations"):
        al.computeOnDevice()


def test_fresh_context_first_call_is_a_device_loop(ctx):
    """A new context whose first call is the device-resident loop of one
    alignment (its pinned result mirror still at the initial size), then a
    batch: the results equal those of the shared context, and the batch's
    pair equals the loop."""
    from kd_slam2_amd import Context, batch_align

    w = cfg1()
    p = icp_params()
    fresh = Context(0)
    try:
        F1, M1 = DeviceTree(fresh, w["fixed"]), DeviceTree(fresh, w["moving"])
        al = _aligner(fresh, ICPHIP3D, F1, M1, w, p)
        res = al.computeOnDevice()
        X_fresh, log_fresh = al.getX(), al.stats_log
        b = batch_align(fresh, [dict(fixed=F1, moving=M1, ct=0, X0=w["X0"], Xf=w["Xf"], params=p)])[0]
        F1.close()
        M1.close()
    finally:
        fresh.close()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
    ref = _aligner(ctx, ICPHIP3D, F, M, w, p)
    assert ref.computeOnDevice() == res
    assert np.array_equal(ref.getX(), X_fresh) and ref.stats_log == log_fresh
    assert np.array_equal(b["X"], X_fresh) and b["iterations"] == len(log_fresh)
