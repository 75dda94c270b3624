"""GPU path against the oracle (CPU restatement of ICP_CPU_ / CTICP_CPU_).

Bar: correspondences and every ICPStats_ counter bit-exact at an identical
FP32 cache; H and b within 2e-5 (relative to max |entry|) of the FP64 oracle;
whole alignments within 1 mm / 1e-3 deg of the FP32 oracle's compute().
"""
import numpy as np
import pytest

from kd_slam2_amd import CTICPHIP3D, DeviceTree, ICPHIP3D, synth, update_cache_f32
from oracle import pyoracle as O

from .workloads import cfg1, cfg2, icp_params, rel_err

pytestmark = pytest.mark.gpu

HB_TOL = 2e-5          # H/b relative error against the FP64 oracle
POSE_T_TOL = 1e-3      # metres
POSE_R_TOL = 1e-3      # degrees
COUNTERS = ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn")


def _upload(ctx, w):
    return DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])


def test_icp_cfg1_iteration0_bitexact(ctx):
    w = cfg1()
    F, M = _upload(ctx, w)
    p = icp_params()
    al = ICPHIP3D(ctx)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    Xm, Xf, Xmf = update_cache_f32(w["X0"], w["Xf"])
    assert np.array_equal(Xmf, O.update_cache(Xm, Xf))
    leaf, depth = al.correspondences(Xm, Xf, Xmf)
    ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, O.from_kdh_params(p))
    print("leaves", len(leaf), "mean depth", depth.mean(), "stats", ref["stats"])
    assert np.array_equal(leaf, ref["corr"])
    assert np.array_equal(depth, ref["depth"])
    al.buildQuadraticForm()
    for k in COUNTERS:
        assert al.stats[k] == ref["stats"][k], k
    e_h = rel_err(al.H, ref["H64"])
    e_b = rel_err(al.b, ref["b64"])
    e_h32 = rel_err(ref["H"], ref["H64"])
    print(f"H relerr gpu {e_h:.2e} (fp32 oracle {e_h32:.2e}), b relerr gpu {e_b:.2e}")
    assert e_h < HB_TOL and e_b < HB_TOL
    assert abs(al.stats["chi2"] - ref["chi2_64"]) <= 1e-5 * max(1.0, ref["chi2_64"])


def test_icp_cfg1_align_matches_oracle(ctx):
    w = cfg1()
    F, M = _upload(ctx, w)
    p = icp_params()
    Xo, _, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p))
    # device-resident loop
    al = ICPHIP3D(ctx)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    res = al.computeOnDevice()
    dt, dr = synth.pose_error(al.getX(), Xo)
    print("device loop vs oracle", dt, dr, "vs gt", synth.pose_error(al.getX(), w["X_gt"]))
    assert res == res_o == p.max_iterations
    assert dt < POSE_T_TOL and dr < POSE_R_TOL
    # host-driven per-call loop (the plain path)
    al2 = ICPHIP3D(ctx)
    al2.params = p
    al2.setFixed(F)
    al2.setMoving(M)
    al2.setX(w["X0"])
    al2.setFixedX(w["Xf"])
    res2 = al2.compute()
    dt2, dr2 = synth.pose_error(al2.getX(), Xo)
    print("per-call loop vs oracle", dt2, dr2)
    assert res2 == p.max_iterations
    assert dt2 < POSE_T_TOL and dr2 < POSE_R_TOL


def test_icp_device_loop_lockstep(ctx):
    """At every iteration of the device loop, the correspondences the device
    used equal the oracle's at the FP32 cache the device formed."""
    w = cfg1()
    F, M = _upload(ctx, w)
    p = icp_params()
    al = ICPHIP3D(ctx)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    al.computeOnDevice(cache_log=True)
    op = O.from_kdh_params(p)
    for it, c in enumerate(al.cache_log):
        Xm, Xf, Xmf = c[:12], c[12:24], c[24:36]
        leaf, _ = al.correspondences(Xm, Xf, Xmf)
        ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, op)
        assert np.array_equal(leaf, ref["corr"]), it
        for k in COUNTERS:
            assert al.stats_log[it][k] == ref["stats"][k], (it, k)


def test_cticp_cfg2_iteration0_bitexact(ctx):
    w = cfg2()
    F, M = _upload(ctx, w)
    p = icp_params(max_iterations=15, velocity_damping=10.0)
    al = CTICPHIP3D(ctx)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    al.setVelocities(w["v0"])
    Xm, Xf, Xmf = update_cache_f32(w["X0"], w["Xf"])
    v = np.asarray(w["v0"], dtype=np.float32)
    leaf, depth = al.correspondences(Xm, Xf, Xmf, v)
    ref = O.build_form(w["fixed"], w["moving"], Xm, Xf, Xmf, O.from_kdh_params(p), ct=1, v=v)
    print("leaves", len(leaf), "mean depth", depth.mean(), "stats", ref["stats"])
    assert np.array_equal(leaf, ref["corr"])
    al.buildQuadraticForm()
    for k in COUNTERS:
        assert al.stats[k] == ref["stats"][k], k
    e_h = rel_err(al.H, ref["H64"])
    e_b = rel_err(al.b, ref["b64"])
    print(f"CT H relerr {e_h:.2e} b relerr {e_b:.2e} (fp32 oracle {rel_err(ref['H'], ref['H64']):.2e})")
    assert e_h < HB_TOL and e_b < HB_TOL


def test_cticp_cfg2_align_matches_oracle(ctx):
    w = cfg2()
    F, M = _upload(ctx, w)
    p = icp_params(max_iterations=15, velocity_damping=10.0)
    Xo, vo, log_o, res_o = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 8), ct=1,
                                     v0=w["v0"])
    al = CTICPHIP3D(ctx)
    al.params = p
    al.setFixed(F)
    al.setMoving(M)
    al.setX(w["X0"])
    al.setFixedX(w["Xf"])
    al.setVelocities(w["v0"])
    res = al.computeOnDevice()
    dt, dr = synth.pose_error(al.getX(), Xo)
    print("CT device vs oracle", dt, dr, "v", al.velocities(), vo, "vs gt", synth.pose_error(al.getX(), w["X_gt"]),
          w["v_gt"])
    assert res == res_o
    assert dt < POSE_T_TOL and dr < POSE_R_TOL
    assert np.abs(al.velocities() - vo).max() < 1e-2
