"""Pins the oracle (oracle/oracle.cpp) to what the mathematics and the
reference's own checks fix, so that it can serve as the checker of the GPU
path.  CPU only."""
import ctypes
import math

import numpy as np
import pytest

from oracle import pyoracle as O

L = O.lib()


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------------------
# shared sine / cosine (Cephes single-precision scheme)


def _ulp_err(a, ref):
    a = np.asarray(a, dtype=np.float32)
    ref64 = np.asarray(ref, dtype=np.float64)
    spacing = np.spacing(np.abs(ref64).astype(np.float32)).astype(np.float64)
    spacing = np.maximum(spacing, np.float64(np.finfo(np.float32).tiny))
    return np.abs(a.astype(np.float64) - ref64) / spacing


def test_sincos_against_libm():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-4, 4, 200_000), rng.uniform(-100, 100, 50_000),
                        np.linspace(-1e-3, 1e-3, 2001), np.array([0.0, -0.0, np.pi / 4, np.pi / 2, np.pi, 1e-9])])
    x = f32(x)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    L.orc_sincos(O.p32(x), len(x), O.p32(s), O.p32(c))
    es = _ulp_err(s, np.sin(x.astype(np.float64)))
    ec = _ulp_err(c, np.cos(x.astype(np.float64)))
    # absolute error near zeros of sin/cos is bounded by the reduction error
    abs_s = np.abs(s.astype(np.float64) - np.sin(x.astype(np.float64)))
    abs_c = np.abs(c.astype(np.float64) - np.cos(x.astype(np.float64)))
    ok_s = (es <= 2.0) | (abs_s < 2e-7)
    ok_c = (ec <= 2.0) | (abs_c < 2e-7)
    assert ok_s.all(), (x[~ok_s][:5], es[~ok_s][:5])
    assert ok_c.all(), (x[~ok_c][:5], ec[~ok_c][:5])
    # a dropped sign or quadrant shows up as an O(1) error
    assert np.abs(s.astype(np.float64) - np.sin(x.astype(np.float64))).max() < 1e-5
    # special values
    assert s[len(x) - 6] == 0.0 and c[len(x) - 6] == 1.0


def test_rodrigues_is_a_rotation():
    rng = np.random.default_rng(1)
    for _ in range(200):
        w = f32(rng.normal(size=3) * rng.choice([1e-12, 1e-4, 0.1, 1.0, 3.0]))
        R = np.zeros(9, dtype=np.float32)
        L.orc_rodrigues_f(O.p32(w), O.p32(R))
        R = R.reshape(3, 3).astype(np.float64)
        assert np.abs(R @ R.T - np.eye(3)).max() < 2e-6
        assert abs(np.linalg.det(R) - 1) < 2e-6
        # FP64 series reference
        a = np.linalg.norm(w.astype(np.float64))
        if a < 1e-9:
            Rref = np.eye(3)
        else:
            n = w.astype(np.float64) / a
            S = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
            Rref = np.eye(3) + math.sin(a) * S + (1 - math.cos(a)) * S @ S
        assert np.abs(R - Rref).max() < 2e-6
        # R w = w (the axis is fixed)
        assert np.abs(R @ w.astype(np.float64) - w.astype(np.float64)).max() < 2e-6 * max(1.0, a)


# ---------------------------------------------------------------------------
# StableAdder_


def test_stable_adder_exact_on_integers():
    rng = np.random.default_rng(2)
    x = f32(rng.integers(-1000, 1000, size=300_000))
    assert L.orc_stable_sum_f(O.p32(x), len(x)) == float(x.astype(np.int64).sum())


def test_stable_adder_beats_naive_sum():
    """The 1024-batch cascade keeps the FP32 sum within a few ulp of the exact
    sum, where a running FP32 sum drifts by orders of magnitude more."""
    rng = np.random.default_rng(3)
    x = f32(rng.uniform(0, 1, size=4_000_000))
    exact = math.fsum(x.astype(np.float64))
    s = L.orc_stable_sum_f(O.p32(x), len(x))
    naive = float(np.cumsum(x, dtype=np.float32)[-1])  # sequential running sum
    ulp = float(np.spacing(np.float32(exact)))
    assert abs(s - exact) <= 4 * ulp
    assert abs(naive - exact) > 100 * ulp


# ---------------------------------------------------------------------------
# LDLT


@pytest.mark.parametrize("n", [6, 12])
def test_ldlt_matches_numpy(n):
    rng = np.random.default_rng(4 + n)
    for _ in range(50):
        A = rng.normal(size=(n, n))
        A = A @ A.T + 0.5 * np.eye(n)
        b = rng.normal(size=n)
        x = np.zeros(n)
        L.orc_ldlt_solve_d(n, O.p64(np.ascontiguousarray(A)), O.p64(b), O.p64(x))
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
        xf = np.zeros(n, dtype=np.float32)
        L.orc_ldlt_solve_f(n, O.p32(f32(A)), O.p32(f32(b)), O.p32(xf))
        assert np.abs(xf - np.linalg.solve(A, b)).max() < 1e-3 * max(1.0, np.abs(np.linalg.solve(A, b)).max())


# ---------------------------------------------------------------------------
# errorAndJacobian (d3/icp_.h:41-60)


def _rot(w):
    a = np.linalg.norm(w)
    if a < 1e-15:
        return np.eye(3)
    n = w / a
    S = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    return np.eye(3) + math.sin(a) * S + (1 - math.cos(a)) * S @ S


def test_error_jacobian_against_finite_differences():
    """J is de/d(dx) for the left perturbation X <- exp(dx) X of the moving
    pose: p_m -> R(dw) p_m + dt, n_m -> R(dw) n_m (pattern of
    apps/tests/jacobians/pose_error_jacobian_3d_test.cpp)."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        nf = rng.normal(size=3); nf /= np.linalg.norm(nf)
        nm = nf + 0.2 * rng.normal(size=3); nm /= np.linalg.norm(nm)
        pf = rng.uniform(-5, 5, size=3)
        pm = pf + 0.05 * rng.normal(size=3)
        e = np.zeros(1, dtype=np.float32)
        J = np.zeros(6, dtype=np.float32)
        rc = L.orc_error_and_jacobian(O.p32(f32(pf)), O.p32(f32(nf)), O.p32(f32(pm)), O.p32(f32(nm)),
                                      -2.0, 100.0, O.p32(e), O.p32(J))
        assert rc == 0

        def err(dx):
            R = _rot(dx[3:])
            return (R @ pm + dx[:3] - pf) @ (R @ nm + nf)

        h = 1e-6
        Jn = np.array([(err(np.eye(6)[k] * h) - err(-np.eye(6)[k] * h)) / (2 * h) for k in range(6)])
        assert abs(e[0] - err(np.zeros(6))) < 1e-5
        assert np.abs(J - Jn).max() < 1e-4 * max(1.0, np.abs(Jn).max())


def test_error_jacobian_gates_and_symmetry():
    pf = f32([1, 2, 3]); nf = f32([0, 0, 1])
    pm = f32([1, 2, 3.2]); nm = f32([0, 0, 1])
    e = np.zeros(1, dtype=np.float32); J = np.zeros(6, dtype=np.float32)
    # normal gate: cos(angle) below threshold -> -1
    nm_bad = f32([1, 0, 0])
    assert L.orc_error_and_jacobian(O.p32(pf), O.p32(nf), O.p32(pm), O.p32(nm_bad), 0.9, 1.0, O.p32(e), O.p32(J)) == -1
    # error gate: |e| = 0.4 > 0.3 -> -2
    assert L.orc_error_and_jacobian(O.p32(pf), O.p32(nf), O.p32(pm), O.p32(nm), 0.9, 0.3, O.p32(e), O.p32(J)) == -2
    assert L.orc_error_and_jacobian(O.p32(pf), O.p32(nf), O.p32(pm), O.p32(nm), 0.9, 1.0, O.p32(e), O.p32(J)) == 0
    assert abs(e[0] - 0.4) < 1e-6  # symmetric error = 2x plane distance
    # Swapping the clouds negates e and keeps J, so de/d(fixed pose) = -J
    # (J_B = -J_A, icp/icp_.h:117-121).
    J2 = np.zeros(6, dtype=np.float32); e2 = np.zeros(1, dtype=np.float32)
    L.orc_error_and_jacobian(O.p32(pm), O.p32(nm), O.p32(pf), O.p32(nf), 0.9, 1.0, O.p32(e2), O.p32(J2))
    assert abs(e2[0] + e[0]) < 1e-6
    assert np.allclose(J2[:3], J[:3]) and np.allclose(J2[3:], J[3:])


# ---------------------------------------------------------------------------
# errorAndJacobianVelocityA: the reference's own check, replayed


def _glibc():
    return ctypes.CDLL("libc.so.6")


def test_velocity_jacobian_reference_check_replayed():
    """apps/tests/jacobians/motion_jacobian_3d_test.cpp:68-127: 20 cases from
    std::srand(42), central differences with h = 1e-3 in FP32, max-abs
    tolerance 5e-3.  Inputs are drawn with glibc rand() in source order
    (the reference's argument evaluation order is unspecified in C++)."""
    libc = _glibc()
    libc.srand(42)
    RAND_MAX = np.float32(2147483647)

    def rnd(r=1.0):
        v = np.float32(2) * np.float32(libc.rand()) / RAND_MAX - np.float32(1)
        return np.float32(r) * np.float32(v)

    def unit():
        v = np.array([rnd(), rnd(), rnd()], dtype=np.float32)
        return (v / np.linalg.norm(v)).astype(np.float32)

    def transform():
        ang = rnd(math.pi)
        ax = unit().astype(np.float64)
        T = np.eye(4)
        T[:3, :3] = _rot(ax * ang)
        T[:3, 3] = [rnd(2), rnd(2), rnd(2)]
        return T

    h, tol = np.float32(1e-3), 5e-3
    n_pass = 0
    for _ in range(20):
        Ta, Tb = transform(), transform()
        pa_l = np.array([rnd(3), rnd(3), rnd(3)], dtype=np.float32)
        na_l = unit()
        pb_l = np.array([rnd(3), rnd(3), rnd(3)], dtype=np.float32)
        nb_l = unit()
        pb_g = f32(Tb[:3, :3] @ pb_l + Tb[:3, 3])
        nb_g = f32(Tb[:3, :3] @ nb_l)
        va = np.array([rnd(.5), rnd(.5), rnd(.5), rnd(.3), rnd(.3), rnd(.3)], dtype=np.float32)
        _vb = [rnd(.5), rnd(.5), rnd(.5), rnd(.3), rnd(.3), rnd(.3)]
        dt = np.float32(0.1)
        Ta34 = f32(Ta[:3, :4].reshape(12))

        def ev(v):
            e = np.zeros(1, dtype=np.float32)
            J = np.zeros(6, dtype=np.float32)
            L.orc_error_and_jacobian_velocity_a(O.p32(Ta34), O.p32(pa_l), O.p32(na_l), O.p32(pb_g), O.p32(nb_g),
                                                O.p32(f32(v)), dt, O.p32(e), O.p32(J))
            return e[0], J

        _, Ja = ev(va)
        Jn = np.zeros(6, dtype=np.float32)
        for i in range(6):
            vp = va.copy(); vp[i] += h
            vm = va.copy(); vm[i] -= h
            Jn[i] = (ev(vp)[0] - ev(vm)[0]) / (np.float32(2) * h)
        n_pass += int(np.abs(Ja - Jn).max() < tol)
    assert n_pass == 20


def test_velocity_jacobian_fp64_finite_differences():
    """Stronger than the replayed check: the FP32 analytic Jacobian against
    FP64 central differences of the same error model, relative 1e-3."""
    rng = np.random.default_rng(6)
    for _ in range(50):
        Ta = np.eye(4); Ta[:3, :3] = _rot(rng.normal(size=3)); Ta[:3, 3] = rng.uniform(-2, 2, 3)
        pa = rng.uniform(-3, 3, 3); na = rng.normal(size=3); na /= np.linalg.norm(na)
        pb = rng.uniform(-3, 3, 3); nb = rng.normal(size=3); nb /= np.linalg.norm(nb)
        va = np.concatenate([rng.uniform(-.5, .5, 3), rng.uniform(-.3, .3, 3)])
        dt = 0.1

        def err(v):
            m = v * dt
            R = _rot(m[3:])
            pg = Ta[:3, :3] @ (R @ pa + m[:3]) + Ta[:3, 3]
            ng = Ta[:3, :3] @ (R @ na)
            return (pg - pb) @ (ng + nb)

        e = np.zeros(1, dtype=np.float32); J = np.zeros(6, dtype=np.float32)
        L.orc_error_and_jacobian_velocity_a(O.p32(f32(Ta[:3, :4].reshape(12))), O.p32(f32(pa)), O.p32(f32(na)),
                                            O.p32(f32(pb)), O.p32(f32(nb)), O.p32(f32(va)), np.float32(dt),
                                            O.p32(e), O.p32(J))
        hh = 1e-6
        Jn = np.array([(err(va + np.eye(6)[k] * hh) - err(va - np.eye(6)[k] * hh)) / (2 * hh) for k in range(6)])
        assert np.abs(J - Jn).max() < 1e-3 * max(1.0, np.abs(Jn).max())


def _small_pair():
    from kd_slam2_amd import default_params

    from .workloads import cfg1

    w = cfg1(n_map=50_000, n_scan=4096)
    p = default_params()
    p.max_error = 0.5
    p.max_iterations = 8
    return w, p


def test_fp64_state_twin_equals_cast_twin_at_fp32_exact_state():
    """build_form_at evaluates the FP64 twin at the state; for a state that
    FP32 holds exactly that is the twin of the rounded cache, bit for bit."""
    from kd_slam2_amd import update_cache_f32

    w, p = _small_pair()
    X = w["X0"].astype(np.float32).astype(np.float64)
    Xf = w["Xf"].astype(np.float32).astype(np.float64)
    a = O.build_form_at(w["fixed"], w["moving"], X, Xf, O.from_kdh_params(p, 4))
    Xm, Xfc, _ = update_cache_f32(X, Xf)
    Xmf = np.linalg.inv(Xf) @ X
    b = O.build_form(w["fixed"], w["moving"], Xm, Xfc, Xmf[:3, :4].astype(np.float32).reshape(12),
                     O.from_kdh_params(p, 4))
    assert np.array_equal(a["corr"], b["corr"])
    assert np.array_equal(a["H64"], b["H64"]) and np.array_equal(a["b64"], b["b64"])


def test_fp64_loop_agrees_with_fp32_loop_near_origin():
    """Near the origin the FP64 yardstick loop and the reference's FP32 loop
    take the same decisions and land within FP32 noise of each other, and
    both move towards the generating pose."""
    from kd_slam2_amd import synth

    w, p = _small_pair()
    X64, _, n = O.compute64(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 4))
    X32, _, _, res = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], O.from_kdh_params(p, 4))
    assert n == p.max_iterations and res == p.max_iterations
    t, r = synth.pose_error(X64, X32)
    assert t < 1e-4 and r < 1e-3
    t, r = synth.pose_error(X64, w["X_gt"])
    t0, r0 = synth.pose_error(w["X0"], w["X_gt"])
    assert t < t0 / 5 and r < r0 / 5


def test_forest_of_one_is_the_plain_compute_and_blocks_recover_the_pose():
    """A fixed forest of one entry is the plain compute bit for bit; the
    config-1 map cut into three re-framed slabs moves the pose most of the way
    to the generating one (every moving leaf is matched in every slab, so
    slabs add false matches near their cuts: the forest lands a few cm off,
    where the whole map lands within 2 mm; cutting in the anchor frame instead
    gives the same result to 1 mm, so the framing is not the cause)."""
    from kd_slam2_amd import synth

    from .workloads import forest_cfg1, icp_params

    f = forest_cfg1()
    w = f["w"]
    p = O.from_kdh_params(icp_params(), 8)
    a = O.compute(w["fixed"], w["moving"], w["X0"], w["Xf"], p, fast=True)
    b = O.compute_forest([w["fixed"]], [w["Xf"]], w["moving"], w["X0"], p, fast=True)
    assert np.array_equal(a[0], b[0]) and a[2] == b[2] and a[3] == b[3]
    X, _, log, res, ent = O.compute_forest(f["trees"], f["poses"], w["moving"], w["X0"], p, fast=True)
    assert res == 10 and len(ent) == 3
    for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn"):
        assert sum(e[k] for e in ent) == log[-1][k]
    t, r = synth.pose_error(X, w["X_gt"])
    t0, r0 = synth.pose_error(w["X0"], w["X_gt"])
    assert t < t0 / 5 and r < r0 / 5


def test_velocity_b_jacobian_fp64_finite_differences():
    """errorAndJacobianVelocityB: the FP32 analytic Jacobian of the error with
    respect to the fixed side's velocities against FP64 central differences."""
    rng = np.random.default_rng(16)
    for _ in range(50):
        Tb = np.eye(4); Tb[:3, :3] = _rot(rng.normal(size=3)); Tb[:3, 3] = rng.uniform(-2, 2, 3)
        pa = rng.uniform(-3, 3, 3); na = rng.normal(size=3); na /= np.linalg.norm(na)
        pb = rng.uniform(-3, 3, 3); nb = rng.normal(size=3); nb /= np.linalg.norm(nb)
        vb = np.concatenate([rng.uniform(-.5, .5, 3), rng.uniform(-.3, .3, 3)])
        dt = 0.07

        def err(v):
            m = v * dt
            R = _rot(m[3:])
            pg = Tb[:3, :3] @ (R @ pb + m[:3]) + Tb[:3, 3]
            ng = Tb[:3, :3] @ (R @ nb)
            return (pa - pg) @ (na + ng)

        J = np.zeros(6, dtype=np.float32)
        L.orc_jacobian_velocity_b(O.p32(f32(Tb[:3, :4].reshape(12))), O.p32(f32(pa)), O.p32(f32(na)),
                                  O.p32(f32(pb)), O.p32(f32(nb)), O.p32(f32(vb)), np.float32(dt), O.p32(J))
        hh = 1e-6
        Jn = np.array([(err(vb + np.eye(6)[k] * hh) - err(vb - np.eye(6)[k] * hh)) / (2 * hh) for k in range(6)])
        assert np.abs(J - Jn).max() < 1e-3 * max(1.0, np.abs(Jn).max())


def test_dual_form_a_side_is_the_ct_form():
    """Outside stats mode the dual's weight rule equals quadraticTerm's, so its
    moving-side blocks are the CT form (FP64 twins to rounding, counts exact);
    its dual blocks are symmetric where they must be."""
    from .workloads import cfg2, icp_params

    w = cfg2(n_map=200_000, n_scan=16_384)
    p = O.from_kdh_params(icp_params(max_error=1.0, velocity_damping=10.0), 8)
    vf = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.2])
    d = O.build_form_dual(w["fixed"], w["moving"], w["X0"], w["v0"], w["Xf"], vf, p)
    ref = O.build_form_at(w["fixed"], w["moving"], w["X0"], w["Xf"], p, ct=1, v=w["v0"])
    assert np.array_equal(d["corr"], ref["corr"])
    assert np.abs(d["H64"] - ref["H64"]).max() <= 1e-12 * np.abs(ref["H64"]).max()
    assert np.abs(d["b64"] - ref["b64"]).max() <= 1e-12 * np.abs(ref["b64"]).max()
    assert np.allclose(d["H_vel_B64"], d["H_vel_B64"].T)
    assert d["stats"]["num_inliers"] > 100
    # stats mode: inliers drop out of H, outliers stay weighted
    s = O.build_form_dual(w["fixed"], w["moving"], w["X0"], w["v0"], w["Xf"], vf, p, stats_mode=True)
    assert s["stats"]["num_inliers"] == d["stats"]["num_inliers"]
    assert np.abs(s["H64"]).max() < np.abs(d["H64"]).max()


def _dual_full(d, suffix=""):
    """The 18x18 normal matrix over [pose A, vel A, vel B] and its b."""
    G = np.zeros((18, 18))
    G[:12, :12] = d["H" + suffix]
    G[:6, 12:] = d["H_cross_pose_vel_B" + suffix]
    G[6:12, 12:] = d["H_cross_vel_AB" + suffix]
    G[12:, :12] = G[:12, 12:].T
    G[12:, 12:] = d["H_vel_B" + suffix]
    return G, np.concatenate([d["b" + suffix], d["b_vel_B" + suffix]])


def test_dual_single_term_is_rank_one_and_b_vel_b_is_the_chi2_gradient():
    """One moving leaf gives one term: the 18x18 dual normal matrix is
    g u u^T and b = g e u for u = [J_pose, J_vel_A, J_vel_B], so
    b b^T = g e^2 G entry for entry -- a transposed or misplaced block
    breaks it.  The residual takes the fixed leaf as stored; J_vel_B is
    the error's derivative when the fixed cloud moves with velocities v_B over
    its stamps.  At v_B = 0 that motion is the tree warp treeApplyVelocities
    (an independent routine), so for an inlier (g = 1) the central difference
    of chi2 over fixed trees warped by +-h e_k is 2 b_vel_B: this pins the
    leaf stamp, the sides and the sign J_vel_B enters the form with."""
    import types

    from .workloads import cfg2, icp_params

    w = cfg2(n_map=200_000, n_scan=16_384)
    p = O.from_kdh_params(icp_params(max_error=1.0, velocity_damping=10.0), 8)
    fixed = w["moving"]  # a stamped cloud as the fixed side, so J_vel_B is live
    moving = w["moving"]
    X_fixed = w["X0"]
    zero = np.zeros(6)
    checked = 0
    for leaf in moving.leaves[::11]:
        one = types.SimpleNamespace(nodes=moving.nodes, leaves=np.array([leaf], dtype=np.uint64))
        d = O.build_form_dual(fixed, one, w["X_gt"], w["v0"], X_fixed, zero, p)
        if d["stats"]["num_inliers"] != 1:
            continue
        G, b = _dual_full(d, "64")
        assert np.abs(G - G.T).max() == 0.0
        # b b^T = (g e^2) G: the ratio from the largest entry fits every entry
        # (FP64 twin), and it is chi2_kernelised (FP32 e, so 1e-4)
        bb = np.outer(b, b)
        m = np.unravel_index(np.abs(G).argmax(), G.shape)
        c = bb[m] / G[m]
        assert np.abs(c * G - bb).max() <= 1e-9 * np.abs(bb).max()
        assert abs(c - d["stats"]["chi2_kernelized"]) <= 1e-4 * c
        assert np.abs(d["b_vel_B64"]).max() > 0.0
        h = 1e-2
        fd = np.zeros(6)
        for k in range(6):
            chi = []
            for s in (1.0, -1.0):
                warped = types.SimpleNamespace(nodes=O.tree_apply_velocities(fixed.nodes, s * h * np.eye(6)[k]),
                                               leaves=fixed.leaves)
                ds = O.build_form_dual(warped, one, w["X_gt"], w["v0"], X_fixed, zero, p)
                assert ds["stats"]["num_inliers"] == 1 and np.array_equal(ds["corr"], d["corr"])
                chi.append(ds["stats"]["chi2"])
            fd[k] = (chi[0] - chi[1]) / (2 * h)
        assert np.abs(fd - 2 * d["b_vel_B64"]).max() <= 2e-2 * np.abs(fd).max(), (fd, d["b_vel_B64"])
        checked += 1
        if checked == 8:
            break
    assert checked >= 4
