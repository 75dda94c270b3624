"""ctypes wrapper of oracle/liboracle.so (test infrastructure only).

See oracle/oracle.cpp for the reference functions each entry point restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


class OrcParams(C.Structure):
    _fields_ = [
        ("kernel_threshold", C.c_float),
        ("damping", C.c_float),
        ("max_error", C.c_float),
        ("min_normal_cos", C.c_float),
        ("min_good_measurements", C.c_int32),
        ("max_iterations", C.c_int32),
        ("velocity_damping", C.c_float),
        ("use_term_crit", C.c_int32),
        ("term_d_pert", C.c_float),
        ("term_d_score", C.c_float),
        ("num_threads", C.c_int32),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("num_inliers", C.c_int32),
        ("num_outliers", C.c_int32),
        ("num_bad", C.c_int32),
        ("num_err_normal", C.c_int32),
        ("num_err_fn", C.c_int32),
        ("pad", C.c_int32),
        ("chi2", C.c_double),
        ("chi2_kernelized", C.c_double),
        ("pert_pose_norm", C.c_double),
        ("pert_vel_norm", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class OrcTreeGenParams(C.Structure):
    """kdh_tree_gen_params / TreeGenerator_ parameters (tree_generator_.h:18-25)."""

    _fields_ = [
        ("propagate_normal", C.c_int32),
        ("tree_min_good_leaves", C.c_int32),
        ("leaf_min_points", C.c_int32),
        ("leaf_max_points", C.c_int32),
        ("leaf_density", C.c_float),
        ("leaf_min_normal_incidence", C.c_float),
        ("leaf_eigenvalue_threshold", C.c_float),
        ("leaf_normal_eigenratio", C.c_float),
        ("parallel_level", C.c_int32),
        ("label_mode", C.c_int32),
    ]


def tree_gen_params(**kw):
    """The reference defaults (tree_generator_.h:18-25), overridden by kw."""
    p = OrcTreeGenParams(propagate_normal=0, tree_min_good_leaves=10, leaf_min_points=5, leaf_max_points=30,
                         leaf_density=100.0, leaf_min_normal_incidence=0.1, leaf_eigenvalue_threshold=1e-3,
                         leaf_normal_eigenratio=0.1, parallel_level=0, label_mode=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(fast=False):
    """The parity build (liboracle.so), or the timing build when fast; the
    environment variable KDH_ORACLE_LIB names another parity build to load
    (the sanitizer build of tests/test_sanitizers.py)."""
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not fast and os.environ.get("KDH_ORACLE_LIB"):
            path = os.environ["KDH_ORACLE_LIB"]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, f32, f64, i32 = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.orc_build.restype = C.c_int
        L.orc_build.argtypes = [C.c_int, vp, C.c_int64, vp, vp, C.c_int64, f32, f32, f32, f32, C.POINTER(OrcParams),
                                C.c_int, f32, f32, f64, f64, f64, C.POINTER(OrcStats), i32, i32]
        L.orc_compute.restype = C.c_int
        L.orc_compute.argtypes = [C.c_int, vp, C.c_int64, vp, vp, C.c_int64, f32, f32, f32, C.POINTER(OrcParams), f32,
                                  f32, C.c_int, C.POINTER(OrcStats), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_compute_forest.restype = C.c_int
        L.orc_compute_forest.argtypes = [C.c_int, C.c_int, C.POINTER(vp), C.POINTER(C.c_int64), f32, vp, vp,
                                         C.c_int64, f32, f32, C.POINTER(OrcParams), f32, f32, C.c_int,
                                         C.POINTER(OrcStats), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(OrcStats)]
        L.orc_build_dual.restype = C.c_int
        L.orc_build_dual.argtypes = [vp, C.c_int64, vp, vp, C.c_int64, f64, f64, f64, f64, C.POINTER(OrcParams),
                                     C.c_int, f32, f64, C.POINTER(OrcStats), i32]
        L.orc_tree_build.restype = C.c_int
        L.orc_tree_build.argtypes = [f32, f64, C.c_int64, vp, vp, C.POINTER(C.c_int64), vp, C.POINTER(C.c_int64),
                                     f32, f64]
        L.orc_jacobian_velocity_b.argtypes = [f32, f32, f32, f32, f32, f32, C.c_float, f32]
        L.orc_tree_apply_velocities.argtypes = [vp, C.c_int64, f32, C.c_int]
        L.orc_tree_apply_isometry.argtypes = [vp, C.c_int64, f32]
        L.orc_extract_leaves.argtypes = [vp, vp, C.c_int64, f32, f32]
        L.orc_build_at.restype = C.c_int
        L.orc_build_at.argtypes = [C.c_int, vp, C.c_int64, vp, vp, C.c_int64, f64, f64, f64, C.POINTER(OrcParams),
                                   f32, f32, f64, f64, i32]
        L.orc_compute64.restype = C.c_int
        L.orc_compute64.argtypes = [C.c_int, vp, C.c_int64, vp, vp, C.c_int64, f64, f64, f64, C.POINTER(OrcParams),
                                    C.POINTER(C.c_int)]
        L.orc_update_cache.argtypes = [f32, f32, f32]
        L.orc_ldlt_solve_f.argtypes = [C.c_int, f32, f32, f32]
        L.orc_ldlt_solve_d.argtypes = [C.c_int, f64, f64, f64]
        L.orc_sincos.argtypes = [f32, C.c_int64, f32, f32]
        L.orc_rodrigues_f.argtypes = [f32, f32]
        L.orc_expmap_f.argtypes = [f32, f32]
        L.orc_error_and_jacobian.restype = C.c_int
        L.orc_error_and_jacobian.argtypes = [f32, f32, f32, f32, C.c_float, C.c_float, f32, f32]
        L.orc_error_and_jacobian_velocity_a.argtypes = [f32, f32, f32, f32, f32, f32, C.c_float, f32, f32]
        L.orc_stable_sum_f.restype = C.c_float
        L.orc_stable_sum_f.argtypes = [f32, C.c_int64]
        L.orc_prior_form.argtypes = [f32, f32, f32, f32, f32]
        L.orc_pose_to_pose_error.argtypes = [f32, f32, f32, f32]
        L.orc_generate_scene.restype = C.c_int64
        L.orc_generate_scene.argtypes = [C.c_float, C.c_float, C.c_int, f32, C.c_int64]
        L.orc_make_scan.restype = C.c_int64
        L.orc_make_scan.argtypes = [f32, C.c_int64, f32, f32, C.c_double, f32, f64, C.c_int64]
        L.orc_scene_expmap.argtypes = [f32, f32]
        L.orc_eig_direct.argtypes = [f32, f32, f32]
        _libs[name] = L
    return _libs[name]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def p32(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def pi32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_params(kernel_threshold=0.1, damping=1.0, max_error=1.0, min_normal_cos=None, min_good_measurements=10,
                max_iterations=20, velocity_damping=100.0, use_term_crit=0, term_d_pert=1e-3, term_d_score=1e-3,
                num_threads=0):
    p = OrcParams()
    p.kernel_threshold = kernel_threshold
    p.damping = damping
    p.max_error = max_error
    p.min_normal_cos = np.float32(np.cos(25.0 / 180.0 * np.pi)) if min_normal_cos is None else min_normal_cos
    p.min_good_measurements = min_good_measurements
    p.max_iterations = max_iterations
    p.velocity_damping = velocity_damping
    p.use_term_crit = use_term_crit
    p.term_d_pert = term_d_pert
    p.term_d_score = term_d_score
    p.num_threads = num_threads
    return p


def from_kdh_params(kp, num_threads=0):
    return make_params(kp.kernel_threshold, kp.damping, kp.max_error, kp.min_normal_cos, kp.min_good_measurements,
                       kp.max_iterations, kp.velocity_damping, kp.use_term_crit, kp.term_d_pert, kp.term_d_score,
                       num_threads)


def build_form(fixed, moving, Xm34, Xf34, Xmf34, params, ct=0, v=None, stats_mode=False, fast=False):
    """One quadratic form at a given FP32 cache.  fixed/moving: HostTree-like
    objects with .nodes (64-byte Node_ records) and .leaves."""
    D = 12 if ct else 6
    H = np.zeros(D * D, dtype=np.float32)
    b = np.zeros(D, dtype=np.float32)
    H64 = np.zeros(D * D)
    b64 = np.zeros(D)
    c64 = np.zeros(2)
    st = OrcStats()
    n = len(moving.leaves)
    corr = np.zeros(max(n, 1), dtype=np.int32)
    depth = np.zeros(max(n, 1), dtype=np.int32)
    vv = _f32(np.zeros(6) if v is None else v)
    Xm, Xf, Xmf = _f32(Xm34), _f32(Xf34), _f32(Xmf34)
    lib(fast).orc_build(ct, fixed.nodes.ctypes.data, len(fixed.nodes), moving.nodes.ctypes.data,
                        moving.leaves.ctypes.data, n, p32(Xm), p32(Xf), p32(Xmf), p32(vv), C.byref(params),
                        int(stats_mode), p32(H), p32(b), p64(H64), p64(b64), p64(c64), C.byref(st), pi32(corr),
                        pi32(depth))
    return dict(H=H.reshape(D, D), b=b, H64=H64.reshape(D, D), b64=b64, chi2_64=c64[0], chi2k_64=c64[1],
                stats=st.as_dict(), corr=corr[:n], depth=depth[:n])


def compute(fixed, moving, X0, Xf, params, ct=0, v0=None, prior_ZXi=None, prior_info=None, stats_mode=False,
            fast=False):
    """TreeAlignerBase_::compute in FP32.  Returns (X 4x4, v, stats_log, result)."""
    X = _f32(np.asarray(X0, dtype=np.float64)[:3, :4].reshape(12))
    v = _f32(np.zeros(6) if v0 is None else v0)
    Xf34 = _f32(np.asarray(Xf, dtype=np.float64)[:3, :4].reshape(12))
    n = max(1, params.max_iterations)
    log = (OrcStats * n)()
    nlog, res = C.c_int(), C.c_int()
    pz = None if prior_ZXi is None else _f32(prior_ZXi)
    pinf = None if prior_info is None else _f32(prior_info)
    lib(fast).orc_compute(ct, fixed.nodes.ctypes.data, len(fixed.nodes), moving.nodes.ctypes.data,
                          moving.leaves.ctypes.data, len(moving.leaves), p32(X), p32(v), p32(Xf34), C.byref(params),
                          None if pz is None else p32(pz), None if pinf is None else p32(pinf), int(stats_mode), log,
                          C.byref(nlog), C.byref(res))
    Xo = np.eye(4)
    Xo[:3, :4] = X.reshape(3, 4).astype(np.float64)
    return Xo, v.astype(np.float64), [log[i].as_dict() for i in range(nlog.value)], res.value


def compute_forest(fixed_list, Xf_list, moving, X0, params, ct=0, v0=None, prior_ZXi=None, prior_info=None,
                   stats_mode=False, fast=False):
    """compute() against a fixed forest (one tree and pose per entry), FP32.
    Returns (X 4x4, v, stats_log, result, per-entry stats of the last round)."""
    n = len(fixed_list)
    X = _f32(np.asarray(X0, dtype=np.float64)[:3, :4].reshape(12))
    v = _f32(np.zeros(6) if v0 is None else v0)
    Xf = _f32(np.concatenate([np.asarray(x, dtype=np.float64)[:3, :4].reshape(12) for x in Xf_list]))
    nodes = (C.c_void_p * n)(*[t.nodes.ctypes.data for t in fixed_list])
    counts = (C.c_int64 * n)(*[len(t.nodes) for t in fixed_list])
    m = max(1, params.max_iterations)
    log = (OrcStats * m)()
    ent = (OrcStats * n)()
    nlog, res = C.c_int(), C.c_int()
    pz = None if prior_ZXi is None else _f32(prior_ZXi)
    pinf = None if prior_info is None else _f32(prior_info)
    lib(fast).orc_compute_forest(ct, n, nodes, counts, p32(Xf), moving.nodes.ctypes.data, moving.leaves.ctypes.data,
                                 len(moving.leaves), p32(X), p32(v), C.byref(params),
                                 None if pz is None else p32(pz), None if pinf is None else p32(pinf),
                                 int(stats_mode), log, C.byref(nlog), C.byref(res), ent)
    Xo = np.eye(4)
    Xo[:3, :4] = X.reshape(3, 4).astype(np.float64)
    return (Xo, v.astype(np.float64), [log[i].as_dict() for i in range(nlog.value)], res.value,
            [ent[i].as_dict() for i in range(n)])


DUAL_BLOCKS = (("H", 0, (12, 12)), ("b", 144, (12,)), ("H_vel_B", 156, (6, 6)), ("b_vel_B", 192, (6,)),
               ("H_cross_pose_vel_B", 198, (6, 6)), ("H_cross_vel_AB", 234, (6, 6)))


def build_form_dual(fixed, moving, Xm, vm, Xf, vf, params, stats_mode=False):
    """CTICP_::buildQuadraticFormDual at an FP64 state (moving pose and
    velocities, fixed pose and velocities): the FP32 reference blocks, the
    FP64 twin's (suffix 64), the stats and the correspondences."""
    out32 = np.zeros(270, dtype=np.float32)
    out64 = np.zeros(270)
    st = OrcStats()
    n = len(moving.leaves)
    corr = np.zeros(max(n, 1), dtype=np.int32)
    a = lambda X: np.array(np.asarray(X, dtype=np.float64)[:3, :4].reshape(12))  # noqa: E731
    Xm34, Xf34 = a(Xm), a(Xf)
    vm6, vf6 = np.array(vm, dtype=np.float64), np.array(vf, dtype=np.float64)
    lib().orc_build_dual(fixed.nodes.ctypes.data, len(fixed.nodes), moving.nodes.ctypes.data,
                         moving.leaves.ctypes.data, n, p64(Xm34), p64(vm6), p64(Xf34), p64(vf6), C.byref(params),
                         int(stats_mode), p32(out32), p64(out64), C.byref(st), pi32(corr))
    d = dict(stats=st.as_dict(), corr=corr[:n])
    for name, off, shape in DUAL_BLOCKS:
        size = int(np.prod(shape))
        d[name] = out32[off:off + size].reshape(shape).copy()
        d[name + "64"] = out64[off:off + size].reshape(shape).copy()
    return d


def tree_apply_velocities(nodes, v, skip_leaves=False):
    """treeApplyVelocities_kernel on a copy of a Node_ array; returns it."""
    out = np.array(nodes, copy=True)
    lib().orc_tree_apply_velocities(out.ctypes.data, len(out), p32(_f32(v)), int(skip_leaves))
    return out


def tree_apply_isometry(nodes, X):
    """treeApplyIsometry_kernel on a copy of a Node_ array; returns it."""
    out = np.array(nodes, copy=True)
    m = _f32(np.asarray(X, dtype=np.float64)[:3, :4].reshape(12))
    lib().orc_tree_apply_isometry(out.ctypes.data, len(out), p32(m))
    return out


def extract_leaves(nodes, leaves, v=None):
    """extractCompressedLeaves(v): (n_leaves, 2, 3) float32."""
    n = len(leaves)
    out = np.zeros((n, 2, 3), dtype=np.float32)
    vv = None if v is None else _f32(v)
    lib().orc_extract_leaves(nodes.ctypes.data, leaves.ctypes.data, n, None if vv is None else p32(vv), p32(out))
    return out


def build_form_at(fixed, moving, X, Xf, params, ct=0, v=None, fast=False):
    """build_form at an FP64 state: the FP32 form at the cache rounded from
    (X, Xf, v), and the FP64 twin evaluated at the state itself."""
    D = 12 if ct else 6
    H = np.zeros(D * D, dtype=np.float32)
    b = np.zeros(D, dtype=np.float32)
    H64 = np.zeros(D * D)
    b64 = np.zeros(D)
    n = len(moving.leaves)
    corr = np.zeros(max(n, 1), dtype=np.int32)
    X34 = np.array(np.asarray(X, dtype=np.float64)[:3, :4].reshape(12))
    Xf34 = np.array(np.asarray(Xf, dtype=np.float64)[:3, :4].reshape(12))
    vv = np.array(np.zeros(6) if v is None else v, dtype=np.float64)
    lib(fast).orc_build_at(ct, fixed.nodes.ctypes.data, len(fixed.nodes), moving.nodes.ctypes.data,
                           moving.leaves.ctypes.data, n, p64(X34), p64(Xf34), p64(vv), C.byref(params), p32(H),
                           p32(b), p64(H64), p64(b64), pi32(corr))
    return dict(H=H.reshape(D, D), b=b, H64=H64.reshape(D, D), b64=b64, corr=corr[:n])


def compute64(fixed, moving, X0, Xf, params, ct=0, v0=None, fast=False):
    """FP64 yardstick of a whole alignment (FP32 decisions, FP64 terms, sums,
    solve and state).  Returns (X 4x4, v, iterations)."""
    X = np.array(np.asarray(X0, dtype=np.float64)[:3, :4].reshape(12))  # own copy: written in place
    v = np.array(np.zeros(6) if v0 is None else v0, dtype=np.float64)
    Xf34 = np.array(np.asarray(Xf, dtype=np.float64)[:3, :4].reshape(12))
    nit = C.c_int()
    lib(fast).orc_compute64(ct, fixed.nodes.ctypes.data, len(fixed.nodes), moving.nodes.ctypes.data,
                            moving.leaves.ctypes.data, len(moving.leaves), p64(X), p64(v), p64(Xf34),
                            C.byref(params), C.byref(nit))
    Xo = np.eye(4)
    Xo[:3, :4] = X.reshape(3, 4)
    return Xo, v, nit.value


def update_cache(Xm34, Xf34):
    Xm, Xf = _f32(Xm34), _f32(Xf34)
    out = np.zeros(12, dtype=np.float32)
    lib().orc_update_cache(p32(Xm), p32(Xf), p32(out))
    return out


def tree_build(points, stamps=None, **params):
    """TreeGenerator_::makeTree: a dict with the BFS-ordered Node_ array
    (NODE_DTYPE), the leaf list, the permuted points and stamps; None where
    makeTree returns nullptr."""
    from kd_slam2_amd._lib import NODE_DTYPE

    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    n = len(pts)
    st = None if stamps is None else np.ascontiguousarray(stamps, dtype=np.float64)
    nodes = np.zeros(max(2 * n, 1), dtype=NODE_DTYPE)
    leaves = np.zeros(max(n, 1), dtype=np.uint64)
    xyz = np.zeros((max(n, 1), 3), dtype=np.float32)
    sto = np.zeros(max(n, 1))
    nn, nl = C.c_int64(), C.c_int64()
    p = tree_gen_params(**params)
    ok = lib().orc_tree_build(p32(pts), None if st is None else p64(st), n, C.byref(p), nodes.ctypes.data,
                              C.byref(nn), leaves.ctypes.data, C.byref(nl), p32(xyz), p64(sto))
    if not ok:
        return None
    return dict(nodes=nodes[:nn.value].copy(), leaves=leaves[:nl.value].copy(), points=xyz[:n].copy(),
                stamps=sto[:n].copy())


def eig_direct(A):
    """SelfAdjointEigenSolver<Matrix3f>::computeDirect of a symmetric 3x3:
    (ascending eigenvalues (3,), eigenvectors as columns (3, 3)), FP32."""
    A = np.ascontiguousarray(A, dtype=np.float32).reshape(3, 3)
    w = np.zeros(3, dtype=np.float32)
    V = np.zeros((3, 3), dtype=np.float32)
    lib().orc_eig_direct(p32(A), p32(w), p32(V))
    return w, V


# ---- the reference tests' scenes (oracle/scenes.cpp) -------------------------

def generate_scene(range_=20.0, resolution=0.05, num_planes=10):
    """generateScene (d3/typedefs.cpp:51-64) from a fresh process's random
    state: (n, 3) float32."""
    L = lib()
    n = L.orc_generate_scene(range_, resolution, num_planes, None, 0)
    out = np.zeros((n, 3), dtype=np.float32)
    L.orc_generate_scene(range_, resolution, num_planes, p32(out), n)
    return out


def make_scan(scene, start_pose=None, velocities=None, t_start=0.0):
    """Scanner::makeScan (d3/scanner.cpp:5-83), default ScannerParams:
    (points (m, 3) float32 in the sensor frame, stamps (m,) float64)."""
    L = lib()
    sc = _f32(scene)
    X = _f32(np.eye(4)[:3, :4].reshape(12) if start_pose is None else np.asarray(start_pose)[:3, :4].reshape(12))
    v = _f32(np.zeros(6) if velocities is None else velocities)
    cap = 1024 * 128
    xyz = np.zeros((cap, 3), dtype=np.float32)
    st = np.zeros(cap)
    m = L.orc_make_scan(p32(sc), len(sc), p32(X), p32(v), t_start, p32(xyz), p64(st), cap)
    return xyz[:m].copy(), st[:m].copy()


def scene_expmap(v):
    """GeometryTraits_<3>::expmap in FP32 (geometry_3d_impl_.h:179-184): 4x4."""
    out = np.zeros(12, dtype=np.float32)
    lib().orc_scene_expmap(p32(_f32(v)), p32(out))
    X = np.eye(4)
    X[:3, :4] = out.reshape(3, 4)
    return X


def pose_to_pose_error(ZXi, X, jacobian=True):
    """poseToPoseError / dPoseToPoseError in FP32 (geometry_3d_impl_.h:117-140):
    (e (6,), J (6, 6) or None)."""
    z = _f32(np.asarray(ZXi)[:3, :4].reshape(12))
    x = _f32(np.asarray(X)[:3, :4].reshape(12))
    e = np.zeros(6, dtype=np.float32)
    J = np.zeros(36, dtype=np.float32) if jacobian else None
    lib().orc_pose_to_pose_error(p32(z), p32(x), p32(e), None if J is None else p32(J))
    return e, (None if J is None else J.reshape(6, 6))
