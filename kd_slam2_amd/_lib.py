"""ctypes bindings of libkdhip.so (declarations in include/kdh.h, include/kdh_host.h).

The library is built in-tree by ``kd_slam2_amd.build.build()``.  Loading fails
loudly when it is missing: there is no CPU fallback for the device path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkdhip.so")

KDH_OK = 0
KDH_ERR_NOT_FOUND = -5
ERRORS = {
    -1: "KDH_ERR_INVALID",
    -2: "KDH_ERR_HIP",
    -3: "KDH_ERR_NOMEM",
    -4: "KDH_ERR_BUDGET",
    -5: "KDH_ERR_NOT_FOUND",
    -6: "KDH_ERR_CANCELLED",
}


class Node3f(C.Structure):
    """Node_<Point3fTraits> layout (kd_slam2/src/lib/kd_slam/tree/node_.h:12-43)."""

    _fields_ = [
        ("dts_mean", C.c_double),
        ("dts_covariance", C.c_double),
        ("mean", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("idx_start", C.c_uint64),
        ("idx_end", C.c_uint64),
        ("idx_left", C.c_int32),
        ("idx_right", C.c_int32),
    ]


NODE_DTYPE = np.dtype(
    [
        ("dts_mean", "<f8"),
        ("dts_covariance", "<f8"),
        ("mean", "<f4", (3,)),
        ("direction", "<f4", (3,)),
        ("idx_start", "<u8"),
        ("idx_end", "<u8"),
        ("idx_left", "<i4"),
        ("idx_right", "<i4"),
    ]
)
assert NODE_DTYPE.itemsize == 64 and C.sizeof(Node3f) == 64


class Params(C.Structure):
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
        ("recenter", C.c_int32),
    ]


class Stats(C.Structure):
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


class Prior(C.Structure):
    _fields_ = [("X_ref", C.c_double * 12), ("Z", C.c_double * 12), ("info", C.c_double * 36)]


class Pair(C.Structure):
    _fields_ = [
        ("fixed", C.c_void_p),
        ("moving", C.c_void_p),
        ("ct", C.c_int32),
        ("forest", C.c_int32),
        ("X0", C.c_double * 12),
        ("v0", C.c_double * 6),
        ("Xf", C.c_double * 12),
        ("params", Params),
        ("prior", C.c_void_p),
    ]


class DualPair(C.Structure):
    _fields_ = [
        ("fixed", C.c_void_p),
        ("moving", C.c_void_p),
        ("X_moving", C.c_double * 12),
        ("v_moving", C.c_double * 6),
        ("X_fixed", C.c_double * 12),
        ("v_fixed", C.c_double * 6),
        ("params", Params),
    ]


class DualForm(C.Structure):
    _fields_ = [
        ("H", C.c_double * 144),
        ("b", C.c_double * 12),
        ("H_dual_vel_B", C.c_double * 36),
        ("b_dual_vel_B", C.c_double * 6),
        ("H_dual_cross_pose_vel_B", C.c_double * 36),
        ("H_dual_cross_vel_AB", C.c_double * 36),
        ("stats", Stats),
    ]


class KDTree3fInfo(C.Structure):
    _fields_ = [
        ("stamp_sec", C.c_int32),
        ("stamp_nanosec", C.c_uint32),
        ("frame_id", C.c_char_p),
        ("src_topic", C.c_char_p),
        ("ts_start", C.c_double),
        ("root_mean", C.c_float * 3),
        ("root_eigenvalues", C.c_float * 3),
        ("root_eigenvectors", C.c_float * 9),
    ]


class McapMessage(C.Structure):
    _fields_ = [
        ("topic", C.c_char_p),
        ("schema_name", C.c_char_p),
        ("message_encoding", C.c_char_p),
        ("log_time", C.c_uint64),
        ("publish_time", C.c_uint64),
        ("sequence", C.c_uint32),
        ("channel_id", C.c_uint16),
        ("data", C.c_void_p),
        ("size", C.c_uint64),
    ]


class PairResult(C.Structure):
    _fields_ = [
        ("X", C.c_double * 12),
        ("v", C.c_double * 6),
        ("last", Stats),
        ("entry", Stats),
        ("iterations", C.c_int32),
        ("result", C.c_int32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("fixed", C.c_void_p),
        ("moving_nodes", C.c_void_p),
        ("n_moving_nodes", C.c_int64),
        ("moving_leaves", C.c_void_p),
        ("n_moving_leaves", C.c_int64),
        ("ct", C.c_int32),
        ("pad", C.c_int32),
        ("X0", C.c_double * 12),
        ("v0", C.c_double * 6),
        ("Xf", C.c_double * 12),
        ("params", Params),
        ("prior", C.c_void_p),
    ]


class RequestResult(C.Structure):
    _fields_ = [
        ("r", PairResult),
        ("latency_us", C.c_double),
        ("queue_us", C.c_double),
        ("status", C.c_int32),
        ("pad", C.c_int32),
    ]


class ServerInfo(C.Structure):
    _fields_ = [
        ("bytes_budget", C.c_int64),
        ("bytes_resident", C.c_int64),
        ("bytes_inflight", C.c_int64),
        ("inflight", C.c_int32),
        ("queued", C.c_int32),
        ("steps", C.c_int64),
        ("completed", C.c_int64),
    ]


class TreeGenParams(C.Structure):
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


_lib = None

P = C.c_void_p
F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int32)

_SIGS = {
    "kdh_params_default": (None, [C.POINTER(Params)]),
    "kdh_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "kdh_ctx_destroy": (C.c_int, [P]),
    "kdh_last_error": (C.c_char_p, [P]),
    "kdh_tree_upload": (C.c_int, [P, P, C.c_int64, P, C.c_int64, C.POINTER(P)]),
    "kdh_tree_free": (C.c_int, [P, P]),
    "kdh_tree_info": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "kdh_tree_apply_velocities": (C.c_int, [P, P, F32P, C.c_int]),
    "kdh_tree_apply_isometry": (C.c_int, [P, P, F32P]),
    "kdh_tree_extract_leaves": (C.c_int, [P, P, F32P, F32P]),
    "kdh_tree_read_nodes": (C.c_int, [P, P, F32P]),
    "kdh_tree_clone": (C.c_int, [P, P, C.POINTER(P)]),
    "kdh_tree_copy_nodes": (C.c_int, [P, P, P]),
    "kdh_icp_build": (C.c_int, [P, P, P, F32P, F32P, F32P, C.POINTER(Params), C.c_int, F64P, F64P, C.POINTER(Stats)]),
    "kdh_build_at": (
        C.c_int,
        [P, P, P, C.c_int, F64P, F64P, F64P, C.POINTER(Params), C.c_int, F64P, F64P, C.POINTER(Stats)],
    ),
    "kdh_cticp_build": (
        C.c_int,
        [P, P, P, F32P, F32P, F32P, F32P, C.POINTER(Params), C.c_int, F64P, F64P, C.POINTER(Stats)],
    ),
    "kdh_correspondences": (C.c_int, [P, P, P, C.c_int, F32P, F32P, F32P, F32P, C.POINTER(Params), I32P, I32P]),
    "kdh_align": (
        C.c_int,
        [P, P, P, C.c_int, F64P, F64P, F64P, C.POINTER(Params), C.POINTER(Prior), C.c_int, F64P, F64P,
         C.POINTER(Stats), C.POINTER(C.c_int), C.POINTER(C.c_int), F32P],
    ),
    "kdh_align_noalloc": (
        C.c_int,
        [P, P, P, C.c_int, F64P, F64P, F64P, C.POINTER(Params), C.POINTER(Prior), C.c_int, F64P, F64P,
         C.POINTER(Stats), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    ),
    "kdh_align_noalloc_max_leaves": (C.c_int, []),
    "kdh_batch_align": (C.c_int, [P, C.c_int, C.POINTER(Pair), C.POINTER(PairResult)]),
    "kdh_batch_prepare": (C.c_int, [P, C.c_int, C.POINTER(Pair), C.POINTER(P)]),
    "kdh_batch_run": (C.c_int, [P, F64P, F64P, C.POINTER(PairResult)]),
    "kdh_batch_submit": (C.c_int, [P, F64P, F64P]),
    "kdh_batch_collect": (C.c_int, [P, C.POINTER(PairResult)]),
    "kdh_batch_release": (C.c_int, [P]),
    "kdh_batch_build": (C.c_int, [P, C.c_int, C.POINTER(Pair), C.c_int, F64P, F64P, C.POINTER(Stats)]),
    "kdh_batch_build_dual": (C.c_int, [P, C.c_int, C.POINTER(DualPair), C.c_int, C.POINTER(DualForm)]),
    "kdh_last_kernel_time": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "kdh_set_launch_config": (C.c_int, [P, C.c_int, C.c_int]),
    "kdh_set_leaves_per_lane": (C.c_int, [P, C.c_int]),
    "kdh_set_tile_leaves": (C.c_int, [P, C.c_int]),
    "kdh_set_batch_streams": (C.c_int, [P, C.c_int]),
    "kdh_set_profiling": (C.c_int, [P, C.c_int]),
    "kdh_last_kernel_breakdown": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "kdh_server_create": (C.c_int, [C.c_int, C.c_int64, C.c_int32, C.POINTER(P)]),
    "kdh_server_destroy": (C.c_int, [P]),
    "kdh_server_last_error": (C.c_char_p, [P]),
    "kdh_server_add_resident": (C.c_int, [P, P]),
    "kdh_server_remove_resident": (C.c_int, [P, P]),
    "kdh_server_submit": (C.c_int, [P, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "kdh_server_poll": (C.c_int, [P, C.c_int64, C.POINTER(C.c_int), C.POINTER(RequestResult)]),
    "kdh_server_wait": (C.c_int, [P, C.c_int64, C.c_double, C.POINTER(RequestResult)]),
    "kdh_server_cancel": (C.c_int, [P, C.c_int64]),
    "kdh_server_info_get": (C.c_int, [P, C.POINTER(ServerInfo)]),
    "kdh_tree_gen_params_default": (None, [C.POINTER(TreeGenParams)]),
    "kdh_host_tree_build": (C.c_int, [F32P, F64P, C.c_int64, C.POINTER(TreeGenParams), C.POINTER(P)]),
    "kdh_tree_build": (C.c_int, [P, F32P, F64P, C.c_int64, C.POINTER(TreeGenParams), C.c_int, C.POINTER(P)]),
    "kdh_tree_download": (C.c_int, [P, P, P, P, F32P, F64P]),
    "kdh_kdtree3f_encoded_size": (C.c_int, [C.POINTER(KDTree3fInfo), C.c_int64, C.POINTER(C.c_size_t)]),
    "kdh_kdtree3f_encode": (C.c_int, [C.POINTER(KDTree3fInfo), P, C.c_int64, P, C.c_size_t,
                                      C.POINTER(C.c_size_t)]),
    "kdh_kdtree3f_decode": (C.c_int, [P, C.c_size_t, C.POINTER(KDTree3fInfo), C.c_char_p, C.c_size_t, C.c_char_p,
                                      C.c_size_t, P, C.c_int64, C.POINTER(C.c_int64)]),
    "kdh_tree_leaves": (C.c_int, [P, C.c_int64, P, C.c_int64, C.POINTER(C.c_int64)]),
    "kdh_kdtree3f_schema_text": (C.c_char_p, []),
    "kdh_mcap_reader_open": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "kdh_mcap_reader_next": (C.c_int, [P, C.POINTER(McapMessage)]),
    "kdh_mcap_reader_error": (C.c_char_p, [P]),
    "kdh_mcap_reader_close": (C.c_int, [P]),
    "kdh_mcap_writer_open": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "kdh_mcap_writer_add_channel": (C.c_int, [P, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_uint16)]),
    "kdh_mcap_writer_write": (C.c_int, [P, C.c_uint16, C.c_uint64, C.c_uint64, P, C.c_uint64]),
    "kdh_mcap_writer_close": (C.c_int, [P]),
    "kdh_host_tree_info": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "kdh_host_tree_nodes": (P, [P]),
    "kdh_host_tree_leaves": (P, [P]),
    "kdh_host_tree_points": (P, [P]),
    "kdh_host_tree_free": (C.c_int, [P]),
}


_NP_SCALARS = {C.c_void_p: np.uint64, C.c_int32: np.int32, C.c_int64: np.int64, C.c_float: np.float32,
               C.c_double: np.float64}


def np_dtype(struct):
    """numpy dtype with the byte layout of a ctypes Structure (pointers as
    uint64), so arrays of the C structs can be filled and read vectorised."""
    names, formats, offsets = [], [], []
    for name, ctype in struct._fields_:
        names.append(name)
        offsets.append(getattr(struct, name).offset)
        if isinstance(ctype, type) and issubclass(ctype, C.Array):
            formats.append((_NP_SCALARS[ctype._type_], (ctype._length_,)))
        elif isinstance(ctype, type) and issubclass(ctype, C.Structure):
            formats.append(np_dtype(ctype))
        else:
            formats.append(_NP_SCALARS[ctype])
    return np.dtype(dict(names=names, formats=formats, offsets=offsets, itemsize=C.sizeof(struct)))


def lib():
    """Load libkdhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with kd_slam2_amd.build.build() "
                "(there is no CPU fallback for the HIP path)"
            )
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


ADAPTER_PATH = os.path.join(_HERE, "libkd_slam_hip_3d.so")
_adapter = None


def adapter():
    """Load libkd_slam_hip_3d.so (the ICPHIP3D / CTICPHIP3D configurable types)."""
    global _adapter
    if _adapter is None:
        lib()  # the adapter links against libkdhip.so
        if not os.path.exists(ADAPTER_PATH):
            raise RuntimeError(f"{ADAPTER_PATH} is missing: build it with kd_slam2_amd.build.build()")
        A = C.CDLL(ADAPTER_PATH)
        A.kdha_registered_types.restype = C.c_int
        A.kdha_registered_types.argtypes = [C.c_char_p, C.c_int]
        A.kdha_align.restype = C.c_int
        A.kdha_align.argtypes = [C.c_char_p, P, P, P, F64P, F64P, F64P, C.POINTER(Params), C.c_int, C.c_int, F64P,
                                 F64P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(Stats), C.POINTER(Stats)]
        A.kdha_build_dual.restype = C.c_int
        A.kdha_build_dual.argtypes = [P, P, P, F64P, F64P, F64P, F64P, C.POINTER(Params), C.c_int,
                                      C.POINTER(DualForm)]
        _adapter = A
    return _adapter


def f32p(a):
    return a.ctypes.data_as(F32P)


def f64p(a):
    return a.ctypes.data_as(F64P)


def i32p(a):
    return a.ctypes.data_as(I32P)


def check(rc, ctx=None, what=""):
    if rc != KDH_OK:
        msg = ""
        if ctx is not None:
            m = lib().kdh_last_error(ctx)
            msg = m.decode() if m else ""
        raise RuntimeError(f"{what} failed: {ERRORS.get(rc, rc)}: {msg}")
