"""Runs the host code of the library and the oracle under AddressSanitizer
and UndefinedBehaviorSanitizer (CPU only).  Started as a child process by
tests/test_sanitizers.py with the sanitizer runtime preloaded; any report
aborts the process with a non-zero status.

  * the host tree builder (kdh_host_tree_build) on the shared test clouds,
    random clouds and degenerate ones, threaded and not;
  * the KDTree3f CDR codec on valid messages and on a seeded corpus of
    mutated ones (bit flips, truncations, overwritten length fields) --
    bytes from a bag are untrusted input;
  * the MCAP reader on a valid file and on mutated copies of it;
  * the oracle's tree generation, whole alignments and dual form.

Usage: python -m tests.sanitize_driver <libkdhost_san.so> <liboracle_san.so>
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from kd_slam2_amd import _lib as L  # noqa: E402  (structs only; libkdhip.so is not loaded)

P = C.c_void_p


def host_lib(path):
    H = C.CDLL(path)
    sig = {
        "kdh_host_tree_build": (C.c_int, [P, P, C.c_int64, C.POINTER(L.TreeGenParams), C.POINTER(P)]),
        "kdh_host_tree_info": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "kdh_host_tree_nodes": (P, [P]),
        "kdh_host_tree_leaves": (P, [P]),
        "kdh_host_tree_free": (C.c_int, [P]),
        "kdh_kdtree3f_encoded_size": (C.c_int, [C.POINTER(L.KDTree3fInfo), C.c_int64, C.POINTER(C.c_size_t)]),
        "kdh_kdtree3f_encode": (C.c_int, [C.POINTER(L.KDTree3fInfo), P, C.c_int64, P, C.c_size_t,
                                          C.POINTER(C.c_size_t)]),
        "kdh_kdtree3f_decode": (C.c_int, [P, C.c_size_t, C.POINTER(L.KDTree3fInfo), P, C.c_size_t, P, C.c_size_t,
                                          P, C.c_int64, C.POINTER(C.c_int64)]),
        "kdh_tree_leaves": (C.c_int, [P, C.c_int64, P, C.c_int64, C.POINTER(C.c_int64)]),
        "kdh_kdtree3f_schema_text": (C.c_char_p, []),
        "kdh_mcap_reader_open": (C.c_int, [C.c_char_p, C.POINTER(P)]),
        "kdh_mcap_reader_next": (C.c_int, [P, C.POINTER(L.McapMessage)]),
        "kdh_mcap_reader_close": (C.c_int, [P]),
        "kdh_mcap_writer_open": (C.c_int, [C.c_char_p, C.POINTER(P)]),
        "kdh_mcap_writer_add_channel": (C.c_int, [P, C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_uint16)]),
        "kdh_mcap_writer_write": (C.c_int, [P, C.c_uint16, C.c_uint64, C.c_uint64, P, C.c_uint64]),
        "kdh_mcap_writer_close": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(H, name)
        f.restype, f.argtypes = res, args
    return H


def gen_params(**kw):
    p = L.TreeGenParams()
    base = dict(propagate_normal=0, tree_min_good_leaves=10, leaf_min_points=5, leaf_max_points=30,
                leaf_density=100.0, leaf_min_normal_incidence=0.1, leaf_eigenvalue_threshold=1e-3,
                leaf_normal_eigenratio=0.1, parallel_level=0, label_mode=0)
    base.update(kw)
    for k, v in base.items():
        setattr(p, k, v)
    return p


def build_tree(H, pts, stamps, **kw):
    pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
    st = None if stamps is None else np.ascontiguousarray(stamps, dtype=np.float64)
    h = P()
    p = gen_params(**kw)
    rc = H.kdh_host_tree_build(pts.ctypes.data, None if st is None else st.ctypes.data, len(pts), C.byref(p),
                               C.byref(h))
    assert rc == 0, rc
    if not h.value:
        return None
    nn, nl, ts = C.c_int64(), C.c_int64(), C.c_double()
    H.kdh_host_tree_info(h, C.byref(nn), C.byref(nl), C.byref(ts))
    nodes = np.ctypeslib.as_array((L.Node3f * nn.value).from_address(H.kdh_host_tree_nodes(h))).copy() \
        if nn.value else None
    H.kdh_host_tree_free(h)
    return np.frombuffer(nodes, dtype=L.NODE_DTYPE).copy() if nodes is not None else None


def trees(H):
    from tests.workloads import tree_clouds

    out = []
    for name, pts, st, kw in tree_clouds():
        for level in (0, 2):
            for mode in (0, 1):
                t = build_tree(H, pts, st, parallel_level=level, label_mode=mode, **kw)
                if t is not None:
                    out.append(t)
    rng = np.random.default_rng(7)
    for mode in (0, 1):
        for n in (0, 1, 4, 5, 31, 200, 5000):
            build_tree(H, rng.normal(size=(n, 3)), rng.uniform(0, 0.1, n) if n else None, tree_min_good_leaves=0,
                       label_mode=mode)
        build_tree(H, np.full((100, 3), np.nan, dtype=np.float32), None, tree_min_good_leaves=0, label_mode=mode)
        build_tree(H, np.full((100, 3), np.inf, dtype=np.float32), None, tree_min_good_leaves=0, label_mode=mode)
    print(f"tree builder: {len(out)} trees", flush=True)
    return out


def encode(H, nodes, frame=b"map", topic=b"/kd_slam/keyframe"):
    info = L.KDTree3fInfo()
    info.stamp_sec, info.stamp_nanosec = 12, 345
    info.frame_id, info.src_topic = frame, topic
    info.ts_start = 1.5
    size = C.c_size_t()
    assert H.kdh_kdtree3f_encoded_size(C.byref(info), len(nodes), C.byref(size)) == 0
    buf = (C.c_uint8 * size.value)()
    n = C.c_size_t()
    assert H.kdh_kdtree3f_encode(C.byref(info), nodes.ctypes.data, len(nodes), buf, size.value, C.byref(n)) == 0
    return bytes(buf[: n.value])


def decode(H, data, node_cap=None):
    raw = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data if data else b"\0")
    info = L.KDTree3fInfo()
    fid = C.create_string_buffer(64)
    top = C.create_string_buffer(64)
    nn = C.c_int64()
    rc = H.kdh_kdtree3f_decode(raw, len(data), C.byref(info), fid, 64, top, 64, None, 0, C.byref(nn))
    if rc != 0:
        return rc
    cap = nn.value if node_cap is None else node_cap
    nodes = np.zeros(max(cap, 1), dtype=L.NODE_DTYPE)
    rc = H.kdh_kdtree3f_decode(raw, len(data), C.byref(info), fid, 64, top, 64, nodes.ctypes.data, cap, C.byref(nn))
    if rc == 0 and 0 < nn.value <= cap:  # nodes are written only when the capacity suffices
        leaves = np.zeros(nn.value, dtype=np.uint64)
        nl = C.c_int64()
        H.kdh_tree_leaves(nodes.ctypes.data, nn.value, leaves.ctypes.data, nn.value, C.byref(nl))
    return rc


def mutate(rng, data):
    b = bytearray(data)
    k = int(rng.integers(4))
    if k == 0 and b:  # bit flips
        for _ in range(int(rng.integers(1, 8))):
            i = int(rng.integers(len(b)))
            b[i] ^= 1 << int(rng.integers(8))
    elif k == 1:  # truncation
        b = b[: int(rng.integers(len(b) + 1))]
    elif k == 2 and len(b) >= 8:  # a 32-bit field overwritten with an extreme length
        i = int(rng.integers(len(b) - 4)) & ~3
        b[i:i + 4] = int(rng.choice([0, 1, 0x7FFFFFFF, 0xFFFFFFFF, len(b), len(b) + 1])).to_bytes(4, "little")
    else:  # random garbage spliced in
        i = int(rng.integers(len(b) + 1))
        b[i:i] = rng.integers(0, 256, int(rng.integers(1, 64)), dtype=np.uint8).tobytes()
    return bytes(b)


def fuzz_cdr(H, msgs, n=3000):
    rng = np.random.default_rng(11)
    ok = 0
    for m in msgs:
        assert decode(H, m) == 0
        decode(H, m, node_cap=0)  # a node buffer too small: refused, nothing written
    for _ in range(n):
        rc = decode(H, mutate(rng, msgs[int(rng.integers(len(msgs)))]))
        ok += int(rc == 0)
    print(f"CDR corpus: {n} mutated messages, {ok} decoded, the rest refused", flush=True)


def mcap_roundtrip_and_fuzz(H, msgs, n=300):
    d = tempfile.mkdtemp(prefix="kdh_san_")
    path = os.path.join(d, "trees.mcap").encode()
    w = P()
    assert H.kdh_mcap_writer_open(path, C.byref(w)) == 0
    cid = C.c_uint16()
    assert H.kdh_mcap_writer_add_channel(w, b"/trees", b"kd_slam2_msgs/msg/KDTree3f", H.kdh_kdtree3f_schema_text(),
                                         C.byref(cid)) == 0
    for k, m in enumerate(msgs):
        buf = (C.c_uint8 * len(m)).from_buffer_copy(m)
        assert H.kdh_mcap_writer_write(w, cid.value, k, k, buf, len(m)) == 0
    assert H.kdh_mcap_writer_close(w) == 0

    def read_all(p):
        r = P()
        if H.kdh_mcap_reader_open(p, C.byref(r)) != 0:
            return -1
        cnt = 0
        m = L.McapMessage()
        while True:
            rc = H.kdh_mcap_reader_next(r, C.byref(m))
            if rc != 0:
                break
            cnt += 1
            if m.size:
                _ = bytes(C.string_at(m.data, min(int(m.size), 64)))
        H.kdh_mcap_reader_close(r)
        return cnt

    assert read_all(path) == len(msgs)
    data = open(path, "rb").read()
    rng = np.random.default_rng(13)
    mpath = os.path.join(d, "mutated.mcap")
    for _ in range(n):
        with open(mpath, "wb") as fh:
            fh.write(mutate(rng, data))
        read_all(mpath.encode())
    os.remove(mpath)
    os.remove(path)
    os.rmdir(d)
    print(f"MCAP: {len(msgs)} messages round trip, {n} mutated files read", flush=True)


def oracle(H, path):
    """The oracle's sanitized build (pyoracle honours KDH_ORACLE_LIB) on a
    small rigid and CT workload, with trees from the sanitized host builder."""
    import types

    os.environ["KDH_ORACLE_LIB"] = path
    from kd_slam2_amd import synth
    from oracle import pyoracle as O

    w = synth.config1(seed=5, n_map=30_000, n_scan=4_000)
    F = types.SimpleNamespace(nodes=build_tree(H, w["map"], None, leaf_max_points=16))
    M = types.SimpleNamespace(nodes=build_tree(H, w["scan"], w["stamps"], leaf_max_points=16))
    for T in (F, M):
        T.leaves = np.flatnonzero((T.nodes["idx_left"] < 0) & (T.nodes["idx_right"] < 0)).astype(np.uint64)
    p = O.make_params(max_error=0.5, max_iterations=5, num_threads=4)
    O.compute(F, M, w["X0"], w["Xf"], p)
    O.compute(F, M, w["X0"], w["Xf"], p, ct=1, v0=np.zeros(6))
    O.build_form_dual(F, M, w["X0"], np.zeros(6), w["Xf"], np.zeros(6), p)
    O.tree_build(w["scan"], None, leaf_max_points=16)
    O.tree_build(w["scan"], None, leaf_max_points=16, label_mode=1)
    O.compute64(F, M, w["X0"], w["Xf"], p)
    print("oracle: compute, CT compute, dual form, tree build, FP64 loop", flush=True)


def main():
    H = host_lib(sys.argv[1])
    ts = trees(H)
    msgs = [encode(H, t) for t in ts[:6]] + [encode(H, np.zeros(0, dtype=L.NODE_DTYPE), b"", b"")]
    fuzz_cdr(H, msgs)
    mcap_roundtrip_and_fuzz(H, msgs)
    if len(sys.argv) > 2:
        oracle(H, sys.argv[2])
    print("sanitizers: clean", flush=True)


if __name__ == "__main__":
    main()
