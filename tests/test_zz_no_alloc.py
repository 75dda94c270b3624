"""kdh_align_noalloc (the device-resident loop without any allocation)
against kdh_align through the C ABI: the same pose, velocities, result code
and statistics log, bit for bit, on two calls in a row -- a CT-ICP sweep
against a 2M-point map, then a smaller rigid pair -- and the refusal of a
moving tree above the scratch maximum.  GPU."""
import ctypes as C

import numpy as np
import pytest

from kd_slam2_amd import DeviceTree, HostTree, _lib as L
from kd_slam2_amd._lib import NODE_DTYPE

from .workloads import cfg1, cfg2, icp_params

pytestmark = pytest.mark.gpu


def _run(fn, ctx, F, M, ct, w, p):
    X0 = np.ascontiguousarray(np.asarray(w["X0"])[:3, :4].reshape(12))
    Xf = np.ascontiguousarray(np.asarray(w["Xf"])[:3, :4].reshape(12))
    v0 = np.ascontiguousarray(w["v0"] if ct else np.zeros(6), dtype=np.float64)
    Xo, vo = np.zeros(12), np.zeros(6)
    log = (L.Stats * p.max_iterations)()
    nlog, res = C.c_int(), C.c_int()
    args = [ctx.h, F.h, M.h, ct, L.f64p(X0), L.f64p(v0), L.f64p(Xf), C.byref(p), None, 0, L.f64p(Xo), L.f64p(vo),
            log, C.byref(nlog), C.byref(res)]
    if fn == "kdh_align":
        args.append(None)
    rc = getattr(L.lib(), fn)(*args)
    ctx.check(rc, fn)
    return Xo, vo, res.value, bytes(log)[: nlog.value * C.sizeof(L.Stats)], nlog.value


def test_noalloc_equals_align_two_calls(ctx):
    w2, w1 = cfg2(), cfg1()
    cases = [(w2, 1, icp_params(max_iterations=15, velocity_damping=10.0)), (w1, 0, icp_params())]
    assert len(w1["moving"].leaves) < len(w2["moving"].leaves)
    for w, ct, p in cases:
        F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, w["moving"])
        a = _run("kdh_align", ctx, F, M, ct, w, p)
        b = _run("kdh_align_noalloc", ctx, F, M, ct, w, p)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert a[2] == b[2] and a[4] == b[4] == p.max_iterations
        assert a[3] == b[3]
        F.close()
        M.close()


def test_noalloc_refuses_trees_above_its_maximum(ctx):
    cap = L.lib().kdh_align_noalloc_max_leaves()
    assert cap == 4096 * 256
    n = cap + 1
    nodes = np.zeros(n, dtype=NODE_DTYPE)
    nodes["idx_left"] = -1
    nodes["idx_right"] = -1
    nodes["direction"] = [0.0, 0.0, 1.0]
    big = HostTree(nodes, np.arange(n, dtype=np.uint64))
    w = cfg1()
    F, M = DeviceTree(ctx, w["fixed"]), DeviceTree(ctx, big)
    p = icp_params()
    Xo, vo = np.zeros(12), np.zeros(6)
    I = np.ascontiguousarray(np.eye(4)[:3, :4].reshape(12))
    rc = L.lib().kdh_align_noalloc(ctx.h, F.h, M.h, 0, L.f64p(I), L.f64p(np.zeros(6)), L.f64p(I), C.byref(p), None,
                                   0, L.f64p(Xo), L.f64p(vo), None, None, None)
    assert rc == -1
    assert b"scratch maximum" in L.lib().kdh_last_error(ctx.h)
    F.close()
    M.close()
