"""Host-side aligner surface that needs no GPU: the pose Hessian (Schur
complement for CT-ICP), the iteration log header, fixed-forest entries."""
import numpy as np

from kd_slam2_amd import CTICPHIP3D, ICPHIP3D, synth


class _NoCtx:
    h = None


def test_ct_pose_hessian_is_the_inverse_of_the_pose_block_of_the_inverse():
    """Schur complement H_pp - H_pv H_vv^-1 H_pv^T equals ((H^-1)_pp)^-1."""
    rng = np.random.default_rng(7)
    J = rng.normal(size=(40, 12))
    al = CTICPHIP3D(_NoCtx())
    al.H = J.T @ J
    P = al.getPoseHessian()
    Q = np.linalg.inv(np.linalg.inv(al.H)[:6, :6])
    assert np.allclose(P, Q, rtol=1e-6, atol=1e-8)
    assert np.allclose(P, P.T)
    r = ICPHIP3D(_NoCtx())
    r.H = (J.T @ J)[:6, :6]
    assert np.array_equal(r.getPoseHessian(), r.H)


def test_iteration_log_header_reports_start_and_end_poses():
    al = CTICPHIP3D(_NoCtx())
    al.setX(synth.expmap(np.array([1.0, 2.0, 3.0, 0.0, 0.0, 0.5])))
    al.saveStartState()
    al.setX(np.eye(4))
    h = al.iterationLogHeader()
    assert h.startswith("CTICP3D| START ") and " | END " in h
    start, end = h.split(" | END ")
    assert "mp: 1 2 3 0 0 0.5" in start and "mp: 0 0 0 0 0 0" in end
    assert "mv: none" in ICPHIP3D(_NoCtx()).iterationLogHeader()


def test_fixed_entries_keep_tree_and_pose():
    al = ICPHIP3D(_NoCtx())
    X = synth.expmap(np.array([0.5, 0.0, 0.0, 0.0, 0.1, 0.0]))
    e = al.makeFixedEntry("tree", X)
    al._fixed_forest[3] = e
    assert e.fixed_tree == "tree" and np.array_equal(e.fixed_X, X) and e.stats is None


def test_pair_array_marshals_every_pair_field():
    """kdh_pair structs filled from the Python pair dicts (numpy views of the
    C array): handles, kinds, forests, poses, velocities and each pair's own
    parameter struct, with some parameter objects shared between pairs."""
    import ctypes as C

    from kd_slam2_amd import api, default_params

    class _Tree:
        def __init__(self, h):
            self.h = C.c_void_p(h)

    rng = np.random.default_rng(3)
    ps = [default_params() for _ in range(3)]
    for k, p in enumerate(ps):
        p.max_iterations = 5 + k
        p.damping = 0.5 * (k + 1)
    pairs = []
    for i in range(7):
        X0 = synth.expmap(rng.normal(size=6))
        Xf = synth.expmap(rng.normal(size=6))
        pairs.append(dict(fixed=_Tree(1000 + i), moving=_Tree(2000 + i), ct=i % 2, forest=1 if i == 3 else 0,
                          X0=X0, Xf=Xf, v0=rng.normal(size=6), params=ps[i % 3]))
    arr = api._pair_array(pairs)
    for i, (src, dst) in enumerate(zip(pairs, arr)):
        assert dst.fixed == src["fixed"].h.value and dst.moving == src["moving"].h.value
        assert dst.ct == src["ct"] and dst.forest == src["forest"]
        assert np.array_equal(np.array(dst.X0), src["X0"][:3, :4].reshape(12))
        assert np.array_equal(np.array(dst.Xf), src["Xf"][:3, :4].reshape(12))
        assert np.array_equal(np.array(dst.v0), src["v0"])
        assert bytes(dst.params) == bytes(src["params"]), i
        assert not dst.prior


def test_dual_factor_marshalling_round_trip(monkeypatch):
    """batch_build_dual fills every kdh_dual_pair field from the factor dicts
    and returns every kdh_dual_form block and statistic of each factor: the
    library call is replaced by one that checks the structs and writes a
    distinct value into each output."""
    import ctypes as C

    from kd_slam2_amd import _lib as L
    from kd_slam2_amd import api, default_params

    class _Tree:
        def __init__(self, h):
            self.h = C.c_void_p(h)

    class _Ctx:
        h = C.c_void_p(1)

        def check(self, rc, what):
            assert rc == 0, what

    rng = np.random.default_rng(5)
    ps = [default_params() for _ in range(2)]
    ps[1].max_iterations = 3
    facs = [dict(fixed=_Tree(100 + i), moving=_Tree(200 + i), X_moving=synth.expmap(rng.normal(size=6)),
                 v_moving=rng.normal(size=6), X_fixed=synth.expmap(rng.normal(size=6)), v_fixed=rng.normal(size=6),
                 params=ps[i % 2]) for i in range(5)]

    class _Lib:
        @staticmethod
        def kdh_batch_build_dual(h, n, arr, stats_mode, out):
            pairs = C.cast(arr, C.POINTER(L.DualPair))
            forms = C.cast(out, C.POINTER(L.DualForm))
            for i in range(n):
                d, f = pairs[i], facs[i]
                assert d.fixed == f["fixed"].h.value and d.moving == f["moving"].h.value
                for name in ("X_moving", "X_fixed"):
                    assert np.array_equal(np.array(getattr(d, name)), f[name][:3, :4].reshape(12)), (i, name)
                for name in ("v_moving", "v_fixed"):
                    assert np.array_equal(np.array(getattr(d, name)), f[name]), (i, name)
                assert bytes(d.params) == bytes(f["params"]), i
                o = forms[i]
                for k, (name, ct) in enumerate(L.DualForm._fields_[:-1]):
                    getattr(o, name)[:] = (1000.0 * i + 100.0 * k + np.arange(ct._length_)).tolist()
                o.stats.num_inliers = 7 + i
                o.stats.chi2 = 0.5 * i
            return 0

    monkeypatch.setattr(L, "lib", lambda: _Lib)
    got = api.batch_build_dual(_Ctx(), facs, stats_mode=True)
    shapes = dict(H=(12, 12), b=(12,), H_dual_vel_B=(6, 6), b_dual_vel_B=(6,), H_dual_cross_pose_vel_B=(6, 6),
                  H_dual_cross_vel_AB=(6, 6))
    for i, g in enumerate(got):
        for k, (name, ct) in enumerate(L.DualForm._fields_[:-1]):
            want = (1000.0 * i + 100.0 * k + np.arange(ct._length_)).reshape(shapes[name])
            assert np.array_equal(g[name], want), (i, name)
        assert g["stats"]["num_inliers"] == 7 + i and g["stats"]["chi2"] == 0.5 * i
        assert set(g["stats"]) == {f for f, _ in L.Stats._fields_ if f != "pad"}
    assert api.batch_build_dual(_Ctx(), []) == []


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits with status 2 and says
    why, before touching a device."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
