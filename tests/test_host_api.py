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
