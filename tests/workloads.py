"""Cached small and full-size workloads shared by the tests."""
import functools

import numpy as np

from kd_slam2_amd import TreeGenerator3D, synth
from kd_slam2_amd._lib import NODE_DTYPE  # noqa: F401


@functools.lru_cache(maxsize=None)
def cfg1(seed=1, n_map=200_000, n_scan=16_384, leaf_max_points=16):
    w = synth.config1(seed=seed, n_map=n_map, n_scan=n_scan)
    g = TreeGenerator3D(leaf_max_points=leaf_max_points)
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"], w["stamps"])
    return w


@functools.lru_cache(maxsize=None)
def cfg2(seed=2, n_map=2_000_000, n_scan=131_072):
    w = synth.config2(seed=seed, n_map=n_map, n_scan=n_scan)
    g = TreeGenerator3D()
    w["fixed"] = g.makeTree(w["map"])
    w["moving"] = g.makeTree(w["scan"], w["stamps"])
    return w


def icp_params(max_error=0.5, max_iterations=10, **kw):
    from kd_slam2_amd import default_params

    p = default_params()
    p.max_error = max_error
    p.max_iterations = max_iterations
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def rel_err(A, B):
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    return float(np.abs(A - B).max() / max(np.abs(B).max(), 1e-30))
