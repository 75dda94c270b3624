"""Tree builder invariants and the oracle's KD descent against brute force.
CPU only."""
import numpy as np

from kd_slam2_amd import HostTree, TreeGenerator3D
from kd_slam2_amd._lib import NODE_DTYPE
from oracle import pyoracle as O


def _node(mean, direction, left=-1, right=-1, dts=0.0):
    n = np.zeros(1, dtype=NODE_DTYPE)
    n["mean"] = mean
    n["direction"] = direction
    n["idx_left"] = left
    n["idx_right"] = right
    n["dts_mean"] = dts
    return n


def _tiny_axis_tree():
    """Depth-3 tree with axis-aligned split planes and one pruned child.
    BFS order; node 2's right child is pruned (-1)."""
    nodes = np.concatenate([
        _node([0, 0, 0], [1, 0, 0], 1, 2),       # 0: x > 0 -> left
        _node([1, 0, 0], [0, 1, 0], 3, 4),       # 1: y > 0 -> left
        _node([-1, 0, 0], [0, 0, 1], 5, -1),     # 2: z > 0 -> left, right pruned
        _node([1, 1, 0], [0, 0, 1]),             # 3 leaf
        _node([1, -1, 0], [0, 0, 1]),            # 4 leaf
        _node([-1, 0, 1], [1, 0, 0]),            # 5 leaf
    ])
    return nodes


def _brute_region(nodes, q):
    """Leaf whose half-space region contains q, found by checking every
    root-to-leaf path (no descent): returns node index, or -1 when q falls in
    the region of a pruned child."""
    paths = []

    def walk(i, cons):
        n = nodes[i]
        l, r = int(n["idx_left"]), int(n["idx_right"])
        if l < 0 and r < 0:
            paths.append((i, cons))
            return
        for child, side in ((l, True), (r, False)):
            c2 = cons + [(i, side)]
            if child < 0:
                paths.append((-1, c2))
            else:
                walk(child, c2)

    walk(0, [])
    hits = []
    for leaf, cons in paths:
        ok = True
        for i, side in cons:
            n = nodes[i]
            d = np.float32(q - n["mean"]).astype(np.float32)
            s = np.float32(np.float32(n["direction"][0] * d[0] + n["direction"][1] * d[1]) + n["direction"][2] * d[2])
            if (s > 0) != side:
                ok = False
                break
        if ok:
            hits.append(leaf)
    assert len(hits) == 1, hits
    return hits[0]


def test_descent_matches_brute_force_regions():
    fixed = HostTree(_tiny_axis_tree(), np.array([3, 4, 5], dtype=np.uint64))
    rng = np.random.default_rng(0)
    qs = rng.uniform(-2, 2, size=(4000, 3)).astype(np.float32)
    # points exactly on split planes: dot == 0 goes right
    qs[:40, 0] = 0.0
    qs[40:80, 1] = 0.0
    moving_nodes = np.concatenate([_node(q, [0, 0, 1]) for q in qs])
    moving = HostTree(moving_nodes, np.arange(len(qs), dtype=np.uint64))
    I34 = np.eye(4)[:3, :4].reshape(12).astype(np.float32)
    p = O.make_params(max_error=1e9, min_normal_cos=-2.0)
    ref = O.build_form(fixed, moving, I34, I34, I34, p)
    expect = np.array([_brute_region(fixed.nodes, q) for q in qs])
    assert np.array_equal(ref["corr"], expect)
    assert (expect == -1).any() and ref["stats"]["num_bad"] == int((expect == -1).sum())
    # textbook kd descent for axis-aligned planes
    tb = np.where(qs[:, 0] > 0, np.where(qs[:, 1] > 0, 3, 4), np.where(qs[:, 2] > 0, 5, -1))
    assert np.array_equal(ref["corr"], tb)
    assert (ref["depth"] == 2).all()  # root + one internal node on every path


def test_tree_builder_invariants():
    rng = np.random.default_rng(1)
    # two noisy planes and a box corner
    a = np.c_[rng.uniform(-5, 5, 6000), rng.uniform(-5, 5, 6000), 2.0 + 0.01 * rng.normal(size=6000)]
    b = np.c_[4.0 + 0.01 * rng.normal(size=4000), rng.uniform(-5, 5, 4000), rng.uniform(0, 4, 4000)]
    P = np.vstack([a, b]).astype(np.float32)
    stamps = rng.uniform(0, 0.1, len(P))
    g = TreeGenerator3D(leaf_max_points=16)
    t = g.makeTree(P, stamps)
    nodes, leaves = t.nodes, t.leaves.astype(int)
    n = len(nodes)
    L = nodes["idx_left"]
    R = nodes["idx_right"]
    # BFS order: children after parents, sibling pairs adjacent, levels non-decreasing
    for i in range(n):
        for c in (L[i], R[i]):
            assert c == -1 or c > i
        if L[i] >= 0 and R[i] >= 0:
            assert R[i] == L[i] + 1
    level = np.zeros(n, dtype=int)
    for i in range(n):
        for c in (L[i], R[i]):
            if c >= 0:
                level[c] = level[i] + 1
    assert np.all(np.diff(level) >= 0)
    # leaf list == all nodes without children, in order
    assert np.array_equal(leaves, np.where((L < 0) & (R < 0))[0])
    pts = t.points
    for li in leaves:
        nd = nodes[li]
        s, e = int(nd["idx_start"]), int(nd["idx_end"])
        assert 3 <= e - s <= 16
        d = nd["direction"].astype(np.float64)
        assert abs(np.linalg.norm(d) - 1) < 1e-5            # unit normal
        assert abs(nd["mean"] @ d) > 0.1                     # IsGoodLeafPredicate_
        assert nd["mean"] @ d <= 0                          # oriented toward the origin
        q = pts[s:e].astype(np.float64)
        assert np.allclose(q.mean(0), nd["mean"], atol=1e-5)
        C = np.cov(q.T, bias=True)
        w, V = np.linalg.eigh(C)
        # the normal is the smallest-eigenvalue eigenvector of the leaf's points
        assert abs(abs(V[:, 0] @ d) - 1) < 1e-3 or w[1] - w[0] < 1e-9
    # internal nodes split along their largest eigenvector
    for i in range(n):
        if L[i] >= 0 or R[i] >= 0:
            s, e = int(nodes[i]["idx_start"]), int(nodes[i]["idx_end"])
            q = pts[s:e].astype(np.float64)
            w, V = np.linalg.eigh(np.cov(q.T, bias=True))
            assert abs(abs(V[:, 2] @ nodes[i]["direction"]) - 1) < 1e-3 or w[2] - w[1] < 1e-9
    # root stamp moments
    assert abs(nodes[0]["dts_mean"] - (stamps - stamps.min()).mean()) < 1e-9


def test_tree_builder_rigid_equivariance_of_structure():
    """Moving every point by a rigid motion about the origin leaves the
    partition unchanged up to rounding: the same number of points per leaf."""
    rng = np.random.default_rng(2)
    P = np.c_[rng.uniform(-5, 5, 8000), rng.uniform(-5, 5, 8000), 3 + 0.01 * rng.normal(size=8000)]
    ang = 0.3
    Rz = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    g = TreeGenerator3D(leaf_max_points=30)
    t1 = g.makeTree(P.astype(np.float32))
    t2 = g.makeTree((P @ Rz.T).astype(np.float32))
    s1 = np.sort((t1.nodes["idx_end"] - t1.nodes["idx_start"])[t1.leaves.astype(int)])
    s2 = np.sort((t2.nodes["idx_end"] - t2.nodes["idx_start"])[t2.leaves.astype(int)])
    assert abs(len(s1) - len(s2)) <= 0.02 * len(s1)


def test_make_tree_returns_none_like_reference():
    g = TreeGenerator3D()
    assert g.makeTree(np.zeros((3, 3), dtype=np.float32)) is None  # < leaf_min_points
    rng = np.random.default_rng(3)
    assert g.makeTree(rng.normal(size=(12, 3)).astype(np.float32)) is None  # < tree_min_good_leaves


# ---- tree maintenance restated in the oracle (tree/tree_cuda_impl_.h:43-173) ----

def _stamped_tree():
    from kd_slam2_amd import TreeGenerator3D

    rng = np.random.default_rng(11)
    P = np.concatenate([rng.uniform([-10, -10, -0.05], [10, 10, 0.05], size=(4000, 3)),
                        rng.uniform([4, -10, 0], [4.1, 10, 3], size=(2000, 3))]).astype(np.float32)
    stamps = rng.uniform(0.0, 0.1, size=len(P))
    return TreeGenerator3D(leaf_max_points=16).makeTree(P, stamps)


def test_tree_ops_identity_and_translation_are_exact():
    t = _stamped_tree()
    same = O.tree_apply_velocities(t.nodes, np.zeros(6))
    assert np.array_equal(same["mean"], t.nodes["mean"]) and np.array_equal(same["direction"], t.nodes["direction"])
    v = np.array([1.0, -2.0, 0.5, 0.0, 0.0, 0.0], dtype=np.float32)
    moved = O.tree_apply_velocities(t.nodes, v)
    dts = t.nodes["dts_mean"].astype(np.float32)[:, None]
    assert np.array_equal(moved["mean"], t.nodes["mean"] + v[:3][None, :] * dts)
    assert np.array_equal(moved["direction"], t.nodes["direction"])
    skip = O.tree_apply_velocities(t.nodes, v, skip_leaves=True)
    leaf = (t.nodes["idx_left"] < 0) & (t.nodes["idx_right"] < 0)
    assert np.array_equal(skip["mean"][leaf], t.nodes["mean"][leaf])
    assert np.array_equal(skip["mean"][~leaf], moved["mean"][~leaf])


def test_tree_ops_rotation_and_isometry_match_fp64():
    from kd_slam2_amd import synth

    t = _stamped_tree()
    w = np.array([0.0, 0.0, 0.0, 0.3, -0.2, 0.5])
    moved = O.tree_apply_velocities(t.nodes, w)
    for i in range(0, len(t.nodes), 97):
        R = synth.rodrigues(w[3:] * float(np.float32(t.nodes["dts_mean"][i])))
        assert np.allclose(moved["mean"][i], R @ t.nodes["mean"][i].astype(np.float64), atol=1e-5)
    X = synth.expmap(np.array([3.0, -1.0, 0.5, 0.1, 0.2, -0.3]))
    Y = O.tree_apply_isometry(t.nodes, X)
    ref = t.nodes["mean"].astype(np.float64) @ X[:3, :3].T + X[:3, 3]
    assert np.allclose(Y["mean"], ref, atol=1e-5)
    back = O.tree_apply_isometry(Y, np.linalg.inv(X))
    assert np.allclose(back["mean"], t.nodes["mean"], atol=2e-5)
    assert np.allclose(back["direction"], t.nodes["direction"], atol=2e-6)


def test_extract_leaves_is_the_leaf_gather_with_optional_warp():
    t = _stamped_tree()
    plain = O.extract_leaves(t.nodes, t.leaves)
    assert np.array_equal(plain[:, 0], t.nodes["mean"][t.leaves]) and np.array_equal(plain[:, 1], t.nodes["direction"][t.leaves])
    v = np.array([0.4, 0.1, 0.0, 0.0, 0.0, 0.7])
    warped = O.extract_leaves(t.nodes, t.leaves, v)
    moved = O.tree_apply_velocities(t.nodes, v)
    assert np.array_equal(warped[:, 0], moved["mean"][t.leaves])


def _same_tree(a, b):
    """Two trees equal bit for bit: every Node_ field, the leaf list, the
    permuted points."""
    assert (a is None) == (b is None)
    if a is None:
        return
    na, nb = a["nodes"], b["nodes"]
    assert len(na) == len(nb)
    assert na.tobytes() == nb.tobytes()
    assert np.array_equal(a["leaves"], b["leaves"])
    assert np.array_equal(a["points"], b["points"])


def test_oracle_tree_build_equals_host_builder():
    """The oracle's restatement of makeTree and the product's host builder
    (one recursive, one shared with the device build) give the same tree bit
    for bit, threaded or not, on scans, maps and degenerate clouds."""
    from .workloads import tree_clouds

    for name, pts, st, kw in tree_clouds():
        ref = O.tree_build(pts, st, **kw)
        for pl in (0, 2):
            t = TreeGenerator3D(parallel_level=pl, **kw).makeTree(pts, st)
            got = None if t is None else dict(nodes=t.nodes, leaves=t.leaves, points=t.points)
            _same_tree(ref, got)
        if name == "too few points":
            assert ref is None
        else:
            assert ref is not None, name
