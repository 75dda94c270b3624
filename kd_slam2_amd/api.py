"""Python host-side mirror of the reference aligner interface over the C ABI.

Names, parameters and return conventions follow the reference types
``TreeGenerator3D``, ``ICPCUDA3D`` / ``CTICPCUDA3D`` (TreeAlignerBase_,
kd_slam2/src/lib/kd_slam/icp/tree_aligner_base_.h:26-161; ICP_,
icp/icp_.h:53-161; CTICP_, cticp/cticp_.h:20-261), exposed here as
``ICPHIP3D`` / ``CTICPHIP3D``.  Every computation runs in libkdhip.so on the
GPU; this module only marshals arrays.
"""
import ctypes as C
import weakref
import math

import numpy as np

from . import _lib as L

# ---------------------------------------------------------------------------
# Trees


class HostTree:
    """A built PCA KD-tree: Node_ array (NODE_DTYPE), leaf-index list, the
    builder's permuted points and the global start stamp (Tree_ fields
    _nodes_ptr, _leaves_indices_ptr, _points_ptr, _ts_start; tree/tree_.h:27-37)."""

    def __init__(self, nodes, leaves, points=None, ts_start=0.0):
        self.nodes = np.ascontiguousarray(nodes, dtype=L.NODE_DTYPE)
        self.leaves = np.ascontiguousarray(leaves, dtype=np.uint64)
        self.points = points
        self.ts_start = ts_start

    @property
    def num_nodes(self):
        return len(self.nodes)

    @property
    def num_leaves(self):
        return len(self.leaves)


class TreeGenerator3D:
    """TreeGenerator_<NodePoint3f> (tree/tree_generator_.h:11-47)."""

    def __init__(self, **kw):
        p = L.TreeGenParams()
        L.lib().kdh_tree_gen_params_default(C.byref(p))
        self._p = p
        for k, v in kw.items():
            self.set(k, v)

    def set(self, name, value):
        name = name[len("param_"):] if name.startswith("param_") else name
        if not hasattr(self._p, name):
            raise AttributeError(name)
        setattr(self._p, name, value)

    def makeTree(self, points, stamps=None):
        """Returns a HostTree, or None where the reference returns nullptr."""
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        st = None if stamps is None else np.ascontiguousarray(stamps, dtype=np.float64)
        h = C.c_void_p()
        rc = L.lib().kdh_host_tree_build(
            L.f32p(pts), None if st is None else L.f64p(st), len(pts), C.byref(self._p), C.byref(h)
        )
        L.check(rc, what="kdh_host_tree_build")
        if not h.value:
            return None
        try:
            nn, nl, ts = C.c_int64(), C.c_int64(), C.c_double()
            L.check(L.lib().kdh_host_tree_info(h, C.byref(nn), C.byref(nl), C.byref(ts)), what="tree info")
            nodes = np.ctypeslib.as_array(
                C.cast(L.lib().kdh_host_tree_nodes(h), C.POINTER(C.c_uint8)), shape=(nn.value * 64,)
            ).view(L.NODE_DTYPE).copy()
            leaves = np.ctypeslib.as_array(
                C.cast(L.lib().kdh_host_tree_leaves(h), C.POINTER(C.c_uint64)), shape=(nl.value,)
            ).copy()
            pp = np.ctypeslib.as_array(
                C.cast(L.lib().kdh_host_tree_points(h), C.POINTER(C.c_float)), shape=(len(pts) * 3,)
            ).copy().reshape(-1, 3)
        finally:
            L.lib().kdh_host_tree_free(h)
        return HostTree(nodes, leaves, pp, ts.value)

    def makeTreeOnDevice(self, ctx, points, stamps=None, download=True):
        """makeTree on the GPU (kdh_tree_build): a DeviceTree, or None where
        the reference returns nullptr.  The tree is bit-identical to
        makeTree's; with download its full records come back as .host."""
        return DeviceTree.build(ctx, points, stamps, self._p, download)


# ---------------------------------------------------------------------------
# Device context and trees


class Context:
    """One kdh_ctx: a HIP stream on one device plus the aligner workspace."""

    def __init__(self, device=0):
        h = C.c_void_p()
        rc = L.lib().kdh_ctx_create(device, C.byref(h))
        if rc != L.KDH_OK:
            raise RuntimeError(f"kdh_ctx_create({device}) failed: {L.ERRORS.get(rc, rc)}")
        self.h = h
        self.device = device
        self._batches = weakref.WeakSet()  # PairBatch objects prepared on this context

    def close(self):
        if self.h:
            for b in list(self._batches):
                b._release(self)
            L.lib().kdh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, what):
        L.check(rc, self.h, what)

    def set_launch_config(self, grid=0, lds_nodes=256, leaves_per_lane=None, batch_streams=None, tile_leaves=None):
        self.check(L.lib().kdh_set_launch_config(self.h, grid, lds_nodes), "kdh_set_launch_config")
        if tile_leaves is not None and int(tile_leaves) == 256:
            self.check(L.lib().kdh_set_tile_leaves(self.h, 256), "kdh_set_tile_leaves")
        if leaves_per_lane is not None:
            self.check(L.lib().kdh_set_leaves_per_lane(self.h, int(leaves_per_lane)), "kdh_set_leaves_per_lane")
        if tile_leaves is not None and int(tile_leaves) != 256:
            self.check(L.lib().kdh_set_tile_leaves(self.h, int(tile_leaves)), "kdh_set_tile_leaves")
        if batch_streams is not None:
            self.check(L.lib().kdh_set_batch_streams(self.h, int(batch_streams)), "kdh_set_batch_streams")

    def set_profiling(self, on=True):
        self.check(L.lib().kdh_set_profiling(self.h, int(bool(on))), "kdh_set_profiling")

    def kernel_breakdown(self):
        """(tile kernel us, finish kernel us) per iteration of the last
        profiled kdh_align / kdh_batch_align."""
        t, f = C.c_double(), C.c_double()
        self.check(L.lib().kdh_last_kernel_breakdown(self.h, C.byref(t), C.byref(f)), "kdh_last_kernel_breakdown")
        return t.value, f.value

    def last_kernel_time(self):
        us, n = C.c_double(), C.c_int()
        self.check(L.lib().kdh_last_kernel_time(self.h, C.byref(us), C.byref(n)), "kdh_last_kernel_time")
        return us.value, n.value


class DeviceTree:
    """A tree resident in HBM (kdh_tree_upload); freed with close()."""

    def __init__(self, ctx, host_tree):
        self.ctx = ctx
        self.host = host_tree
        h = C.c_void_p()
        nodes = host_tree.nodes
        leaves = host_tree.leaves
        rc = L.lib().kdh_tree_upload(
            ctx.h, nodes.ctypes.data_as(C.c_void_p), len(nodes), leaves.ctypes.data_as(C.c_void_p), len(leaves),
            C.byref(h),
        )
        ctx.check(rc, "kdh_tree_upload")
        self.h = h

    @classmethod
    def build(cls, ctx, points, stamps, gen_params, download=True):
        """Build the tree on the device (kdh_tree_build); None where makeTree
        returns nullptr."""
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        st = None if stamps is None else np.ascontiguousarray(stamps, dtype=np.float64)
        h = C.c_void_p()
        ctx.check(L.lib().kdh_tree_build(ctx.h, L.f32p(pts), None if st is None else L.f64p(st), len(pts),
                                         C.byref(gen_params), 0, C.byref(h)), "kdh_tree_build")
        if not h.value:
            return None
        self = cls.__new__(cls)
        self.ctx = ctx
        self.h = h
        self.host = None
        if download:
            nn, nl = self._counts()
            nodes = np.zeros(nn, dtype=L.NODE_DTYPE)
            leaves = np.zeros(nl, dtype=np.uint64)
            xyz = np.zeros((len(pts), 3), dtype=np.float32)
            ts = np.zeros(len(pts))
            ctx.check(L.lib().kdh_tree_download(ctx.h, h, nodes.ctypes.data_as(C.c_void_p),
                                                leaves.ctypes.data_as(C.c_void_p), L.f32p(xyz), L.f64p(ts)),
                      "kdh_tree_download")
            self.host = HostTree(nodes, leaves, xyz, float(ts.min()) if len(ts) else 0.0)
            self.stamps = ts
        return self

    def _counts(self):
        nn, nl = C.c_int64(), C.c_int64()
        L.lib().kdh_tree_info(self.h, C.byref(nn), C.byref(nl), None)
        return nn.value, nl.value

    @property
    def num_leaves(self):
        return self._counts()[1]

    def device_bytes(self):
        nb = C.c_int64()
        L.lib().kdh_tree_info(self.h, None, None, C.byref(nb))
        return nb.value

    # ---- maintenance on the device (TreeCUDA_, tree/tree_cuda_impl_.h:43-173) ----
    def applyVelocitiesInPlace(self, v, skip_leaves=False):
        """Deskew every node (leaves too unless skip_leaves) by v = [t; w]."""
        v32 = np.ascontiguousarray(v, dtype=np.float32)
        self.ctx.check(L.lib().kdh_tree_apply_velocities(self.ctx.h, self.h, L.f32p(v32), int(skip_leaves)),
                       "kdh_tree_apply_velocities")

    def applyIsometryInPlace(self, X):
        """Move the tree by the isometry X (4x4 or 3x4)."""
        m = np.ascontiguousarray(np.asarray(X, dtype=np.float64)[:3, :4], dtype=np.float32).reshape(12)
        self.ctx.check(L.lib().kdh_tree_apply_isometry(self.ctx.h, self.h, L.f32p(m)), "kdh_tree_apply_isometry")

    def extractCompressedLeaves(self, v=None):
        """(n_leaves, 2, 3) float32: per leaf its mean and direction, warped by
        v when given."""
        out = np.zeros((max(self.num_leaves, 0), 2, 3), dtype=np.float32)
        vp = None if v is None else L.f32p(np.ascontiguousarray(v, dtype=np.float32))
        if len(out):
            self.ctx.check(L.lib().kdh_tree_extract_leaves(self.ctx.h, self.h, vp, L.f32p(out)),
                           "kdh_tree_extract_leaves")
        return out

    def readNodes(self):
        """(n_nodes, 2, 3) float32: every node's mean and direction on the device."""
        out = np.zeros((self._counts()[0], 2, 3), dtype=np.float32)
        if len(out):
            self.ctx.check(L.lib().kdh_tree_read_nodes(self.ctx.h, self.h, L.f32p(out)), "kdh_tree_read_nodes")
        return out

    def clone(self):
        """TreeCUDA_::clone (tree/tree_cuda_impl_.h:176-195): a deep copy on
        the device (the Frame::save backup, map/map_.h:88-94)."""
        h = C.c_void_p()
        self.ctx.check(L.lib().kdh_tree_clone(self.ctx.h, self.h, C.byref(h)), "kdh_tree_clone")
        out = self.__class__.__new__(self.__class__)
        out.ctx = self.ctx
        out.h = h
        out.host = self.host
        if hasattr(self, "stamps"):
            out.stamps = self.stamps
        return out

    def copyNodes(self, other):
        """TreeCUDA_::copyNodes (tree/tree_cuda_impl_.h:60-69): overwrite
        other's nodes with this tree's (Frame::restore, map/map_.h:96-102)."""
        self.ctx.check(L.lib().kdh_tree_copy_nodes(self.ctx.h, self.h, other.h), "kdh_tree_copy_nodes")

    def downloadNodes(self):
        """A device-built tree's full Node_ records as they are now."""
        nn, _ = self._counts()
        nodes = np.zeros(nn, dtype=L.NODE_DTYPE)
        self.ctx.check(L.lib().kdh_tree_download(self.ctx.h, self.h, nodes.ctypes.data_as(C.c_void_p), None, None,
                                                 None), "kdh_tree_download")
        return nodes

    def close(self):
        if self.h and self.ctx.h:
            L.lib().kdh_tree_free(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Aligners


def default_params():
    p = L.Params()
    L.lib().kdh_params_default(C.byref(p))
    return p


def iso34_f32(X):
    return np.ascontiguousarray(np.asarray(X, dtype=np.float64)[:3, :4], dtype=np.float32).reshape(12)


def update_cache_f32(Xm, Xf):
    """updateCache (icp/icp_impl_.h:16-21) in FP32 with the reference's
    evaluation order: Xmf = Xf^-1 * Xm."""
    m = np.asarray(Xm, dtype=np.float64)[:3, :4].astype(np.float32)
    f = np.asarray(Xf, dtype=np.float64)[:3, :4].astype(np.float32)
    Rf, tf = f[:, :3], f[:, 3]
    Rm, tm = m[:, :3], m[:, 3]
    Ri = Rf.T.copy()
    ti = np.zeros(3, dtype=np.float32)
    for r in range(3):
        ti[r] = -(((Ri[r, 0] * tf[0]) + Ri[r, 1] * tf[1]) + Ri[r, 2] * tf[2])
    R = np.zeros((3, 3), dtype=np.float32)
    t = np.zeros(3, dtype=np.float32)
    for r in range(3):
        for c in range(3):
            R[r, c] = ((Ri[r, 0] * Rm[0, c]) + Ri[r, 1] * Rm[1, c]) + Ri[r, 2] * Rm[2, c]
        t[r] = (((Ri[r, 0] * tm[0]) + Ri[r, 1] * tm[1]) + Ri[r, 2] * tm[2]) + ti[r]
    out = np.zeros((3, 4), dtype=np.float32)
    out[:, :3] = R
    out[:, 3] = t
    return m.reshape(12).copy(), f.reshape(12).copy(), out.reshape(12)


class FixedEntry:
    """TreeAlignerBase_::FixedEntry_ (tree_aligner_base_.h:57-74): a fixed
    tree, its pose, and its share (H, b, stats) of the last round."""

    def __init__(self, tree, X):
        self.fixed_tree = tree
        self.fixed_X = np.array(X, dtype=np.float64)
        self.stats = None
        self.H = None
        self.b = None


class _AlignerBase:
    """TreeAlignerBase_ state and parameters shared by ICPHIP3D / CTICPHIP3D."""

    CT = 0
    DIM = 6

    def __init__(self, ctx, **params):
        self.ctx = ctx
        self.params = default_params()
        for k, v in params.items():
            self.set(k, v)
        self._fixed = None
        self._moving = None
        self.X = np.eye(4)
        self.Xf = np.eye(4)
        self.v = np.zeros(6)
        self.prior = None
        self.stats = None
        self.stats_log = []
        self.H = np.zeros((self.DIM, self.DIM))
        self.b = np.zeros(self.DIM)
        self.dx = np.zeros(self.DIM)
        # fixed forest (ICP_::_fixed_forest, key order = iteration order)
        self._fixed_forest = {}

    # ---- parameters (PARAM, tree_aligner_base_.h:39-47) ----
    def set(self, name, value):
        name = name[len("param_"):] if name.startswith("param_") else name
        if not hasattr(self.params, name):
            raise AttributeError(name)
        setattr(self.params, name, value)

    # ---- inputs / state ----
    def setFixed(self, tree):
        self._fixed = tree

    def setMoving(self, tree):
        self._moving = tree

    def setX(self, X):
        self.X = np.array(X, dtype=np.float64)

    def getX(self):
        return self.X.copy()

    def setFixedX(self, X):
        self.Xf = np.array(X, dtype=np.float64)

    def getPoseHessian(self):
        """Pose information of the last round (icp/icp_.h:47; for CT-ICP the
        Schur complement marginalising the velocities, cticp/cticp_.h:169-177)."""
        H = np.asarray(self.H, dtype=np.float64)
        if not self.CT:
            return H.copy()
        Hpv = H[:6, 6:]
        return H[:6, :6] - Hpv @ np.linalg.solve(H[6:, 6:] + 1e-9 * np.eye(6), Hpv.T)

    def saveStartState(self):
        """saveStartState (icp/icp_.h:142-145): the poses at compute() start."""
        self._start = (self.X.copy(), self.Xf.copy(), self.v.copy())

    def iterationLogHeader(self):
        """iterationLogHeader (icp/icp_impl_.h:155-169, cticp/cticp_impl_.h:296-310)."""
        from .synth import logmap

        X0, Xf0, v0 = getattr(self, "_start", (self.X, self.Xf, self.v))
        fmt = lambda a: " ".join(f"{float(x):g}" for x in a)  # noqa: E731
        fv = fmt(np.zeros(6)) if self.CT else "none"
        return (f"{'CTICP' if self.CT else 'ICP'}3D| START  fp: {fmt(logmap(Xf0))} fv: {fv} mp: {fmt(logmap(X0))}"
                f" mv: {fmt(v0) if self.CT else 'none'} | END  fp: {fmt(logmap(self.Xf))} fv: {fv}"
                f" mp: {fmt(logmap(self.X))} mv: {fmt(self.v) if self.CT else 'none'}")

    def makeFixedEntry(self, tree, Xf):
        """FixedEntry of the fixed forest (icp/icp_.h:90-95): a fixed tree at
        its own pose.  Store it as aligner._fixed_forest[ref] = entry."""
        return FixedEntry(tree, Xf)

    def setPosePrior(self, X_ref=None, Z=None, info=None):
        if info is None or not np.any(info):
            self.prior = None
            return
        pr = L.Prior()
        pr.X_ref[:] = list(np.asarray(X_ref if X_ref is not None else np.eye(4))[:3, :4].reshape(12))
        pr.Z[:] = list(np.asarray(Z if Z is not None else np.eye(4))[:3, :4].reshape(12))
        pr.info[:] = list(np.asarray(info, dtype=np.float64).reshape(36))
        self.prior = pr

    def isGPU(self):
        return True

    # ---- the plain per-call path: one kernel build per iteration, host solve
    def buildQuadraticForm(self, stats_mode=False):
        Xm, Xf, Xmf = update_cache_f32(self.X, self.Xf)
        H = np.zeros(self.DIM * self.DIM)
        b = np.zeros(self.DIM)
        st = L.Stats()
        if self.CT:
            v = np.ascontiguousarray(self.v, dtype=np.float32)
            rc = L.lib().kdh_cticp_build(self.ctx.h, self._fixed.h, self._moving.h, L.f32p(Xm), L.f32p(Xf),
                                         L.f32p(Xmf), L.f32p(v), C.byref(self.params), int(stats_mode), L.f64p(H),
                                         L.f64p(b), C.byref(st))
        else:
            rc = L.lib().kdh_icp_build(self.ctx.h, self._fixed.h, self._moving.h, L.f32p(Xm), L.f32p(Xf),
                                       L.f32p(Xmf), C.byref(self.params), int(stats_mode), L.f64p(H), L.f64p(b),
                                       C.byref(st))
        self.ctx.check(rc, "build")
        self.H = H.reshape(self.DIM, self.DIM)
        self.b = b
        self.stats = st.as_dict()
        self.stats["pert_pose_norm"] = 0.0
        self.stats["pert_vel_norm"] = 0.0

    def buildQuadraticFormAt(self, stats_mode=False):
        """buildQuadraticForm at the FP64 state (kdh_build_at): the cache is
        rounded once from (X, Xf, v), as the device loop forms it."""
        X = np.ascontiguousarray(self.X[:3, :4].reshape(12))
        Xf = np.ascontiguousarray(self.Xf[:3, :4].reshape(12))
        v = np.ascontiguousarray(self.v if self.CT else np.zeros(6), dtype=np.float64)
        H = np.zeros(self.DIM * self.DIM)
        b = np.zeros(self.DIM)
        st = L.Stats()
        rc = L.lib().kdh_build_at(self.ctx.h, self._fixed.h, self._moving.h, self.CT, L.f64p(X), L.f64p(Xf),
                                  L.f64p(v), C.byref(self.params), int(stats_mode), L.f64p(H), L.f64p(b),
                                  C.byref(st))
        self.ctx.check(rc, "kdh_build_at")
        self.H = H.reshape(self.DIM, self.DIM)
        self.b = b
        self.stats = st.as_dict()

    def _buildForest(self, stats_mode):
        """buildQuadraticForm(FixedEntry&) per entry, summed in key order
        (icp_impl_.h:55-70)."""
        fixed, Xf = self._fixed, self.Xf
        H = np.zeros((self.DIM, self.DIM))
        b = np.zeros(self.DIM)
        tot = None
        try:
            for ref in sorted(self._fixed_forest):
                e = self._fixed_forest[ref]
                self._fixed, self.Xf = e.fixed_tree, e.fixed_X
                self.buildQuadraticForm(stats_mode)
                e.stats, e.H, e.b = dict(self.stats), self.H.copy(), self.b.copy()
                H += self.H
                b += self.b
                if tot is None:
                    tot = dict(self.stats)
                else:
                    for k in ("num_inliers", "num_outliers", "num_bad", "num_err_normal", "num_err_fn", "chi2",
                              "chi2_kernelized"):
                        tot[k] += self.stats[k]
        finally:
            self._fixed, self.Xf = fixed, Xf
        self.H, self.b, self.stats = H, b, tot

    def oneRound(self, stats_mode=False):
        """ICP_::oneRound / CTICP_::oneRound with an FP64 solve on the host."""
        from .synth import expmap

        if self._fixed_forest:
            self._buildForest(stats_mode)
        else:
            self.buildQuadraticForm(stats_mode)
        H = self.H.copy()
        b = self.b.copy()
        if self.prior is not None:
            Hp, bp = _prior_form(self.prior, self.X)
            H[:6, :6] += Hp
            b[:6] += bp
        good = self.stats["num_outliers"] + self.stats["num_inliers"]
        if good < self.params.min_good_measurements:
            return False
        damp = np.full(self.DIM, float(self.params.damping))
        if self.CT:
            damp[6:] = float(self.params.velocity_damping)
        self.dx = np.linalg.solve(H + np.diag(damp), -b)
        self.stats["pert_pose_norm"] = float(np.linalg.norm(self.dx[:6]))
        if self.CT:
            self.stats["pert_vel_norm"] = float(np.linalg.norm(self.dx[6:]))
        if stats_mode:
            return True
        self.X = expmap(self.dx[:6]) @ self.X
        if self.CT:
            self.v = self.v + self.dx[6:]
        return True

    def compute(self, stats_mode=False):
        """TreeAlignerBase_::compute (tree_aligner_base_.h:100-115), host loop."""
        self.stats_log = []
        self.saveStartState()
        max_iter = 1 if stats_mode else self.params.max_iterations
        prev_score = np.float32(0.0)
        for i in range(max_iter):
            ok = self.oneRound(stats_mode)
            self.stats_log.append(dict(self.stats))
            if not ok:
                return -i
            if self.params.use_term_crit:
                # ICPTerminationCriteriaScore_::hasToStop (icp_stats_.h:85-102):
                # the score in FP32 as ICPStats_::score(), the relative change
                # divided in double (the reference's 1e-6 is a double literal)
                f = np.float32
                s = self.stats
                nl = s["num_inliers"] + s["num_outliers"] + s["num_bad"] + s["num_err_normal"] + s["num_err_fn"]
                ratio = f(s["num_inliers"]) / f(nl) if nl else f(0.0)
                with np.errstate(divide="ignore", invalid="ignore"):
                    score = f(ratio / (f(s["chi2_kernelized"]) / f(s["num_inliers"]) + f(1e-6)))
                if i > 0:
                    d = f(score - prev_score)
                    if f(s["pert_pose_norm"]) < f(self.params.term_d_pert) and d > 0 and \
                            float(d) / (float(prev_score) + 1e-6) < float(f(self.params.term_d_score)):
                        return i
                prev_score = score
        return self.params.max_iterations

    # ---- the device-resident loop ----
    def _computeForestOnDevice(self, stats_mode):
        if stats_mode:
            # one round without update: the per-call form of every entry
            return self.compute(stats_mode=True)
        refs = sorted(self._fixed_forest)
        p = L.Params.from_buffer_copy(self.params)
        pairs = []
        for j, ref in enumerate(refs):
            e = self._fixed_forest[ref]
            pairs.append(dict(fixed=e.fixed_tree, moving=self._moving, ct=self.CT, X0=self.X, v0=self.v,
                              Xf=e.fixed_X, params=p, prior=self.prior if j == 0 else None,
                              forest=len(refs) - 1 if j == 0 else 0))
        out = batch_align(self.ctx, pairs)
        for ref, r in zip(refs, out):
            self._fixed_forest[ref].stats = r["entry"]
        r = out[0]
        self.X = r["X"]
        if self.CT:
            self.v = r["v"].copy()
        self.stats = r["last"]
        self.stats_log = [self.stats]
        return r["result"]

    def computeOnDevice(self, stats_mode=False, cache_log=False):
        """compute() with the whole Gauss-Newton loop on the GPU (kdh_align;
        with a fixed forest, one kdh_batch_align forest)."""
        if self._fixed_forest:
            return self._computeForestOnDevice(stats_mode)
        X0 = np.ascontiguousarray(self.X[:3, :4].reshape(12))
        Xf = np.ascontiguousarray(self.Xf[:3, :4].reshape(12))
        v0 = np.ascontiguousarray(self.v, dtype=np.float64)
        Xo = np.zeros(12)
        vo = np.zeros(6)
        n = max(1, int(self.params.max_iterations))
        log = (L.Stats * n)()
        nlog, res = C.c_int(), C.c_int()
        clog = np.zeros((n, 36), dtype=np.float32) if cache_log else None
        rc = L.lib().kdh_align(self.ctx.h, self._fixed.h, self._moving.h, self.CT, L.f64p(X0), L.f64p(v0),
                               L.f64p(Xf), C.byref(self.params),
                               C.byref(self.prior) if self.prior is not None else None, int(stats_mode),
                               L.f64p(Xo), L.f64p(vo), log, C.byref(nlog), C.byref(res),
                               L.f32p(clog) if clog is not None else None)
        self.ctx.check(rc, "kdh_align")
        self.X = np.eye(4)
        self.X[:3, :4] = Xo.reshape(3, 4)
        if self.CT:
            self.v = vo.copy()
        self.stats_log = [log[i].as_dict() for i in range(nlog.value)]
        self.stats = self.stats_log[-1] if self.stats_log else None
        self.cache_log = clog[: nlog.value] if clog is not None else None
        return res.value

    def correspondences(self, Xm34=None, Xf34=None, Xmf34=None, v=None):
        """Per moving leaf: fixed leaf index (-1 = num_bad) and descent depth at
        the given FP32 cache (default: the cache of the current state)."""
        if Xm34 is None:
            Xm34, Xf34, Xmf34 = update_cache_f32(self.X, self.Xf)
        n = self._moving.num_leaves
        leaf = np.zeros(max(n, 1), dtype=np.int32)
        depth = np.zeros(max(n, 1), dtype=np.int32)
        vv = np.ascontiguousarray(self.v if v is None else v, dtype=np.float32)
        rc = L.lib().kdh_correspondences(self.ctx.h, self._fixed.h, self._moving.h, self.CT,
                                         L.f32p(np.ascontiguousarray(Xm34, dtype=np.float32)),
                                         L.f32p(np.ascontiguousarray(Xf34, dtype=np.float32)),
                                         L.f32p(np.ascontiguousarray(Xmf34, dtype=np.float32)), L.f32p(vv),
                                         C.byref(self.params), L.i32p(leaf), L.i32p(depth))
        self.ctx.check(rc, "kdh_correspondences")
        return leaf[:n], depth[:n]


def _prior_form(prior, X):
    """Pose prior H/b in FP64 (tree_aligner_base_.h:139-152)."""
    from .synth import from34

    Xr = from34(np.array(prior.X_ref[:]))
    Z = from34(np.array(prior.Z[:]))
    info = np.array(prior.info[:]).reshape(6, 6)
    ZXi = np.linalg.inv(Xr @ Z)

    def err(Xc):
        P = ZXi @ Xc
        R = P[:3, :3]
        c = max(-1.0, min(1.0, 0.5 * (np.trace(R) - 1)))
        a = math.acos(c)
        ax = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
        w = ax * 0.5 if a < 1e-7 else ax * (a / (2 * math.sin(a)))
        return np.concatenate([P[:3, 3], w])

    from .synth import expmap

    e = err(X)
    J = np.zeros((6, 6))
    h = 1e-6
    for k in range(6):
        d = np.zeros(6)
        d[k] = h
        J[:, k] = (err(expmap(d) @ X) - err(expmap(-d) @ X)) / (2 * h)
    return J.T @ info @ J, J.T @ info @ e


class ICPHIP3D(_AlignerBase):
    """Rigid point-to-plane ICP on the GPU (drop-in for ICPCPU3D / ICPCUDA3D)."""

    CT = 0
    DIM = 6


class CTICPHIP3D(_AlignerBase):
    """Continuous-time ICP (pose + velocity) on the GPU (drop-in for
    CTICPCPU3D / CTICPCUDA3D)."""

    CT = 1
    DIM = 12

    def setVelocities(self, v):
        self.v = np.array(v, dtype=np.float64)

    def velocities(self):
        return self.v.copy()

    def resetMovingState(self):
        self.v = np.zeros(6)

    def setFixedX(self, X):
        self.Xf = np.array(X, dtype=np.float64)

    def setFixedVelocities(self, v):
        """fixed_state.velocities: the fixed cloud's own motion, used only by
        the dual form."""
        self.vf = np.array(v, dtype=np.float64)

    def buildQuadraticFormDual(self, stats_mode=False):
        """CTICP_::buildQuadraticFormDual (cticp/cticp_impl_.h:42-45): H, b
        and the fixed side's velocity blocks H_dual_vel_B, b_dual_vel_B,
        H_dual_cross_pose_vel_B, H_dual_cross_vel_AB (cticp/cticp_.h:179-186)
        at the FP64 state (X, v) / (Xf, vf)."""
        f = batch_build_dual(self.ctx, [dict(fixed=self._fixed, moving=self._moving, X_moving=self.X, v_moving=self.v,
                                             X_fixed=self.Xf, v_fixed=getattr(self, "vf", np.zeros(6)),
                                             params=self.params)], stats_mode)[0]
        for k in ("H", "b", "H_dual_vel_B", "b_dual_vel_B", "H_dual_cross_pose_vel_B", "H_dual_cross_vel_AB",
                  "stats"):
            setattr(self, k, f[k])


def adapter_align(type_name, ctx, fixed, moving, X0, Xf, params, v0=None, on_device=False, stats_mode=False,
                  log=None):
    """Create `type_name` ("ICPHIP3D" / "CTICPHIP3D") through the C++ plugin
    registry and run compute() (plain per-call path) or computeOnDevice().
    Returns (X 4x4, v, result, iterations, last stats); with a list as `log`
    it also receives the stats_log entries as dicts."""
    A = L.adapter()
    X0a = np.ascontiguousarray(np.asarray(X0, dtype=np.float64)[:3, :4].reshape(12))
    Xfa = np.ascontiguousarray(np.asarray(Xf, dtype=np.float64)[:3, :4].reshape(12))
    v0a = np.ascontiguousarray(np.zeros(6) if v0 is None else v0, dtype=np.float64)
    Xo = np.zeros(12)
    vo = np.zeros(6)
    res, nit = C.c_int(), C.c_int()
    st = L.Stats()
    cap = max(int(params.max_iterations), 1)
    lg = (L.Stats * cap)() if log is not None else None
    rc = A.kdha_align(type_name.encode(), ctx.h, fixed.h, moving.h, L.f64p(X0a), L.f64p(v0a), L.f64p(Xfa),
                      C.byref(params), int(on_device), int(stats_mode), L.f64p(Xo), L.f64p(vo), C.byref(res),
                      C.byref(nit), C.byref(st), lg)
    if rc != L.KDH_OK:
        raise RuntimeError(f"kdha_align({type_name}) failed: {L.ERRORS.get(rc, rc)}")
    if log is not None:
        log.extend(lg[i].as_dict() for i in range(min(nit.value, cap)))
    X = np.eye(4)
    X[:3, :4] = Xo.reshape(3, 4)
    return X, vo, res.value, nit.value, st.as_dict()


def adapter_build_dual(ctx, fixed, moving, X_moving, v_moving, X_fixed, v_fixed, params, stats_mode=False):
    """CTICPHIP3D::buildQuadraticFormDual through the C++ plugin registry;
    returns the kdh_dual_form fields as a dict."""
    A = L.adapter()
    a34 = lambda X: np.ascontiguousarray(np.asarray(X, dtype=np.float64)[:3, :4].reshape(12))  # noqa: E731
    a6 = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(6))  # noqa: E731
    Xm, vm, Xf, vf = a34(X_moving), a6(v_moving), a34(X_fixed), a6(v_fixed)
    out = L.DualForm()
    rc = A.kdha_build_dual(ctx.h, fixed.h, moving.h, L.f64p(Xm), L.f64p(vm), L.f64p(Xf), L.f64p(vf),
                           C.byref(params), int(stats_mode), C.byref(out))
    if rc != L.KDH_OK:
        raise RuntimeError(f"kdha_build_dual failed: {L.ERRORS.get(rc, rc)}")
    return {name: np.array(getattr(out, name)) for name, _ in L.DualForm._fields_ if name != "stats"} | dict(
        stats=out.stats.as_dict())


_PAIR_DT = None
_RES_DT = None
_DUAL_DT = None
_FORM_DT = None
_I4 = np.eye(4)
_Z6 = np.zeros(6)


_STAT_NAMES = None


def _stats_dicts(rec):
    global _STAT_NAMES
    if _STAT_NAMES is None:
        _STAT_NAMES = [f for f, _ in L.Stats._fields_ if f != "pad"]
    names = _STAT_NAMES
    return [dict(zip(names, row)) for row in zip(*[rec[f].tolist() for f in names])]


def _params_column(items, dt):
    """The params structs of items as one array: each distinct struct
    converted once, then gathered."""
    slot, blobs, idx = {}, [], []
    for p in items:
        k = id(p["params"])
        if k not in slot:
            slot[k] = len(blobs)
            blobs.append(bytes(p["params"]))
        idx.append(slot[k])
    return np.frombuffer(b"".join(blobs), dtype=dt)[idx]


def _pair_array(pairs):
    global _PAIR_DT, _RES_DT
    if _PAIR_DT is None:
        _PAIR_DT, _RES_DT = L.np_dtype(L.Pair), L.np_dtype(L.PairResult)
    n = len(pairs)
    arr = (L.Pair * n)()
    A = np.frombuffer(arr, dtype=_PAIR_DT)
    A["fixed"] = [p["fixed"].h.value for p in pairs]
    A["moving"] = [p["moving"].h.value for p in pairs]
    A["ct"] = [int(p.get("ct", 0)) for p in pairs]
    A["forest"] = [int(p.get("forest", 0)) for p in pairs]
    A["X0"] = np.asarray([p["X0"] for p in pairs], dtype=np.float64)[:, :3, :4].reshape(n, 12)
    A["Xf"] = np.asarray([p["Xf"] if "Xf" in p else _I4 for p in pairs], dtype=np.float64)[:, :3, :4].reshape(n, 12)
    A["v0"] = np.asarray([p["v0"] if "v0" in p else _Z6 for p in pairs], dtype=np.float64).reshape(n, 6)
    A["params"] = _params_column(pairs, _PAIR_DT["params"])
    if any(p.get("prior") is not None for p in pairs):
        A["prior"] = [C.addressof(p["prior"]) if p.get("prior") is not None else 0 for p in pairs]
    return arr


def batch_build(ctx, pairs, stats_mode=False):
    """One quadratic form per pair at its X0 / v0 / Xf in one launch
    (kdh_batch_build).  Returns a list of (H, b, stats)."""
    n = len(pairs)
    arr = _pair_array(pairs)
    H = np.zeros((n, 144))
    b = np.zeros((n, 12))
    st = (L.Stats * n)()
    ctx.check(L.lib().kdh_batch_build(ctx.h, n, arr, int(stats_mode), L.f64p(H), L.f64p(b), st), "kdh_batch_build")
    out = []
    for i, p in enumerate(pairs):
        D = 12 if int(p.get("ct", 0)) else 6
        out.append((H[i, :D * D].reshape(D, D).copy(), b[i, :D].copy(), st[i].as_dict()))
    return out


def batch_build_dual(ctx, factors, stats_mode=False):
    """The CT dual forms of many bundle-adjustment factors in one launch
    (kdh_batch_build_dual).  factors: dicts with fixed, moving (DeviceTree),
    X_moving, v_moving, X_fixed, v_fixed, params.  Returns dicts with H
    (12x12), b, H_dual_vel_B (6x6), b_dual_vel_B, H_dual_cross_pose_vel_B
    (6x6, rows: pose), H_dual_cross_vel_AB (6x6, rows: velocities A), stats."""
    global _DUAL_DT, _FORM_DT
    if _DUAL_DT is None:
        _DUAL_DT, _FORM_DT = L.np_dtype(L.DualPair), L.np_dtype(L.DualForm)
    n = len(factors)
    if n == 0:
        return []
    arr = (L.DualPair * n)()
    A = np.frombuffer(arr, dtype=_DUAL_DT)
    A["fixed"] = [f["fixed"].h.value for f in factors]
    A["moving"] = [f["moving"].h.value for f in factors]
    A["X_moving"] = np.asarray([f["X_moving"] for f in factors], dtype=np.float64)[:, :3, :4].reshape(n, 12)
    A["v_moving"] = np.asarray([f.get("v_moving", _Z6) for f in factors], dtype=np.float64).reshape(n, 6)
    A["X_fixed"] = np.asarray([f.get("X_fixed", _I4) for f in factors], dtype=np.float64)[:, :3, :4].reshape(n, 12)
    A["v_fixed"] = np.asarray([f.get("v_fixed", _Z6) for f in factors], dtype=np.float64).reshape(n, 6)
    A["params"] = _params_column(factors, _DUAL_DT["params"])
    out = (L.DualForm * n)()
    ctx.check(L.lib().kdh_batch_build_dual(ctx.h, n, arr, int(stats_mode), out), "kdh_batch_build_dual")
    R = np.frombuffer(out, dtype=_FORM_DT).copy()
    H = R["H"].reshape(n, 12, 12)
    Hv, Hcp, Hca = (R[k].reshape(n, 6, 6) for k in ("H_dual_vel_B", "H_dual_cross_pose_vel_B",
                                                      "H_dual_cross_vel_AB"))
    stats = _stats_dicts(R["stats"])
    return [dict(H=H[i], b=R["b"][i], H_dual_vel_B=Hv[i], b_dual_vel_B=R["b_dual_vel_B"][i],
                 H_dual_cross_pose_vel_B=Hcp[i], H_dual_cross_vel_AB=Hca[i], stats=stats[i]) for i in range(n)]


class BatchResult:
    """Results of one kdh_batch_align call as arrays: X (n, 4, 4), v (n, 6),
    iterations (n,), result (n,) (compute()'s return value), last / entry
    (structured arrays of kdh_stats: the last round's statistics and, for
    fixed-forest entries, the entry's share).  Indexing gives pair i as a
    dict (X, v, iterations, result, last, entry)."""

    def __init__(self, R):
        # one byte copy of the result records (a field-wise structured copy
        # costs ten times more); the arrays are views of it
        self._R = np.frombuffer(R.tobytes(), dtype=R.dtype)
        self._X = None

    @property
    def X(self):
        if self._X is None:
            n = len(self._R)
            X = np.zeros((n, 4, 4))
            X[:, :3, :4] = self._R["X"].reshape(n, 3, 4)
            X[:, 3, 3] = 1.0
            self._X = X
        return self._X

    @property
    def v(self):
        return self._R["v"]

    @property
    def iterations(self):
        return self._R["iterations"]

    @property
    def result(self):
        return self._R["result"]

    @property
    def last(self):
        return self._R["last"]

    @property
    def entry(self):
        return self._R["entry"]

    def __len__(self):
        return len(self._R)

    def __getitem__(self, i):
        return dict(X=self.X[i], v=self.v[i], iterations=int(self.iterations[i]), result=int(self.result[i]),
                    last=_stats_dicts(self.last[i:i + 1])[0], entry=_stats_dicts(self.entry[i:i + 1])[0])

    def __iter__(self):
        last, entry = _stats_dicts(self.last), _stats_dicts(self.entry)
        its, rs = self.iterations.tolist(), self.result.tolist()
        for i in range(len(self.X)):
            yield dict(X=self.X[i], v=self.v[i], iterations=its[i], result=rs[i], last=last[i], entry=entry[i])


class PairBatch:
    """A batch of alignments marshalled once into the kdh_pair array: trees,
    fixed poses, parameters, priors and forests (the `batch_align` pair
    dicts).  The first run() on a context prepares it there
    (kdh_batch_prepare: validated, uploaded, one graph for the whole run);
    set_initial() replaces the initial poses / velocities of every pair, and
    each run() is one kdh_batch_run returning a BatchResult.  The serving
    loop's shape: a resident set of alignments whose initial states change
    from call to call.  Results equal kdh_batch_align's bit for bit.  The
    trees must stay open while the batch runs."""

    def __init__(self, pairs):
        global _PAIR_DT, _RES_DT
        self.n = len(pairs)
        self._arr = _pair_array(pairs)
        self._A = np.frombuffer(self._arr, dtype=_PAIR_DT)
        self._res = (L.PairResult * max(self.n, 1))()
        self._R = np.frombuffer(self._res, dtype=_RES_DT)[:self.n]
        self._keep = list(pairs)  # the trees and priors stay alive with the batch
        self._X0 = np.array(self._A["X0"], dtype=np.float64)  # (n, 12), the run inputs
        self._v0 = np.array(self._A["v0"], dtype=np.float64)  # (n, 6)
        self._prep = {}  # id(ctx) -> (ctx, kdh_batch handle)

    def set_initial(self, X0=None, v0=None):
        """X0: (n, 4, 4) or (n, 3, 4) poses; v0: (n, 6) velocities.  Every
        run passes them to kdh_batch_run; the pair array takes them only when
        a context prepares the batch."""
        if X0 is not None:
            X0 = np.asarray(X0)
            if X0.shape != (self.n, 4, 4) and X0.shape != (self.n, 3, 4):
                raise ValueError(f"X0 must be ({self.n}, 4, 4) or ({self.n}, 3, 4), got {X0.shape}")
            self._X0.reshape(self.n, 3, 4)[:] = X0[:, :3, :4]  # one strided copy, converted to float64
        if v0 is not None:
            self._v0[:] = np.asarray(v0, dtype=np.float64).reshape(self.n, 6)

    def _handle(self, ctx):
        e = self._prep.get(id(ctx))
        if e is None or e[0] is not ctx:
            self._A["X0"] = self._X0
            self._A["v0"] = self._v0
            h = C.c_void_p()
            ctx.check(L.lib().kdh_batch_prepare(ctx.h, self.n, self._arr, C.byref(h)), "kdh_batch_prepare")
            e = (ctx, h)
            self._prep[id(ctx)] = e
            ctx._batches.add(self)
        return e[1]

    def run(self, ctx):
        if self.n == 0:
            return BatchResult(self._R)
        h = self._handle(ctx)
        rc = L.lib().kdh_batch_run(h, self._X0.ctypes.data_as(L.F64P), self._v0.ctypes.data_as(L.F64P), self._res)
        ctx.check(rc, "kdh_batch_run")
        return BatchResult(self._R)

    def submit(self, ctx):
        """Enqueue one run from the current initial states and return at once
        (kdh_batch_submit).  At most two runs are in flight; collect() returns
        their results in submission order, each equal to run()'s bit for bit."""
        if self.n == 0:
            return
        h = self._handle(ctx)
        rc = L.lib().kdh_batch_submit(h, self._X0.ctypes.data_as(L.F64P), self._v0.ctypes.data_as(L.F64P))
        ctx.check(rc, "kdh_batch_submit")

    def collect(self, ctx):
        """Wait for the oldest submitted run and return its BatchResult
        (kdh_batch_collect)."""
        if self.n == 0:
            return BatchResult(self._R)
        ctx.check(L.lib().kdh_batch_collect(self._handle(ctx), self._res), "kdh_batch_collect")
        return BatchResult(self._R)

    def _release(self, ctx):
        e = self._prep.pop(id(ctx), None)
        if e is not None and e[0].h:
            L.lib().kdh_batch_release(e[1])

    def close(self):
        for ctx, _ in list(self._prep.values()):
            self._release(ctx)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_align(ctx, pairs):
    """Run many alignments in one batched launch (kdh_batch_align).

    pairs: list of dicts with keys fixed, moving (DeviceTree), ct, X0, Xf, v0,
    params (L.Params), prior, forest (k > 0: this pair leads a fixed forest
    with the next k pairs).  Returns a list of dicts (X, v, iterations,
    result, last, entry).  The C structs are filled and read through numpy
    views, column by column (PairBatch keeps them between calls)."""
    n = len(pairs)
    if n == 0:
        return []
    arr = _pair_array(pairs)
    res = (L.PairResult * n)()
    ctx.check(L.lib().kdh_batch_align(ctx.h, n, arr, res), "kdh_batch_align")
    return list(BatchResult(np.frombuffer(res, dtype=_RES_DT)[:n]))


# ---------------------------------------------------------------------------
# Serving


class Server:
    """Continuous-batching alignment server (kdh_server_*): an engine thread
    runs steps of Gauss-Newton iterations (four, or one while a request
    waits) over every in-flight request and admits queued requests between
    steps while slots and the HBM budget allow."""

    def __init__(self, device=0, budget_bytes=64 << 30, max_inflight=64):
        h = C.c_void_p()
        rc = L.lib().kdh_server_create(device, int(budget_bytes), int(max_inflight), C.byref(h))
        if rc != L.KDH_OK:
            raise RuntimeError(f"kdh_server_create failed: {L.ERRORS.get(rc, rc)}")
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            L.lib().kdh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != L.KDH_OK:
            m = L.lib().kdh_server_last_error(self.h)
            raise RuntimeError(f"{what} failed: {L.ERRORS.get(rc, rc)}: {m.decode() if m else ''}")

    def add_resident(self, tree):
        self._check(L.lib().kdh_server_add_resident(self.h, tree.h), "kdh_server_add_resident")

    def submit(self, fixed, moving_host, X0, Xf, params, ct=0, v0=None, prior=None, raise_on_error=True):
        """Queue an alignment of host tree `moving_host` against resident `fixed`.
        Returns the request id (or the negative error code if not raise_on_error)."""
        rq = L.Request()
        rq.fixed = fixed.h.value
        rq.moving_nodes = moving_host.nodes.ctypes.data
        rq.n_moving_nodes = len(moving_host.nodes)
        rq.moving_leaves = moving_host.leaves.ctypes.data
        rq.n_moving_leaves = len(moving_host.leaves)
        rq.ct = int(ct)
        rq.X0[:] = list(np.asarray(X0)[:3, :4].reshape(12))
        rq.v0[:] = list(np.asarray(np.zeros(6) if v0 is None else v0, dtype=np.float64))
        rq.Xf[:] = list(np.asarray(Xf)[:3, :4].reshape(12))
        rq.params = params
        rq.prior = C.cast(C.pointer(prior), C.c_void_p).value if prior is not None else None
        rid = C.c_int64()
        rc = L.lib().kdh_server_submit(self.h, C.byref(rq), C.byref(rid))
        if rc != L.KDH_OK:
            if raise_on_error:
                self._check(rc, "kdh_server_submit")
            return rc
        return rid.value

    @staticmethod
    def _result(rr):
        X = np.eye(4)
        X[:3, :4] = np.array(rr.r.X[:]).reshape(3, 4)
        return dict(X=X, v=np.array(rr.r.v[:]), iterations=rr.r.iterations, result=rr.r.result,
                    last=rr.r.last.as_dict(), latency_us=rr.latency_us, queue_us=rr.queue_us, status=rr.status)

    def wait(self, rid, timeout_s=0.0):
        rr = L.RequestResult()
        self._check(L.lib().kdh_server_wait(self.h, int(rid), float(timeout_s), C.byref(rr)), "kdh_server_wait")
        return self._result(rr)

    def poll(self, rid):
        rr = L.RequestResult()
        done = C.c_int()
        self._check(L.lib().kdh_server_poll(self.h, int(rid), C.byref(done), C.byref(rr)), "kdh_server_poll")
        return self._result(rr) if done.value else None

    def cancel(self, rid):
        self._check(L.lib().kdh_server_cancel(self.h, int(rid)), "kdh_server_cancel")

    def info(self):
        i = L.ServerInfo()
        self._check(L.lib().kdh_server_info_get(self.h, C.byref(i)), "kdh_server_info_get")
        return {f: getattr(i, f) for f, _ in i._fields_}
