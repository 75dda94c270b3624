"""Seeded synthetic scenes and LiDAR scans for the alignment workloads.

Random numbers come from numpy's PCG64 bit generator (``np.random.Generator(
np.random.PCG64(seed))``); Gaussians from its ``normal``.  Every generator is a
pure function of its seed and sizes.

Scenes are analytic (axis-aligned boxes, a ground plane, vertical cylinders),
so maps are sampled on the surfaces by area and scans are ray-cast exactly.
The spinning-LiDAR model follows the reference's ``Scanner::makeScan``
(kd_slam2/src/lib/kd_slam/d3/scanner.cpp:5-83): points are expressed in the
sensor frame at their capture time, the column stamp is ``t_start + column *
scan_time / columns``, and the sensor moves with constant velocity
``[v_t; w]`` as ``X(t) = X_start * [R(w t) | v_t t]`` -- the same model the
CT aligner's leaf deskew inverts (d3/cticp_impl_.h:76-85).
"""
import math

import numpy as np

# --------------------------------------------------------------------------
# SE(3) helpers (FP64)


def rodrigues(w):
    w = np.asarray(w, dtype=np.float64)
    a = float(np.linalg.norm(w))
    if a < 1e-12:
        return np.eye(3)
    n = w / a
    S = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    return np.eye(3) + math.sin(a) * S + (1 - math.cos(a)) * (S @ S)


def pose(R=None, t=None):
    X = np.eye(4)
    if R is not None:
        X[:3, :3] = R
    if t is not None:
        X[:3, 3] = t
    return X


def expmap(v):
    """expmap of the reference: translation = v[:3], rotation = rodrigues(v[3:])."""
    return pose(rodrigues(v[3:]), v[:3])


def logmap(X):
    """logmap of the reference (geometry_3d_impl_.h:48-115): [t; axis*angle]."""
    R = np.asarray(X)[:3, :3]
    c = min(1.0, max(-1.0, 0.5 * (np.trace(R) - 1.0)))
    a = math.acos(c)
    ax = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    w = 0.5 * ax if a < 1e-7 else ax * (a / (2.0 * math.sin(a)))
    return np.concatenate([np.asarray(X)[:3, 3], w])


def to34(X):
    return np.ascontiguousarray(np.asarray(X)[:3, :4], dtype=np.float64).reshape(12)


def from34(a):
    X = np.eye(4)
    X[:3, :4] = np.asarray(a, dtype=np.float64).reshape(3, 4)
    return X


def pose_error(Xa, Xb):
    """(translation error in m, rotation error in deg) between two poses."""
    D = np.linalg.inv(Xa) @ Xb
    R = D[:3, :3]
    # atan2(|vee(R - R^T)| / 2, (tr R - 1) / 2) is well conditioned at small
    # angles, where acos of the trace term loses half the digits.
    s = 0.5 * math.sqrt((R[2, 1] - R[1, 2]) ** 2 + (R[0, 2] - R[2, 0]) ** 2 + (R[1, 0] - R[0, 1]) ** 2)
    c = (np.trace(R) - 1) / 2
    ang = math.degrees(math.atan2(s, c))
    return float(np.linalg.norm(Xa[:3, 3] - Xb[:3, 3])), ang


def random_unit(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def perturb(X, rng, rot_deg, trans_m, uniform=False):
    """X perturbed in its own (sensor) frame by a rotation of rot_deg about a
    random axis and a translation of trans_m in a random direction (magnitudes
    drawn U(0, max) when uniform): X * D, so the sensor moves by exactly
    trans_m and turns by rot_deg."""
    a = math.radians(rot_deg) * (rng.uniform() if uniform else 1.0)
    t = trans_m * (rng.uniform() if uniform else 1.0)
    D = pose(rodrigues(random_unit(rng) * a), random_unit(rng) * t)
    return X @ D


# --------------------------------------------------------------------------
# Geometry primitives


class Scene:
    """Boxes seen from outside, an optional room seen from inside, an optional
    ground plane z = 0 and vertical cylinders."""

    def __init__(self):
        self.boxes = np.zeros((0, 2, 3))       # (lo, hi)
        self.room = None                        # (lo, hi) seen from inside
        self.ground = None                      # half-extent of the ground square
        self.cyl = np.zeros((0, 4))             # cx, cy, r, h

    # ---- surface sampling ------------------------------------------------
    def _faces(self):
        faces = []  # (origin, u, v) rectangles
        def box_faces(lo, hi, with_bottom):
            x0, y0, z0 = lo
            x1, y1, z1 = hi
            out = [
                ((x0, y0, z1), (x1 - x0, 0, 0), (0, y1 - y0, 0)),  # top
                ((x0, y0, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0)),  # y = y0
                ((x0, y1, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0)),  # y = y1
                ((x0, y0, z0), (0, y1 - y0, 0), (0, 0, z1 - z0)),  # x = x0
                ((x1, y0, z0), (0, y1 - y0, 0), (0, 0, z1 - z0)),  # x = x1
            ]
            if with_bottom:
                out.append(((x0, y0, z0), (x1 - x0, 0, 0), (0, y1 - y0, 0)))
            return out
        if self.room is not None:
            faces += box_faces(self.room[0], self.room[1], True)
        for lo, hi in self.boxes:
            faces += box_faces(lo, hi, False)
        if self.ground is not None:
            g = self.ground
            faces.append(((-g, -g, 0.0), (2 * g, 0, 0), (0, 2 * g, 0)))
        return [tuple(np.asarray(f, dtype=np.float64) for f in face) for face in faces]

    def sample(self, n, rng, sigma):
        faces = self._faces()
        areas = [np.linalg.norm(np.cross(u, v)) for _, u, v in faces]
        cyl_areas = [2 * math.pi * r * h for _, _, r, h in self.cyl]
        areas = np.array(areas + cyl_areas)
        counts = rng.multinomial(n, areas / areas.sum())
        out = []
        nf = len(faces)
        for i, c in enumerate(counts):
            if c == 0:
                continue
            if i < nf:
                o, u, v = faces[i]
                a = rng.uniform(size=(c, 1))
                b = rng.uniform(size=(c, 1))
                out.append(o + a * u + b * v)
            else:
                cx, cy, r, h = self.cyl[i - nf]
                th = rng.uniform(0, 2 * math.pi, size=c)
                z = rng.uniform(0, h, size=c)
                out.append(np.stack([cx + r * np.cos(th), cy + r * np.sin(th), z], axis=1))
        P = np.concatenate(out, axis=0)
        P = P[rng.permutation(len(P))]
        P += rng.normal(scale=sigma, size=P.shape)
        return P

    # ---- ray casting -----------------------------------------------------
    def cast(self, o, d, max_range):
        """Nearest hit distance along unit rays d (N,3) from origins o (N,3) or (3,).
        Returns t (N,), inf where nothing is hit within max_range."""
        d = np.asarray(d, dtype=np.float64)
        o = np.broadcast_to(np.asarray(o, dtype=np.float64), d.shape)
        N = len(d)
        best = np.full(N, np.inf)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / d
            for lo, hi in self.boxes:
                t1 = (lo - o) * inv
                t2 = (hi - o) * inv
                tmin = np.nanmax(np.minimum(t1, t2), axis=1)
                tmax = np.nanmin(np.maximum(t1, t2), axis=1)
                hit = (tmax >= tmin) & (tmin > 1e-6)
                best = np.where(hit & (tmin < best), tmin, best)
            if self.room is not None:
                lo, hi = self.room
                t1 = (lo - o) * inv
                t2 = (hi - o) * inv
                tmax = np.nanmin(np.maximum(t1, t2), axis=1)
                hit = tmax > 1e-6
                best = np.where(hit & (tmax < best), tmax, best)
            if self.ground is not None:
                t = -o[:, 2] / d[:, 2]
                x = o[:, 0] + t * d[:, 0]
                y = o[:, 1] + t * d[:, 1]
                g = self.ground
                hit = (d[:, 2] < 0) & (t > 1e-6) & (np.abs(x) <= g) & (np.abs(y) <= g)
                best = np.where(hit & (t < best), t, best)
            if len(self.cyl):
                dxy = d[:, :2]
                a = (dxy * dxy).sum(1)
                for cx, cy, r, h in self.cyl:
                    oc = o[:, :2] - np.array([cx, cy])
                    bq = 2 * (oc * dxy).sum(1)
                    cq = (oc * oc).sum(1) - r * r
                    disc = bq * bq - 4 * a * cq
                    ok = (disc >= 0) & (a > 1e-12)
                    sq = np.sqrt(np.where(ok, disc, 0))
                    t = (-bq - sq) / (2 * np.where(a > 1e-12, a, 1))
                    z = o[:, 2] + t * d[:, 2]
                    hit = ok & (t > 1e-6) & (z >= 0) & (z <= h)
                    best = np.where(hit & (t < best), t, best)
        best[best > max_range] = np.inf
        return best


# --------------------------------------------------------------------------
# Scenes


def room_scene(rng, size=(20.0, 15.0, 4.0), n_boxes=20):
    """Room of 6 walls (size metres) with n_boxes axis-aligned boxes of sides
    U(0.5, 3) m standing on the floor."""
    S = Scene()
    lo = np.zeros(3)
    hi = np.array(size, dtype=np.float64)
    S.room = (lo, hi)
    boxes = []
    for _ in range(n_boxes):
        sides = rng.uniform(0.5, 3.0, size=3)
        sides[2] = min(sides[2], hi[2] - 0.5)
        cx = rng.uniform(1.0 + sides[0] / 2, hi[0] - 1.0 - sides[0] / 2)
        cy = rng.uniform(1.0 + sides[1] / 2, hi[1] - 1.0 - sides[1] / 2)
        blo = np.array([cx - sides[0] / 2, cy - sides[1] / 2, 0.0])
        boxes.append((blo, blo + sides))
    S.boxes = np.array(boxes)
    return S


def outdoor_scene(rng, extent=100.0, n_buildings=50, n_poles=200, road_half_width=8.0):
    """Ground square of half-side `extent`, box buildings with sides and
    heights U(5, 20) m off a road along x, and vertical poles of radius
    U(0.1, 0.3) m and height U(3, 8) m."""
    S = Scene()
    S.ground = extent
    boxes = []
    while len(boxes) < n_buildings:
        sx, sy, sz = rng.uniform(5.0, 20.0, size=3)
        cx = rng.uniform(-extent + sx / 2, extent - sx / 2)
        side = 1 if rng.uniform() < 0.5 else -1
        cy = side * rng.uniform(road_half_width + 2 + sy / 2, 0.8 * extent)
        blo = np.array([cx - sx / 2, cy - sy / 2, 0.0])
        boxes.append((blo, blo + np.array([sx, sy, sz])))
    S.boxes = np.array(boxes)
    cyl = []
    for i in range(n_poles):
        r = rng.uniform(0.1, 0.3)
        h = rng.uniform(3.0, 8.0)
        cx = rng.uniform(-extent * 0.9, extent * 0.9)
        if i % 2 == 0:
            cy = (road_half_width + rng.uniform(0.3, 1.5)) * (1 if rng.uniform() < 0.5 else -1)
        else:
            cy = rng.uniform(-extent * 0.9, extent * 0.9)
        cyl.append((cx, cy, r, h))
    S.cyl = np.array(cyl)
    return S


def free_space_position(S, rng, lo, hi, z, margin=0.7):
    for _ in range(1000):
        p = np.array([rng.uniform(lo[0], hi[0]), rng.uniform(lo[1], hi[1]), z])
        inside = False
        for blo, bhi in S.boxes:
            if np.all(p >= blo - margin) and np.all(p <= bhi + margin):
                inside = True
                break
        if not inside:
            for cx, cy, r, h in S.cyl:
                if (p[0] - cx) ** 2 + (p[1] - cy) ** 2 < (r + margin) ** 2:
                    inside = True
                    break
        if not inside:
            return p
    raise RuntimeError("no free sensor position")


# --------------------------------------------------------------------------
# LiDAR


def lidar_scan(S, rng, X_start, n_beams, vfov_deg, azimuth_step_deg=None, n_columns=None, range_sigma=0.02,
               max_range=200.0, min_range=0.3, velocity=None, scan_time=0.1, n_points=None, t_start=0.0):
    """Spinning LiDAR: n_beams rows over the vertical FOV, columns over 360
    degrees.  Returns (points in the sensor frame at capture time (N,3),
    stamps (N,)).  With n_points, hits are subsampled to exactly n_points, or
    topped up with extra rays at random azimuths when the grid yields fewer."""
    if n_columns is None:
        n_columns = int(round(360.0 / azimuth_step_deg))
    v = np.zeros(6) if velocity is None else np.asarray(velocity, dtype=np.float64)
    elev = np.radians(np.linspace(vfov_deg[0], vfov_deg[1], n_beams))

    def shoot(col_pos):
        # col_pos: fractional column index in [0, n_columns); every ray of a
        # column shares the sensor pose at that column's time.
        t = scan_time * col_pos / n_columns
        az = -math.pi + 2 * math.pi * col_pos / n_columns
        ce, se = np.cos(elev), np.sin(elev)
        nc, nb = len(col_pos), len(elev)
        W = np.empty((nc, 4, 4))
        for k in range(nc):
            W[k] = X_start @ pose(rodrigues(v[3:] * t[k]), v[:3] * t[k])
        dirs_s = np.stack([ce[None, :] * np.cos(az)[:, None], ce[None, :] * np.sin(az)[:, None],
                           np.broadcast_to(se[None, :], (nc, nb))], axis=2).reshape(-1, 3)
        Rw = np.repeat(W[:, :3, :3], nb, axis=0)
        dirs_w = np.einsum("nij,nj->ni", Rw, dirs_s)
        orig = np.repeat(W[:, :3, 3], nb, axis=0)
        r = S.cast(orig, dirs_w, max_range)
        ok = np.isfinite(r) & (r > min_range)
        rr = r[ok] + rng.normal(scale=range_sigma, size=int(ok.sum()))
        stamps = np.repeat(t_start + t, nb)[ok]
        return dirs_s[ok] * rr[:, None], stamps

    P, T = shoot(np.arange(n_columns, dtype=np.float64))
    if n_points is not None:
        while len(P) < n_points:
            extra = rng.uniform(0, n_columns, size=max(64, (n_points - len(P)) // max(1, n_beams // 2) + 8))
            P2, T2 = shoot(np.sort(extra))
            P = np.concatenate([P, P2])
            T = np.concatenate([T, T2])
        idx = np.sort(rng.choice(len(P), size=n_points, replace=False))
        P, T = P[idx], T[idx]
    return P.astype(np.float32), T


# --------------------------------------------------------------------------
# Workloads


def anchor_frame(rng, position, offset=0.5):
    """Frame the map tree is expressed in: world-aligned axes at `position`
    plus a random horizontal offset.  The reference orients every tree normal
    toward its frame origin (label_policy_base_impl_.h:88-95) and aligns
    against keyframe trees in their own sensor frames, so a map must be
    anchored near the viewpoint for its normals to face the scanner."""
    d = random_unit(rng)
    d[2] = 0.0
    d /= max(np.linalg.norm(d), 1e-9)
    return pose(None, np.asarray(position, dtype=np.float64) + offset * d)


def to_frame(points, X):
    """World points expressed in frame X (FP64 math, FP32 result)."""
    P = np.asarray(points, dtype=np.float64)
    return ((P - X[:3, 3]) @ X[:3, :3]).astype(np.float32)


def config1(seed=1, n_map=200_000, n_scan=16_384, rot_deg=2.0, trans_m=0.3, uniform_init=False):
    """Rigid ICP: room map (n_map points, sigma 0.01 m) and one 32-beam scan
    (-15..+15 deg, 0.35 deg azimuth step, range sigma 0.02 m) subsampled to
    n_scan points; initial pose perturbed by rot_deg about a random axis and
    trans_m in a random direction.  The map is returned in the anchor frame
    Xf; poses (X_gt, X0, Xf) are world poses."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    S = room_scene(rng)
    Mw = S.sample(n_map, rng, 0.01)
    p = free_space_position(S, rng, (3.0, 3.0), (17.0, 12.0), 1.5)
    X_gt = pose(rodrigues([0, 0, rng.uniform(-math.pi, math.pi)]), p)
    P, T = lidar_scan(S, rng, X_gt, 32, (-15.0, 15.0), azimuth_step_deg=0.35, range_sigma=0.02,
                      max_range=100.0, n_points=n_scan)
    X0 = perturb(X_gt, rng, rot_deg, trans_m, uniform=uniform_init)
    Xf = anchor_frame(rng, p)
    return dict(map=to_frame(Mw, Xf), scan=P, stamps=T, X_gt=X_gt, X0=X0, Xf=Xf, v_gt=np.zeros(6),
                v0=np.zeros(6))


def config2(seed=2, n_map=2_000_000, n_scan=131_072, extent=100.0):
    """CT-ICP with deskew: outdoor map (n_map points, sigma 0.02 m) and one
    128-beam sweep (+-22.5 deg, 1024 columns, 0.1 s) of n_scan points taken
    while moving at 10 m/s forward and 30 deg/s yaw."""
    rng = np.random.default_rng(np.random.PCG64(seed))
    S = outdoor_scene(rng, extent=extent)
    Mw = S.sample(n_map, rng, 0.02)
    p = free_space_position(S, rng, (-20.0, -3.0), (20.0, 3.0), 1.8)
    X_gt = pose(rodrigues([0, 0, rng.uniform(-0.2, 0.2)]), p)
    v_gt = np.array([10.0, 0.0, 0.0, 0.0, 0.0, math.radians(30.0)])
    P, T = lidar_scan(S, rng, X_gt, 128, (-22.5, 22.5), n_columns=1024, range_sigma=0.02, max_range=200.0,
                      velocity=v_gt, scan_time=0.1, n_points=n_scan)
    X0 = perturb(X_gt, rng, 1.0, 0.3)
    v0 = v_gt * np.array([0.9, 1, 1, 1, 1, 0.9]) + np.array([0.0, 0.3, 0.0, 0.0, 0.0, 0.0])
    Xf = anchor_frame(rng, p, offset=2.0)
    return dict(map=to_frame(Mw, Xf), scan=P, stamps=T, X_gt=X_gt, X0=X0, Xf=Xf, v_gt=v_gt, v0=v0)


def config3_sizes(n_pairs=256, seed=3):
    rng = np.random.default_rng(np.random.PCG64(seed))
    return np.exp(rng.uniform(math.log(4000), math.log(120000), size=n_pairs)).astype(int)


def config3_pair(i, n_scan):
    """Loop-closure pair i: config-1 generator with seed 3000+i, map 3x the
    scan, initial error U(0, 5 deg) / U(0, 1 m)."""
    return config1(seed=3000 + i, n_map=3 * int(n_scan), n_scan=int(n_scan), rot_deg=5.0, trans_m=1.0,
                   uniform_init=True)


def config5_sequence(s, n_map=4_000_000, n_scan=65_536, n_frames=4, extent=120.0):
    """Odometry sequence s of the serving mix: a resident n_map-point outdoor
    map (seed 5000+s) and n_frames CT sweeps of n_scan points taken 0.1 s
    apart while driving at 10 m/s with a 30 deg/s yaw rate."""
    rng = np.random.default_rng(np.random.PCG64(5000 + s))
    S = outdoor_scene(rng, extent=extent)
    Mw = S.sample(n_map, rng, 0.02)
    p0 = free_space_position(S, rng, (-30.0, -3.0), (30.0, 3.0), 1.8)
    X = pose(rodrigues([0, 0, rng.uniform(-0.2, 0.2)]), p0)
    v_gt = np.array([10.0, 0.0, 0.0, 0.0, 0.0, math.radians(30.0)])
    Xf = anchor_frame(rng, p0, offset=2.0)
    frames = []
    for f in range(n_frames):
        P, T = lidar_scan(S, rng, X, 128, (-22.5, 22.5), n_columns=512, range_sigma=0.02, max_range=200.0,
                          velocity=v_gt, scan_time=0.1, n_points=n_scan)
        X0 = perturb(X, rng, 1.0, 0.3)
        v0 = v_gt * np.array([0.9, 1, 1, 1, 1, 0.9]) + np.array([0.0, 0.3, 0.0, 0.0, 0.0, 0.0])
        frames.append(dict(scan=P, stamps=T, X_gt=X.copy(), X0=X0, v0=v0, v_gt=v_gt))
        # the next sweep starts where this one ended
        X = X @ pose(rodrigues(v_gt[3:] * 0.1), v_gt[:3] * 0.1)
    return dict(map=to_frame(Mw, Xf), Xf=Xf, frames=frames)


def config5_loop_pair(j):
    """Loop-closure request j of the serving mix: config-1 generator with seed
    5100+j, scan size log-uniform in [8k, 200k], map 3x the scan."""
    rng = np.random.default_rng(np.random.PCG64(5100 + j))
    n = int(math.exp(rng.uniform(math.log(8000), math.log(200000))))
    return config1(seed=5100 + j, n_map=3 * n, n_scan=n, rot_deg=5.0, trans_m=1.0, uniform_init=True)


def poisson_arrivals(rate_hz, duration_s, rng):
    """Arrival times of a Poisson process on [0, duration_s)."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t >= duration_s:
            return out
        out.append(t)


# --------------------------------------------------------------------------
# Config 4: precision stress (large scene far from the origin)

CONFIG4_OFFSET = np.array([10_000.0, 0.0, 0.0])


def config4_scene(seed=4, extent=220.0):
    rng = np.random.default_rng(np.random.PCG64(seed))
    S = outdoor_scene(rng, extent=extent, n_buildings=120, n_poles=300)
    p = free_space_position(S, rng, (-20.0, -3.0), (20.0, 3.0), 1.8)
    X_local = pose(rodrigues([0, 0, rng.uniform(-0.2, 0.2)]), p)
    Xf_local = anchor_frame(rng, p, offset=2.0)
    X0_local = perturb(X_local, rng, 0.5, 0.2)
    return S, X_local, Xf_local, X0_local


def config4_map_block(block, n_blocks=8, n_map=8_000_000, seed=4):
    """Block of the 8M-point map (sigma 0.02 m), in the anchor frame."""
    S, _, Xf_local, _ = config4_scene(seed)
    rng = np.random.default_rng(np.random.PCG64(40_000 + 100 * seed + block))
    return to_frame(S.sample(n_map // n_blocks, rng, 0.02), Xf_local)


def config4_scan_block(block, n_blocks=8, n_columns=22000, seed=4):
    """Columns [block/n_blocks of a 128-beam +-22.5 deg sweep) of the static
    sensor, ranges up to 200 m, range sigma 0.02 m; sensor-frame points."""
    S, X_local, _, _ = config4_scene(seed)
    rng = np.random.default_rng(np.random.PCG64(41_000 + 100 * seed + block))
    c0 = block * n_columns // n_blocks
    c1 = (block + 1) * n_columns // n_blocks
    elev = np.radians(np.linspace(-22.5, 22.5, 128))
    az = -math.pi + 2 * math.pi * (np.arange(c0, c1) + 0.5) / n_columns
    ce, se = np.cos(elev), np.sin(elev)
    dirs = np.stack([ce[None, :] * np.cos(az)[:, None], ce[None, :] * np.sin(az)[:, None],
                     np.broadcast_to(se[None, :], (len(az), len(elev)))], axis=2).reshape(-1, 3)
    r = S.cast(X_local[:3, 3], dirs @ X_local[:3, :3].T, 200.0)
    ok = np.isfinite(r) & (r > 0.3)
    rr = r[ok] + rng.normal(scale=0.02, size=int(ok.sum()))
    return (dirs[ok] * rr[:, None]).astype(np.float32)


def config4_poses(seed=4):
    """World poses with the 10 km offset: ground truth, initial guess, map anchor."""
    _, X_local, Xf_local, X0_local = config4_scene(seed)
    O = pose(None, CONFIG4_OFFSET)
    return O @ X_local, O @ X0_local, O @ Xf_local
