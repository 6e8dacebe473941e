"""Float64 geometry used to pin the oracle's rotated-shape queries.

Independent of the oracle's own formulas: rotations are applied as 3x3
matrices built from the normalised quaternion (not the quaternion sandwich
the oracle uses), a point's closest point on a box is the clamp of its
box-frame coordinates, and a segment pair's closest points come from the
exact minimisation of the 2-variable quadratic over the unit square
(interior stationary point, else the best of the four edges).  The distance
between two separated convex boxes is attained between a vertex of one and
the other box, or between two edges, so the minimum over those feature
pairs is the exact distance; `sat_gap` (the largest projected gap over the
15 separating axes) is a lower bound on it and the closest of densely
sampled surface points an upper bound.
"""
import numpy as np

EDGES = [(0, 4), (1, 5), (2, 6), (3, 7), (0, 2), (1, 3), (4, 6), (5, 7), (0, 1), (2, 3), (4, 5), (6, 7)]


def quat_mat(q):
    """3x3 rotation of the normalised quaternion q = (x, y, z, w)."""
    x, y, z, w = np.asarray(q, np.float64) / np.linalg.norm(np.asarray(q, np.float64))
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def corners(pos, half, R):
    """Corner k: bit 2 -> -x, bit 1 -> -y, bit 0 -> -z (0 = +)."""
    out = []
    for k in range(8):
        s = np.array([-1.0 if (k >> 2) & 1 else 1.0, -1.0 if (k >> 1) & 1 else 1.0, -1.0 if k & 1 else 1.0])
        out.append(np.asarray(pos, np.float64) + R @ (s * half))
    return np.array(out)


def point_box(p, pos, half, R):
    """Closest point of the solid box to p and whether p is inside."""
    loc = R.T @ (np.asarray(p, np.float64) - pos)
    cl = np.clip(loc, -half, half)
    return pos + R @ cl, bool(np.all(np.abs(loc) <= half))


def seg_seg_many(a0, a1, b0, b1):
    """Closest points of segment pairs (M x 3 arrays of end points), exact:
    the interior stationary point of |a0 + s d1 - b0 - t d2|^2 when it lies
    in the unit square, else the best of the four edge minimisers; returns
    squared distances and the two closest points."""
    d1, d2, r = a1 - a0, b1 - b0, a0 - b0
    a = (d1 * d1).sum(-1)
    e = (d2 * d2).sum(-1)
    b = (d1 * d2).sum(-1)
    c = (d1 * r).sum(-1)
    f = (d2 * r).sum(-1)
    den = a * e - b * b
    with np.errstate(divide="ignore", invalid="ignore"):
        si = np.where(den > 1e-300, (b * f - c * e) / den, -1.0)
        ti = np.where(den > 1e-300, (a * f - b * c) / den, -1.0)
        ss = [si]
        ts = [ti]
        for s0 in (0.0, 1.0):
            ss.append(np.full_like(a, s0))
            ts.append(np.where(e > 0, np.clip((b * s0 + f) / e, 0.0, 1.0), 0.0))
        for t0 in (0.0, 1.0):
            ts.append(np.full_like(a, t0))
            ss.append(np.where(a > 0, np.clip((b * t0 - c) / a, 0.0, 1.0), 0.0))
    S, Tt = np.stack(ss, 1), np.stack(ts, 1)
    X = a0[:, None, :] + S[..., None] * d1[:, None, :]
    Y = b0[:, None, :] + Tt[..., None] * d2[:, None, :]
    D2 = ((X - Y) ** 2).sum(-1)
    ok = (S >= 0) & (S <= 1) & (Tt >= 0) & (Tt <= 1)
    D2 = np.where(ok, D2, np.inf)
    k = D2.argmin(1)
    idx = np.arange(len(a))
    return D2[idx, k], X[idx, k], Y[idx, k]


def _points_box(P, pos, half, R):
    loc = (P - pos) @ R
    return pos + np.clip(loc, -half, half) @ R.T


_EI = np.array(EDGES)


def box_box_distance(pa, ha, Ra, pb, hb, Rb):
    """Exact distance and witnesses of two separated boxes."""
    ca, cb = corners(pa, ha, Ra), corners(pb, hb, Rb)
    xs, ys = [ca, _points_box(cb, pa, ha, Ra)], [_points_box(ca, pb, hb, Rb), cb]
    i, j = np.meshgrid(np.arange(12), np.arange(12), indexing="ij")
    i, j = i.ravel(), j.ravel()
    _, x, y = seg_seg_many(ca[_EI[i, 0]], ca[_EI[i, 1]], cb[_EI[j, 0]], cb[_EI[j, 1]])
    xs.append(x)
    ys.append(y)
    X, Y = np.concatenate(xs), np.concatenate(ys)
    d2 = ((X - Y) ** 2).sum(-1)
    k = int(d2.argmin())
    return np.sqrt(d2[k]), X[k], Y[k]


def sat_gap(pa, ha, Ra, pb, hb, Rb):
    """Largest projected gap over the 15 separating axes (negative: the
    smallest overlap); axes from near-parallel edge pairs are skipped."""
    axes = [Ra[:, i] for i in range(3)] + [Rb[:, i] for i in range(3)]
    for i in range(3):
        for j in range(3):
            c = np.cross(Ra[:, i], Rb[:, j])
            if c @ c > 1e-12:
                axes.append(c / np.linalg.norm(c))
    d = np.asarray(pb, np.float64) - pa
    best = -np.inf
    for n in axes:
        ea = np.sum(ha * np.abs(n @ Ra))
        eb = np.sum(hb * np.abs(n @ Rb))
        best = max(best, abs(d @ n) - ea - eb)
    return best


def surface_samples(pos, half, R, k=9):
    """Points on the six faces of a box, k x k per face."""
    u = np.linspace(-1.0, 1.0, k)
    uu, vv = np.meshgrid(u, u)
    pts = []
    for ax in range(3):
        o = [a for a in range(3) if a != ax]
        for s in (-1.0, 1.0):
            loc = np.zeros((k * k, 3))
            loc[:, ax] = s * half[ax]
            loc[:, o[0]] = uu.ravel() * half[o[0]]
            loc[:, o[1]] = vv.ravel() * half[o[1]]
            pts.append(loc)
    loc = np.concatenate(pts)
    return pos + loc @ R.T


def on_box_surface(p, pos, half, R, tol):
    loc = R.T @ (np.asarray(p, np.float64) - pos)
    return bool(np.all(np.abs(loc) <= half + tol) and np.min(half - np.abs(loc)) <= tol)


def sphere_box_distance(c, r, pos, half, R):
    """Signed distance of a sphere to a solid box (inside: minus the
    distance to the nearest face)."""
    loc = R.T @ (np.asarray(c, np.float64) - pos)
    if np.all(np.abs(loc) <= half):
        return -np.min(half - np.abs(loc)) - r
    cl = np.clip(loc, -half, half)
    return np.linalg.norm(loc - cl) - r


def box_plane_distance(pos, half, R, p0, n):
    """Signed distance of a box to the half-space below the plane (p0, n)."""
    return float(np.min((corners(pos, half, R) - p0) @ n))


def shape_dist64(kind_a, sp_a, pa, qa, kind_b, sp_b, pb, qb):
    """Signed distance for the box / sphere pairs (box-box overlap: the SAT
    gap, which is what the query reports for overlapping boxes)."""
    Ra, Rb = quat_mat(qa), quat_mat(qb)
    pa, pb = np.asarray(pa, np.float64), np.asarray(pb, np.float64)
    sp_a, sp_b = np.asarray(sp_a, np.float64), np.asarray(sp_b, np.float64)
    if kind_a == 1 and kind_b == 1:
        g = sat_gap(pa, sp_a, Ra, pb, sp_b, Rb)
        return g if g <= 0 else box_box_distance(pa, sp_a, Ra, pb, sp_b, Rb)[0]
    if kind_a == 0 and kind_b == 1:
        return sphere_box_distance(pa, sp_a[0], pb, sp_b, Rb)
    if kind_a == 1 and kind_b == 0:
        return sphere_box_distance(pb, sp_b[0], pa, sp_a, Ra)
    return np.linalg.norm(pb - pa) - sp_a[0] - sp_b[0]


def first_reach(fn, level, lo, hi, iters=40):
    """Smallest t in [lo, hi] with fn(t) <= level for a function convex on
    the interval (distance of a translating convex shape), or None."""
    a, b = lo, hi
    for _ in range(iters):  # golden-section search for the minimum
        m1, m2 = a + (b - a) * 0.381966, a + (b - a) * 0.618034
        if fn(m1) <= fn(m2):
            b = m2
        else:
            a = m1
    tmin = 0.5 * (a + b)
    if fn(lo) <= level:
        return lo
    if fn(tmin) > level:
        return None
    a, b = lo, tmin
    for _ in range(iters):
        m = 0.5 * (a + b)
        if fn(m) <= level:
            b = m
        else:
            a = m
    return b
