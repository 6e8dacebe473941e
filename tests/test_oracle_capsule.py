"""Pins of the oracle's capsule shape (Shape::Capsule: a core segment along
the body's y axis swept by a radius).

  - the reference's own capsule distance tests
    (crates/physics/src/collision/distance.rs:1140-1276) and capsule inertia
    test (crates/physics/src/collision/shape.rs:2233-2243), restated;
  - float64 mathematics computed another way (tests/geom64.py matrices), on
    seeded random rotated configurations:
      sphere-capsule (distance.rs:218-230): the minimum over the core
        segment of the distance to the sphere's centre, by bounded scalar
        minimisation, less both radii;
      capsule-capsule (distance.rs:232-245): the exact closest points of the
        two cores (quadratic over the unit square), less both radii;
      box-capsule (distance_box.rs:199-265), separated and penetrating: the
        signed distance of two convex bodies is the largest gap between
        their projections over all unit directions, found by dense sampling
        of the sphere and local refinement;
      half-space-capsule (distance.rs:607-641 with gjk.rs:68-79): the lower
        end cap's depth below the plane;
  - an engine-level restatement of
    contained_boxes_and_face_crossing_capsules_are_not_missing_overlaps
    (crates/physics/tests/solver_lifecycle.rs:465-492);
  - an oracle built with a planted mistake (the spine's extent dropped from
    the penetration axes) fails the box-capsule pin.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from scipy.optimize import minimize, minimize_scalar

from oracle import pyoracle as po
from ornis_amd import scenes
import geom64 as g64

F = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SPHERE, BOX, HALFSPACE, CAPSULE = 0, 1, 2, 3
IDENT = (0.0, 0.0, 0.0, 1.0)
DT = 1.0 / 60.0


def _p(a):
    return a.ctypes.data_as(po._f32p)


def _bind(L):
    L.orc_shape_distance_f32.argtypes = [ctypes.c_int, po._f32p, po._f32p, po._f32p, ctypes.c_int, po._f32p,
                                         po._f32p, po._f32p, po._f32p]
    return L


def f3(v):
    return np.ascontiguousarray(np.asarray(v, F).reshape(3))


def dist(L, ka, spa, pa, qa, kb, spb, pb, qb):
    """(dist, point_a, point_b) of the oracle's shape_distance."""
    out = np.zeros(8, F)
    L.orc_shape_distance_f32(ka, _p(f3(spa)), _p(f3(pa)), _p(np.asarray(qa, F)), kb, _p(f3(spb)), _p(f3(pb)),
                             _p(np.asarray(qb, F)), _p(out))
    return float(out[0]), out[1:4].astype(np.float64), out[4:7].astype(np.float64)


def rot_z(angle):
    return (0.0, 0.0, float(np.sin(angle / 2.0)), float(np.cos(angle / 2.0)))


def rand_quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(F)


def core(pos, hh, q):
    """float64 end points of a capsule's core segment (matrix rotation)."""
    a = g64.quat_mat(q) @ np.array([0.0, 1.0, 0.0])
    p = np.asarray(pos, np.float64)
    return p - a * float(hh), p + a * float(hh)


def seg_point_dist(p, s0, s1):
    """Distance from p to the segment by bounded minimisation over its
    parameter (no projection formula)."""
    f = lambda t: np.linalg.norm(s0 + (s1 - s0) * t - p)
    r = minimize_scalar(f, bounds=(0.0, 1.0), method="bounded", options={"xatol": 1e-12})
    return min(r.fun, f(0.0), f(1.0))


def _fib_sphere(n):
    k = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / n)
    th = np.pi * (1.0 + 5.0 ** 0.5) * k
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)


DIRS = _fib_sphere(40000)


def box_capsule_sd(pb, half, qb, pc, r, hh, qc):
    """Signed distance of a box and a capsule: max over unit n of
    |d.n| - sum_i half_i |(R^T n)_i| - hh |a.n| - r (the projections' gap)."""
    R = g64.quat_mat(qb)
    a = g64.quat_mat(qc) @ np.array([0.0, 1.0, 0.0])
    d = np.asarray(pc, np.float64) - np.asarray(pb, np.float64)
    h = np.asarray(half, np.float64)

    def gap(n):
        return abs(d @ n) - np.abs(R.T @ n) @ h - float(hh) * abs(a @ n) - float(r)

    g = np.abs(DIRS @ d) - np.abs(DIRS @ R) @ h - float(hh) * np.abs(DIRS @ a) - float(r)
    best = -np.inf
    for i in np.argsort(g)[-6:]:
        n0 = DIRS[i]
        x0 = np.array([np.arccos(np.clip(n0[2], -1, 1)), np.arctan2(n0[1], n0[0])])
        f = lambda x: -gap(np.array([np.sin(x[0]) * np.cos(x[1]), np.sin(x[0]) * np.sin(x[1]), np.cos(x[0])]))
        res = minimize(f, x0, method="Nelder-Mead", options={"xatol": 1e-11, "fatol": 1e-13, "maxiter": 4000})
        best = max(best, -res.fun, g[i])
    return best


# --- the reference's own cases ----------------------------------------------

def test_reference_capsule_distance_cases():
    """distance.rs:1140-1276 (EPS 1e-5)."""
    L = _bind(po.lib())
    cap, sph, cube = (0.5, 1.0, 0.0), (0.5, 0.0, 0.0), (1.0, 1.0, 1.0)
    o = (0.0, 0.0, 0.0)

    def close(got, want):
        assert np.linalg.norm(np.asarray(got) - np.asarray(want)) < 1e-5, (got, want)

    d, pa, pb = dist(L, SPHERE, sph, (3, 0, 0), IDENT, CAPSULE, cap, o, IDENT)  # sphere_capsule_side
    assert abs(d - 2.0) < 1e-5
    close(pa, (2.5, 0, 0))
    close(pb, (0.5, 0, 0))
    d, pa, pb = dist(L, SPHERE, sph, (0, 3, 0), IDENT, CAPSULE, cap, o, IDENT)  # sphere_capsule_cap
    assert abs(d - 1.0) < 1e-5
    close(pb, (0, 1.5, 0))
    d, _, _ = dist(L, SPHERE, sph, (1, 0, 0), IDENT, CAPSULE, cap, o, IDENT)  # sphere_capsule_touching
    assert abs(d) < 1e-5
    d, pa, pb = dist(L, CAPSULE, cap, o, IDENT, SPHERE, sph, (3, 0, 0), IDENT)  # capsule_sphere_swapped
    assert abs(d - 2.0) < 1e-5
    close(pa, (0.5, 0, 0))
    close(pb, (2.5, 0, 0))
    d, pa, pb = dist(L, CAPSULE, cap, o, IDENT, CAPSULE, cap, (3, 0, 0), IDENT)  # capsule_capsule_parallel
    assert abs(d - 2.0) < 1e-5
    close(pa, (0.5, pa[1], 0))
    close(pb, (2.5, pb[1], 0))
    d, pa, pb = dist(L, CAPSULE, cap, o, IDENT, CAPSULE, cap, (0, 0, 2), rot_z(np.pi / 2))  # perpendicular
    assert abs(d - 1.0) < 1e-5
    close(pa, (0, 0, 0.5))
    close(pb, (0, 0, 1.5))
    d, _, _ = dist(L, CAPSULE, cap, o, IDENT, CAPSULE, cap, (0.5, 0, 0), IDENT)  # overlapping
    assert abs(d + 0.5) < 1e-5
    d, _, _ = dist(L, CAPSULE, (0.5, 0, 0), o, IDENT, CAPSULE, (0.5, 0, 0), (2, 0, 0), IDENT)  # zero half height
    assert abs(d - 1.0) < 1e-5
    d, pa, pb = dist(L, BOX, cube, o, IDENT, CAPSULE, cap, (3, 0, 0), IDENT)  # box_capsule_side
    assert abs(d - 1.5) < 1e-5
    close(pa, (1.0, pa[1], 0))
    close(pb, (2.5, pb[1], 0))
    d, _, _ = dist(L, BOX, cube, o, IDENT, CAPSULE, cap, (0, 4, 0), IDENT)  # box_capsule_above_cap
    assert abs(d - 1.5) < 1e-5
    d, pa, pb = dist(L, CAPSULE, cap, (3, 0, 0), IDENT, BOX, cube, o, IDENT)  # capsule_box_swapped
    assert abs(d - 1.5) < 1e-5
    close(pa, (2.5, pa[1], 0))
    close(pb, (1.0, pb[1], 0))


def test_capsule_inertia_closed_form():
    """capsule_inertia_matches_closed_form_and_is_symmetric_about_axis
    (shape.rs:2233-2243): radius 1, half height 2, mass 3 -> (5.75, 1.5,
    5.75), symmetric about the core axis; the bound radius (hh + r) lets
    two capsules end to end along y at 2(hh + r) + 0.05 apart pair up."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.capsule((0.0, 0.0, 0.0), 1.0, 2.0, 3.0)
    e = po.OracleEngine(sb.build())
    i = e.inertia(0)
    assert np.allclose(i, (5.75, 1.5, 5.75), atol=1e-5) and i[0] == i[2]
    e64 = po.OracleEngine(sb.build(), precision=64)
    assert np.allclose(e64.inertia(0), (5.75, 1.5, 5.75), atol=1e-12)


# --- float64 pins on rotated configurations ------------------------------------

def test_sphere_capsule_rotated():
    L = _bind(po.lib())
    rng = np.random.default_rng(31)
    for _ in range(300):
        r, cr, hh = (float(x) for x in rng.uniform(0.1, 1.0, 3).astype(F))
        pc = rng.uniform(-5, 5, 3).astype(F)
        qc = rand_quat(rng)
        dirn = rng.normal(size=3)
        ps = (pc + dirn / np.linalg.norm(dirn) * rng.uniform(0.0, 1.5) * (hh + r + cr)).astype(F)
        d, pa, pb = dist(L, SPHERE, (r, 0, 0), ps, IDENT, CAPSULE, (cr, hh, 0), pc, qc)
        s0, s1 = core(pc, hh, qc)
        core_d = seg_point_dist(ps.astype(np.float64), s0, s1)
        want = core_d - r - cr
        assert abs(d - want) < 2e-5 * (1 + np.abs(pc).max()), (d, want)
        if core_d > 1e-3:  # witnesses: on each surface, the signed distance apart
            assert abs(np.linalg.norm(pa - ps) - r) < 3e-5
            assert abs(seg_point_dist(pb, s0, s1) - cr) < 3e-5
            assert abs(np.dot(pb - pa, (pb - pa)) ** 0.5 - abs(d)) < 5e-5


def test_capsule_capsule_rotated():
    L = _bind(po.lib())
    rng = np.random.default_rng(32)
    A0, A1, B0, B1, ds, ws = [], [], [], [], [], []
    for _ in range(400):
        ra, ha, rb, hb = (float(x) for x in rng.uniform(0.1, 1.0, 4).astype(F))
        pa_ = rng.uniform(-5, 5, 3).astype(F)
        qa, qb = rand_quat(rng), rand_quat(rng)
        dirn = rng.normal(size=3)
        pb_ = (pa_ + dirn / np.linalg.norm(dirn) * rng.uniform(0.0, 1.2) * (ha + ra + hb + rb)).astype(F)
        d, wa, wb = dist(L, CAPSULE, (ra, ha, 0), pa_, qa, CAPSULE, (rb, hb, 0), pb_, qb)
        a0, a1 = core(pa_, ha, qa)
        b0, b1 = core(pb_, hb, qb)
        A0.append(a0)
        A1.append(a1)
        B0.append(b0)
        B1.append(b1)
        ds.append((d, ra + rb))
        ws.append((wa, wb, ra, rb))
    sq, _, _ = g64.seg_seg_many(*(np.array(x) for x in (A0, A1, B0, B1)))
    for k, (d, rr) in enumerate(ds):
        want = np.sqrt(sq[k]) - rr
        assert abs(d - want) < 3e-5 * 6, (k, d, want)
        if sq[k] > 1e-6:
            wa, wb, ra, rb = ws[k]
            assert abs(seg_point_dist(wa, A0[k], A1[k]) - ra) < 5e-5
            assert abs(seg_point_dist(wb, B0[k], B1[k]) - rb) < 5e-5


def box_capsule_cases(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        half = rng.uniform(0.2, 1.0, 3).astype(F)
        r, hh = (float(x) for x in rng.uniform(0.05, 0.8, 2).astype(F))
        pb = rng.uniform(-5, 5, 3).astype(F)
        dirn = rng.normal(size=3)
        reach = np.linalg.norm(half) + hh + r
        pc = (pb + dirn / np.linalg.norm(dirn) * rng.uniform(0.0, 1.2) * reach).astype(F)
        out.append((half, pb, rand_quat(rng), r, hh, pc, rand_quat(rng)))
    return out


def box_capsule_failures(L, cases):
    bad = 0
    for half, pb, qb, r, hh, pc, qc in cases:
        want = box_capsule_sd(pb, half, qb, pc, r, hh, qc)
        d, wa, wb = dist(L, BOX, half, pb, qb, CAPSULE, (r, hh, 0), pc, qc)
        if abs(d - want) > 4e-5 * (1 + np.abs(pb).max()):
            bad += 1
    return bad


def test_box_capsule_rotated():
    """Separated and penetrating (spine through the box) configurations:
    the distance equals the largest projection gap; separated witnesses
    lie on their surfaces the distance apart; the swapped order reports
    the same distance with swapped witnesses."""
    L = _bind(po.lib())
    cases = box_capsule_cases(300, 33)
    n_spine = n_sep = 0
    for half, pb, qb, r, hh, pc, qc in cases:
        want = box_capsule_sd(pb, half, qb, pc, r, hh, qc)
        d, wa, wb = dist(L, BOX, half, pb, qb, CAPSULE, (r, hh, 0), pc, qc)
        assert abs(d - want) < 4e-5 * (1 + np.abs(pb).max()), (d, want)
        d2, wa2, wb2 = dist(L, CAPSULE, (r, hh, 0), pc, qc, BOX, half, pb, qb)
        assert d2 == d and np.array_equal(wa2, wb) and np.array_equal(wb2, wa)
        R = g64.quat_mat(qb)
        s0, s1 = core(pc, hh, qc)
        core_gap = want + r
        tol = 6e-5 * (1 + np.abs(pb).max())
        # point_a lies on the box's surface in both paths
        loc = R.T @ (wa - pb.astype(np.float64))
        assert np.all(np.abs(loc) <= half + tol) and np.any(np.abs(np.abs(loc) - half) <= tol), loc
        if core_gap > 1e-3:
            n_sep += 1
            assert abs(seg_point_dist(wb, s0, s1) - r) < tol
            assert abs(np.linalg.norm(wb - wa) - abs(d)) < tol
        elif core_gap < -1e-3:
            n_spine += 1
    assert n_spine > 30 and n_sep > 100, (n_spine, n_sep)


def test_halfspace_capsule_rotated():
    L = _bind(po.lib())
    rng = np.random.default_rng(34)
    for _ in range(300):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(F)
        p0 = rng.uniform(-3, 3, 3).astype(F)
        r, hh = (float(x) for x in rng.uniform(0.1, 1.0, 2).astype(F))
        pc = (p0 + rng.uniform(-2, 2, 3)).astype(F)
        qc = rand_quat(rng)
        d, wa, wb = dist(L, HALFSPACE, n, p0, IDENT, CAPSULE, (r, hh, 0), pc, qc)
        s0, s1 = core(pc, hh, qc)
        n64, p64 = n.astype(np.float64), p0.astype(np.float64)
        want = min((s0 - p64) @ n64, (s1 - p64) @ n64) - r
        assert abs(d - want) < 2e-5 * 4, (d, want)
        assert abs((wa - p64) @ n64) < 3e-5  # on the plane
        assert abs(seg_point_dist(wb, s0, s1) - r) < 3e-5  # on the capsule


# --- engine level ------------------------------------------------------------------

@pytest.mark.parametrize("order", [po.ORDER_REFERENCE, po.ORDER_COLOUR])
def test_contained_and_face_crossing_capsules_enter_the_trigger(order):
    """solver_lifecycle.rs:465-492: a static trigger box (half 1) and a body
    at its centre, as a contained box, a contained capsule and a capsule
    whose core crosses two faces: one Entered each."""
    for kind, sp in ((BOX, (0.2, 0.2, 0.2)), (CAPSULE, (0.1, 0.2, 0.0)), (CAPSULE, (0.1, 3.0, 0.0))):
        sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
        t = sb.box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0)
        if kind == BOX:
            sb.box((0.0, 0.0, 0.0), sp, 1.0)
        else:
            sb.capsule((0.0, 0.0, 0.0), sp[0], sp[1], 1.0)
        sc = sb.build()
        sc["trigger"][t] = 1
        e = po.OracleEngine(sc, order=order)
        e.step(DT)
        kinds, ab, *_ = e.drain_events(trigger=True)
        assert kinds.tolist() == [3] and ab.tolist() == [[0, 1]], (kind, sp, kinds)


def test_capsule_rests_on_its_side():
    """A capsule dropped lying along x onto a static box floor comes to
    rest on its side: core height = radius above the floor, axis level,
    at rest (mechanics, not a restatement)."""
    sb = scenes.SceneBuilder()
    sb.box((0.0, -0.5, 0.0), (5.0, 0.5, 5.0), 0.0)
    c = sb.capsule((0.0, 0.6, 0.0), 0.25, 0.6, 2.0, quat=rot_z(np.pi / 2))
    e = po.OracleEngine(sb.build())
    e.step(DT, 240)
    st = e.state()[c]
    axis = g64.quat_mat(st[3:7]) @ np.array([0.0, 1.0, 0.0])
    assert abs(st[1] - 0.25) < 0.02, st[1]
    assert abs(axis[1]) < 0.02, axis
    assert np.linalg.norm(st[7:10]) < 0.05 and np.linalg.norm(st[10:13]) < 0.1


# --- a planted mistake must fail the pins ---------------------------------------

@pytest.fixture(scope="module")
def planted_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("planted_capsule")
    so = str(d / "liboavbd_oracle_planted.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-shared", "-DORC_PLANT_CAPSULE",
                    "-o", so, os.path.join(ROOT, "oracle", "avbd_oracle.cpp")], check=True)
    return _bind(ctypes.CDLL(so))


def test_planted_spine_mistake_fails_the_box_capsule_pin(planted_lib):
    cases = box_capsule_cases(200, 35)
    assert box_capsule_failures(_bind(po.lib()), cases) == 0
    assert box_capsule_failures(planted_lib, cases) > 15
