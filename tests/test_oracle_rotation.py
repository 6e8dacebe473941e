"""Pins of the oracle's geometry and contact rows under non-identity
rotations, where the reference's golden file (identity rotations only) says
nothing.

Every check compares the oracle (float32, the restatement of
crates/physics/src/collision/distance.rs, distance_box.rs and
crates/physics/src/avbd/rows.rs) with float64 mathematics computed another
way (tests/geom64.py), on seeded random rotated configurations:
  - box_box_signed_gap (distance_box.rs:54-79) = the largest projected gap
    over the 15 separating axes;
  - separated box-box shape_distance (distance_box.rs:100-147) = the exact
    feature-pair minimum, between the SAT lower bound and a surface-sample
    upper bound, witnesses on their own surfaces and the distance apart;
  - sphere-box (distance.rs:192-216) and box-half-space (:607-641) signed
    distances;
  - cast_shape's hit distance (distance.rs:841-891) inside the interval that
    conservative advancement guarantees: between the first reach of the
    touch distance and of half of it, from float64 bisection of the swept
    distance;
  - world_inertia (rows.rs:194-207) = R diag(I) R^T.
Then the reference's own contact-row unit tests
(crates/physics/src/avbd/contact_tests.rs:5-170) are restated, and an
oracle built with a planted mistake (a box frame taken with the rotation
instead of its conjugate) must fail the distance pins.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes
import geom64 as g64

F = np.float32
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHAPE_TOUCH = 1e-3


def _p(a):
    return a.ctypes.data_as(po._f32p)


def _bind(L):
    L.orc_shape_distance_f32.argtypes = [ctypes.c_int, po._f32p, po._f32p, po._f32p, ctypes.c_int, po._f32p,
                                         po._f32p, po._f32p, po._f32p]
    L.orc_cast_f32.argtypes = [ctypes.c_int, po._f32p, po._f32p, po._f32p, po._f32p, ctypes.c_int, po._f32p,
                               po._f32p, po._f32p, po._f32p]
    L.orc_world_inertia_f32.argtypes = [po._f32p, po._f32p, po._f32p]
    return L


def _lib():
    return _bind(po.lib())


def rand_quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(F)


def oracle_dist(L, ka, spa, pa, qa, kb, spb, pb, qb):
    out = np.zeros(8, F)
    L.orc_shape_distance_f32(ka, _p(f3(spa)), _p(f3(pa)), _p(np.asarray(qa, F)), kb, _p(f3(spb)), _p(f3(pb)),
                             _p(np.asarray(qb, F)), _p(out))
    return out


def f3(v):
    return np.ascontiguousarray(np.asarray(v, F).reshape(3))


def box_pairs(n, seed):
    """Random rotated box pairs, about half separated, half overlapping."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ha = rng.uniform(0.2, 1.0, 3).astype(F)
        hb = rng.uniform(0.2, 1.0, 3).astype(F)
        pa = rng.uniform(-20, 20, 3).astype(F)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        reach = np.linalg.norm(ha) + np.linalg.norm(hb)
        pb = (pa + d * rng.uniform(0.3, 1.1) * reach).astype(F)
        out.append((ha, pa, rand_quat(rng), hb, pb, rand_quat(rng)))
    return out


def distance_failures(L, cases):
    """Count of separated box pairs whose oracle distance or witnesses
    disagree with the float64 geometry."""
    bad = 0
    for ha, pa, qa, hb, pb, qb in cases:
        Ra, Rb = g64.quat_mat(qa), g64.quat_mat(qb)
        ha64, hb64, pa64, pb64 = (x.astype(np.float64) for x in (ha, hb, pa, pb))
        sat = g64.sat_gap(pa64, ha64, Ra, pb64, hb64, Rb)
        if sat <= 1e-3:
            continue
        o = oracle_dist(L, 1, ha, pa, qa, 1, hb, pb, qb)
        d64, _, _ = g64.box_box_distance(pa64, ha64, Ra, pb64, hb64, Rb)
        tol = 3e-5 * (1.0 + np.abs(pa64).max())
        wa, wb = o[1:4].astype(np.float64), o[4:7].astype(np.float64)
        if not (abs(o[0] - d64) <= tol and g64.on_box_surface(wa, pa64, ha64, Ra, tol)
                and g64.on_box_surface(wb, pb64, hb64, Rb, tol) and abs(np.linalg.norm(wa - wb) - o[0]) <= tol):
            bad += 1
    return bad


def test_box_box_sat_gap_rotated():
    L = _lib()
    n_sep = n_over = 0
    for ha, pa, qa, hb, pb, qb in box_pairs(1000, 1):
        o = oracle_dist(L, 1, ha, pa, qa, 1, hb, pb, qb)
        Ra, Rb = g64.quat_mat(qa), g64.quat_mat(qb)
        want = g64.sat_gap(pa.astype(np.float64), ha.astype(np.float64), Ra, pb.astype(np.float64),
                           hb.astype(np.float64), Rb)
        assert abs(o[7] - want) <= 3e-5 * (1.0 + np.abs(pa).max()), (o[7], want)
        n_sep += want > 0
        n_over += want <= 0
    assert n_sep > 200 and n_over > 200


def test_box_box_separated_distance_rotated():
    """Separated pairs: exact distance, bracketed by the SAT bound and the
    sampled surfaces, witnesses on their surfaces."""
    L = _lib()
    cases = box_pairs(1000, 2)
    assert distance_failures(L, cases) == 0
    rng = np.random.default_rng(3)
    for ha, pa, qa, hb, pb, qb in cases[:150]:
        Ra, Rb = g64.quat_mat(qa), g64.quat_mat(qb)
        args = (pa.astype(np.float64), ha.astype(np.float64), Ra, pb.astype(np.float64), hb.astype(np.float64), Rb)
        sat = g64.sat_gap(*args)
        if sat <= 1e-3:
            continue
        o = oracle_dist(L, 1, ha, pa, qa, 1, hb, pb, qb)
        sa = g64.surface_samples(args[0], args[1], Ra)
        sb = g64.surface_samples(args[3], args[4], Rb)
        upper = np.sqrt(((sa[:, None, :] - sb[None, :, :]) ** 2).sum(-1).min())
        tol = 3e-5 * (1.0 + np.abs(pa).max())
        assert sat - tol <= o[0] <= upper + tol
    del rng


def test_box_box_overlapping_reports_sat_gap():
    L = _lib()
    for ha, pa, qa, hb, pb, qb in box_pairs(400, 4):
        Ra, Rb = g64.quat_mat(qa), g64.quat_mat(qb)
        sat = g64.sat_gap(pa.astype(np.float64), ha.astype(np.float64), Ra, pb.astype(np.float64),
                          hb.astype(np.float64), Rb)
        if sat > 0:
            continue
        o = oracle_dist(L, 1, ha, pa, qa, 1, hb, pb, qb)
        assert abs(o[0] - sat) <= 3e-5 * (1.0 + np.abs(pa).max())


def test_sphere_box_and_box_plane_rotated():
    L = _lib()
    rng = np.random.default_rng(5)
    ident = np.array([0, 0, 0, 1], F)
    for _ in range(1000):
        h = rng.uniform(0.2, 1.0, 3).astype(F)
        pb = rng.uniform(-10, 10, 3).astype(F)
        q = rand_quat(rng)
        r = F(rng.uniform(0.1, 0.8))
        c = (pb + rng.normal(size=3) * rng.uniform(0.0, 2.0)).astype(F)
        R = g64.quat_mat(q)
        tol = 3e-5 * (1.0 + np.abs(pb).max())
        want = g64.sphere_box_distance(c.astype(np.float64), float(r), pb.astype(np.float64), h.astype(np.float64),
                                       R)
        o = oracle_dist(L, 0, (r, 0, 0), c, ident, 1, h, pb, q)
        assert abs(o[0] - want) <= tol, (o[0], want)
        o2 = oracle_dist(L, 1, h, pb, q, 0, (r, 0, 0), c, ident)
        assert abs(o2[0] - want) <= tol
        # box against the half-space below a tilted plane through p0
        nrm = rng.normal(size=3)
        nrm = (nrm / np.linalg.norm(nrm)).astype(F)
        p0 = (pb + rng.normal(size=3)).astype(F)
        o3 = oracle_dist(L, 2, nrm, p0, ident, 1, h, pb, q)
        want3 = g64.box_plane_distance(pb.astype(np.float64), h.astype(np.float64), R, p0.astype(np.float64),
                                       nrm.astype(np.float64))
        assert abs(o3[0] - want3) <= tol, (o3[0], want3)
        # the witness on the box is a box point on the plane side
        assert g64.on_box_surface(o3[4:7].astype(np.float64), pb.astype(np.float64), h.astype(np.float64), R, tol)


def test_cast_shape_hit_inside_the_advancement_interval():
    """A reported hit lies between the first reach of SHAPE_TOUCH and of
    SHAPE_TOUCH / 2 along the sweep (float64 bisection), and sweeps that
    really reach a target head-on are reported."""
    L = _lib()
    rng = np.random.default_rng(6)
    ident = np.array([0, 0, 0, 1], F)
    checked = reached = reported = 0
    for i in range(250):
        hb = rng.uniform(0.3, 1.0, 3).astype(F)
        pb = rng.uniform(-5, 5, 3).astype(F)
        qb = rand_quat(rng)
        sphere = i % 4 == 0
        ha = (F(rng.uniform(0.2, 0.5)), F(0), F(0)) if sphere else rng.uniform(0.2, 0.6, 3).astype(F)
        qa = ident if sphere else rand_quat(rng)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pa = (pb - d * rng.uniform(2.5, 4.0)).astype(F)
        aim = (pb + rng.normal(size=3) * 0.4 - pa).astype(np.float64)
        delta = (aim / np.linalg.norm(aim) * rng.uniform(1.0, 5.0)).astype(F)
        out = np.zeros(5, F)
        L.orc_cast_f32(0 if sphere else 1, _p(f3(ha)), _p(f3(pa)), _p(qa), _p(f3(delta)), 1, _p(f3(hb)), _p(f3(pb)),
                       _p(qb), _p(out))
        ln = float(np.linalg.norm(delta.astype(np.float64)))
        dirv = delta.astype(np.float64) / ln

        def dist(t):
            return g64.shape_dist64(0 if sphere else 1, ha, pa.astype(np.float64) + dirv * t, qa, 1, hb, pb, qb)

        if dist(0.0) <= 2 * SHAPE_TOUCH:
            continue
        t1 = g64.first_reach(dist, SHAPE_TOUCH, 0.0, ln)
        t05 = g64.first_reach(dist, 0.5 * SHAPE_TOUCH, 0.0, ln)
        tol = 5e-5 * (1.0 + np.abs(pa).max())
        if out[4] == 1.0:
            checked += 1
            assert t1 is not None and t1 - tol <= out[0], (out[0], t1)
            assert t05 is None or out[0] <= t05 + tol, (out[0], t05)
            assert dist(float(out[0])) <= SHAPE_TOUCH + tol
        if t05 is not None and t05 < 0.98 * ln:
            reached += 1
            reported += out[4] == 1.0
    assert checked > 40
    assert reported >= 0.95 * reached


def test_world_inertia_rotated():
    L = _lib()
    rng = np.random.default_rng(7)
    for _ in range(1000):
        inertia = rng.uniform(0.01, 50.0, 3).astype(F)
        q = rand_quat(rng)
        out = np.zeros(9, F)
        L.orc_world_inertia_f32(_p(inertia), _p(q), _p(out))
        R = g64.quat_mat(q)
        want = R @ np.diag(inertia.astype(np.float64)) @ R.T
        assert np.abs(out.reshape(3, 3) - want).max() <= 2e-6 * inertia.max()


# --- the reference's contact-row unit tests (contact_tests.rs) ------------

def _engine(scene, order=po.ORDER_REFERENCE):
    e = po.OracleEngine(scene, order=order)
    L = e._L
    L.orc_prepare_pairs.argtypes = [ctypes.c_void_p]
    L.orc_gap_c0.restype = ctypes.c_double
    L.orc_gap_c0.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.orc_row_c.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, po._f32p, ctypes.c_float, po._f64p]
    L.orc_set_pose.argtypes = [ctypes.c_void_p, ctypes.c_int, po._f32p, po._f32p]
    L.orc_set_step_start.argtypes = [ctypes.c_void_p, ctypes.c_int, po._f32p, po._f32p]
    L.orc_set_point.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, po._f32p, po._f32p]
    L.orc_set_single_pair.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, po._f32p, po._f32p,
                                      ctypes.c_float, po._f32p, po._f32p]
    return e, L


def _row_c(e, L, p, q, axis, c0):
    out = np.zeros(7)
    L.orc_row_c(e.h, p, q, _p(f3(axis)), ctypes.c_float(c0), out.ctypes.data_as(po._f64p))
    return out[0], out[1:4], out[4:7]


def _floor_and(sb):
    sb.box((0.0, -1.0, 0.0), (10.0, 1.0, 10.0), 0.0)
    return sb


def test_rolling_sphere_keeps_support_without_gaining_speed():
    """contact_tests.rs:5-38 (both insertion orders, 1500 steps)."""
    for reversed_ in (False, True):
        sb = scenes.SceneBuilder()
        if reversed_:
            sb.sphere((0.0, 0.5, 0.0), 0.5, 1.0, vel=(1.0, 0.0, 0.0), angvel=(0.0, 0.0, -2.0))
            sb.box((0.0, -1.0, 0.0), (50.0, 1.0, 50.0), 0.0)
            h = 0
        else:
            sb.box((0.0, -1.0, 0.0), (50.0, 1.0, 50.0), 0.0)
            sb.sphere((0.0, 0.5, 0.0), 0.5, 1.0, vel=(1.0, 0.0, 0.0), angvel=(0.0, 0.0, -2.0))
            h = 1
        e = po.OracleEngine(sb.build(), order=po.ORDER_REFERENCE)
        for step in range(1500):
            e.step(1.0 / 60.0, 1)
            b = e.state()[h]
            assert abs(b[1] - 0.5) < 0.03, (reversed_, step, b[1])
            assert abs(b[7] - 1.0) < 0.02 and abs(b[12] + 2.0) < 0.04, (step, b[7:13])


def test_refreshed_face_contacts_preserve_penetration_and_anchors():
    """contact_tests.rs:40-90: a box buried 0.1 m in a floor keeps C0 <
    -0.08 on every fresh point; the anchors lie on their own faces."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    _floor_and(sb)
    sb.box((-3.0, 0.3, 0.0), (0.4, 0.4, 0.4), 1.0)
    e, L = _engine(sb.build())
    L.orc_prepare_pairs(e.h)
    pts = e.points()
    assert len(pts) >= 3
    for qi in range(len(pts)):
        assert L.orc_gap_c0(e.h, 0, qi) < -0.08
        assert abs(pts[qi, 1] - 1.0) < 1e-6  # floor anchor (body frame)
        assert abs(pts[qi, 4] + 0.4) < 1e-6  # box anchor
        assert abs(pts[qi, 3]) <= 0.401 and abs(pts[qi, 5]) <= 0.401


def test_sphere_contact_normal_is_independent_of_spin():
    """contact_tests.rs:92-125."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    _floor_and(sb)
    sb.sphere((0.0, 0.49, 0.0), 0.5, 1.0)
    e, L = _engine(sb.build())
    L.orc_prepare_pairs(e.h)
    rz = np.array([0.0, 0.0, np.sin(0.15), np.cos(0.15)], F)  # Quat::from_rotation_z(0.3)
    rb = (g64.quat_mat(rz) @ np.array([0.0, -0.5, 0.0])).astype(F)
    pts = e.points()
    L.orc_set_point(e.h, 0, 0, _p(f3(pts[0, 0:3])), _p(f3(rb)))
    n = e.pairs()[3][0]
    c0 = L.orc_gap_c0(e.h, 0, 0)
    before, _, _ = _row_c(e, L, 0, 0, n, c0)
    L.orc_set_pose(e.h, 1, _p(f3((0.0, 0.49, 0.0))), _p(np.array([0.0, 0.0, np.sin(0.05), np.cos(0.05)], F)))
    after, _, r = _row_c(e, L, 0, 0, n, c0)
    assert np.linalg.norm(np.cross(r, n)) < 1e-7
    assert abs(after - before) < 1e-7
    assert abs(c0) < 1e-6


def test_contact_rotation_enters_taylor_residual_once():
    """contact_tests.rs:127-170: a pure rotation delta moves the row value by
    exactly the angular Jacobian at the lever (within 1e-7)."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    sb.box((0.0, 1.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    e, L = _engine(sb.build())
    ident = np.array([0, 0, 0, 1], F)
    L.orc_set_step_start(e.h, 0, _p(f3((0, 0, 0))), _p(ident))
    L.orc_set_step_start(e.h, 1, _p(f3((0, 1, 0))), _p(ident))
    L.orc_set_single_pair(e.h, 0, 1, _p(f3((0, 1, 0))), _p(np.array([0.5, 0.5], F)), ctypes.c_float(0.0),
                          _p(f3((0.3, -0.5, 0.2))), _p(f3((-0.2, 0.5, -0.1))))
    delta = np.array([0.0001, -0.0002, 0.0003], F)
    qd = np.zeros(4, F)
    L.orc_quat_integrate_f32(_p(ident), _p(delta), po.MATH_LIBM, _p(qd))
    for side in range(2):
        pos = f3((0, 0, 0)) if side == 0 else f3((0, 1, 0))
        for axis in np.eye(3, dtype=F):
            L.orc_set_pose(e.h, side, _p(pos), _p(qd))
            c, ra, rb = _row_c(e, L, 0, 0, axis, 0.0)
            want = np.cross(ra, axis) @ delta if side == 0 else -(np.cross(rb, axis) @ delta)
            assert abs(c - want) < 1e-7, (side, axis, c, want)
            L.orc_set_pose(e.h, side, _p(pos), _p(ident))


# --- a planted mistake must fail the pins ---------------------------------

@pytest.fixture(scope="module")
def planted_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("planted")
    so = str(d / "liboavbd_oracle_planted.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-shared", "-DORC_PLANT_CONJ", "-o",
                    so, os.path.join(ROOT, "oracle", "avbd_oracle.cpp")], check=True)
    return _bind(ctypes.CDLL(so))


def test_planted_conjugation_fails_the_distance_pins(planted_lib):
    """The same checks that the real oracle passes reject a build whose
    point_obb_closest uses the rotation instead of its conjugate."""
    cases = box_pairs(300, 2)
    assert distance_failures(_lib(), cases) == 0
    assert distance_failures(planted_lib, cases) > 20
    rng = np.random.default_rng(5)
    ident = np.array([0, 0, 0, 1], F)
    wrong = 0
    for _ in range(200):
        h = rng.uniform(0.2, 1.0, 3).astype(F)
        pb = rng.uniform(-10, 10, 3).astype(F)
        q = rand_quat(rng)
        c = (pb + rng.normal(size=3) * 1.5).astype(F)
        want = g64.sphere_box_distance(c.astype(np.float64), 0.3, pb.astype(np.float64), h.astype(np.float64),
                                       g64.quat_mat(q))
        o = oracle_dist(planted_lib, 0, (0.3, 0, 0), c, ident, 1, h, pb, q)
        wrong += abs(o[0] - want) > 1e-3
    assert wrong > 50
