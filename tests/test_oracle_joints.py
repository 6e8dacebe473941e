"""The oracle's Revolute, Prismatic and Fixed joints (assembly.rs:279-327,
:589-648, :787-808; joints_ext.rs:543-822; sleep.rs:30-66) against what the
reference's own tests pin:

- unit pins of the joint rows (crates/physics/src/avbd/joint_tests.rs:
  253-349): a zero violation keeps the warm reaction (linear and angular);
  the Fixed angular correction has the closed form 0.3 I / (I + K) about
  each body axis under any world rotation; a warm torque that balances the
  inertial target leaves the body still, for small and large inertia;
- the hinge rows' correction is invariant under a world rotation (the two
  rows span the plane across the hinge axis, whatever tangent basis the
  step-start axis picks);
- behaviour (crates/physics/tests/avbd_engine.rs): a hinged arm keeps its
  axis and anchor (:139-183), a welded pair survives a drop (:891-935), and
  a slider moves freely along its axis while holding its line (the
  slider test of :598-646 without its limit, which this build does not
  model)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes

DT = 1.0 / 60.0
E = np.eye(3)


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def conj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def axis_angle(axis, ang):
    axis = np.asarray(axis, np.float64)
    return np.concatenate([axis * math.sin(ang / 2), [math.cos(ang / 2)]])


def rotvec(q):
    """Rotation vector of a unit quaternion (shortest arc)."""
    q = np.asarray(q, np.float64)
    if q[3] < 0:
        q = -q
    v = q[:3]
    s = np.linalg.norm(v)
    return np.zeros(3) if s == 0 else v / s * 2.0 * math.atan2(s, q[3])


def row_scene(kind, radius=0.25, mass=1.0, axis=(0.0, 1.0, 0.0)):
    """joint_tests.rs:4-28: a static sphere and a dynamic one at the origin,
    joined at their centres, no gravity, step-start and inertial poses at
    rest; `axis` is the hinge / slide axis in both bodies."""
    b = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    b.sphere((0.0, 0.0, 0.0), 0.25, 0.0)
    b.sphere((0.0, 0.0, 0.0), radius, mass)
    s = b.build()
    e = po.OracleEngine(s, order=po.ORDER_REFERENCE)
    e.add_joint(0, 1, (0, 0, 0), (0, 0, 0), kind=kind, axis_a=axis, axis_b=axis)
    for i in range(2):
        e.set_step_start(i, (0, 0, 0), (0, 0, 0, 1))
        e.set_inertial(i, (0, 0, 0), (0, 0, 0, 1))
    return e


def duals(lam_l=(0, 0, 0), pen_l=(1, 1, 1), lam_a=(0, 0, 0), pen_a=(1, 1, 1)):
    return np.array([*lam_l, *pen_l, *lam_a, *pen_a], np.float32)


def test_linear_zero_violation_keeps_warm_reaction():
    e = row_scene(scenes.JOINT_BALL)
    e.set_joint_duals(0, duals(lam_l=(10, 0, 0), pen_l=(100, 1, 1)))
    e.dual_update()
    assert e.joints(1)[0, 0] == 10.0 and e.joints(1)[0, 3] == 100.0
    e.solve_body(1)
    expected = 10.0 / (1.0 / (DT * DT) + 100.0)
    assert abs(e.state()[1, 0] - expected) < 1e-8


def test_angular_zero_violation_keeps_warm_reaction():
    e = row_scene(scenes.JOINT_FIXED)
    e.set_joint_duals(0, duals(lam_a=(10, 0, 0), pen_a=(100, 1, 1)))
    e.dual_update()
    assert e.joints(1)[0, 6] == 10.0 and e.joints(1)[0, 9] == 100.0
    e.solve_body(1)
    assert e.state()[1, 3] < -1e-4


@pytest.mark.parametrize("pitch", [0.0, 0.7, math.pi * 0.5, math.pi])
def test_fixed_angular_primal_is_invariant_under_world_rotation(pitch):
    for i in range(3):
        e = row_scene(scenes.JOINT_FIXED)
        qa = axis_angle((0, 1, 0), pitch)
        qb = qmul(qa, axis_angle(E[i], 0.3))
        e.set_pose(0, (0, 0, 0), qa)
        e.set_pose(1, (0, 0, 0), qb)
        for k in range(2):
            e.set_step_start(k, (0, 0, 0), qa)
        e.set_inertial(1, (0, 0, 0), qb)
        pen_a = [1.0, 1.0, 1.0]
        pen_a[i] = 1000.0
        e.set_joint_duals(0, duals(pen_a=pen_a))
        e.solve_body(1)
        err = rotvec(qmul(conj(qa), e.state()[1, 3:7]))
        inertia = e.inertia(1)[i] / (DT * DT)
        expected = 0.3 * inertia / (inertia + 1000.0)
        assert np.linalg.norm(err - E[i] * expected) < 2e-6, (pitch, i, err, expected)


def test_small_inertia_keeps_balanced_warm_torque():
    for mass in (0.01, 1.0, 100.0):
        for radius in (0.01, 0.1, 1.0):
            for i in range(3):
                e = row_scene(scenes.JOINT_FIXED, radius=radius, mass=mass)
                inertia_dt2 = e.inertia(1)[i] / (DT * DT)
                torque = 1e-4
                lam_a = [0.0, 0.0, 0.0]
                lam_a[i] = -torque
                pen_a = [1.0, 1.0, 1.0]
                pen_a[i] = 1e4
                e.set_joint_duals(0, duals(lam_a=lam_a, pen_a=pen_a))
                target = po.quat_integrate(np.array([0, 0, 0, 1], np.float32), -E[i] * (torque / inertia_dt2))
                e.set_inertial(1, (0, 0, 0), target)
                e.solve_body(1)
                assert np.linalg.norm(rotvec(e.state()[1, 3:7])) < 1e-7, (mass, radius, i)


@pytest.mark.parametrize("kind", [scenes.JOINT_REVOLUTE, scenes.JOINT_PRISMATIC])
def test_hinge_rows_are_invariant_under_world_rotation(kind):
    """Body B tilted 0.3 rad across the hinge axis (Y in both frames): the
    correction, seen in A's frame, is the same for every world rotation,
    reduces the tilt and leaves the spin about the hinge alone."""
    results = []
    for qw in (axis_angle((0, 0, 1), 0.0), axis_angle((0, 0, 1), 0.9), axis_angle((1, 1, 0) / np.sqrt(2), 2.1),
               axis_angle((0, 1, 0), math.pi)):
        e = row_scene(kind)
        qb = qmul(qw, axis_angle((1, 0, 0), 0.3))
        e.set_pose(0, (0, 0, 0), qw)
        e.set_pose(1, (0, 0, 0), qb)
        for k in range(2):
            e.set_step_start(k, (0, 0, 0), qw)
        e.set_inertial(1, (0, 0, 0), qb)
        e.set_joint_duals(0, duals(pen_a=(1000, 1000, 1)))
        e.solve_body(1)
        results.append(rotvec(qmul(conj(qw), e.state()[1, 3:7])))
    ref = results[0]
    for r in results[1:]:
        assert np.linalg.norm(r - ref) < 2e-6, results
    assert 0.0 < ref[0] < 0.3 and abs(ref[1]) < 1e-6 and abs(ref[2]) < 1e-6


def test_revolute_hinge_keeps_axis():
    b = scenes.SceneBuilder()
    a = b.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    arm = b.box((0.3, -1.0, 0.0), (0.2, 1.0, 0.2), 1.0)
    lb = (-0.3, 1.0, 0.0)
    b.joint(scenes.JOINT_REVOLUTE, a, arm, (0, 0, 0), lb, (0, 0, 1), (0, 0, 1))
    for order in (po.ORDER_REFERENCE, po.ORDER_COLOUR):
        e = po.OracleEngine(b.build(), order=order)
        e.step(DT, 300)
        st = e.state()
        axa = qrot(st[a, 3:7], (0, 0, 1))
        axb = qrot(st[arm, 3:7], (0, 0, 1))
        assert np.dot(axa, axb) > 0.99
        anchor = st[arm, 0:3] + qrot(st[arm, 3:7], lb)
        assert np.linalg.norm(anchor) < 0.15
        # the arm swung: the hinge carried a real motion
        assert abs(st[arm, 0]) < 0.3


def test_fixed_weld_holds():
    b = scenes.SceneBuilder()
    b.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0)
    a = b.box((-0.5, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    c = b.box((0.5, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    b.joint(scenes.JOINT_FIXED, a, c, (0.5, 0, 0), (-0.5, 0, 0))
    for order in (po.ORDER_REFERENCE, po.ORDER_COLOUR):
        e = po.OracleEngine(b.build(), order=order)
        e.step(DT, 300)
        st = e.state()
        assert abs(np.linalg.norm(st[c, :3] - st[a, :3]) - 1.0) < 0.15
        assert np.dot(qrot(st[a, 3:7], (1, 0, 0)), qrot(st[c, 3:7], (1, 0, 0))) > 0.98
        assert st[a, 1] > 0.2 and st[c, 1] > 0.2


def test_prismatic_slides_freely_on_its_line():
    b = scenes.SceneBuilder()
    anchor = b.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bob = b.sphere((0.0, -2.9, 0.0), 0.25, 1.0)
    b.joint(scenes.JOINT_PRISMATIC, anchor, bob, (0, 0, 0), (0, 1, 0), (0, 1, 0), (0, 1, 0))
    for order in (po.ORDER_REFERENCE, po.ORDER_COLOUR):
        e = po.OracleEngine(b.build(), order=order)
        for _ in range(60):
            e.step(DT, 1)
            p = e.state()[bob]
            assert abs(p[0]) < 0.08 and abs(p[2]) < 0.08
        # implicit Euler free fall along the slide: y_n = y_0 - g dt^2 n (n + 1) / 2
        assert abs(e.state()[bob, 1] - (-2.9 - 9.81 * DT * DT * 60 * 61 / 2)) < 0.05


def test_fixed_joint_bodies_keep_contact_but_hinges_do_not():
    """The no-collide rule (rows.rs:437-450) skips every jointed pair except
    Fixed ones: a welded pair in contact keeps its pair, a hinged one not."""
    for kind, want in ((scenes.JOINT_FIXED, True), (scenes.JOINT_REVOLUTE, False)):
        b = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
        a = b.box((-0.5, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
        c = b.box((0.49, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
        b.joint(kind, a, c, (0.5, 0, 0), (-0.49, 0, 0), (0, 0, 1), (0, 0, 1))
        e = po.OracleEngine(b.build(), order=po.ORDER_COLOUR)
        e.step_inner(1)
        assert (len(e.pairs()[0]) > 0) == want


def qrot(q, v):
    q = np.asarray(q, np.float64)
    u = q[:3]
    v = np.asarray(v, np.float64)
    return v * (q[3] ** 2 - u @ u) + 2 * u * (u @ v) + 2 * q[3] * np.cross(u, v)
