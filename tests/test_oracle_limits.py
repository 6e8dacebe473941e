"""The oracle's Revolute / Prismatic travel limits and velocity motors
(limit rows crates/physics/src/avbd/assembly.rs:809-874, their one-sided
accumulators joints_ext.rs:820-890, motor_impulse joints_ext.rs:237-356,
the hinge twist of crates/physics/src/sequential_impulse/joints.rs:38-48)
against the reference's own tests:

- unit pins (crates/physics/src/avbd/joint_tests.rs): a warm limit reaction
  at zero violation still moves the body, linear and angular (:102-143);
  the slider's limit dual carries its force between iterations, holds it at
  zero violation, unwinds inside the slop band and clears outside the
  window (:222-250); a slider hanging on its limit under 1, 10 and 100 g
  in both directions holds its rest position, its dual equals the load,
  and once released it leaves without a catapult or a phantom hold
  (:145-220);
- behaviour (crates/physics/tests/avbd_engine.rs): a pendulum clamps at its
  +-0.5 rad hinge window (:507-549), a hinge motor spins a bob up to its
  target speed (:552-595), a slider keeps its line and rests on its lower
  limit (:598-646) without ratcheting over 1500 steps (:648-700), a slider
  motor drives the bob along the slide (:782-815)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes

DT = 1.0 / 60.0
REV, PRISM = scenes.JOINT_REVOLUTE, scenes.JOINT_PRISMATIC
SLOP_LIN = 0.002
IDQ = (0.0, 0.0, 0.0, 1.0)


def z_twist(q):
    return 2.0 * math.atan2(float(q[2]), float(q[3]))


def row_scene(kind, limit, motor=None):
    """joint_tests.rs:4-28: static and dynamic spheres at the origin, no
    gravity, the joint at their centres with axis y, step-start and
    inertial poses at rest."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.sphere((0.0, 0.0, 0.0), 0.25, 0.0)
    sb.sphere((0.0, 0.0, 0.0), 0.25, 1.0)
    sb.joint(kind, 0, 1, (0, 0, 0), (0, 0, 0), (0, 1, 0), (0, 1, 0), limit=limit, motor=motor)
    e = po.OracleEngine(sb.build())
    for i in range(2):
        e.set_step_start(i, (0.0, 0.0, 0.0), IDQ)
        e.set_inertial(i, (0.0, 0.0, 0.0), IDQ)
    return e


@pytest.mark.parametrize("kind", [PRISM, REV])
def test_warm_limit_reaction_kept_at_zero_violation(kind):
    e = row_scene(kind, (0.0, 1.0))
    e.set_joint_limit_state(0, -10.0, -10.0)
    e.solve_body(1)
    st = e.state()[1]
    movement = np.abs(st[3:6]).sum() if kind == REV else np.linalg.norm(st[0:3])
    assert movement > 1e-4, movement


@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_prismatic_dual_carries_force_between_iterations(sign):
    e = row_scene(PRISM, (-1.0, 1.0))
    e.set_joint_duals(0, [0, 0, 0, 1, 1, 100, 0, 0, 0, 1, 1, 1])
    e.set_joint_limit_state(0, 0.0, sign * 10.0)
    y = np.float32(sign * 1.01)
    e.set_pose(1, (0.0, float(y), 0.0), IDQ)
    expected = sign * 10.0 + 100.0 * (float(y) - sign)
    e.dual_update()
    carried = float(e.joint_limit_state(0)[1])
    assert abs(carried - expected) < 1e-5, (carried, expected)
    e.set_pose(1, (0.0, sign, 0.0), IDQ)  # zero C, the reaction stays
    e.dual_update()
    assert e.joint_limit_state(0)[1] == np.float32(carried)
    e.set_pose(1, (0.0, sign * (1.0 - SLOP_LIN * 0.5), 0.0), IDQ)  # inside the slop band: unwinds
    e.dual_update()
    assert abs(e.joint_limit_state(0)[1]) < abs(carried)
    e.set_pose(1, (0.0, 0.0, 0.0), IDQ)
    e.dual_update()
    assert e.joint_limit_state(0)[1] == 0.0


@pytest.mark.parametrize("sign", [-1.0, 1.0])
@pytest.mark.parametrize("load", [9.81, 98.1, 981.0])
def test_prismatic_high_load_release_no_catapult(sign, load):
    sb = scenes.SceneBuilder(gravity=(0.0, sign * load, 0.0))
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    b = sb.sphere((0.0, sign * 2.9, 0.0), 0.25, 1.0)
    sb.joint(PRISM, 0, b, (0, 0, 0), (0, -sign, 0), (0, 1, 0), (0, 1, 0), limit=(-2.0, 2.0))
    e = po.OracleEngine(sb.build())
    rest = sign * 4.9
    for step in range(3000):
        e.wake_body(b)
        e.step(DT)
        if step >= 200:
            st = e.state()[b]
            assert abs(st[1] - rest) < 0.02, (step, st[:3])
            assert np.linalg.norm(st[7:10]) < 1.0, (step, st[7:10])
    force = float(e.joint_limit_state(0)[1])
    assert abs(force - sign * load) < load * 0.02, (force, sign * load)
    e.set_gravity((0.0, 0.0, 0.0))
    e.set_vel(b, (0.0, -sign, 0.0))
    for step in range(120):
        e.wake_body(b)
        e.step(DT)
        assert np.linalg.norm(e.state()[b, 7:10]) < 2.0
    assert sign * (rest - e.state()[b, 1]) > 0.5
    assert e.joint_limit_state(0)[1] == 0.0


def test_hinge_limit_holds_bound():
    sb = scenes.SceneBuilder()
    sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    bob = sb.sphere((1.0, -1.0, 0.0), 0.25, 1.0)
    lb = np.array([-1.0, 1.0, 0.0])
    sb.joint(REV, 0, bob, (0, 0, 0), tuple(lb), (0, 0, 1), (0, 0, 1), limit=(-0.5, 0.5))
    e = po.OracleEngine(sb.build())
    max_abs = 0.0
    for _ in range(300):
        e.step(DT)
        st = e.state()[bob]
        q = st[3:7].astype(np.float64)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        assert np.linalg.norm(st[0:3] + R @ lb) < 0.2
        max_abs = max(max_abs, abs(z_twist(q)))
    assert 0.2 < max_abs <= 0.62, max_abs


def test_hinge_motor_spins_up():
    sb = scenes.SceneBuilder()
    a = sb.sphere((0.0, 0.0, 0.0), 0.05, 0.0)
    bob = sb.sphere((0.0, 0.0, 0.0), 0.1, 1.0)
    sb.joint(REV, a, bob, (0, 0, 0), (0, 0, 0), (0, 0, 1), (0, 0, 1), motor=(3.0, 20.0))
    sc = sb.build()
    sc["trigger"][a] = 1
    e = po.OracleEngine(sc)
    e.step(DT, 300)
    mean_w = 0.0
    for _ in range(60):
        e.step(DT)
        mean_w += float(e.state()[bob, 12])
    assert abs(mean_w / 60.0 - 3.0) < 0.6, mean_w / 60.0


def _slider():
    sb = scenes.SceneBuilder()
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bob = sb.sphere((0.0, -2.9, 0.0), 0.25, 1.0)
    sb.joint(PRISM, 0, bob, (0, 0, 0), (0, 1, 0), (0, 1, 0), (0, 1, 0), limit=(-2.0, 0.0))
    return po.OracleEngine(sb.build()), bob


def test_slider_holds_line_and_limit():
    e, bob = _slider()
    for _ in range(300):
        e.step(DT)
        p = e.state()[bob]
        assert abs(p[0]) < 0.08 and abs(p[2]) < 0.08
    assert -5.4 < e.state()[bob, 1] < -4.4


def test_slider_limit_sustained_no_ratchet():
    e, bob = _slider()
    for i in range(1500):
        e.step(DT)
        if i >= 200:
            assert np.linalg.norm(e.state()[bob, 7:10]) < 10.0
    assert -5.6 < e.state()[bob, 1] < -4.2


def test_slider_motor_drives():
    sb = scenes.SceneBuilder()
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bob = sb.sphere((1.0, 0.0, 0.0), 0.25, 1.0)
    sb.joint(PRISM, 0, bob, (0, 0, 0), (-1, 0, 0), (1, 0, 0), (1, 0, 0), motor=(2.0, 50.0))
    e = po.OracleEngine(sb.build())
    e.step(DT, 300)
    st = e.state()[bob]
    assert abs(st[7] - 2.0) < 0.4 and abs(st[1]) < 0.15 and abs(st[2]) < 0.1, st


def test_bad_limit_and_motor_rejected():
    sb = scenes.SceneBuilder()
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    sb.sphere((1.0, 0.0, 0.0), 0.25, 1.0)
    e = po.OracleEngine(sb.build())
    for prm in ((1.0, -1.0, 0, 0, 1, 0), (float("nan"), 1.0, 0, 0, 1, 0), (0, 0, 1.0, -1.0, 0, 1),
                (0, 0, float("inf"), 1.0, 0, 1)):
        with pytest.raises(ValueError):
            e.add_joint(0, 1, (0, 0, 0), (0, 0, 0), kind=REV, axis_a=(0, 0, 1), axis_b=(0, 0, 1), prm=prm)
