"""The oracle's free Motor joint (JointKind::Motor: a 6-DOF relative
velocity drive, free_motor_impulse crates/physics/src/avbd/joints_ext.rs:
368-421; no position rows, no dual state; its rest residual and sleep vote,
sleep.rs:80-82 and joint.rs:474-477):

- avbd_motor_drives_box_to_target_against_friction
  (crates/physics/tests/motor_joint.rs:90-110): a resting box driven at
  2 m/s along x against floor friction reaches it within 0.4 m/s;
- the budget is a hard cap: a force budget below the sliding friction
  (mu m g) leaves the box at rest, and the first step's speed gain of a
  free body is exactly budget * dt / m;
- an angular target spins a free ball up to it about the target axis;
- a zero-budget motor lets its bodies sleep, a live one keeps them awake;
- a non-finite target or a negative budget is rejected."""
import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes

DT = 1.0 / 60.0
MOTOR = 10


def _box_scene(lin, ang, max_force, max_torque, gravity=(0.0, -9.81, 0.0)):
    sb = scenes.SceneBuilder(gravity=gravity)
    sb.box((0.0, -0.5, 0.0), (10.0, 0.5, 10.0), 0.0)
    block = sb.box((0.0, 0.5, 0.0), (0.5, 0.5, 0.5), 1.0)
    anchor = sb.sphere((0.0, 5.0, 0.0), 0.1, 0.0)
    sb.motor(anchor, block, lin, ang, max_force, max_torque)
    return sb.build(), block


def test_motor_drives_box_against_friction():
    sc, block = _box_scene((2.0, 0.0, 0.0), (0.0, 0.0, 0.0), 8.0, 10.0)
    e = po.OracleEngine(sc)
    e.step(DT, 300)
    v = e.state()[block, 7:10]
    assert abs(v[0] - 2.0) < 0.4, v


def test_motor_budget_caps_the_drive():
    # box friction 0.5 on the floor's 0.5: sliding needs 0.5 * 9.81 N > 3 N
    sc, block = _box_scene((2.0, 0.0, 0.0), (0.0, 0.0, 0.0), 3.0, 0.0)
    e = po.OracleEngine(sc)
    e.step(DT, 120)
    assert np.linalg.norm(e.state()[block, 7:10]) < 0.05
    # a free 2 kg ball in zero gravity gains exactly budget * dt / m per step
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    a = sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    b = sb.sphere((3.0, 0.0, 0.0), 0.3, 2.0)
    sb.motor(a, b, (10.0, 0.0, 0.0), (0.0, 0.0, 0.0), 4.0, 0.0)
    e = po.OracleEngine(sb.build())
    e.step(DT)
    assert abs(e.state()[b, 7] - 4.0 * DT / 2.0) < 1e-5, e.state()[b, 7]


def test_motor_spins_free_ball_to_angular_target():
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    a = sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    ball = sb.sphere((0.0, 2.0, 0.0), 0.3, 1.0)
    sb.motor(a, ball, (0.0, 0.0, 0.0), (0.0, 0.0, 3.0), 0.0, 10.0)
    e = po.OracleEngine(sb.build())
    e.step(DT, 120)
    w = e.state()[ball, 10:13]
    assert abs(w[2] - 3.0) < 0.05 and abs(w[0]) < 1e-3 and abs(w[1]) < 1e-3, w


def test_motor_sleep_vote():
    for max_force, expect_sleep in ((0.0, True), (8.0, False)):
        sc, block = _box_scene((1.0, 0.0, 0.0), (0.0, 0.0, 0.0), max_force, 0.0)
        e = po.OracleEngine(sc)
        e.step(DT, 300)
        assert bool(e.sleep()[0][block]) == expect_sleep


def test_bad_motor_rejected():
    sc, block = _box_scene((1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0, 1.0)
    e = po.OracleEngine(sc)
    for prm in ((float("nan"), 0, 0, 0, 0, 0, 1.0, 1.0), (0, 0, 0, 0, 0, 0, -1.0, 1.0),
                (0, 0, 0, 0, 0, 0, 1.0, float("inf"))):
        with pytest.raises(ValueError):
            e.add_joint(0, block, (0, 0, 0), (0, 0, 0), kind=MOTOR, prm=prm)
