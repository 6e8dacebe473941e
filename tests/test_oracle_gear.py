"""The oracle's Gear joint (gear sides and their unwrap memory
crates/physics/src/avbd/joints_ext.rs:38-161, rows assembly.rs:426-449 and
:650-679, dual joints_ext.rs:897-917, creation mod.rs:1329-1379) against the
reference's own tests (crates/physics/tests/avbd_engine.rs):

- creation is refused for dangling joint references, a non-finite ratio
  and a gear over joints that are neither Revolute nor Prismatic
  (:187-255);
- a motor-driven hinge coupled 2:1 to a free hinge turns past pi while the
  gear holds a + 2 b = 0 on the unwrapped twists (:1360-1450);
- two sliders coupled with ratio -1 travel metres together, also across a
  host teleport of one slider by 4 m (a linear side must never be wrapped
  to +-pi, :824-890).

Closed-form pins: a gear at its assembly pose has zero residual and leaves
a resting scene's poses untouched; the residual the hinge test reads is
the oracle's own constraint value."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes

DT = 1.0 / 60.0
REV, PRISM, GEAR = scenes.JOINT_REVOLUTE, scenes.JOINT_PRISMATIC, scenes.JOINT_GEAR


def wrap(x):
    return (x + math.pi) % (2.0 * math.pi) - math.pi


def atan_twist(q):
    return 2.0 * math.atan2(float(q[2]), float(q[3]))


def hinge_pair_scene(motor=(1.5, 20.0)):
    """avbd_engine.rs:1360-1420: a static ground, two arms on z hinges, A
    driven at 1.5 rad/s within 20 N m, B free, a 2:1 gear, no gravity."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    ground = sb.box((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 0.0)
    arm_a = sb.box((0.0, 0.6, 0.0), (0.1, 0.6, 0.1), 1.0)
    arm_b = sb.box((1.0, 0.6, 0.0), (0.1, 0.6, 0.1), 1.0)
    ja = sb.joint(REV, ground, arm_a, (0, 0, 0), (0, -0.6, 0), (0, 0, 1), (0, 0, 1), motor=motor)
    jb = sb.joint(REV, ground, arm_b, (1.0, 0, 0), (0, -0.6, 0), (0, 0, 1), (0, 0, 1))
    sb.gear(arm_a, arm_b, ja, jb, 2.0)
    return sb, arm_a, arm_b


def slider_pair_scene():
    """avbd_engine.rs:824-866: a static anchor, two spheres on x slides at
    x = 1 and 3, the first driven at 2 m/s within 50 N, a ratio -1 gear."""
    sb = scenes.SceneBuilder(gravity=(0.0, -9.81, 0.0))
    anchor = sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bobs, js = [], []
    for x, motor in ((1.0, (2.0, 50.0)), (3.0, None)):
        bob = sb.sphere((x, 0.0, 0.0), 0.25, 1.0)
        js.append(sb.joint(PRISM, anchor, bob, (x, 0, 0), (0, 0, 0), (1, 0, 0), (1, 0, 0), motor=motor))
        bobs.append(bob)
    sb.gear(bobs[0], bobs[1], js[0], js[1], -1.0)
    return sb, bobs


def test_gear_spec_validation():
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    g = sb.box((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 0.0)
    a = sb.box((0.0, 1.0, 0.0), (0.1, 0.1, 0.1), 1.0)
    b = sb.box((1.0, 1.0, 0.0), (0.1, 0.1, 0.1), 1.0)
    ja = sb.joint(REV, g, a, (0, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, 1))
    jball = sb.joint(0, g, b, (1, 0, 0), (0, -1, 0))
    e = po.OracleEngine(sb.build())
    z = (0.0, 0.0, 0.0)
    for prm in ((ja, 7, 1.0),            # dangling reference
                (ja, ja, float("nan")),  # non-finite ratio
                (ja, ja, float("inf")),
                (ja, jball, 1.0)):       # a Ball joint cannot be geared
        with pytest.raises(ValueError):
            e.add_joint(a, b, z, z, GEAR, prm=prm)
    assert e.add_joint(a, b, z, z, GEAR, prm=(ja, ja, 1.0)) == 2


def test_gear_at_assembly_pose_is_inert():
    """No drive, no gravity: the gear's residual is zero at its assembly
    pose, so every row is dead and no body moves by a single bit."""
    sb, arm_a, arm_b = hinge_pair_scene(motor=None)
    e = po.OracleEngine(sb.build())
    before = e.state().copy()
    e.step_inner(30)
    assert np.array_equal(before.view(np.uint32), e.state().view(np.uint32))


def test_gear_ratio_couples_hinges():
    sb, arm_a, arm_b = hinge_pair_scene()
    e = po.OracleEngine(sb.build())
    ua = ub = pa = pb = 0.0
    for i in range(180):
        e.step_inner(1)
        st = e.state()
        xa, xb = atan_twist(st[arm_a][3:7]), atan_twist(st[arm_b][3:7])
        if i == 0:
            pa, pb = xa, xb
        ua += wrap(xa - pa)
        ub += wrap(xb - pb)
        pa, pb = xa, xb
    ta = atan_twist(e.state()[arm_a][3:7])
    assert abs(ta) > 0.5, ta
    assert abs(ua) > math.pi, ua  # the drive turned A past pi: the unwrap was exercised
    assert abs(ua + 2.0 * ub) < 0.35 * max(abs(ua), 1.0), (ua, ub)


def test_gear_linear_sides_no_wrap_teleport():
    sb, (bob_a, bob_b) = slider_pair_scene()
    e = po.OracleEngine(sb.build())
    e.step_inner(120)
    st = e.state()
    e.set_pose(bob_a, (float(st[bob_a][0]) + 4.0, float(st[bob_a][1]), float(st[bob_a][2])),
               tuple(float(x) for x in st[bob_a][3:7]))
    e.step_inner(120)
    st = e.state()
    sa, sb_ = float(st[bob_a][0]) - 1.0, float(st[bob_b][0]) - 3.0
    assert sa > 2.0 and sb_ > 2.0, (sa, sb_)
    assert abs(sa - sb_) < 0.6, (sa, sb_)
