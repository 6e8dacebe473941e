"""The oracle's SixDof joint (JointKind::SixDof: per axis of body A's live
frame free, locked or limited; rows crates/physics/src/avbd/assembly.rs:
452-519 and :682-740, duals joints_ext.rs:646-714, rest residual
sleep.rs:86-103) against the reference's own tests:

- sixdof_angular_rows_commit_duals_in_local_frame
  (crates/physics/src/avbd/joint_tests.rs:31-68): a 0.3 rad local-z twist
  under a common world rotation commits 0.3 to a locked axis' dual and
  0.1 to a [-0.2, 0.2] limited axis' accumulator, which clears once the
  twist is gone;
- sixdof_locked_dual_is_invariant_under_world_rotation (:313-326) for every
  axis and pitch;
- a4_limit_zero_violation_keeps_warm_reaction's SixDof cases (:102-143);
- avbd_sixdof_angular_limit_blocks_spin
  (crates/physics/tests/avbd_engine.rs:1514-1546): an 8 rad/s spin clamps
  near its +-0.2 rad window;
- a fully locked SixDof holds like a weld, a free one lets the bodies part,
  and its bodies keep colliding (SixDof is no no-collide joint)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes

DT = 1.0 / 60.0
F_, L_, M_ = scenes.AXIS_FREE, scenes.AXIS_LOCKED, scenes.AXIS_LIMITED
IDQ = (0.0, 0.0, 0.0, 1.0)


def aa(axis, ang):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.concatenate([axis * math.sin(ang / 2), [math.cos(ang / 2)]])


def qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def row_scene(kinds, lo=(0.0,) * 6, hi=(0.0,) * 6):
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.sphere((0.0, 0.0, 0.0), 0.25, 0.0)
    sb.sphere((0.0, 0.0, 0.0), 0.25, 1.0)
    sb.sixdof(0, 1, (0, 0, 0), (0, 0, 0), kinds, lo, hi)
    e = po.OracleEngine(sb.build())
    for i in range(2):
        e.set_step_start(i, (0.0, 0.0, 0.0), IDQ)
        e.set_inertial(i, (0.0, 0.0, 0.0), IDQ)
    return e


@pytest.mark.parametrize("locked", [False, True])
def test_angular_rows_commit_duals_in_local_frame(locked):
    kinds = [F_] * 5 + [L_ if locked else M_]
    e = row_scene(kinds, (0,) * 5 + (-0.2,), (0,) * 5 + (0.2,))
    qa = aa((0, 1, 0), 0.7)
    qb = qmul(qa, aa((0, 0, 1), 0.3))
    e.set_pose(0, (0.0, 0.0, 0.0), tuple(np.float32(qa)))
    e.set_pose(1, (0.0, 0.0, 0.0), tuple(np.float32(qb)))
    e.dual_update()
    force = e.joints(1)[0, 8] if locked else e.joint_sacc(0)[5]
    assert abs(force - (0.3 if locked else 0.1)) < 1e-6, force
    assert e.joints(1)[0, 11] > 1.0
    e.set_pose(1, (0.0, 0.0, 0.0), tuple(np.float32(qa)))
    e.dual_update()
    if not locked:
        assert e.joint_sacc(0)[5] == 0.0


@pytest.mark.parametrize("pitch", [0.0, 0.7, math.pi * 0.5, math.pi])
def test_locked_dual_invariant_under_world_rotation(pitch):
    for i, axis in enumerate(np.eye(3)):
        kinds = [F_] * 6
        kinds[3 + i] = L_
        e = row_scene(kinds)
        qa = aa((0, 1, 0), pitch)
        e.set_pose(0, (0.0, 0.0, 0.0), tuple(np.float32(qa)))
        e.set_pose(1, (0.0, 0.0, 0.0), tuple(np.float32(qmul(qa, aa(axis, 0.3)))))
        e.dual_update()
        assert abs(e.joints(1)[0, 6 + i] - 0.3) < 1e-6, (pitch, i, e.joints(1)[0, 6:9])


def test_spin_clamps_at_the_angular_window():
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    a = sb.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), 0.0)
    b = sb.box((0.0, 0.0, 0.0), (0.4, 0.4, 0.4), 1.0, angvel=(0.0, 0.0, 8.0))
    sb.sixdof(a, b, (0, 0, 0), (0, 0, 0), [L_] * 5 + [M_], (0,) * 5 + (-0.2,), (0,) * 5 + (0.2,))
    e = po.OracleEngine(sb.build())
    e.step(DT, 120)
    q = e.state()[b, 3:7]
    tw = 2.0 * math.atan2(float(q[2]), float(q[3]))
    assert abs(tw) < 0.4, tw


def test_locked_holds_free_parts_and_bodies_collide():
    for kinds, expect_held in (([L_] * 6, True), ([F_] * 6, False)):
        sb = scenes.SceneBuilder()
        sb.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0)
        a = sb.box((-0.5, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0)
        c = sb.box((0.5, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0, vel=(0.0, 0.0, 2.0))
        sb.sixdof(a, c, (0.5, 0, 0), (-0.5, 0, 0), kinds)
        e = po.OracleEngine(sb.build())
        e.step(DT, 120)
        st = e.state()
        d = np.linalg.norm(st[c, :3] - st[a, :3])
        assert (abs(d - 1.0) < 0.15) == expect_held, (kinds, d)
    # SixDof is no no-collide joint (rows.rs:437-466): overlapping boxes
    # joined by one keep their contact pair, joined by a Ball joint they lose it
    for kind, expect_pair in ((scenes.JOINT_SIXDOF, True), (scenes.JOINT_BALL, False)):
        sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
        a = sb.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
        c = sb.box((0.9, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
        if kind == scenes.JOINT_SIXDOF:
            sb.sixdof(a, c, (0, 0, 0), (0, 0, 0), [F_] * 6)
        else:
            sb.ball_joint(a, c, (0.45, 0, 0), (-0.45, 0, 0))
        e = po.OracleEngine(sb.build())
        e.step(DT)
        assert ([a, c] in e.pairs()[0].tolist()) == expect_pair, kind


def test_bad_sixdof_rejected():
    sb = scenes.SceneBuilder()
    sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    sb.sphere((1.0, 0.0, 0.0), 0.2, 1.0)
    e = po.OracleEngine(sb.build())
    for prm in ((2, 0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 0, -1.0, 0, 0, 0, 0, 0),
                (3, 0, 0, 0, 0, 0) + (0,) * 12, (2, 0, 0, 0, 0, 0, float("nan")) + (0,) * 11):
        with pytest.raises(ValueError):
            e.add_joint(0, 1, (0, 0, 0), (0, 0, 0), kind=9, prm=prm)
