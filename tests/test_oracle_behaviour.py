"""The reference's behavioural AVBD tests (crates/physics/tests/
avbd_engine.rs), restated against the CPU oracle in the reference's own
order (reverse-handle sweep, history-ordered pairs, libm)."""
import numpy as np

from oracle import pyoracle as po
from ornis_amd import scenes

DT = 1.0 / 60.0


def _single_box():
    sb = scenes.SceneBuilder()
    sb.box((0.0, 5.0, 0.0), (0.25, 0.25, 0.25), 1.0)
    return sb.build()


def test_substeps_honor_host_dt():
    """avbd_substeps_honor_host_dt (avbd_engine.rs:1132-1189): 120 Hz host
    frames and 1/30 s hitches reproduce 60 Hz within 1e-4; a 1 s debt runs
    at most four substeps."""
    ref = po.OracleEngine(_single_box())
    ref.step(DT, 60)
    y_ref = ref.state()[0, 1]
    fast = po.OracleEngine(_single_box())
    fast.step(DT / 2.0, 120)
    assert abs(fast.state()[0, 1] - y_ref) < 1e-4
    hitch = po.OracleEngine(_single_box())
    hitch.step(2.0 * DT, 30)
    assert abs(hitch.state()[0, 1] - y_ref) < 1e-4
    stall = po.OracleEngine(_single_box())
    stall.step(1.0, 1)
    assert stall.state()[0, 1] > 4.9


def test_fast_box_stops_at_thin_wall():
    """avbd_fast_box_stops_at_thin_wall (avbd_engine.rs:1290-1313)."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.box((5.0, 0.0, 0.0), (0.02, 2.0, 2.0), 0.0)
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 1.0, vel=(10.0, 0.0, 0.0), restitution=0.0)
    e = po.OracleEngine(sb.build())
    e.step(DT, 120)
    assert 4.0 < e.state()[1, 0] < 5.0


def test_bounced_retouch_catches_at_offset():
    """avbd_bounced_retouch_catches_at_offset (avbd_engine.rs:1770-1808)."""
    for x0 in (-3.0, 0.0):
        sb = scenes.SceneBuilder()
        sb.box((0.0, -1.0, 0.0), (10.0, 1.0, 10.0), 0.0)
        sb.box((x0, 8.0, 0.0), (0.4, 0.4, 0.4), 1.0, vel=(0.0, -30.0, 0.0))
        e = po.OracleEngine(sb.build())
        min_y = np.inf
        for _ in range(300):
            e.step(DT, 1)
            min_y = min(min_y, e.state()[1, 1])
        b = e.state()[1]
        assert min_y > 0.0
        assert abs(b[1] - 0.4) < 0.05 and abs(b[0] - x0) < 0.15 and np.linalg.norm(b[7:10]) < 0.5


def test_sleep_freezes_settled_stack_and_wakes_on_touch():
    """avbd_sleep_freezes_settled_stack_and_wakes_on_touch
    (avbd_engine.rs:1065-1131): a settled stack sleeps; a box added later
    and dropped on it wakes it; the pile holds."""
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0)
    for i in range(3):
        sb.box((0.0, 0.25 + 0.5 * i, 0.0), (0.25, 0.25, 0.25), 1.0)
    e = po.OracleEngine(sb.build())
    e.step(DT, 600)
    assert e.stats()["sleeping"] == 3
    tops = e.state()[1:4, 1].copy()
    add = scenes.SceneBuilder()
    add.box((0.0, 3.0, 0.0), (0.25, 0.25, 0.25), 1.0)
    e.add_bodies(add.build())
    woke = False
    for _ in range(180):
        e.step(DT, 1)
        if e.stats()["sleeping"] < 3:
            woke = True
            break
    assert woke
    e.step(DT, 600)
    st = e.state()
    assert np.all(np.abs(st[1:4, 1] - tops) < 0.05)
    assert abs(st[4, 1] - 1.75) < 0.15


def test_discovery_thread_count_does_not_change_bits():
    """avbd_engine.rs:1661-1768: 1-thread and many-thread pair discovery give
    the same bodies bit for bit (the sort after discovery fixes the order)."""
    s = scenes.box_drop_scene(600, np.random.Generator(np.random.PCG64(8)), floor=14.0)
    outs = []
    for th in (1, 8):
        e = po.OracleEngine(s, order=po.ORDER_REFERENCE)
        e.set_threads(th)
        e.step_inner(30)
        outs.append((e.state(), e.pairs()[0]))
    assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
    assert np.array_equal(outs[0][1], outs[1][1])


def test_removing_an_unrelated_body_keeps_the_moved_tails_joint():
    """removing_an_unrelated_body_preserves_the_moved_tails_joint
    (solver_lifecycle.rs:41-78), with a Ball joint in place of the Distance
    joint: the tail moves into the removed spare's slot with its joint, and
    the joint pulls it back from a 1 m teleport within 600 steps."""
    for order in (po.ORDER_REFERENCE, po.ORDER_COLOUR):
        b = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
        b.sphere((0.0, 0.0, 0.0), 0.2, 0.0)
        b.sphere((20.0, 0.0, 0.0), 0.2, 1.0)
        b.sphere((0.0, 2.0, 0.0), 0.2, 1.0)
        b.ball_joint(0, 2, (0.0, 0.0, 0.0), (0.0, -2.0, 0.0))
        s = b.build()
        s["mask"][:] = 0
        e = po.OracleEngine(s, order=order)
        e.remove_body(1)
        assert e.n == 2
        st = e.state()
        assert np.array_equal(st[1], po.OracleEngine(s, order=order).state()[2])
        e.set_pose(1, (0.0, 3.0, 0.0), st[1, 3:7])
        e.step(DT, 600)
        assert abs(e.state()[1, 1] - 2.0) < 0.04
        assert np.abs(e.joints(1)).max() > 0.0


def test_remove_body_drops_renumbers_and_wakes():
    """remove_body (mod.rs:1241-1300) on a settled pile: the removed body's
    pairs go, the last body's pairs follow it to the freed index with their
    warm starts (while the pair keeps a < b), the rest are untouched, and
    every sleeping body wakes."""
    s = scenes.box_drop_scene(40, np.random.Generator(np.random.PCG64(12)))
    e = po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    e.step_inner(240)
    assert e.stats()["sleeping"] > 0
    n = e.n
    last = n - 1
    ab, npts, _, _ = e.pairs()
    pts = e.points()
    off = np.concatenate([[0], np.cumsum(npts)]).astype(np.int64)
    before = {tuple(p): pts[off[i]:off[i + 1]] for i, p in enumerate(ab.tolist())}
    touched = [p for p in before if last in p]
    assert touched, "the last body should rest on something"
    h = int(next(p for p in before if last not in p and p[0] != 0)[0])
    st = e.state()
    e.remove_body(h)
    assert e.n == n - 1
    assert np.array_equal(e.state()[h], st[last])
    assert not e.sleep()[0].any()
    ab2, npts2, _, _ = e.pairs()
    pts2 = e.points()
    off2 = np.concatenate([[0], np.cumsum(npts2)]).astype(np.int64)
    after = {tuple(p): pts2[off2[i]:off2[i + 1]] for i, p in enumerate(ab2.tolist())}
    remap = lambda x: h if x == last else x  # noqa: E731
    want = {(remap(a), remap(b)): v for (a, b), v in before.items() if h not in (a, b)}
    assert set(after) == set(want)
    for k, v in want.items():
        assert np.array_equal(after[k].view(np.uint32), v.view(np.uint32))


def test_heavy_column_sinks_under_the_penalty_ramp():
    """Heavy stacks sink in the reference's own algorithm: a column of 40
    one-metre 1,000 kg cubes on a plane, 300 steps in the reference's order
    (reverse-handle sweep), ends with the bottom cube about 15 cm into the
    plane and the column about 5 m shorter than stacked (penalties start at
    PENALTY_INIT = 1 and grow by BETA |C| per iteration, mod.rs:154-247, so
    contacts under a heavy load stay penetrated for many steps); the GPU's
    colour order sinks the same amount within 1 cm at the bottom.  The
    100,000-box pile's floor and wall penetrations under 40 layers of such
    boxes are this behaviour."""
    sb = scenes.SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    for i in range(40):
        sb.box((0.0, 0.5 + i * 1.02, 0.0), (0.5, 0.5, 0.5), 1000.0)
    sc = sb.build()
    ref = po.OracleEngine(sc)
    col = po.OracleEngine(sc, order=po.ORDER_COLOUR)
    ref.step_inner(300)
    col.step_inner(300)
    y_ref, y_col = ref.state()[1:, 1], col.state()[1:, 1]
    assert 0.1 < 0.5 - y_ref[0] < 0.2, y_ref[0]
    assert 39.8 - (y_ref[-1] + 0.5) > 3.0  # the column's total compression
    assert abs(y_col[0] - y_ref[0]) < 0.01
    assert np.isfinite(y_ref).all() and np.isfinite(y_col).all()


def _aniso_slide(direction, quat=(0.0, 0.0, 0.0, 1.0), order=po.ORDER_REFERENCE):
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0, friction=0.1)
    d = np.asarray(direction, np.float32)
    sb.box((0.0, 0.5, 0.0), (0.5, 0.5, 0.5), 1.0, friction=0.1, vel=tuple(5.0 * d), quat=quat)
    e = po.OracleEngine(sb.build(), order=order)
    e.set_friction_axes(np.array([[0, 0, 0], [1, 0, 0]], np.float32), np.array([0.1, 1.0], np.float32))
    for _ in range(10):
        e.step(DT)
    return 5.0 - float(e.state()[1, 7:10] @ d)


def test_anisotropic_friction_channels():
    """avbd_anisotropic_friction_channels (avbd_engine.rs:466-495): a box
    with friction 0.1 along its x axis and 1.0 across it slides 10 steps at
    5 m/s; the transverse slide decays more than 4x faster.  Each channel
    loses mu g t of speed (Coulomb: 0.164 and 1.64 m/s) within 5 %, in both
    sweep orders, and a box turned 90 degrees about y swaps the channels:
    the axis turns with the body (friction_frame, rows.rs:32-52)."""
    for order in (po.ORDER_REFERENCE, po.ORDER_COLOUR):
        dx, dz = _aniso_slide((1, 0, 0), order=order), _aniso_slide((0, 0, 1), order=order)
        assert dz > 4.0 * max(dx, 1e-6)
        assert abs(dx - 0.1 * 9.81 * 10 * DT) < 0.05 * 0.1 * 9.81 * 10 * DT
        assert abs(dz - 1.0 * 9.81 * 10 * DT) < 0.05 * 1.0 * 9.81 * 10 * DT
    c = float(np.cos(np.pi / 4))
    turned = (0.0, c, 0.0, c)
    assert abs(_aniso_slide((1, 0, 0), turned) - dz) < 1e-3
    assert abs(_aniso_slide((0, 0, 1), turned) - dx) < 1e-3
