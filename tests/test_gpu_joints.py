"""Revolute, Prismatic (with travel limits and velocity motors), Fixed,
Distance, Rope and Spring joints on the GPU
against the oracle (float32, colour order), bit for bit: bodies, pairs,
colourings and the joints' linear and angular duals.  The scenes are the
reference's own joint tests (crates/physics/tests/avbd_engine.rs: hinge
:139-183, weld :891-935, slider :598-646 without its limit, rod :939-966;
crates/physics/tests/p6_rope_spring_motor.rs: rope :45-105, springs
:124-216), chains whose links cycle through the first four kinds with
tilted axes, falling into a field of spheres and onto each other, with a
body removed mid-run, and a field of tethered bodies (ropes sharing bodies,
rods, springs of both motor models and a damper) over a floor; the
reference's hinge / slider limit and motor tests (avbd_engine.rs:507-815)
and chains of limited and motorised hinges and sliders."""
import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes
from ornis_amd.avbd import AvbdEngine

pytestmark = pytest.mark.gpu


def _same(g, o, nj, what):
    sg, so = g.state(), o.state()
    bad = np.where((sg.view(np.uint32) != so.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} bodies differ (first {bad[:5]}), max |d| {np.abs(sg - so).max():.3e}"
    assert np.array_equal(g.pairs(), o.pairs()[0]), f"{what}: pair lists differ"
    assert np.array_equal(g.colours(), o.colours()), f"{what}: colourings differ"
    jg, jo = g.joints(), o.joints(nj)
    assert np.array_equal(jg.view(np.uint32), jo.view(np.uint32)), f"{what}: joint duals differ"


def _run(scene, steps, every=10):
    g = AvbdEngine(scene=scene)
    o = po.OracleEngine(scene, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    nj = len(scene["joints"])
    for s in range(steps):
        g.step_n(1)
        o.step_inner(1)
        if s % every == every - 1:
            _same(g, o, nj, f"step {s}")
    return g, o


def test_revolute_hinge_bitwise(gpu_available):
    b = scenes.SceneBuilder()
    a = b.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    arm = b.box((0.3, -1.0, 0.0), (0.2, 1.0, 0.2), 1.0)
    b.joint(scenes.JOINT_REVOLUTE, a, arm, (0, 0, 0), (-0.3, 1.0, 0.0), (0, 0, 1), (0, 0, 1))
    g, _ = _run(b.build(), 300, every=20)
    st = g.state()
    assert np.abs(g.joints()[0, 6:9]).max() > 0 or np.abs(g.joints()[0, 0:3]).max() > 0


def test_fixed_weld_drop_bitwise(gpu_available):
    b = scenes.SceneBuilder()
    b.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0)
    a = b.box((-0.5, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    c = b.box((0.5, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    b.joint(scenes.JOINT_FIXED, a, c, (0.5, 0, 0), (-0.5, 0, 0))
    g, _ = _run(b.build(), 300, every=20)
    st = g.state()
    assert abs(np.linalg.norm(st[c, :3] - st[a, :3]) - 1.0) < 0.15 and st[a, 1] > 0.2


def test_prismatic_slide_bitwise(gpu_available):
    b = scenes.SceneBuilder()
    anchor = b.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bob = b.sphere((0.0, -2.9, 0.0), 0.25, 1.0, vel=(0.3, 0.0, -0.2))
    b.joint(scenes.JOINT_PRISMATIC, anchor, bob, (0, 0, 0), (0, 1, 0), (0, 1, 0), (0, 1, 0))
    g, _ = _run(b.build(), 120, every=10)
    p = g.state()[bob]
    assert abs(p[0]) < 0.08 and abs(p[2]) < 0.08


def mixed_chains(n_chains=8, links=8, seed=11):
    """Chains of boxes whose consecutive links are joined by Ball,
    Revolute, Prismatic and Fixed joints in turn (hinge / slide axes tilted
    at random), the first link of every other chain pinned to a static
    anchor, over a floor with static spheres."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = scenes.SceneBuilder()
    b.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=0.5, restitution=0.0)
    for i in range(12):
        b.sphere((rng.uniform(-4, 4), 0.3, rng.uniform(-4, 4)), 0.3, 0.0, friction=0.5, restitution=0.0)
    hy = 0.2
    for c in range(n_chains):
        x0, z0 = -3.5 + c, rng.uniform(-1.0, 1.0)
        prev = None
        if c % 2 == 0:
            prev = b.box((x0, 6.0 + hy, z0), (0.1, 0.1, 0.1), 0.0)
        for k in range(links):
            q = rng.normal(size=4) * np.array([0.1, 0.1, 0.1, 1.0])
            q = (q / np.linalg.norm(q)).astype(np.float32)
            h = b.box((x0, 6.0 - 2 * hy * k, z0), (0.08, hy, 0.08), float(rng.uniform(0.5, 2.0)),
                      quat=tuple(q.tolist()), friction=0.5, restitution=0.0)
            if prev is not None:
                kind = (c + k) % 4
                axis = rng.normal(size=3)
                axis = tuple((axis / np.linalg.norm(axis)).tolist())
                b.joint(kind, prev, h, (0.0, -hy, 0.0), (0.0, hy, 0.0), axis, axis)
            prev = h
    return b.build()


def test_mixed_joint_chains_bitwise(gpu_available):
    s = mixed_chains()
    kinds = [j[4] for j in s["joints"]]
    assert set(kinds) == {0, 1, 2, 3}
    g, o = _run(s, 150, every=10)
    # a link in the middle of a chain goes: its joints go, the rest renumber
    victim = int(s["joints"][len(s["joints"]) // 2][1])
    g.remove_body(victim)
    o.remove_body(victim)
    for s_ in range(60):
        g.step_n(1)
        o.step_inner(1)
    _same(g, o, g.njoints, "after a removal")
    assert g.njoints < len(s["joints"])
    assert np.isfinite(g.state()).all()


def test_joint_spec_validation(gpu_available):
    b = scenes.SceneBuilder()
    b.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    b.box((1.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    g = AvbdEngine(scene=b.build())
    from ornis_amd.avbd import AvbdError
    for args in ((0, 0, (0, 0, 0), (0, 0, 0), 1, (0, 0, 1), (0, 0, 1)),       # self joint
                 (0, 5, (0, 0, 0), (0, 0, 0), 0),                          # body out of range
                 (0, 1, (0, 0, 0), (0, 0, 0), 8),                          # Gear needs its own entry
                 (0, 1, (0, 0, 0), (0, 0, 0), 11),                         # unknown kind
                 (0, 1, (0, 0, 0), (0, 0, 0), 1, (0, 0, 0), (0, 0, 1)),    # degenerate hinge axis
                 (0, 1, (0, 0, 0), (0, 0, 0), 2, (0, 0, 1), (1e-7, 0, 0))):  # degenerate slide axis
        with pytest.raises(AvbdError):
            g.add_joint(*args)
    assert g.stats()["joints"] == 0


def test_reference_rod_rope_spring_scenes_bitwise(gpu_available):
    """The reference's rod pendulum, hanging rope and spring oscillators
    (both motor models, and a pure damper) over their whole runs."""
    sb = scenes.SceneBuilder()
    a = sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    b = sb.sphere((2.0, 0.0, 0.0), 0.25, 1.0)
    sb.joint(scenes.JOINT_DISTANCE, a, b, (0, 0, 0), (0, 0, 0))
    g, _ = _run(sb.build(), 300, every=20)
    assert abs(np.linalg.norm(g.state()[b, 0:3]) - 2.0) < 0.25
    sb = scenes.SceneBuilder()
    a = sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    b = sb.sphere((0.0, -1.0, 0.0), 0.2, 2.0)
    sb.rope(a, b, (0, 0, 0), (0, 1, 0), 2.0)
    g, _ = _run(sb.build(), 240, every=10)
    assert abs(np.linalg.norm(g.state()[b, 0:3] + np.float32([0, 1, 0])) - 2.0) < 0.2
    for model, kind, k in ((scenes.MOTOR_ACCELERATION_BASED, scenes.MOTOR_POSITION, 30.0),
                           (scenes.MOTOR_FORCE_BASED, scenes.MOTOR_POSITION, 30.0),
                           (scenes.MOTOR_FORCE_BASED, scenes.MOTOR_VELOCITY, 0.0)):
        sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
        a = sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
        b = sb.sphere((0.0, -2.5, 0.0), 0.2, 1.0, vel=(0.0, -1.0, 0.0))
        sb.spring(a, b, (0, 0, 0), (0, 0.5, 0), 1.0, k, 6.0, motor_kind=kind, model=model)
        g, _ = _run(sb.build(), 600 if k else 120, every=30)


def tether_field(seed=5, n=40):
    """Boxes and spheres hung from static anchors and from each other by
    ropes (several on one body, so the catch order matters), rods and
    springs (both motor models, one a damper), dropped over a floor with
    static spheres."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = scenes.SceneBuilder()
    b.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=0.5, restitution=0.0)
    for i in range(8):
        b.sphere((rng.uniform(-4, 4), 0.4, rng.uniform(-4, 4)), 0.4, 0.0)
    anchors = [b.box((x, 6.0, z), (0.1, 0.1, 0.1), 0.0) for x in (-3.0, 0.0, 3.0) for z in (-2.0, 2.0)]
    bodies = []
    for i in range(n):
        q = rng.normal(size=4)
        q = tuple(float(v) for v in (q / np.linalg.norm(q)).astype(np.float32))
        p = (float(rng.uniform(-4, 4)), float(rng.uniform(2.5, 5.5)), float(rng.uniform(-3, 3)))
        v = tuple(float(x) for x in rng.uniform(-3, 3, 3))
        h = b.box(p, tuple(rng.uniform(0.1, 0.3, 3)), float(rng.uniform(0.5, 3.0)), quat=q, vel=v) if i % 2 else \
            b.sphere(p, float(rng.uniform(0.1, 0.3)), float(rng.uniform(0.5, 3.0)), vel=v)
        bodies.append(h)
    for i, h in enumerate(bodies):
        other = anchors[i % len(anchors)] if i % 3 == 0 else bodies[(i * 7 + 3) % n]
        if other == h:
            other = anchors[0]
        la = tuple(float(x) for x in rng.uniform(-0.1, 0.1, 3))
        lb = tuple(float(x) for x in rng.uniform(-0.1, 0.1, 3))
        kind = i % 5
        if kind in (0, 1):
            b.rope(other, h, la, lb, float(rng.uniform(1.0, 3.0)))
        elif kind == 2:
            b.joint(scenes.JOINT_DISTANCE, other, h, la, lb)
        elif kind == 3:
            b.spring(other, h, la, lb, float(rng.uniform(0.5, 2.0)), float(rng.uniform(5, 60)),
                     float(rng.uniform(0.5, 5)), model=i % 2)
        else:
            b.spring(other, h, la, lb, 1.0, 0.0, 3.0, motor_kind=scenes.MOTOR_VELOCITY,
                     model=scenes.MOTOR_FORCE_BASED)
    # a second rope on three bodies that already hang from one
    for i in (0, 5, 10):
        b.rope(anchors[(i + 1) % len(anchors)], bodies[i], (0, 0, 0), (0, 0, 0), 4.0)
    return b.build()


def test_tether_field_bitwise(gpu_available):
    s = tether_field()
    kinds = {j[4] for j in s["joints"]}
    assert {scenes.JOINT_DISTANCE, scenes.JOINT_ROPE, scenes.JOINT_SPRING} <= kinds
    g, o = _run(s, 200, every=10)
    assert g.stats()["breakdowns"] == 0 and np.isfinite(g.state()).all()
    victim = int(s["joints"][7][1])
    g.remove_body(victim)
    o.remove_body(victim)
    for s_ in range(40):
        g.step_n(1)
        o.step_inner(1)
    _same(g, o, g.njoints, "after a removal")


def test_rope_spring_spec_validation(gpu_available):
    b = scenes.SceneBuilder()
    b.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    b.box((1.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    g = AvbdEngine(scene=b.build())
    from ornis_amd.avbd import AvbdError
    for m in (0.0, -1.0, float("inf")):
        with pytest.raises(AvbdError):
            g.add_joint(0, 1, (0, 0, 0), (0, 0, 0), scenes.JOINT_ROPE, prm=(m,))
    for prm in ((1.0, -1.0, 1.0, 1.0, 0.0, 1, 0), (1.0, 1.0, 1.0, -1.0, 0.0, 1, 0),
                (float("nan"), 1.0, 1.0, 1.0, 0.0, 1, 0), (1.0, 1.0, 1.0, 1.0, 0.0, 3, 0)):
        with pytest.raises(AvbdError):
            g.add_joint(0, 1, (0, 0, 0), (0, 0, 0), scenes.JOINT_SPRING, prm=prm)
    from ornis_amd.avbd import _f32p
    z = np.zeros(3, np.float32)
    zp = z.ctypes.data_as(_f32p)
    for kind in (scenes.JOINT_ROPE, scenes.JOINT_SPRING):  # only through their own entry points
        assert g._L.oavbd_add_joint(g._ctx, kind, 0, 1, zp, zp, None, None, None) != 0
    assert g.stats()["joints"] == 0


def test_reference_limit_and_motor_scenes_bitwise(gpu_available):
    """avbd_engine.rs:507-549 (hinge window), :552-595 (hinge motor, trigger
    anchor), :598-646 (slider resting on its lower limit), :782-815 (slider
    motor), and a slider hanging on an upper limit under 10 g."""
    R, P = scenes.JOINT_REVOLUTE, scenes.JOINT_PRISMATIC
    sb = scenes.SceneBuilder()
    sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    bob = sb.sphere((1.0, -1.0, 0.0), 0.25, 1.0)
    sb.joint(R, 0, bob, (0, 0, 0), (-1.0, 1.0, 0.0), (0, 0, 1), (0, 0, 1), limit=(-0.5, 0.5))
    _run(sb.build(), 300, every=20)
    sb = scenes.SceneBuilder()
    a = sb.sphere((0.0, 0.0, 0.0), 0.05, 0.0)
    bob = sb.sphere((0.0, 0.0, 0.0), 0.1, 1.0)
    sb.joint(R, a, bob, (0, 0, 0), (0, 0, 0), (0, 0, 1), (0, 0, 1), motor=(3.0, 20.0))
    sc = sb.build()
    sc["trigger"][a] = 1
    g, _ = _run(sc, 360, every=30)
    assert abs(g.state()[bob, 12] - 3.0) < 1.5
    sb = scenes.SceneBuilder()
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bob = sb.sphere((0.0, -2.9, 0.0), 0.25, 1.0)
    sb.joint(P, 0, bob, (0, 0, 0), (0, 1, 0), (0, 1, 0), (0, 1, 0), limit=(-2.0, 0.0))
    g, _ = _run(sb.build(), 300, every=20)
    assert -5.4 < g.state()[bob, 1] < -4.4
    sb = scenes.SceneBuilder()
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bob = sb.sphere((1.0, 0.0, 0.0), 0.25, 1.0)
    sb.joint(P, 0, bob, (0, 0, 0), (-1, 0, 0), (1, 0, 0), (1, 0, 0), motor=(2.0, 50.0))
    _run(sb.build(), 300, every=20)
    sb = scenes.SceneBuilder(gravity=(0.0, 98.1, 0.0))
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 0.0)
    bob = sb.sphere((0.0, 2.9, 0.0), 0.25, 1.0)
    sb.joint(P, 0, bob, (0, 0, 0), (0, -1, 0), (0, 1, 0), (0, 1, 0), limit=(-2.0, 2.0))
    _run(sb.build(), 400, every=20)


def limited_chains(n_chains=6, links=10, seed=23):
    """Chains of boxes joined by hinges and sliders with random tilted axes,
    travel windows and velocity motors (some with no budget, some turning
    the wrong way into their limit), the first link of every chain pinned
    to a static anchor, swinging into static spheres and a floor."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = scenes.SceneBuilder()
    b.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=0.5, restitution=0.0)
    for i in range(10):
        b.sphere((rng.uniform(-4, 4), 0.4, rng.uniform(-4, 4)), 0.4, 0.0)
    hy = 0.2
    for c in range(n_chains):
        x0, z0 = -3.0 + 1.2 * c, rng.uniform(-1.0, 1.0)
        prev = b.box((x0, 5.0 + hy, z0), (0.1, 0.1, 0.1), 0.0)
        for k in range(links):
            h = b.box((x0 + 0.3 * k, 5.0 - 2 * hy * k, z0), (0.08, hy, 0.08), float(rng.uniform(0.5, 2.0)),
                      friction=0.5, restitution=0.0)
            kind = scenes.JOINT_REVOLUTE if (c + k) % 2 else scenes.JOINT_PRISMATIC
            axis = rng.normal(size=3)
            axis = tuple((axis / np.linalg.norm(axis)).tolist())
            lo = float(rng.uniform(-0.6, -0.05))
            limit = (lo, lo + float(rng.uniform(0.1, 0.8))) if k % 3 != 2 else None
            motor = (float(rng.uniform(-3, 3)), float(rng.choice([0.0, 5.0, 40.0]))) if k % 2 == 0 else None
            b.joint(kind, prev, h, (0.0, -hy, 0.0), (0.0, hy, 0.0), axis, axis, limit=limit, motor=motor)
            prev = h
    return b.build()


def test_limited_motor_chains_bitwise(gpu_available):
    s = limited_chains()
    assert any(len(j) == 8 and j[7][4] for j in s["joints"]) and any(len(j) == 8 and j[7][5] for j in s["joints"])
    g, o = _run(s, 240, every=10)
    assert np.isfinite(g.state()).all() and g.stats()["breakdowns"] == 0


def test_limit_motor_spec_validation(gpu_available):
    b = scenes.SceneBuilder()
    b.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    b.box((1.0, 0.0, 0.0), (0.5, 0.5, 0.5), 1.0)
    g = AvbdEngine(scene=b.build())
    from ornis_amd.avbd import AvbdError
    for prm in ((1.0, -1.0, 0, 0, 1, 0), (float("nan"), 1.0, 0, 0, 1, 0), (0, 0, 1.0, -1.0, 0, 1),
                (0, 0, float("inf"), 1.0, 0, 1)):
        with pytest.raises(AvbdError):
            g.add_joint(0, 1, (0, 0, 0), (0, 0, 0), scenes.JOINT_REVOLUTE, (0, 0, 1), (0, 0, 1), prm=prm)
    assert g.stats()["joints"] == 0


def test_motor_joint_scenes_bitwise(gpu_available):
    """Free Motor joints (joints_ext.rs:368-421): a box driven along x
    against floor friction (motor_joint.rs:90-110), a budget below the
    friction, a free ball spun up about z, and drives between pairs of
    dynamic bodies tumbling onto a floor."""
    for lin, ang, fmax, tmax in (((2.0, 0.0, 0.0), (0.0, 0.0, 0.0), 8.0, 10.0),
                                 ((2.0, 0.0, 0.0), (0.0, 0.0, 0.0), 3.0, 0.0)):
        sb = scenes.SceneBuilder()
        sb.box((0.0, -0.5, 0.0), (10.0, 0.5, 10.0), 0.0)
        block = sb.box((0.0, 0.5, 0.0), (0.5, 0.5, 0.5), 1.0)
        anchor = sb.sphere((0.0, 5.0, 0.0), 0.1, 0.0)
        sb.motor(anchor, block, lin, ang, fmax, tmax)
        g, _ = _run(sb.build(), 300, every=20)
        if fmax > 5.0:
            assert abs(g.state()[block, 7] - 2.0) < 0.4
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    a = sb.sphere((0.0, 0.0, 0.0), 0.1, 0.0)
    ball = sb.sphere((0.0, 2.0, 0.0), 0.3, 1.0)
    sb.motor(a, ball, (0.0, 0.0, 0.0), (0.0, 0.0, 3.0), 0.0, 10.0)
    _run(sb.build(), 120, every=10)
    rng = np.random.Generator(np.random.PCG64(4))
    sb = scenes.SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    for i in range(12):
        q = rng.normal(size=4)
        q = tuple(float(v) for v in (q / np.linalg.norm(q)).astype(np.float32))
        p1 = (float(rng.uniform(-3, 3)), 1.0 + 0.8 * i, float(rng.uniform(-3, 3)))
        h1 = sb.box(p1, tuple(rng.uniform(0.15, 0.35, 3)), float(rng.uniform(0.5, 2)), quat=q)
        h2 = sb.sphere((p1[0] + 0.9, p1[1], p1[2]), float(rng.uniform(0.15, 0.3)), float(rng.uniform(0.5, 2)))
        sb.motor(h1, h2, tuple(rng.uniform(-2, 2, 3)), tuple(rng.uniform(-4, 4, 3)), float(rng.uniform(0, 20)),
                 float(rng.uniform(0, 5)))
    g, o = _run(sb.build(), 200, every=10)
    assert np.isfinite(g.state()).all()
    with pytest.raises(Exception):
        AvbdEngine(scene=sb.build()).add_joint(1, 2, (0, 0, 0), (0, 0, 0), 10,
                                               prm=(float("nan"), 0, 0, 0, 0, 0, 1.0, 1.0))


def test_wheel_scenes_bitwise(gpu_available):
    """Wheel joints: the reference's sprung chassis and axle motor
    (avbd_engine.rs:966-1060), a hanging sprung wheel and a rigid one, an
    oblique axle, and a four-wheeled cart with driven rear axles rolling on
    a floor into static spheres."""
    W = scenes.JOINT_WHEEL
    sb = scenes.SceneBuilder()
    chassis = sb.box((0.0, 1.0, 0.0), (0.5, 0.2, 0.3), 2.0)
    wheel = sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.15), 1.0)
    sb.wheel(chassis, wheel, (0, 0, 0), (0, 0, 0), (0, 1, 0), (0, 1, 0), (0, 0, 1), (0, 0, 1), 3.0, 0.7)
    _run(sb.build(), 300, every=20)
    for hz, motor, axle in ((2.0, (6.0, 50.0), (0, 0, 1)), (0.0, None, (0, 0, 1)), (2.0, (4.0, 50.0), (0, 1, 1))):
        sb = scenes.SceneBuilder(gravity=(0.0, -9.81 if motor is None else 0.0, 0.0))
        anchor = sb.box((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 0.0)
        wheel = sb.box((0.0, -0.5, 0.0), (0.2, 0.2, 0.1), 1.0)
        sb.wheel(anchor, wheel, (0, 0, 0), (0, 0, 0), (0, 1, 0), (0, 1, 0), axle, axle, hz, 0.5, motor=motor)
        _run(sb.build(), 240, every=20)
    sb = scenes.SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=0.9, restitution=0.0)
    for k in range(4):
        sb.sphere((3.0 + 1.5 * k, 0.3, -1.0 + 0.7 * k), 0.3, 0.0)
    body = sb.box((0.0, 0.9, 0.0), (1.0, 0.15, 0.5), 4.0)
    for i, (x, z) in enumerate(((-0.8, -0.6), (-0.8, 0.6), (0.8, -0.6), (0.8, 0.6))):
        w = sb.sphere((x, 0.35, z), 0.3, 0.8, friction=0.9, restitution=0.0)
        sb.wheel(body, w, (x, -0.15, z), (0, 0, 0), (0, 1, 0), (0, 1, 0), (0, 0, 1), (0, 0, 1), 4.0, 0.6,
                 motor=(-6.0, 20.0) if i < 2 else None)
    g, o = _run(sb.build(), 240, every=10)
    assert np.isfinite(g.state()).all() and g.stats()["breakdowns"] == 0
    from ornis_amd.avbd import AvbdError
    with pytest.raises(AvbdError):
        g.add_joint(body, 0, (0, 0, 0), (0, 0, 0), W, (0, 1, 0), (0, 1, 0),
                    prm=(0, 0, 1, 0, 0, 1, -1.0, 0.5, 0.0, 0.0, 0.0))


def test_sixdof_scenes_bitwise(gpu_available):
    """SixDof joints: the reference's angular window against an 8 rad/s spin
    (avbd_engine.rs:1514-1546), a locked weld dropped on a floor, a free one
    whose boxes keep colliding, and chains of links mixing free, locked and
    limited linear and angular axes tumbling onto spheres."""
    F_, L_, M_ = scenes.AXIS_FREE, scenes.AXIS_LOCKED, scenes.AXIS_LIMITED
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    a = sb.box((0.0, 0.0, 0.0), (0.5, 0.5, 0.5), 0.0)
    b = sb.box((0.0, 0.0, 0.0), (0.4, 0.4, 0.4), 1.0, angvel=(0.0, 0.0, 8.0))
    sb.sixdof(a, b, (0, 0, 0), (0, 0, 0), [L_] * 5 + [M_], (0,) * 5 + (-0.2,), (0,) * 5 + (0.2,))
    _run(sb.build(), 120, every=10)
    for kinds in ([L_] * 6, [F_] * 6):
        sb = scenes.SceneBuilder()
        sb.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0)
        a = sb.box((-0.5, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0)
        c = sb.box((0.45, 3.0, 0.0), (0.5, 0.5, 0.5), 1.0, vel=(0.0, 0.0, 1.0))
        sb.sixdof(a, c, (0.5, 0, 0), (-0.5, 0, 0), kinds)
        _run(sb.build(), 200, every=10)
    rng = np.random.Generator(np.random.PCG64(9))
    sb = scenes.SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    for i in range(6):
        sb.sphere((rng.uniform(-3, 3), 0.4, rng.uniform(-2, 2)), 0.4, 0.0)
    for c in range(4):
        prev = sb.box((-2.0 + 1.3 * c, 4.5, 0.0), (0.1, 0.1, 0.1), 0.0)
        for k in range(8):
            h = sb.box((-2.0 + 1.3 * c + 0.25 * k, 4.2 - 0.4 * k, 0.0), (0.08, 0.18, 0.08), float(rng.uniform(0.5, 2)))
            # ragdoll-like: locked linear axes (one limited on every third
            # link), random angular axes; fully random configurations of
            # these chains blow up in the oracle and on the GPU alike (the
            # two stay bitwise equal until NaNs appear)
            kinds = [L_, L_, L_] + [int(x) for x in rng.integers(0, 3, 3)]
            if k % 3 == 1:
                kinds[int(rng.integers(0, 3))] = M_
            lo = tuple(float(x) for x in rng.uniform(-0.4, -0.05, 6))
            hi = tuple(float(x) for x in rng.uniform(0.05, 0.4, 6))
            sb.sixdof(prev, h, (0.0, -0.18, 0.0), (0.0, 0.18, 0.0), kinds, lo, hi)
            prev = h
    g, o = _run(sb.build(), 200, every=10)
    assert np.isfinite(g.state()).all()


def four_member_gears(n=6, seed=31):
    """Pairs of mechanisms under gravity, released from horizontal: a
    two-link pendulum (ground hinge, then a hinge between the links) and an
    arm on a ground hinge carrying a slider on a limited slide.  The gear
    couples the links' hinge to the slide, so its row reaches four distinct
    dynamic bodies (its colouring links all of them)."""
    rng = np.random.default_rng(seed)
    b = scenes.SceneBuilder()
    g = b.box((0.0, -1.0, 0.0), (30.0, 0.5, 30.0), 0.0)
    R, P = scenes.JOINT_REVOLUTE, scenes.JOINT_PRISMATIC
    for i in range(n):
        x = -12.0 + 4.0 * i
        l1 = b.box((x + 0.5, 5.5, 0.0), (0.5, 0.1, 0.1), 1.0)
        l2 = b.box((x + 1.5, 5.5, 0.0), (0.5, 0.1, 0.1), 1.0)
        b.joint(R, g, l1, (x, 5.5, 0.0), (-0.5, 0, 0), (0, 0, 1), (0, 0, 1))
        j1 = b.joint(R, l1, l2, (0.5, 0, 0), (-0.5, 0, 0), (0, 0, 1), (0, 0, 1))
        arm = b.box((x + 0.5, 3.0, 0.6), (0.5, 0.1, 0.1), 1.0)
        slider = b.sphere((x + 1.0, 3.0, 0.6), 0.15, 1.0)
        b.joint(R, g, arm, (x, 3.0, 0.6), (-0.5, 0, 0), (0, 0, 1), (0, 0, 1))
        j3 = b.joint(P, arm, slider, (0.5, 0, 0), (0, 0, 0), (1, 0, 0), (1, 0, 0), limit=(-0.4, 0.4))
        b.gear(l2, slider, j1, j3, float(rng.uniform(-2.0, 2.0)))
    return b.build()


def test_gear_scenes_bitwise(gpu_available):
    """The reference's gear scenes (avbd_engine.rs:1360-1450, :824-890,
    without the host teleport) and pendulum pairs whose gears reach four
    bodies: every pose, pair, colour and joint dual bit-identical."""
    import test_oracle_gear as og
    sb, _, _ = og.hinge_pair_scene()
    _run(sb.build(), 180, every=20)
    sb, _ = og.slider_pair_scene()
    _run(sb.build(), 240, every=20)
    print("gear: four-member pendulum pairs", flush=True)
    _run(four_member_gears(), 240, every=20)


def test_gear_spec_validation(gpu_available):
    from ornis_amd.avbd import AvbdError
    b = scenes.SceneBuilder()
    g = b.box((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 0.0)
    a = b.box((0.0, 1.0, 0.0), (0.1, 0.1, 0.1), 1.0)
    c = b.box((1.0, 1.0, 0.0), (0.1, 0.1, 0.1), 1.0)
    e = AvbdEngine(scene=b.build())
    ja = e.add_joint(g, a, (0, 0, 0), (0, -1, 0), scenes.JOINT_REVOLUTE, (0, 0, 1), (0, 0, 1))
    jball = e.add_joint(g, c, (1, 0, 0), (0, -1, 0), 0)
    z = (0.0, 0.0, 0.0)
    for prm in ((ja, 7, 1.0), (ja, ja, float("nan")), (ja, jball, 1.0)):
        with pytest.raises(AvbdError):
            e.add_joint(a, c, z, z, scenes.JOINT_GEAR, prm=prm)
    assert e.add_joint(a, c, z, z, scenes.JOINT_GEAR, prm=(ja, ja, 1.0)) == 2
