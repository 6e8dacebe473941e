"""GPU parity: the HIP step (through the C ABI) against the CPU oracle run
in the same colour order on the same inputs.  Bodies, pair lists and
colourings are compared bit for bit."""
import ctypes
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes
from ornis_amd.avbd import AvbdEngine
from scenes_ref import read_golden, snapshot_scene

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "avbd_snapshot.hex")


def run_pair(scene, steps, iters=10, check_every=0, scene_ids=None):
    g = AvbdEngine(iters=iters)
    g.load(scene, scene_ids=scene_ids)
    o = po.OracleEngine(scene, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE, iters=iters)
    for s in range(steps):
        g.step_n(1)
        o.step_inner(1)
        if (check_every and s % check_every == 0) or s == steps - 1:
            assert_same(g, o, s)
    return g, o


def assert_same(g, o, step):
    pg, po_ = g.pairs(), o.pairs()[0]
    assert pg.shape == po_.shape and np.array_equal(pg, po_), f"pair list differs at step {step}"
    cg, co = g.colours(), o.colours()
    assert np.array_equal(cg, co), f"colouring differs at step {step}"
    tg, ttg = g.toi()
    to, tto = o.toi()
    assert np.array_equal(tg.view(np.uint32), to.view(np.uint32)) and np.array_equal(ttg, tto), \
        f"TOI clamps differ at step {step}"
    sg, so = g.state(), o.state()
    bad = np.where((sg.view(np.uint32) != so.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"step {step}: {bad.size} bodies differ, first {bad[:5]}, " \
                          f"max |d| {np.abs(sg - so).max():.3e}"


def test_snapshot_scene_bitwise(gpu_available):
    g, o = run_pair(snapshot_scene(), 120, check_every=10)
    # the colour-ordered step stays close to the reference's own golden
    _, want = read_golden(GOLDEN)
    assert np.max(np.abs(g.state()[:, :7] - want[:, :7])) < 1e-3
    assert g.stats()["breakdowns"] == 0


def test_pyramid_bitwise(gpu_available):
    """The 210-cube pyramid over its whole 300-step trajectory."""
    run_pair(scenes.box_pyramid(), 300, check_every=20)


def test_random_box_drop_bitwise(gpu_available):
    s = scenes.box_drop_scene(200, np.random.Generator(np.random.PCG64(7)))
    run_pair(s, 150, check_every=10)


def test_mixed_shapes_bitwise(gpu_available):
    rng = np.random.Generator(np.random.PCG64(21))
    sb = scenes.SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    for i in range(60):
        p = (rng.uniform(-3, 3), 0.6 + 0.9 * i / 6.0, rng.uniform(-3, 3))
        if i % 3 == 0:
            sb.sphere(p, rng.uniform(0.2, 0.5), 1.0 + i % 5)
        else:
            sb.box(p, tuple(rng.uniform(0.2, 0.5, 3)), 2.0, quat=tuple(
                (lambda q: q / np.linalg.norm(q))(rng.normal(size=4)).astype(np.float32).tolist()))
    sb.box((0.0, 1.0, 6.0), (4.0, 1.0, 0.3), 0.0)
    sb.sphere((2.0, 0.3, -5.0), 0.3, 0.0)
    run_pair(sb.build(), 150, check_every=10)


def test_ball_joint_chains_bitwise(gpu_available):
    s = scenes.chains(n_chains=6, links=8, grid=3, n_spheres=20)
    g, o = run_pair(s, 120, iters=12, check_every=10)
    assert np.array_equal(g.joints().view(np.uint32), o.joints(len(s["joints"])).view(np.uint32))


def test_sleep_and_fast_impact_bitwise(gpu_available):
    """Settling boxes fall asleep; a fast box triggers the TOI clamp and
    wakes nothing it does not hit."""
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (8.0, 1.0, 8.0), 0.0)
    for i in range(5):
        sb.box((-3.0 + 1.5 * i, 0.5, 0.0), (0.5, 0.5, 0.5), 1.0)
    sb.box((2.0, 6.0, 3.0), (0.3, 0.3, 0.3), 1.0, vel=(0.0, -40.0, 0.0))
    g, o = run_pair(sb.build(), 200, check_every=10)
    a_g, t_g = g.sleep()
    a_o, t_o = o.sleep()
    assert np.array_equal(a_g, a_o) and np.array_equal(t_g.view(np.uint32), t_o.view(np.uint32))
    assert a_g.sum() >= 3


def test_fast_rain_toi_bitwise(gpu_available):
    """Fast rotated boxes and spheres shot at a box floor and at each other:
    many box-box casts, clamps compared every step."""
    rng = np.random.Generator(np.random.PCG64(31))
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (8.0, 1.0, 8.0), 0.0)
    for i in range(160):
        p = (rng.uniform(-6, 6), rng.uniform(0.8, 5.0), rng.uniform(-6, 6))
        v = (rng.uniform(-8, 8), -rng.uniform(20, 70), rng.uniform(-8, 8))
        if i % 5 == 0:
            sb.sphere(p, rng.uniform(0.15, 0.4), 1.0, vel=v)
        else:
            q = rng.normal(size=4)
            sb.box(p, tuple(rng.uniform(0.15, 0.5, 3)), 1.0, vel=v,
                   quat=tuple((q / np.linalg.norm(q)).astype(np.float32).tolist()))
    g, o = run_pair(sb.build(), 40, check_every=1)
    assert g.stats()["toi_hits"] >= 0


def test_unnormalised_rotations_bitwise(gpu_available):
    """Rotations given off unit length (static boxes keep them; dynamic ones
    are renormalised after their first step): the box-box feature scans take
    no shortcut for them, and the results stay those of the full scan."""
    rng = np.random.Generator(np.random.PCG64(45))
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (8.0, 1.0, 8.0), 0.0)
    for i in range(6):
        q = rng.normal(size=4)
        q = q / np.linalg.norm(q) * (0.95 if i % 2 else 1.05)
        sb.box((-4.0 + 1.6 * i, 0.6, 0.0), (0.6, 0.4, 0.5), 0.0, quat=tuple(q.astype(np.float32).tolist()))
    for i in range(40):
        q = rng.normal(size=4)
        q = q / np.linalg.norm(q) * rng.uniform(0.97, 1.03)
        p = (rng.uniform(-4.5, 4.5), rng.uniform(1.5, 6.0), rng.uniform(-1.0, 1.0))
        sb.box(p, tuple(rng.uniform(0.2, 0.5, 3)), 1.0, quat=tuple(q.astype(np.float32).tolist()),
               vel=(0.0, -rng.uniform(0.0, 30.0), 0.0))
    g, _ = run_pair(sb.build(), 80, check_every=5)
    assert g.stats()["pairs"] >= 20


def test_batched_scenes_equal_independent_runs(gpu_available):
    """Scenes stepped together in one context equal each scene stepped
    alone by the oracle."""
    parts, allsc = scenes.batched_drops(n_scenes=6, nmin=8, nmax=64, seed=99)
    g = AvbdEngine()
    g.load(allsc, scene_ids=allsc["scene_id"])
    orcs = [po.OracleEngine(p, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE) for p in parts]
    for _ in range(60):
        g.step_n(1)
        for o in orcs:
            o.step_inner(1)
    sg = g.state()
    off = 0
    for o in orcs:
        so = o.state()
        assert np.array_equal(sg[off:off + len(so)].view(np.uint32), so.view(np.uint32))
        off += len(so)


def test_large_context_equals_small_bitwise(gpu_available):
    """Over 2^19 bodies in one context, so the hash grid's bucket scan spans
    more than one round of tile offsets: the first scenes step exactly as in
    a context of their own."""
    parts, allsc = scenes.batched_drops(n_scenes=4700, nmin=8, nmax=512, seed=5)
    assert len(allsc["shape"]) > 2 ** 19
    few = scenes.concat(parts[:4])
    few_ids = np.concatenate([np.full(len(p["shape"]), i, np.uint32) for i, p in enumerate(parts[:4])])
    big, small = AvbdEngine(), AvbdEngine()
    try:
        big.load(allsc, scene_ids=allsc["scene_id"])
        small.load(few, scene_ids=few_ids)
        for _ in range(30):
            big.step_n(1)
            small.step_n(1)
        k = len(few["shape"])
        assert big.stats()["pairs"] > 0
        assert np.array_equal(big.state()[:k].view(np.uint32), small.state().view(np.uint32))
    finally:
        big.close()
        small.close()


def test_gpu_run_to_run_bitwise(gpu_available):
    s = scenes.box_drop_scene(300, np.random.Generator(np.random.PCG64(13)))
    outs = []
    for _ in range(2):
        g = AvbdEngine(scene=s)
        g.step_n(60)
        outs.append(g.state())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("shape", [(16, 1), (16, 4), (16, 8)])
def test_launch_shape_does_not_change_bits(gpu_available, shape):
    """The primal launch shape (wavefronts per workgroup) is a pure speed
    knob."""
    s = scenes.box_drop_scene(150, np.random.Generator(np.random.PCG64(17)))
    ref = AvbdEngine(scene=s)
    ref.set_primal_tail(0)
    ref.step_n(40)
    g = AvbdEngine(scene=s)
    g.set_primal_tail(0)
    g.set_primal_shape(*shape)
    g.step_n(40)
    assert np.array_equal(ref.state().view(np.uint32), g.state().view(np.uint32))


def _plate_scene():
    """A heavy plate on the floor carrying a 6 x 6 grid of cubes: the plate
    solves 185 contact point-sides (more than the LDS variant stages), every
    body with rolling and torsion friction."""
    sb = scenes.SceneBuilder()
    sb.box((0.0, -0.5, 0.0), (10.0, 0.5, 10.0), 0.0, friction=0.6, restitution=0.0)
    sb.box((0.0, 0.2, 0.0), (3.0, 0.2, 3.0), 50.0, friction=0.6, restitution=0.0)
    for i in range(36):
        sb.box((-2.5 + (i % 6), 0.7, -2.5 + (i // 6)), (0.3, 0.3, 0.3), 1.0, friction=0.5, restitution=0.0)
    s = sb.build()
    s["rolling"][:] = 0.02
    s["torsion"][:] = 0.01
    return s


@pytest.mark.parametrize("variant", range(10))
def test_primal_variant_does_not_change_bits(gpu_available, variant):
    """Every primal launch variant (sixteen lanes per body at 1-8
    wavefronts per workgroup, one body per wavefront with its records staged
    in LDS, or sixteen lanes with the next item's records prefetched) gives
    the oracle's bits: a random drop, and the plate
    scene whose plate has more items than the LDS holds, with rolling and
    torsion rows."""
    for s in (scenes.box_drop_scene(300, np.random.Generator(np.random.PCG64(19))), _plate_scene()):
        g = AvbdEngine(scene=s)
        g.set_primal_tail(0)  # every class its own launch, of the variant under test
        g.set_primal_table([variant] * 24)
        o = po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
        g.step_n(60)
        o.step_inner(60)
        assert np.array_equal(g.state().view(np.uint32), o.state().view(np.uint32))
        assert np.array_equal(g.points().view(np.uint32), o.points().view(np.uint32))


@pytest.mark.parametrize("tail", [(0, 0), (40, 0), (4096, 0), (1 << 30, 0), (1 << 30, 8)])
def test_fused_tail_does_not_change_bits(gpu_available, tail):
    """The last colour classes fused into one launch, where a body waits for
    its own lower-class partners only, give the oracle's colour-by-colour
    bits: no fusion (the default), a small tail, 4,096 bodies, the whole
    sweep in one launch, and the whole sweep on eight wavefronts (each
    solving many batches in list order); a random drop, the plate scene (rolling and torsion
    rows), and jointed chains."""
    rows = []
    from test_gpu_joints import mixed_chains
    for s in (scenes.box_drop_scene(600, np.random.Generator(np.random.PCG64(29))), _plate_scene(),
              mixed_chains(n_chains=6, links=8, seed=5)):
        g = AvbdEngine(scene=s)
        g.set_primal_tail(*tail)
        o = po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
        fused = 0
        for k in range(6):
            g.step_n(10)
            o.step_inner(10)
            fused = max(fused, g.stats()["tail_colours"])
            assert np.array_equal(g.state().view(np.uint32), o.state().view(np.uint32)), f"step {10 * k + 9}"
        assert np.array_equal(g.points().view(np.uint32), o.points().view(np.uint32))
        if s.get("joints") is not None and len(s["joints"]):
            nj = len(s["joints"])
            assert np.array_equal(g.joints().view(np.uint32), o.joints(nj).view(np.uint32))
        rows.append(fused)
    if tail[0] == 0:
        assert max(rows) == 0
    else:
        assert max(rows) >= 2


def test_autotune_restores_state_and_keeps_bits(gpu_available):
    """oavbd_autotune leaves the trajectory untouched: a tuned context steps
    to the same bits as an untuned one, and its table names real variants."""
    s = scenes.box_drop_scene(400, np.random.Generator(np.random.PCG64(23)))
    ref = AvbdEngine(scene=s)
    ref.step_n(50)
    g = AvbdEngine(scene=s)
    g.step_n(30)
    before = (g.state(), g.points(), g.sleep()[1])
    table = g.autotune(3)
    after = (g.state(), g.points(), g.sleep()[1])
    for x, y in zip(before, after):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert ((table >= 0) & (table < 10)).all()
    assert np.array_equal(g.primal_table(), table)
    g.step_n(20)
    assert np.array_equal(ref.state().view(np.uint32), g.state().view(np.uint32))


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def test_accumulator_host_dt_bitwise(gpu_available):
    """oavbd_step(dt) is PhysicsEngine::step (mod.rs:1212-1235): 120 Hz host
    frames (a substep every second call), 1/30 s hitches (two substeps), a
    1 s stall (capped at four), ignored non-positive and non-finite dts, and
    odd fractions; bodies, pairs and TOI clamps equal the oracle's."""
    s = scenes.box_drop_scene(60, _rng(41))
    g = AvbdEngine(scene=s)
    o = po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    dts = [1.0 / 120.0] * 40 + [1.0 / 30.0] * 10 + [1.0, 0.0, -1.0, float("nan"), float("inf")] + \
        [1.0 / 60.0] * 5 + [0.3 / 60.0] * 25 + [2.5 / 60.0] * 6
    for i, dt in enumerate(dts):
        g.step(dt)
        o.step(dt, 1)
        if i % 10 == 0 or i == len(dts) - 1:
            assert_same(g, o, i)


def test_rolling_torsion_friction_bitwise(gpu_available):
    """Rolling and torsion friction rows (assembly.rs:209-250) and their
    duals (joints_ext.rs:549-571): spheres rolling and boxes spinning on a
    floor, with per-body rolling / torsion coefficients."""
    rng = _rng(52)
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (12.0, 1.0, 12.0), 0.0)
    for i in range(24):
        p = (rng.uniform(-8, 8), rng.uniform(0.4, 2.0), rng.uniform(-8, 8))
        if i % 2 == 0:
            sb.sphere(p, rng.uniform(0.2, 0.5), 1.0 + i % 3, vel=(rng.uniform(-3, 3), 0.0, rng.uniform(-3, 3)),
                      angvel=tuple(rng.uniform(-6, 6, 3)))
        else:
            q = rng.normal(size=4)
            sb.box(p, tuple(rng.uniform(0.2, 0.5, 3)), 2.0, quat=tuple((q / np.linalg.norm(q)).astype(np.float32)),
                   angvel=(0.0, rng.uniform(-8, 8), 0.0))
    s = sb.build()
    n = len(s["shape"])
    s["rolling"] = rng.uniform(0.0, 0.1, n).astype(np.float32)
    s["torsion"] = rng.uniform(0.0, 0.05, n).astype(np.float32)
    s["rolling"][::5] = 0.0
    g, o = run_pair(s, 150, check_every=10)
    pg = g.points()
    po_ = o.points()
    assert np.array_equal(pg[:, 12:15].view(np.uint32), po_[:, 12:15].view(np.uint32))
    assert np.abs(pg[:, 12:15]).max() > 0.0


def test_kinematic_pusher_bitwise(gpu_available):
    """A kinematic box teleported by its driver every step at 2 m/s pushes a
    row of boxes (the driver rewind, mod.rs:993-1005: its step-start pose is
    its previous pose, so contact rows see the motion within the step)."""
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (12.0, 1.0, 4.0), 0.0)
    for i in range(5):
        sb.box((-1.0 + 1.05 * i, 0.5, 0.0), (0.5, 0.5, 0.5), 1.0)
    k = sb.box((-3.0, 0.6, 0.0), (0.5, 0.6, 1.0), 1.0, body_type=2)
    s = sb.build()
    g = AvbdEngine(scene=s)
    o = po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    L = o._L
    L.orc_set_pose.argtypes = [ctypes.c_void_p, ctypes.c_int, po._f32p, po._f32p]
    x = np.float32(-3.0)
    for step in range(150):
        x = np.float32(x + np.float32(2.0 / 60.0))
        st = g.state()
        st[k, 0] = x
        g.set_state(st)
        pos = np.ascontiguousarray(st[k, 0:3])
        quat = np.ascontiguousarray(st[k, 3:7])
        L.orc_set_pose(o.h, k, pos.ctypes.data_as(po._f32p), quat.ctypes.data_as(po._f32p))
        g.step_n(1)
        o.step_inner(1)
        if step % 10 == 0 or step == 149:
            assert_same(g, o, step)
    assert g.state()[5, 0] > 4.5  # the row was pushed


def test_layer_mask_trigger_filters_bitwise(gpu_available):
    """pair_allowed layer/mask filtering and trigger bodies (rows.rs:366-368,
    :511-523): boxes on three layers with masks that exclude one another,
    and a trigger sphere that forms no contact."""
    rng = _rng(63)
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (8.0, 1.0, 8.0), 0.0)
    for i in range(60):
        p = (rng.uniform(-3, 3), 0.6 + 0.5 * i / 6.0, rng.uniform(-3, 3))
        sb.box(p, tuple(rng.uniform(0.2, 0.45, 3)), 1.0)
    trig = sb.sphere((0.0, 1.0, 0.0), 1.5, 0.0)
    s = sb.build()
    n = len(s["shape"])
    s["layer"] = np.array([1] + [1 << (i % 3) for i in range(60)] + [1], np.uint32)
    s["mask"] = np.array([0xFFFFFFFF] + [0xFFFFFFFF if i % 3 == 0 else (1 | (1 << (i % 3))) for i in range(60)] +
                         [0xFFFFFFFF], np.uint32)
    s["trigger"] = np.zeros(n, np.uint8)
    s["trigger"][trig] = 1
    g, o = run_pair(s, 150, check_every=10)
    ab = g.pairs()
    assert not ((ab == trig).any(axis=1)).any()
    lay, msk = s["layer"], s["mask"]
    assert ((lay[ab[:, 0]] & msk[ab[:, 1]]) != 0).all() and ((lay[ab[:, 1]] & msk[ab[:, 0]]) != 0).all()


def test_thin_wall_and_bounced_retouch_bitwise(gpu_available):
    """avbd_fast_box_stops_at_thin_wall (avbd_engine.rs:1290-1313) and
    avbd_bounced_retouch_catches_at_offset (:1770-1808) on the GPU, bit for
    bit against the oracle, with the reference's behavioural bounds."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.box((5.0, 0.0, 0.0), (0.02, 2.0, 2.0), 0.0)
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 1.0, vel=(10.0, 0.0, 0.0), restitution=0.0)
    g, _ = run_pair(sb.build(), 120, check_every=10)
    assert 4.0 < g.state()[1, 0] < 5.0
    for x0 in (-3.0, 0.0):
        sb = scenes.SceneBuilder()
        sb.box((0.0, -1.0, 0.0), (10.0, 1.0, 10.0), 0.0)
        sb.box((x0, 8.0, 0.0), (0.4, 0.4, 0.4), 1.0, vel=(0.0, -30.0, 0.0))
        g = AvbdEngine(scene=sb.build())
        o = po.OracleEngine(sb.build(), order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
        min_y = np.inf
        for step in range(300):
            g.step_n(1)
            o.step_inner(1)
            min_y = min(min_y, g.state()[1, 1])
            if step % 20 == 0:
                assert_same(g, o, step)
        assert_same(g, o, 299)
        b = g.state()[1]
        assert min_y > 0.0 and abs(b[1] - 0.4) < 0.05 and abs(b[0] - x0) < 0.15 and np.linalg.norm(b[7:10]) < 0.5


def test_anisotropic_friction_bitwise(gpu_available):
    """Anisotropic friction axes (friction_frame, rows.rs:32-52): boxes with
    random body-frame axes and transverse frictions, randomly turned, slide
    and tumble over a floor and over each other; the friction frame follows
    the live orientations in every solve and dual update.  Bodies, pairs and
    TOI clamps equal the oracle's bit for bit; the reference's channel test
    (avbd_engine.rs:466-495) holds on the GPU; a zero-length axis is
    rejected."""
    rng = _rng(77)
    sb = scenes.SceneBuilder()
    sb.box((0.0, -1.0, 0.0), (12.0, 1.0, 12.0), 0.0, friction=0.2)
    n = 48
    for i in range(n):
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        sb.box((rng.uniform(-6, 6), 0.6 + 1.1 * (i // 16), rng.uniform(-6, 6)), tuple(rng.uniform(0.2, 0.5, 3)),
               2.0, friction=float(rng.uniform(0.05, 0.3)), quat=tuple(float(x) for x in q),
               vel=(float(rng.uniform(-4, 4)), 0.0, float(rng.uniform(-4, 4))))
    s = sb.build()
    axis = np.zeros((n + 1, 3), np.float32)
    trans = s["friction"].copy()
    for i in range(1, n + 1):
        if i % 4:
            axis[i] = rng.normal(size=3) * rng.uniform(0.5, 2.0)
            trans[i] = rng.uniform(0.5, 1.2)
    g = AvbdEngine(scene=s)
    g.set_friction_axes(axis, trans)
    o = po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    o.set_friction_axes(axis, trans)
    for step in range(120):
        g.step_n(1)
        o.step_inner(1)
        if step % 10 == 0 or step == 119:
            assert_same(g, o, step)

    def channel(direction):
        sb = scenes.SceneBuilder()
        sb.box((0.0, -1.0, 0.0), (5.0, 1.0, 5.0), 0.0, friction=0.1)
        d = np.asarray(direction, np.float32)
        sb.box((0.0, 0.5, 0.0), (0.5, 0.5, 0.5), 1.0, friction=0.1, vel=tuple(5.0 * d))
        e = AvbdEngine(scene=sb.build())
        e.set_friction_axes(np.array([[0, 0, 0], [1, 0, 0]], np.float32), np.array([0.1, 1.0], np.float32))
        e.step_n(10)
        return 5.0 - float(e.state()[1, 7:10] @ d)

    assert channel((0, 0, 1)) > 4.0 * channel((1, 0, 0))
    with pytest.raises(Exception):
        g.set_friction_axes(np.array([[0, 0, 0], [1e-9, 0, 0]], np.float32), np.array([0.1, 0.1], np.float32))


def _capsule_scene(seed):
    """Capsules of every proportion (squat to long, radius 0.08-0.4) with
    boxes and spheres, randomly turned, dropped on a half-space floor beside
    a static box ledge and a static capsule rail; four fast capsules thrown
    at a thin wall take the TOI clamp."""
    rng = _rng(seed)
    sb = scenes.SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    sb.box((-4.0, 0.5, 0.0), (1.5, 0.5, 3.0), 0.0)
    sb.capsule((3.0, 0.4, 0.0), 0.3, 2.5, 0.0, quat=(0.0, 0.0, float(np.sin(np.pi / 4)), float(np.cos(np.pi / 4))))
    sb.box((9.0, 2.0, 0.0), (0.03, 2.0, 3.0), 0.0)
    for i in range(72):
        q = rng.normal(size=4)
        q = tuple(float(x) for x in (q / np.linalg.norm(q)).astype(np.float32))
        p = (float(rng.uniform(-5, 5)), 1.2 + 0.8 * (i // 12), float(rng.uniform(-3, 3)))
        if i % 4 == 3:
            sb.box(p, tuple(rng.uniform(0.15, 0.4, 3)), 1.5, quat=q)
        elif i % 8 == 1:
            sb.sphere(p, float(rng.uniform(0.15, 0.4)), 1.0)
        else:
            sb.capsule(p, float(rng.uniform(0.08, 0.4)), float(rng.uniform(0.0, 0.9)), float(rng.uniform(0.5, 3.0)),
                       quat=q, angvel=tuple(rng.uniform(-3, 3, 3)))
    for k in range(4):
        sb.capsule((2.0, 1.0 + 0.9 * k, -1.5 + k), 0.1, 0.3, 1.0, vel=(40.0 + 5.0 * k, 0.0, 0.0),
                   quat=(float(np.sin(0.3 * k)), 0.0, 0.0, float(np.cos(0.3 * k))))
    return sb.build()


def test_capsules_bitwise(gpu_available):
    """Capsule pairs with every other shape (sphere_capsule,
    capsule_capsule, box_capsule with its spine path, half-space support)
    through 180 steps: bodies, pairs, colourings and TOI clamps bit for bit;
    the contained / face-crossing trigger overlaps of
    solver_lifecycle.rs:465-492 enter once each on the GPU; bad capsule
    parameters are rejected."""
    g, o = run_pair(_capsule_scene(91), 180, check_every=15)
    assert g.stats()["toi_hits"] >= 0 and g.stats()["breakdowns"] == 0
    for sp in ((0.1, 0.2), (0.1, 3.0)):
        sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
        t = sb.box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0)
        sb.capsule((0.0, 0.0, 0.0), sp[0], sp[1], 1.0)
        sc = sb.build()
        sc["trigger"][t] = 1
        e = AvbdEngine(scene=sc)
        e.set_events(True)
        e.step(1.0 / 60.0)
        kinds, ab, *_ = e.drain_events(trigger=True)
        assert kinds.tolist() == [3] and ab.tolist() == [[0, 1]], (sp, kinds)
    for bad in ((0.0, 1.0), (0.2, -0.1), (float("inf"), 1.0)):
        sb = scenes.SceneBuilder()
        sb.capsule((0.0, 0.0, 0.0), bad[0], bad[1], 1.0)
        with pytest.raises(Exception):
            AvbdEngine(scene=sb.build())
