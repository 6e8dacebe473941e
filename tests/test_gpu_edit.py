"""Incremental edits and the contact state carried between steps, on the GPU
against the oracle (float32, colour order), bit for bit.

- oavbd_add_bodies / oavbd_remove_body mid-run (PhysicsEngine::add_body and
  remove_body, crates/physics/src/avbd/mod.rs:1236-1300): appended bodies
  leave the others' pairs and warm starts alone; a swap-remove drops the
  removed body's pairs and joints, renumbers the rest in place and wakes
  every sleeping body.  The oracle applies the same edits.
- The reference's lifecycle test (crates/physics/tests/solver_lifecycle.rs:
  41-78) with a Ball joint in place of its Distance joint: removing an
  unrelated body keeps the moved tail's joint, which then pulls the tail
  back from a teleport.
- oavbd_get_contacts / oavbd_set_contacts / oavbd_set_sleep: a round trip
  through a shuffled table, rejected tables, and the 100k-box workload
  continued for 10 consecutive steps from step 150 with its full contact
  history, by the same context, by a fresh context loaded from the export,
  and by the oracle loaded from it."""
import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes
from ornis_amd.avbd import AvbdEngine, AvbdError

pytestmark = pytest.mark.gpu


def _oracle(scene):
    return po.OracleEngine(scene, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)


def _same(g, o, what):
    sg, so = g.state(), o.state()
    assert sg.shape == so.shape, what
    bad = np.where((sg.view(np.uint32) != so.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} bodies differ (first {bad[:5]})"
    # right after a removal the oracle keeps the reference's in-place order
    # until its next merge sorts it; the library stores the table sorted
    # (the merge looks pairs up by key, so the order carries no state)
    pg = g.pairs()
    po_, npts, _, _ = o.pairs()
    order = np.lexsort((po_[:, 1], po_[:, 0]))
    assert pg.shape == po_.shape and np.array_equal(pg, po_[order]), f"{what}: pair lists differ"
    off = np.concatenate([[0], np.cumsum(npts)]).astype(np.int64)
    op = o.points()
    op = np.concatenate([op[off[i]:off[i + 1]] for i in order]) if len(order) else op
    gp = g.points()
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), f"{what}: points differ"
    ga, gt = g.sleep()
    oa, ot = o.sleep()
    assert np.array_equal(ga, oa) and np.array_equal(gt.view(np.uint32), ot.view(np.uint32)), f"{what}: sleep"


def _boxes(rng, n, y0):
    """n unit-mass boxes in a loose column above the origin (a scene dict)."""
    s = scenes.SceneBuilder()
    for i in range(n):
        pos = (rng.uniform(-1.0, 1.0), y0 + 1.1 * i, rng.uniform(-1.0, 1.0))
        s.box(pos, (0.3, 0.25, 0.35), 1.0, friction=0.6)
    return s.build()


def test_add_and_remove_bodies_bitwise(gpu_available):
    rng = np.random.Generator(np.random.PCG64(41))
    scene = scenes.box_drop_scene(120, rng)
    g = AvbdEngine(scene=scene, max_bodies=200)
    o = _oracle(scene)
    g.step_n(250)  # four bodies asleep by then
    o.step_inner(250)
    _same(g, o, "before edits")
    assert g.stats()["sleeping"] > 0, "the scene should hold sleeping bodies before the removal"
    extra = _boxes(rng, 6, 3.0)
    first = g.add_bodies(extra)
    o.add_bodies(extra)
    assert first == len(scene["shape"])
    _same(g, o, "after add")
    g.step_n(20)
    o.step_inner(20)
    _same(g, o, "20 steps after add")
    # a body with contacts, then the last body (which moves onto itself)
    ab = g.pairs()
    victim = int(ab[len(ab) // 2, 1])
    for h in (victim, None, 0):
        h = g.n - 1 if h is None else h
        g.remove_body(h)
        o.remove_body(h)
        _same(g, o, f"after removing {h}")
        assert not g.sleep()[0].any(), "every body wakes on a removal"
        g.step_n(15)
        o.step_inner(15)
        _same(g, o, f"15 steps after removing {h}")
    g.remove_body(10 ** 6)  # out of range: no-op
    assert g.n == o.n


def test_remove_renumbers_joints_bitwise(gpu_available):
    s = scenes.chains(n_chains=4, links=6, grid=2, height=3.0, n_spheres=3)
    g = AvbdEngine(scene=s)
    o = _oracle(s)
    g.step_n(30)
    o.step_inner(30)
    nj = len(s["joints"])
    assert np.array_equal(g.joints().view(np.uint32), o.joints(nj).view(np.uint32))
    link = int(s["joints"][3][1])  # a link inside the first chain
    g.remove_body(link)
    o.remove_body(link)
    g.step_n(40)
    o.step_inner(40)
    _same(g, o, "chains after a removal")
    nj2 = g.njoints
    assert nj2 < nj
    assert np.array_equal(g.joints().view(np.uint32), o.joints(nj2).view(np.uint32))


def test_removing_an_unrelated_body_keeps_the_moved_tails_joint(gpu_available):
    b = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    b.sphere((0.0, 0.0, 0.0), 0.2, 0.0)   # anchor (static)
    b.sphere((20.0, 0.0, 0.0), 0.2, 1.0)  # spare
    b.sphere((0.0, 2.0, 0.0), 0.2, 1.0)   # tail
    b.ball_joint(0, 2, (0.0, 0.0, 0.0), (0.0, -2.0, 0.0))
    s = b.build()
    s["mask"][:] = 0  # isolated: collides with nothing
    g = AvbdEngine(scene=s)
    o = _oracle(s)
    g.remove_body(1)
    o.remove_body(1)
    assert g.n == 2 and g.njoints == 1
    st = g.state()
    st[1, 1] = 3.0  # the tail, now at the spare's index, is pulled up a metre
    g.set_state(st)
    o.set_pose(1, st[1, 0:3], st[1, 3:7])
    g.step_n(600)
    o.step_inner(600)
    _same(g, o, "lifecycle")
    assert abs(g.state()[1, 1] - 2.0) < 0.04


def test_contacts_round_trip_and_validation(gpu_available):
    scene = scenes.box_drop_scene(80, np.random.Generator(np.random.PCG64(5)))
    g = AvbdEngine(scene=scene)
    g.step_n(60)
    pairs, points = g.contacts()
    assert len(pairs) > 20 and points.shape[0] == pairs["npoints"].sum()
    perm = np.random.default_rng(1).permutation(len(pairs))
    g.set_contacts(pairs[perm], points)  # any pair order; stored sorted
    p2, q2 = g.contacts()
    assert p2.tobytes() == pairs.tobytes() and q2.tobytes() == points.tobytes()
    for bad in ("index", "self", "dup", "count", "range"):
        p = pairs.copy()
        if bad == "index":
            p[0]["b"] = g.n
        elif bad == "self":
            p[0]["b"] = p[0]["a"]
        elif bad == "dup":
            p[1]["a"], p[1]["b"] = p[0]["a"], p[0]["b"]
        elif bad == "count":
            p[0]["npoints"] = 9
        else:
            p[-1]["first_point"], p[-1]["npoints"] = len(points), 1
        with pytest.raises(AvbdError):
            g.set_contacts(p, points)
    p2, q2 = g.contacts()  # a rejected table leaves the state alone
    assert p2.tobytes() == pairs.tobytes()


@pytest.mark.timeout(900)
def test_100k_consecutive_steps_with_contact_history_bitwise(gpu_available):
    scene = scenes.box_drop_lattice()
    g = AvbdEngine(scene=scene)
    g.step_n(150)
    st = g.state()
    asleep, timer = g.sleep()
    pairs, points = g.contacts()
    assert len(pairs) > 100000
    mid = dict(scene)
    mid["pos"], mid["quat"] = np.ascontiguousarray(st[:, 0:3]), np.ascontiguousarray(st[:, 3:7])
    mid["vel"], mid["angvel"] = np.ascontiguousarray(st[:, 7:10]), np.ascontiguousarray(st[:, 10:13])
    o = _oracle(mid)
    o.set_threads(16)
    o.set_sleep(asleep, timer)
    o.set_contacts(pairs, points)
    f = AvbdEngine(scene=mid)
    f.set_sleep(asleep, timer)
    f.set_contacts(pairs, points)
    assert np.array_equal(f.state().view(np.uint32), st.view(np.uint32))
    for step in range(10):
        g.step_n(1)
        f.step_n(1)
        o.step_inner(1)
        sg, sf, so = g.state(), f.state(), o.state()
        assert np.array_equal(sg.view(np.uint32), sf.view(np.uint32)), f"fresh context differs at step {step}"
        bad = np.where((sg.view(np.uint32) != so.view(np.uint32)).any(axis=1))[0]
        assert bad.size == 0, f"step {step}: {bad.size} bodies differ from the oracle"
        assert np.array_equal(g.pairs(), o.pairs()[0]), f"pair list differs at step {step}"
    assert np.array_equal(g.points().view(np.uint32), o.points().view(np.uint32))
    ga, gt = g.sleep()
    oa, ot = o.sleep()
    assert np.array_equal(ga, oa) and np.array_equal(gt.view(np.uint32), ot.view(np.uint32))
