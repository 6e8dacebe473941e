"""The 100k-box workload at its real size.

Bit-exact parity against the oracle is checked from a dense mid-trajectory
state (the GPU trajectory's bodies at step 150 loaded into a fresh GPU
context and into the oracle, both without contact history); the 300-step
trajectory itself is checked through properties that hold at any size."""
import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes
from ornis_amd.avbd import AvbdEngine

pytestmark = pytest.mark.gpu


def _with_state(scene, st):
    s = dict(scene)
    s["pos"] = np.ascontiguousarray(st[:, 0:3])
    s["quat"] = np.ascontiguousarray(st[:, 3:7])
    s["vel"] = np.ascontiguousarray(st[:, 7:10])
    s["angvel"] = np.ascontiguousarray(st[:, 10:13])
    return s


def test_100k_dense_state_bitwise(gpu_available):
    scene = scenes.box_drop_lattice()
    g = AvbdEngine(scene=scene)
    g.step_n(150)
    mid = _with_state(scene, g.state())
    g.close()
    g = AvbdEngine(scene=mid)
    o = po.OracleEngine(mid, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    for step in range(2):
        g.step_n(1)
        o.step_inner(1)
        pg, po_ = g.pairs(), o.pairs()[0]
        assert pg.shape == po_.shape and np.array_equal(pg, po_), f"pair list differs at step {step}"
        assert np.array_equal(g.colours(), o.colours()), f"colouring differs at step {step}"
        tg, ttg = g.toi()
        to, tto = o.toi()
        assert np.array_equal(tg.view(np.uint32), to.view(np.uint32)) and np.array_equal(ttg, tto)
        sg, so = g.state(), o.state()
        bad = np.where((sg.view(np.uint32) != so.view(np.uint32)).any(axis=1))[0]
        assert bad.size == 0, f"step {step}: {bad.size} bodies differ, max |d| {np.abs(sg - so).max():.3e}"
    assert g.stats()["pairs"] > 100000


def test_100k_trajectory_properties(gpu_available):
    scene = scenes.box_drop_lattice()
    g = AvbdEngine(scene=scene)
    g.step_n(300)
    st = g.state()
    assert np.isfinite(st).all()
    dyn = scene["body_type"] == 1
    # no body falls through the floor.  Under the pile's weight the penalty
    # ramp lets bodies sink into the plane a little, most at the container's
    # edges, where they are wedged against a wall (at step 300 the lowest
    # centre is about 2.8 m down, about 20 of 100,000 are below -0.6 m);
    # walls are 40 m high and the lattice reaches 80 m, so a few ejected
    # bodies may leave the container
    y = st[dyn, 1]
    assert (y > -5.0).all() and (y < -1.0).mean() < 5e-4
    # the pile spills over the walls (about 1.3% of the bodies by step 300),
    # but nothing explodes: every body stays near the container and no body
    # moves faster than a free fall from the top of the lattice (~40 m/s)
    inside = (np.abs(st[dyn, 0]) < 56.0) & (np.abs(st[dyn, 2]) < 56.0)
    assert inside.mean() > 0.95
    assert (np.abs(st[dyn][:, [0, 2]]) < 300.0).all()
    assert np.linalg.norm(st[dyn, 7:10], axis=1).max() < 50.0
    q = np.linalg.norm(st[:, 3:7], axis=1)
    assert np.abs(q - 1.0).max() < 1e-5
    ab = g.pairs()
    assert (ab[:, 0] < ab[:, 1]).all()
    key = ab[:, 0].astype(np.uint64) << np.uint64(32) | ab[:, 1].astype(np.uint64)
    assert (np.diff(key) > 0).all(), "pair list not strictly sorted"
    s = g.stats()
    assert s["breakdowns"] == 0 and s["colours"] >= 2
    col = g.colours()
    assert (col >= 0).sum() == s["solved"] <= dyn.sum()
    assert (col[~dyn] < 0).all()


def _ball_joint_gaps(st, joints):
    """World distance between the two anchors of every ball joint."""
    a = np.array([j[0] for j in joints])
    b = np.array([j[1] for j in joints])
    la = np.array([j[2] for j in joints], np.float64)
    lb = np.array([j[3] for j in joints], np.float64)

    def world(i, lv):
        q = st[i, 3:7].astype(np.float64)
        u, w = q[:, :3], q[:, 3:4]
        t = 2.0 * np.cross(u, lv)
        return st[i, 0:3] + lv + w * t + np.cross(u, t)

    return np.linalg.norm(world(a, la) - world(b, lb), axis=1)


@pytest.mark.timeout(900)
def test_chains_full_size(gpu_available):
    """The contacts + joints workload at its real size (2,000 chains of 32
    boxes, 62,000 ball joints, 5,000 static spheres, 12 iterations): its
    240-step trajectory through size-independent properties (finite state,
    unit rotations, no link detached -- anchors stay within 1 m; on a reduced
    copy the oracle's own anchors part by up to ~0.5 m while chains pile up
    -- no body under the plane, no runaway speed), then two steps from the
    trajectory's state at step 120 bit for bit against the oracle (both
    loaded fresh: no contact history, joint duals from zero)."""
    scene = scenes.chains()
    g = AvbdEngine(scene=scene, iters=12)
    g.step_n(120)
    mid = _with_state(scene, g.state())
    g.step_n(120)
    st = g.state()
    dyn = scene["body_type"] == 1
    assert np.isfinite(st).all()
    assert np.abs(np.linalg.norm(st[:, 3:7], axis=1) - 1.0).max() < 1e-5
    gaps = _ball_joint_gaps(st, scene["joints"])
    assert gaps.max() < 1.0, gaps.max()
    assert (st[dyn, 1] > -0.1).all(), st[dyn, 1].min()
    assert np.linalg.norm(st[dyn, 7:10], axis=1).max() < 30.0
    s = g.stats()
    assert s["breakdowns"] == 0 and s["joint_sides"] > 0 and s["pairs"] > 10000
    g.close()
    print("chains: 240 GPU steps checked; oracle steps from step 120 follow", flush=True)
    g = AvbdEngine(scene=mid, iters=12)
    o = po.OracleEngine(mid, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE, iters=12)
    o.set_threads(16)
    for step in range(2):
        g.step_n(1)
        o.step_inner(1)
        print(f"chains: oracle step {step} done", flush=True)
        assert np.array_equal(g.pairs(), o.pairs()[0]), f"pair list differs at step {step}"
        assert np.array_equal(g.colours(), o.colours()), f"colouring differs at step {step}"
        sg, so = g.state(), o.state()
        bad = np.where((sg.view(np.uint32) != so.view(np.uint32)).any(axis=1))[0]
        assert bad.size == 0, f"step {step}: {bad.size} bodies differ, max |d| {np.abs(sg - so).max():.3e}"
        assert np.array_equal(g.joints().view(np.uint32),
                              o.joints(len(scene["joints"])).view(np.uint32)), f"joint duals differ at step {step}"
