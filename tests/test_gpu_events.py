"""Contact and trigger events on the GPU (through the C ABI) against the CPU
oracle in the same colour order: the same events in the same order with
the same bits, on the reference's own event tests and on scenes with
resting contacts, impacts, TOI clamps and trigger volumes."""
import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes
from ornis_amd.avbd import AvbdEngine

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
BEGIN, END, HIT, ENTERED, EXITED = 0, 1, 2, 3, 4


def _pair(scene, iters=10):
    g = AvbdEngine(scene=scene, iters=iters)
    g.set_events(True)
    o = po.OracleEngine(scene, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE, iters=iters)
    return g, o


def _same_events(g, o, where):
    """Drains both queues of both engines and compares them bit for bit;
    returns the event kinds seen."""
    seen = []
    for trig in (False, True):
        kg, abg, ptg, ng, ag = g.drain_events(trigger=trig)
        ko, abo, pto, no, ao = o.drain_events(trigger=trig)
        label = "trigger" if trig else "contact"
        assert kg.tolist() == ko.tolist() and abg.tolist() == abo.tolist(), \
            f"{label} events differ {where}: gpu {list(zip(kg.tolist(), abg.tolist()))} " \
            f"oracle {list(zip(ko.tolist(), abo.tolist()))}"
        for x, y in ((ptg, pto), (ng, no), (ag, ao)):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{label} event data differ {where}"
        seen += kg.tolist()
    return seen


def _teleport(g, o, i, pos):
    st = g.state()
    st[i, 0:3] = pos
    g.set_state(st)
    o.set_pose(i, pos, tuple(float(x) for x in st[i, 3:7]))


def test_trigger_and_contact_roundtrips(gpu_available):
    """avbd_trigger_enter_exit_roundtrip and avbd_contact_begin_end_roundtrip
    (crates/physics/tests/avbd_engine.rs:275-319) on the GPU."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    s = sb.sphere((0.0, 0.0, 0.0), 1.0, 0.0)
    m = sb.sphere((5.0, 0.0, 0.0), 0.5, 1.0)
    sc = sb.build()
    sc["trigger"][s] = 1
    g, o = _pair(sc)
    g.step(DT)
    o.step(DT)
    _same_events(g, o, "before the overlap")
    for pos, kind in (((0.0, 0.0, 0.0), ENTERED), ((5.0, 0.0, 0.0), EXITED)):
        _teleport(g, o, m, pos)
        g.step(DT)
        o.step(DT)
        kg, abg, *_ = g.drain_events(trigger=True)
        ko, abo, *_ = o.drain_events(trigger=True)
        assert kg.tolist() == [kind] == ko.tolist() and abg.tolist() == [[0, 1]] == abo.tolist()
        _same_events(g, o, f"after the move to {pos}")

    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.sphere((0.0, 0.0, 0.0), 0.5, 0.0)
    b = sb.sphere((5.0, 0.0, 0.0), 0.5, 1.0)
    g, o = _pair(sb.build())
    g.step(DT)
    o.step(DT)
    _same_events(g, o, "apart")
    for pos, kind in (((0.9, 0.0, 0.0), BEGIN), ((5.0, 0.0, 0.0), END)):
        _teleport(g, o, b, pos)
        g.step(DT)
        o.step(DT)
        kg, abg, pt, nrm, app = g.drain_events()
        ko, *_ = o.drain_events()
        assert kind in kg.tolist() and kg.tolist() == ko.tolist()
        assert np.array_equal(g.event_state(), o.event_state())


def _event_scene(seed):
    """A floor, a static trigger box the boxes fall through, 60 boxes of
    mixed mass (heavy ones land hard enough to report hits), and fast boxes
    thrown at a thin wall (clamp impacts)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sb = scenes.SceneBuilder()
    sb.box((0.0, -0.5, 0.0), (12.0, 0.5, 12.0), 0.0)
    trig = sb.box((0.0, 2.0, 0.0), (2.5, 0.6, 2.5), 0.0)
    sb.box((9.0, 2.0, 0.0), (0.03, 2.0, 3.0), 0.0)
    for i in range(60):
        h = tuple(rng.uniform(0.2, 0.45, 3))
        m = float(1000.0 * 8.0 * h[0] * h[1] * h[2]) if i % 2 else 1.0
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(np.float32)
        sb.box((rng.uniform(-3, 3), 3.5 + 0.9 * (i // 6), rng.uniform(-3, 3)), h, m,
               quat=tuple(float(x) for x in q))
    for k in range(4):
        sb.box((2.0, 1.0 + 0.9 * k, -1.5 + k), (0.2, 0.2, 0.2), 2.0, vel=(35.0 + 5.0 * k, 0.0, 0.0))
    sc = sb.build()
    sc["trigger"][trig] = 1
    return sc


def test_event_stream_equals_oracle(gpu_available):
    """Every call's contact and trigger events over 150 steps: single steps,
    120 Hz frames and 1/30 s hitches (two substeps, the last reporting)."""
    g, o = _pair(_event_scene(11))
    kinds_seen = set()
    dts = [DT] * 60 + [DT / 2.0] * 40 + [2.0 * DT] * 35 + [DT] * 10
    for i, dt in enumerate(dts):
        g.step(dt)
        o.step(dt)
        kinds_seen.update(_same_events(g, o, f"after call {i} ({dt:.4f} s)"))
    assert {BEGIN, END, HIT, ENTERED, EXITED} <= kinds_seen, kinds_seen
    g2, o2 = _pair(_event_scene(11))
    for i in range(90):
        g2.step_n(1)
        o2.step_inner(1)
        _same_events(g2, o2, f"at step {i}")
    assert np.array_equal(g2.event_state(), o2.event_state())
    assert np.array_equal(g2.event_state(trigger=True), o2.event_state(trigger=True))


def test_remove_body_and_event_state(gpu_available):
    """oavbd_remove_body's event semantics (mod.rs:1240-1290) and the event
    baseline's export and import (mod.rs:516-556)."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    t = sb.sphere((0.0, 0.0, 0.0), 1.0, 0.0)
    sb.sphere((0.3, 0.0, 0.0), 0.2, 1.0)
    sb.sphere((10.0, 0.0, 0.0), 1.0, 0.0)
    sb.sphere((10.3, 0.0, 0.0), 0.2, 1.0)
    sb.sphere((20.0, 0.0, 0.0), 0.5, 0.0)
    sb.sphere((20.9, 0.0, 0.0), 0.5, 1.0)
    sc = sb.build()
    sc["trigger"][t] = 1
    sc["trigger"][2] = 1
    g, o = _pair(sc)
    g.step(DT)
    o.step(DT)
    _same_events(g, o, "first step")
    g.remove_body(1)
    o.remove_body(1)
    _same_events(g, o, "after the removal")
    assert np.array_equal(g.event_state(trigger=True), o.event_state(trigger=True))
    assert g.event_state().shape[0] == 0
    g.step(DT)
    o.step(DT)
    _same_events(g, o, "after the removal's step")
    # a rebuilt context seeded with the baseline reports no repeated Begin / Entered
    h = AvbdEngine(iters=10)
    st = g.state()
    sc2 = {k: (v.copy() if hasattr(v, "copy") else v) for k, v in sc.items()}
    keep = [0, 5, 2, 3, 4]  # the swap-removed order
    for k in ("shape", "shape_param", "mass", "friction", "restitution", "body_type", "layer", "mask",
              "trigger", "rolling", "torsion"):
        sc2[k] = sc[k][keep]
    sc2["pos"], sc2["quat"], sc2["vel"], sc2["angvel"] = st[:, 0:3], st[:, 3:7], st[:, 7:10], st[:, 10:13]
    h.load(sc2)
    h.set_events(True)
    h.set_event_state(g.event_state())
    h.set_event_state(g.event_state(trigger=True), trigger=True)
    h.step(DT)
    kinds = h.drain_events()[0].tolist() + h.drain_events(trigger=True)[0].tolist()
    assert BEGIN not in kinds and ENTERED not in kinds
    with pytest.raises(Exception):
        h.set_event_state(np.array([[3, 1]], np.uint32))
