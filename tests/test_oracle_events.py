"""Contact and trigger events of the CPU oracle (emit_events,
crates/physics/src/avbd/mod.rs:1606-1669, and the TOI clamp's impact events,
mod.rs:943-960 / :1126-1136): the reference's own event tests restated, and
checks of what an event must carry (its pair, the point on the right body,
the normal's direction, the closing speed) against the scene's mechanics."""
import numpy as np

from oracle import pyoracle as po
from ornis_amd import scenes

DT = 1.0 / 60.0
BEGIN, END, HIT, ENTERED, EXITED = 0, 1, 2, 3, 4


def test_trigger_enter_exit_roundtrip():
    """avbd_trigger_enter_exit_roundtrip (crates/physics/tests/avbd_engine.rs:275-293)."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    s = sb.sphere((0.0, 0.0, 0.0), 1.0, 0.0)
    m = sb.sphere((5.0, 0.0, 0.0), 0.5, 1.0)
    sc = sb.build()
    sc["trigger"][s] = 1
    e = po.OracleEngine(sc)
    e.step(DT)
    assert len(e.drain_events(trigger=True)[0]) == 0
    e.set_pose(m, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0))
    e.step(DT)
    kinds, ab, *_ = e.drain_events(trigger=True)
    assert kinds.tolist() == [ENTERED] and ab.tolist() == [[min(s, m), max(s, m)]]
    assert e.event_state(trigger=True).tolist() == [[min(s, m), max(s, m)]]
    e.set_pose(m, (5.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0))
    e.step(DT)
    kinds, ab, *_ = e.drain_events(trigger=True)
    assert kinds.tolist() == [EXITED] and ab.tolist() == [[min(s, m), max(s, m)]]
    # a trigger never forms a contact pair
    assert len(e.drain_events()[0]) == 0 and e.pairs()[0].shape[0] == 0


def test_contact_begin_end_roundtrip():
    """avbd_contact_begin_end_roundtrip (crates/physics/tests/avbd_engine.rs:296-319)."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.sphere((0.0, 0.0, 0.0), 0.5, 0.0)
    b = sb.sphere((5.0, 0.0, 0.0), 0.5, 1.0)
    e = po.OracleEngine(sb.build())
    e.step(DT)
    assert len(e.drain_events()[0]) == 0
    e.set_pose(b, (0.9, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0))
    e.step(DT)
    kinds, ab, *_ = e.drain_events()
    assert BEGIN in kinds.tolist() and [0, 1] in ab.tolist()
    assert e.event_state().tolist() == [[0, 1]]
    e.set_pose(b, (5.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0))
    e.step(DT)
    kinds, ab, *_ = e.drain_events()
    assert END in kinds.tolist()
    assert e.event_state().shape[0] == 0


def test_landing_box_hits_once_with_its_impact_speed():
    """A 125 kg half-metre box dropped onto a static floor (slow enough
    that no TOI clamp applies): the step it first sinks more than 5 mm
    reports Begin then Hit for (floor, box); the hit's approach speed is the
    box's downward speed at the start of that step (the floor is static);
    its normal -n points from body a (floor) towards body b (box), i.e. +y,
    n being the pair's B -> A normal; its point is the deepest anchor on
    body a, on the floor's top face under the box.  No second Begin or Hit
    follows while the box rests.  (A 1 kg box never sinks 5 mm at impact:
    its first Begin comes later, from creep, with no Hit.)"""
    sb = scenes.SceneBuilder()
    sb.box((0.0, -0.5, 0.0), (5.0, 0.5, 5.0), 0.0)
    box = sb.box((0.0, 2.0, 0.0), (0.25, 0.25, 0.25), 125.0, restitution=0.0)
    e = po.OracleEngine(sb.build())
    events, v_before = [], []
    for _ in range(90):
        v_before.append(e.state()[box, 7:10].copy())
        e.step(DT)
        events.append(e.drain_events())
    first = next(i for i, ev in enumerate(events) if len(ev[0]))
    kinds, ab, pt, nrm, app = events[first]
    assert kinds.tolist() == [BEGIN, HIT] and ab.tolist() == [[0, box], [0, box]]
    vy = float(v_before[first][1])
    assert vy < -1.0
    assert abs(app[1] + vy) < 1e-5, (app[1], vy)
    assert np.allclose(nrm[1], (0.0, 1.0, 0.0), atol=1e-6)
    assert abs(pt[1][1]) < 1e-6 and np.all(np.abs(pt[1][[0, 2]]) <= 0.25 + 1e-4)
    later = [k for ev in events[first + 1:] for k in ev[0].tolist()]
    assert BEGIN not in later and HIT not in later
    assert e.event_state().tolist() == [[0, box]]


def test_thin_wall_impact_reported_by_the_clamp():
    """The TOI clamp's impact event (mod.rs:943-960, :1126-1136): a 10 m/s
    box stopped at a 4 cm wall (avbd_engine.rs:1290-1313's scene) reports
    a Hit for (wall, box) with approach 10 m/s, normal -x (the clamp's
    normal, from the mover's witness to the wall's, kept as is since the
    mover has the larger handle: it points from the wall at a to the box at
    b) and the contact point on the wall's -x face."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.box((5.0, 0.0, 0.0), (0.02, 2.0, 2.0), 0.0)
    sb.box((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), 1.0, vel=(10.0, 0.0, 0.0), restitution=0.0)
    for order in (po.ORDER_REFERENCE, po.ORDER_COLOUR):
        e = po.OracleEngine(sb.build(), order=order)
        hits = []
        for _ in range(60):
            e.step(DT)
            kinds, ab, pt, nrm, app = e.drain_events()
            hits += [(ab[i].tolist(), pt[i], nrm[i], app[i]) for i in range(len(kinds)) if kinds[i] == HIT]
        assert len(hits) >= 1
        ab, pt, nrm, app = hits[0]
        assert ab == [0, 1]
        assert abs(app - 10.0) < 1e-4
        assert np.allclose(nrm, (-1.0, 0.0, 0.0), atol=1e-5)  # from the wall (a) to the box on its -x side (b)
        assert abs(pt[0] - 4.98) < 2e-3  # the wall's -x face
        assert e.state()[1, 0] < 5.0  # stopped before the wall


def test_remove_body_reports_trigger_exits_and_clears_touches():
    """remove_body (mod.rs:1240-1290): overlapping trigger pairs with the
    removed body report Exited, the touch baseline is dropped and the
    remaining trigger baseline follows the swap-remove renumbering."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    t = sb.sphere((0.0, 0.0, 0.0), 1.0, 0.0)
    sb.sphere((0.3, 0.0, 0.0), 0.2, 1.0)
    sb.sphere((10.0, 0.0, 0.0), 1.0, 0.0)
    sb.sphere((10.3, 0.0, 0.0), 0.2, 1.0)
    sc = sb.build()
    sc["trigger"][t] = 1
    sc["trigger"][2] = 1
    e = po.OracleEngine(sc)
    e.step(DT)
    kinds, ab, *_ = e.drain_events(trigger=True)
    assert kinds.tolist() == [ENTERED, ENTERED] and ab.tolist() == [[0, 1], [2, 3]]
    e.remove_body(1)  # body 3 becomes body 1
    kinds, ab, *_ = e.drain_events(trigger=True)
    assert kinds.tolist() == [EXITED] and ab.tolist() == [[0, 1]]
    assert e.event_state(trigger=True).tolist() == [[1, 2]]
    e.step(DT)
    assert len(e.drain_events(trigger=True)[0]) == 0  # (1, 2) still overlaps: no new Entered
