"""Unwaited stepping: oavbd_step_n running its steps without reading any
count back mid-step gives the same bits as the waited path (which the
parity tests hold to the oracle), including batches repeated after a count
outgrew its buffer."""
import numpy as np
import pytest

from ornis_amd import scenes
from ornis_amd.avbd import AvbdEngine

pytestmark = [pytest.mark.gpu, pytest.mark.skip(
    reason="unwaited stepping is switched off: a 3,000-box batch that outgrew its buffers (and was to be "
           "repeated) ended in a GPU memory fault whose cause has not been found; the 400-box batches were "
           "bitwise equal to the waited path before that")]


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _same(a, b, where):
    assert np.array_equal(a.pairs(), b.pairs()), f"pair lists differ {where}"
    assert np.array_equal(a.colours(), b.colours()), f"colourings differ {where}"
    ta, tta = a.toi()
    tb, ttb = b.toi()
    assert np.array_equal(ta.view(np.uint32), tb.view(np.uint32)) and np.array_equal(tta, ttb), \
        f"TOI clamps differ {where}"
    sa, sb = a.state(), b.state()
    bad = np.where((sa.view(np.uint32) != sb.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{where}: {bad.size} bodies differ, first {bad[:5]}"
    for k in ("pairs", "points", "candidates", "colours", "sleeping", "toi_casts", "toi_hits", "point_sides"):
        assert a.stats()[k] == b.stats()[k], f"stat {k} differs {where}"


def _engines(scene, iters=10, spare=1):
    waited = AvbdEngine(scene=scene, iters=iters)
    waited.set_unwaited(False)
    unwaited = AvbdEngine(scene=scene, iters=iters)
    unwaited.set_unwaited(True, spare)
    return waited, unwaited


def test_unwaited_batches_equal_waited(gpu_available):
    """A 400-box drop in batches of 10 steps, compared after every batch."""
    s = scenes.box_drop_scene(400, _rng(3))
    w, u = _engines(s)
    for b in range(15):
        w.step_n(10)
        u.step_n(10)
        assert u.stats()["unwaited"] == (1 if b else 0)
        assert w.stats()["unwaited"] == 0
        _same(w, u, f"after batch {b}")


def test_overflow_replays_keep_bits(gpu_available):
    """Counts outgrowing their buffers inside a batch (a 3,000-box drop that
    lands during a 90-step batch, no spare colour launches) make the batch
    repeat with waited steps; the result is the waited path's bits."""
    s = scenes.box_drop_scene(3000, _rng(5), floor=25.0, height=2.0)
    w, u = _engines(s, spare=0)
    for e in (w, u):
        e.step_n(1)
    u.step_n(90)
    w.step_n(90)
    assert u.stats()["replays"] >= 1
    _same(w, u, "after the landing batch")
    for b in range(4):  # settled counts: batches run unwaited again
        w.step_n(15)
        u.step_n(15)
        _same(w, u, f"after settled batch {b}")
    assert u.stats()["unwaited"] == 1


def test_unwaited_joint_chains_and_scenes(gpu_available):
    """Ball-joint chains dropped on spheres, and several independent scenes
    in one context, in batches of 6 steps."""
    c = scenes.chains(n_chains=40, links=12, grid=8, n_spheres=60)
    w, u = _engines(c, iters=12)
    for b in range(10):
        w.step_n(6)
        u.step_n(6)
        _same(w, u, f"chains, after batch {b}")
    parts = [scenes.box_drop_scene(int(k), _rng(100 + i)) for i, k in enumerate((12, 40, 7, 90, 25))]
    allsc = scenes.concat(parts)
    allsc["scene_id"] = np.concatenate([np.full(len(p["shape"]), i, np.uint32) for i, p in enumerate(parts)])
    w = AvbdEngine(iters=10)
    w.set_unwaited(False)
    w.load(allsc, scene_ids=allsc["scene_id"])
    u = AvbdEngine(iters=10)
    u.load(allsc, scene_ids=allsc["scene_id"])
    for b in range(8):
        w.step_n(12)
        u.step_n(12)
        _same(w, u, f"batched scenes, after batch {b}")
