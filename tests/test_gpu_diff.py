"""Differentiable AVBD on the GPU: device checkpoints repeat steps bit for
bit, and the adjoint's gradients of the scene-centroid loss with respect to
friction, mass and initial velocity agree with FP64 central finite
differences of the CPU oracle stepping through the GPU's own recorded active
sets (pairs, anchors, colouring, sleep and every clamp branch), on the throw
workload: six scenes at 60 and 120 steps, and scenes drawn across the full
1,024-scene x 64-box batch.

The reference for the finite differences: the float32 oracle in colour order
steps each scene bit for bit as the GPU did and records its active sets; the
float64 oracle then replays them, so the derivative is that of one smooth
map.  At 60 steps h = 1e-4 is converged on these six scenes (not on all:
scene 511 of the batch needs h = 1e-6); at 120 steps it is not (h = 1e-5 to
1e-7 agree to about 1e-6 relative, h = 1e-4 is off by up to 4e-3), so those
checks use h = 1e-6.

The criterion is each scene's gradient-vector relative error |g_adjoint -
g_FD| / |g_FD| over its checked derivatives (friction, masses, velocity
components and a random velocity direction), with every single component
also within ten times the bound (against max(|FD|, 1% of the scene's largest
derivative)).  The adjoint runs in float32 about the float32 trajectory,
which ends up micrometres (60 steps) to millimetres (120 steps) from the
float64 replay of the same branches, so the two derivatives are of nearby
points; a component that is a small net sum moving fast with the state
differs by more than its share.  On scene 511 at 60 steps the friction
derivative is -0.0025 (it passes from +0.066 at 30 steps to -0.37 at 120):
the float32 / float64 gap there is worth about 1e-5 of friction, and the FD
derivative at friction - 1e-5 is -0.002486, against the adjoint's -0.002488
and the FD's -0.002499 at the nominal friction.  Bounds: 1e-3 at 60 steps;
5e-3 at 120 steps, once the boxes have settled against each other."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from ornis_amd import diff, scenes
from ornis_amd.avbd import AvbdEngine

pytestmark = pytest.mark.gpu


def test_checkpoint_load_repeats_steps_bitwise(gpu_available):
    s = scenes.box_drop_scene(300, np.random.Generator(np.random.PCG64(23)))
    g = AvbdEngine(scene=s)
    g.step_n(30)
    g.checkpoint_save(0)
    g.step_n(20)
    a = (g.state(), g.pairs(), g.points(), g.colours())
    g.checkpoint_load(0)
    g.step_n(20)
    b = (g.state(), g.pairs(), g.points(), g.colours())
    for x, y in zip(a, b):
        assert x.shape == y.shape and np.array_equal(x.view(np.uint32), y.view(np.uint32))


def _oracle(scene, prec=64):
    s = {k: v for k, v in scene.items() if k != "target"}
    return po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE, precision=prec)


def _centroid_loss64(e, dyn, target):
    st = e.state_f64()
    c = st[dyn, 0:3].mean(axis=0)
    return float(((c - target.astype(np.float64)) ** 2).sum())


def _fd_checks(p, steps, h, gv, gm, gf, rng):
    """(name, adjoint, FP64 central difference) for one scene: its friction,
    two masses, two initial-velocity components and a random direction over
    its initial velocities.  gv, gm, gf are the scene's slices."""
    n = len(p["shape"])
    dyn = p["body_type"] == 1
    nominal = _oracle(p, prec=32)
    nominal.record_active_sets()
    nominal.step_inner(steps)
    assert all(nominal.recorded_toi_hits(k) == 0 for k in range(steps))
    vel0 = p["vel"].astype(np.float64)
    mass0 = p["mass"].astype(np.float64)
    fric0 = p["friction"].astype(np.float64)

    def L(vel=vel0, mass=mass0, fric=fric0):
        e = _oracle(p)
        e.set_params(vel=vel, mass=mass, friction=fric)
        e.replay_active_sets(nominal)
        e.step_inner(steps)
        return _centroid_loss64(e, dyn, p["target"])

    # the scene's friction (its bodies and its plane)
    checks = [("friction", gf.sum(), (L(fric=fric0 + h) - L(fric=fric0 - h)) / (2 * h))]
    for b in rng.choice(np.where(dyn)[0], 2, replace=False):
        dm = np.zeros(n)
        dm[b] = h
        checks.append((f"mass[{b}]", gm[b], (L(mass=mass0 + dm) - L(mass=mass0 - dm)) / (2 * h)))
        k = int(rng.integers(3))
        dv = np.zeros((n, 3))
        dv[b, k] = h
        checks.append((f"v0[{b},{k}]", gv[b, k], (L(vel=vel0 + dv) - L(vel=vel0 - dv)) / (2 * h)))
    d = rng.normal(size=(n, 3)) * dyn[:, None]
    checks.append(("v0 . d", float((gv * d).sum()), (L(vel=vel0 + h * d) - L(vel=vel0 - h * d)) / (2 * h)))
    return nominal.state(), checks


def _rel_errors(checks, tag):
    """Per-component relative errors, each against max(|FD|, 1% of the
    scene's largest FD derivative), and the scene's gradient-vector relative
    error |g_adjoint - g_FD| / |g_FD| over all its checks."""
    scale = max(abs(c[2]) for c in checks)
    comp = [(abs(ga - gd) / max(abs(gd), 1e-2 * scale), tag, name, ga, gd) for name, ga, gd in checks]
    ga = np.array([c[1] for c in checks], np.float64)
    gd = np.array([c[2] for c in checks], np.float64)
    return comp, float(np.linalg.norm(ga - gd) / np.linalg.norm(gd))


def _assert_close(comp, vec, tol, label):
    comp.sort(reverse=True)
    print(f"{label}: gradient-vector relative errors {np.round(vec, 6).tolist()}, worst components", comp[:4])
    assert max(vec) < tol, (vec, comp[:4])
    assert comp[0][0] < 10 * tol, comp[:4]


def _rollout(eng, allsc, parts, steps):
    targets = torch.tensor(np.stack([p["target"] for p in parts]), dtype=torch.float32)
    v0 = torch.tensor(allsc["vel"], requires_grad=True)
    mass = torch.tensor(allsc["mass"], requires_grad=True)
    fric = torch.tensor(allsc["friction"], requires_grad=True)
    final = diff.rollout(eng, allsc, v0, mass, fric, steps, every=10)
    loss = diff.centroid_loss(final, allsc["scene_id"], allsc["body_type"] == 1, targets)
    loss.backward()
    return final.detach().numpy(), v0.grad.numpy(), mass.grad.numpy(), fric.grad.numpy()


@pytest.mark.parametrize("steps,h,tol", [(60, 1e-4, 1e-3), (120, 1e-6, 5e-3)])
def test_adjoint_matches_fp64_central_differences(gpu_available, steps, h, tol):
    parts, allsc = scenes.diff_throws(n_scenes=6)
    final, gv, gm, gf = _rollout(AvbdEngine(), allsc, parts, steps)
    for g in (gv, gm, gf):
        assert np.isfinite(g).all()
    rng = np.random.default_rng(0)
    comp, vec = [], []
    off = 0
    for si, p in enumerate(parts[:3]):
        n = len(p["shape"])
        sl = slice(off, off + n)
        state, checks = _fd_checks(p, steps, h, gv[sl], gm[sl], gf[sl], rng)
        # the float32 oracle in colour order stepped this scene as the GPU did
        assert np.array_equal(state.view(np.uint32), final[sl].view(np.uint32))
        c, v = _rel_errors(checks, si)
        comp += c
        vec.append(v)
        off += n
    _assert_close(comp, vec, tol, f"{steps} steps, h {h}")


def test_full_size_gradients(gpu_available):
    """1,024 scenes x 64 boxes.  At 120 steps forward and backward: every
    gradient is finite, static bodies get none, and the first scenes'
    gradients equal those of a 6-scene batch (scenes are independent; the
    adjoint's float atomics may reorder sums: within 1e-4 of the largest
    gradient).  At 60 steps: the gradients of scenes drawn across the batch,
    the last one included, match FP64 central differences of their own
    active sets: each scene's gradient vector within 1e-3 relative (h = 1e-6:
    on scene 511, h = 1e-4 is still 5% from the converged friction
    derivative)."""
    parts, allsc = scenes.diff_throws(n_scenes=1024)
    dyn = allsc["body_type"] == 1
    eng = AvbdEngine()
    final, gv, gm, gf = _rollout(eng, allsc, parts, 120)
    for g in (gv, gm, gf):
        assert np.isfinite(g).all()
    assert np.abs(gv[dyn]).max() > 0 and np.abs(gm[dyn]).max() > 0 and np.abs(gf).max() > 0
    assert np.abs(gv[~dyn]).max() == 0 and np.abs(gm[~dyn]).max() == 0

    small, sallsc = scenes.diff_throws(n_scenes=6)
    sfinal, sgv, sgm, sgf = _rollout(AvbdEngine(), sallsc, small, 120)
    m = len(sallsc["shape"])
    assert np.array_equal(sfinal.view(np.uint32), final[:m].view(np.uint32))
    for name, a, b in (("v0", sgv, gv[:m]), ("mass", sgm, gm[:m]), ("friction", sgf, gf[:m])):
        r = float(np.abs(a - b).max() / np.abs(a).max())
        assert r < 1e-4, (name, r)

    steps = 60
    final, gv, gm, gf = _rollout(eng, allsc, parts, steps)
    starts = np.concatenate([[0], np.cumsum([len(p["shape"]) for p in parts])])
    rng = np.random.default_rng(1)
    comp, vec = [], []
    for si in (97, 511, 770, 1023):
        sl = slice(starts[si], starts[si + 1])
        state, checks = _fd_checks(parts[si], steps, 1e-6, gv[sl], gm[sl], gf[sl], rng)
        assert np.array_equal(state.view(np.uint32), final[sl].view(np.uint32))
        c, v = _rel_errors(checks, si)
        comp += c
        vec.append(v)
    _assert_close(comp, vec, 1e-3, f"full size, {steps} steps")
