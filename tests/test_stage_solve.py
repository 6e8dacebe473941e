"""Staged per-body solves of the reference's GPU rungs: rung 1, the lumped
inertial diagonal and its diagonal solve (GpuAvbdStub::stage_diag_solve,
crates/physics/src/gpu/avbd.rs:207-226), and rung 2, GpuAvbdSystem records
with one stamped row and the dense 6x6 LDL (avbd_stage_contact_solve
:415-449).

The oracle's staged solve is pinned against the reference's own known
answers (crates/physics/tests/gpu_batches.rs:537-572 for the lumped Hessian
diagonal, :904-926 and :968-1010 for the staged cases) and against the dense
float64 solution of (D + pen v v^T) x = rhs + force v on random systems.  The
device entry point is then compared with the oracle bit for bit."""
import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd.avbd import AvbdEngine, pack_states, pack_systems

DT = 1.0 / 60.0
INERT = dict(axis=(0.0, 0.0, 0.0), pen=0.0, lever=(0.0, 0.0, 0.0), force=0.0, sign=1.0)


def _inert(inv_mass, inertia, rhs, dt=DT):
    return po.stage_contact_solve(pack_systems(inv_mass, inertia, rhs, **INERT), dt)


def _residual(delta):
    return np.concatenate([delta[:, 0:3], delta[:, 4:7]], axis=1)


def _box_inertia(mass, half):
    # Shape::inertia of a box (shape.rs:713-767): m/12 (y^2 + z^2) per axis
    s = np.float32(2.0) * np.asarray(half, np.float32)
    x, y, z = s * s
    k = np.float32(mass) / np.float32(12.0)
    return np.array([k * (y + z), k * (z + x), k * (x + y)], np.float32)


def _dense64(systems, dt):
    """float64 solve of the lumped diagonal plus one rank-one row."""
    s = np.asarray(systems, np.float64)
    out = []
    for r in s:
        inertia, inv_mass = r[0:3], r[3]
        dt2 = np.float64(np.float32(dt)) ** 2
        d = np.concatenate([np.full(3, 1.0 / (inv_mass * dt2)), inertia / dt2])
        nn = r[20] * r[12:15]
        v = np.concatenate([nn, np.cross(r[16:19], nn)])
        a = np.diag(d) + r[15] * np.outer(v, v)
        b = np.concatenate([r[4:7], r[8:11]]) + r[19] * v
        out.append(np.linalg.solve(a, b))
    return np.array(out)


def _masses(inv_mass, inertia):
    inv_mass = np.atleast_1d(np.asarray(inv_mass, np.float32))
    m = np.zeros((inv_mass.shape[0], 4), np.float32)
    m[:, 0:3] = np.broadcast_to(np.asarray(inertia, np.float32), (inv_mass.shape[0], 3))
    m[:, 3] = inv_mass
    return m


def test_diag_solve_known_answers():
    # rung 1: x = rhs / diag, so a unit residual returns 1 / diag
    d = po.stage_diag_solve(_masses(1.0, (2.0, 3.0, 4.0)), pack_states(np.ones(6)))
    diag = 1.0 / _residual(d)[0].astype(np.float64)
    np.testing.assert_allclose(diag, [3600, 3600, 3600, 7200, 10800, 14400], atol=1e-2, rtol=0)
    d = po.stage_diag_solve(_masses(0.1, (1.0, 1.0, 1.0)), pack_states(np.ones(6)))
    heavy = 1.0 / _residual(d)[0].astype(np.float64)
    assert abs(heavy[0] - 36000.0) < 1.0 and abs(heavy[3] - 3600.0) < 1e-2
    assert d[:, [3, 7]].tolist() == [[0.0, 0.0]]


def test_diag_solve_statics_and_random():
    rng = np.random.default_rng(2)
    n = 300
    inv_mass = rng.uniform(0.05, 2.0, n).astype(np.float32)
    inv_mass[::7] = 0.0  # statics: the roster zeroes their inertia too
    inertia = rng.uniform(0.1, 5.0, (n, 3)).astype(np.float32)
    inertia[::7] = 0.0
    rhs = rng.normal(0, 300.0, (n, 6)).astype(np.float32)
    d = po.stage_diag_solve(_masses(inv_mass, inertia), pack_states(rhs))
    x = _residual(d)
    assert not x[::7].any()
    dt2 = np.float64(np.float32(DT)) ** 2
    dyn = inv_mass > 0
    diag = np.concatenate([np.repeat(1.0 / (inv_mass[dyn, None] * dt2), 3, axis=1), inertia[dyn] / dt2], axis=1)
    np.testing.assert_allclose(x[dyn], rhs[dyn] / diag, rtol=1e-6, atol=0)
    # degenerate steps solve to zero everywhere
    for dt in (0.0, -DT, np.nan, 1e-30):
        assert not po.stage_diag_solve(_masses(inv_mass, inertia), pack_states(rhs), dt).any()


def test_hessian_diagonal_known_answers():
    # inert row: only the lumped diagonal is assembled, so x = 1 / diag
    d, ok = _inert(1.0, (2.0, 3.0, 4.0), np.ones(6, np.float32))
    assert ok[0] == 1.0
    diag = 1.0 / _residual(d)[0].astype(np.float64)
    np.testing.assert_allclose(diag, [3600, 3600, 3600, 7200, 10800, 14400], atol=1e-2, rtol=0)
    d, ok = _inert(0.1, (1.0, 1.0, 1.0), np.ones(6, np.float32))
    heavy = 1.0 / _residual(d)[0].astype(np.float64)
    assert abs(heavy[0] - 36000.0) < 1.0 and abs(heavy[3] - 3600.0) < 1e-2


@pytest.mark.parametrize("inv_mass,dt", [(0.0, DT), (-1.0, DT), (np.inf, DT), (np.nan, DT),
                                         (1.0, 0.0), (1.0, -DT), (1.0, np.nan), (1.0, 1e-30)])
def test_statics_and_degenerate_dt_break_down(inv_mass, dt):
    # no diagonal and an inert row: singular, so a zero delta and ok = 0
    d, ok = _inert(inv_mass, (1.0, 1.0, 1.0), np.full(6, 10.0, np.float32), dt)
    assert ok[0] == 0.0 and not d.any()


def test_reference_staged_cases():
    inertia = _box_inertia(2.0, (0.5, 0.5, 0.5))
    # gpu_batches.rs:904-926: a dynamic box with one normal row, a static box
    # with an inert row
    sy = np.concatenate([
        pack_systems(0.5, inertia, np.full(6, 10.0), (0, 1, 0), 500.0, (0.2, 0, 0), -3.0, 1.0),
        pack_systems(0.0, (0, 0, 0), np.full(6, 10.0), **INERT)])
    d, ok = po.stage_contact_solve(sy, DT)
    assert ok.tolist() == [1.0, 0.0]
    assert np.isfinite(d).all() and (np.abs(d[0]) > 1e-9).any() and not d[1].any()
    np.testing.assert_allclose(_residual(d)[0], _dense64(sy[:1], DT)[0], rtol=1e-5, atol=1e-9)
    # gpu_batches.rs:968-1010: O(1e2) residuals against O(1e3) Hessian entries
    sy = np.concatenate([
        pack_systems(0.5, inertia, [360.0, -720.0, 180.0, 90.0, -45.0, 720.0], (0, 1, 0), 1000.0,
                     (0.3, 0, 0.1), -50.0, 1.0),
        pack_systems(0.0, (0, 0, 0), np.zeros(6), **INERT)])
    d, ok = po.stage_contact_solve(sy, DT)
    assert ok.tolist() == [1.0, 0.0] and not d[1].any()
    np.testing.assert_allclose(_residual(d)[0], _dense64(sy[:1], DT)[0], rtol=1e-5, atol=1e-9)
    # the row moves the body: the delta differs from the bare diagonal solve
    bare, _ = _inert(0.5, inertia, [360.0, -720.0, 180.0, 90.0, -45.0, 720.0])
    assert np.abs(_residual(d)[0] - _residual(bare)[0]).max() > 1e-4


def _random_systems(n, seed):
    rng = np.random.default_rng(seed)
    inv_mass = rng.uniform(0.05, 2.0, n)
    inertia = rng.uniform(0.1, 5.0, (n, 3))
    rhs = rng.normal(0, 300.0, (n, 6))
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    pen = 10.0 ** rng.uniform(0, 5, n)
    lever = rng.uniform(-1, 1, (n, 3))
    force = rng.normal(0, 100.0, n)
    sign = rng.choice([-1.0, 1.0], n)
    return pack_systems(inv_mass, inertia, rhs, axis, pen, lever, force, sign)


def test_random_systems_match_dense_float64():
    sy = _random_systems(400, 7)
    d, ok = po.stage_contact_solve(sy, DT)
    assert (ok == 1.0).all()
    x = _residual(d).astype(np.float64)
    ref = _dense64(sy, DT)
    err = np.linalg.norm(x - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-4, err.max()


def _edge_systems():
    inertia = _box_inertia(2.0, (0.5, 0.5, 0.5))
    return np.concatenate([
        pack_systems(0.5, inertia, np.full(6, 10.0), (0, 1, 0), 500.0, (0.2, 0, 0), -3.0, 1.0),
        pack_systems(0.0, (0, 0, 0), np.full(6, 10.0), **INERT),
        pack_systems(np.nan, (1, 1, 1), np.ones(6), **INERT),
        pack_systems(np.inf, (1, 1, 1), np.ones(6), **INERT),
        pack_systems(-1.0, (1, 1, 1), np.ones(6), (1, 0, 0), 100.0, (0, 0.5, 0), 2.0, -1.0),
        # a static body with a stiff row: rank one, so the LDL breaks down
        pack_systems(0.0, (0, 0, 0), np.ones(6), (0, 0, 1), 1e4, (0.5, 0.5, 0), 1.0, 1.0),
    ])


@pytest.mark.gpu
def test_device_diag_solve_bitwise(gpu_available):
    rng = np.random.default_rng(9)
    n = 4000
    inv_mass = rng.uniform(0.05, 2.0, n).astype(np.float32)
    inv_mass[::5] = 0.0
    inv_mass[1::97] = np.nan
    inertia = rng.uniform(0.1, 5.0, (n, 3)).astype(np.float32)
    inertia[::5] = 0.0
    rhs = pack_states(rng.normal(0, 300.0, (n, 6)))
    g = AvbdEngine()
    try:
        for dt in (DT, 1.0 / 240.0, 0.0, np.nan):
            dg = g.solve_diag(_masses(inv_mass, inertia), rhs, dt)
            do = po.stage_diag_solve(_masses(inv_mass, inertia), rhs, dt)
            assert np.array_equal(dg.view(np.uint32), do.view(np.uint32))
        assert g.solve_diag(np.zeros((0, 4)), np.zeros((0, 8))).shape == (0, 8)
    finally:
        g.close()


@pytest.mark.gpu
def test_device_staged_solve_bitwise(gpu_available):
    g = AvbdEngine()
    try:
        for sy, dt in ((_random_systems(5000, 11), DT), (_edge_systems(), DT), (_random_systems(300, 3), 0.0),
                       (_random_systems(64, 5), 1.0 / 240.0)):
            dg, okg = g.solve_systems(sy, dt)
            do, oko = po.stage_contact_solve(sy, dt)
            assert np.array_equal(okg, oko)
            assert np.array_equal(dg.view(np.uint32), do.view(np.uint32))
        dg, okg = g.solve_systems(np.zeros((0, 24), np.float32))
        assert dg.shape == (0, 8) and okg.shape == (0,)
    finally:
        g.close()
