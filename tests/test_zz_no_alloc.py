"""oavbd_solve_systems_noalloc (the allocation-free staged solve) against
oavbd_solve_systems through the C ABI: two calls in a row, the second
smaller, bit for bit; oversized calls are rejected."""
import numpy as np
import pytest

from ornis_amd.avbd import AvbdEngine, AvbdError, pack_systems

pytestmark = pytest.mark.gpu


def _systems(n, seed):
    rng = np.random.default_rng(seed)
    inv_mass = rng.uniform(0.01, 2.0, n).astype(np.float32)
    inv_mass[::17] = 0.0  # statics
    inertia = rng.uniform(0.01, 5.0, (n, 3)).astype(np.float32)
    inertia[::17] = 0.0
    rhs = rng.normal(size=(n, 6)).astype(np.float32)
    axis = rng.normal(size=(n, 3))
    axis = (axis / np.linalg.norm(axis, axis=1, keepdims=True)).astype(np.float32)
    pen = rng.uniform(1.0, 1e5, n).astype(np.float32)
    lever = rng.normal(size=(n, 3)).astype(np.float32)
    force = rng.normal(size=n).astype(np.float32)
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0).astype(np.float32)
    return pack_systems(inv_mass, inertia, rhs, axis, pen, lever, force, sign)


def test_noalloc_equals_allocating_entry_point(gpu_available):
    g = AvbdEngine()
    for n, seed in ((50000, 1), (1234, 2)):
        sysa = _systems(n, seed)
        d_ref, ok_ref = g.solve_systems(sysa)
        d, ok = g.solve_systems_noalloc(sysa)
        assert np.array_equal(d.view(np.uint32), d_ref.view(np.uint32))
        assert np.array_equal(ok.view(np.uint32), ok_ref.view(np.uint32))
        assert ok.sum() > 0.9 * n
    with pytest.raises(AvbdError):
        g.solve_systems_noalloc(np.zeros((65537, 24), np.float32))
    g.close()
