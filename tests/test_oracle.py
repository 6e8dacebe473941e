"""CPU oracle pins: the reference's golden snapshot, its analytic and
behavioural tests, and invariants of the restatement itself."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from ornis_amd import scenes
from scenes_ref import read_golden, snapshot_scene, stack_scene

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "avbd_snapshot.hex")
F = np.float32


def _p(a):
    return a.ctypes.data_as(po._f32p)


@pytest.mark.parametrize("math", [po.MATH_LIBM, po.MATH_PORTABLE])
def test_snapshot_matches_reference_golden_bitwise(math):
    """avbd_snapshot_matches_canonical (crates/physics/tests/avbd_engine.rs:
    1632-1645): 9 bodies x 13 f32 after 120 steps, bit for bit."""
    header, want = read_golden(GOLDEN)
    assert header.startswith("ornis-avbd-snapshot v1 bodies=9 steps=120")
    e = po.OracleEngine(snapshot_scene(), order=po.ORDER_REFERENCE, math=math)
    e.step(1.0 / 60.0, 120)
    got = e.state()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_snapshot_golden_catches_a_planted_mistake():
    """The golden is sharp: an oracle run with one fewer iteration (a
    plausible off-by-one in the sweep loop) no longer matches."""
    _, want = read_golden(GOLDEN)
    e = po.OracleEngine(snapshot_scene(), order=po.ORDER_REFERENCE, iters=9)
    e.step(1.0 / 60.0, 120)
    assert not np.array_equal(e.state().view(np.uint32), want.view(np.uint32))


def test_stack_stands():
    """avbd_stack_stands_like_si (avbd_engine.rs:31-62)."""
    e = po.OracleEngine(stack_scene(), order=po.ORDER_REFERENCE)
    e.step(1.0 / 60.0, 300)
    st = e.state()
    for level in range(4):
        b = st[1 + level]
        assert abs(b[1] - (0.5 + level)) < 0.1
        assert abs(b[0]) < 0.15 and abs(b[2]) < 0.15
        assert np.linalg.norm(b[7:10]) < 0.08
        assert np.linalg.norm(b[10:13]) < 0.08


def test_speculative_damper_momentum_and_energy():
    """avbd_speculative_pair_does_not_create_momentum (avbd_engine.rs:748-779)."""
    sb = scenes.SceneBuilder(gravity=(0.0, 0.0, 0.0))
    sb.sphere((1.01, 0.0, 0.0), 0.5, 1.0, vel=(-1.0, 0.0, 0.0), friction=0.0, restitution=0.0)
    sb.sphere((0.0, 0.0, 0.0), 0.5, 100.0, friction=0.0, restitution=0.0)
    s = sb.build()
    e = po.OracleEngine(s)
    e.step(1.0 / 60.0, 1)
    st = e.state()
    m = np.array([1.0, 100.0])
    p = m[0] * st[0, 7:10] + m[1] * st[1, 7:10]
    assert np.all(np.isfinite(p))
    assert np.linalg.norm(p - np.array([-1.0, 0.0, 0.0])) < 0.05
    energy = 0.5 * (m[0] * np.dot(st[0, 7:10], st[0, 7:10]) + m[1] * np.dot(st[1, 7:10], st[1, 7:10]))
    assert energy <= 0.51


def test_run_to_run_determinism():
    """avbd_determinism_run_to_run (avbd_engine.rs:256-272)."""
    s = scenes.box_drop_scene(60, np.random.Generator(np.random.PCG64(3)))
    outs = []
    for _ in range(2):
        e = po.OracleEngine(s, order=po.ORDER_REFERENCE)
        e.step(1.0 / 60.0, 40)
        outs.append(e.state())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_grid_discovery_equals_brute_force():
    """The grid-accelerated discovery and TOI candidate search give the
    same bits as the reference's O(n^2) scans (n > 256 uses the grid)."""
    s = scenes.box_drop_scene(300, np.random.Generator(np.random.PCG64(5)))
    runs = []
    for brute in (True, False):
        e = po.OracleEngine(s, order=po.ORDER_REFERENCE, brute=brute)
        e.step_inner(25)
        runs.append((e.state(), e.pairs()[0]))
    assert np.array_equal(runs[0][0].view(np.uint32), runs[1][0].view(np.uint32))
    assert np.array_equal(runs[0][1], runs[1][1])


def test_colour_order_close_to_reference_order_on_golden():
    """Colour-ordered sweeps change the Gauss-Seidel order only: on the
    golden scene every pose stays within 1e-3 of the reference result."""
    _, want = read_golden(GOLDEN)
    e = po.OracleEngine(snapshot_scene(), order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    e.step(1.0 / 60.0, 120)
    assert np.max(np.abs(e.state()[:, :7] - want[:, :7])) < 1e-3


def test_colouring_is_proper_and_greedy():
    """Adjacent swept bodies never share a colour, and each body's colour
    is the smallest one unused by its higher-priority neighbours."""
    s = scenes.box_drop_scene(120, np.random.Generator(np.random.PCG64(11)))
    e = po.OracleEngine(s, order=po.ORDER_COLOUR, math=po.MATH_PORTABLE)
    e.step_inner(30)
    col = e.colours()
    ab, npts, _, _ = e.pairs()
    swept = col >= 0
    adj = {}
    for (a, b), k in zip(ab, npts):
        if k > 0 and swept[a] and swept[b]:
            assert col[a] != col[b]
            adj.setdefault(a, set()).add(b)
            adj.setdefault(b, set()).add(a)

    def prio(h):
        h = int(h)
        x = h & 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 0x7FEB352D) & 0xFFFFFFFF
        x ^= x >> 15
        x = (x * 0x846CA68B) & 0xFFFFFFFF
        x ^= x >> 16
        return (x, -h)
    for h, nb in adj.items():
        higher = {col[o] for o in nb if prio(o) > prio(h)}
        c = 0
        while c in higher:
            c += 1
        assert col[h] == c


def test_fp64_oracle_tracks_fp32_on_stack():
    e32 = po.OracleEngine(stack_scene(), precision=32)
    e64 = po.OracleEngine(stack_scene(), precision=64)
    e32.step(1.0 / 60.0, 120)
    e64.step(1.0 / 60.0, 120)
    assert np.max(np.abs(e32.state()[:, :3] - e64.state_f64()[:, :3])) < 1e-2


# --- analytic pins (crates/physics/tests/gpu_batches.rs) -------------------

def test_solve_6x6_matches_dense_and_breaks_down():
    """avbd_ldl_6x6_cpu_matches_dense_and_signals_breakdown (gpu_batches.rs:863-898)."""
    L = po.lib()
    lhs = np.zeros((6, 6), F)
    for i in range(6):
        for j in range(6):
            lhs[i, j] = 4.0 + i if i == j else 0.1 * ((i + j) + 1.0)
    lhs = np.triu(lhs) + np.triu(lhs, 1).T
    rhs = np.array([1, -2, 3, -4, 5, -6], F)
    x = np.zeros(6, F)
    assert L.orc_solve_6x6_f32(_p(lhs), _p(rhs), _p(x)) == 1
    np.testing.assert_allclose(x, np.linalg.solve(lhs.astype(np.float64), rhs.astype(np.float64)), rtol=1e-5)
    assert L.orc_solve_6x6_f32(_p(np.zeros((6, 6), F)), _p(rhs), _p(x)) == 0
    neg = lhs.copy()
    neg[0, 0] = -1.0
    assert L.orc_solve_6x6_f32(_p(neg), _p(rhs), _p(x)) == 0
    # identity solves exactly (joint_tests.rs:351)
    assert L.orc_solve_6x6_f32(_p(np.eye(6, dtype=F)), _p(rhs), _p(x)) == 1
    assert np.array_equal(x, rhs)


def test_stamp_row_matches_analytic_row():
    """avbd_stamp_row_cpu_matches_analytic_row (gpu_batches.rs:781-856)."""
    L = po.lib()
    for sign in (1.0, -1.0):
        lhs = np.zeros((6, 6), F)
        rhs = np.zeros(6, F)
        L.orc_stamp_row_f32(_p(lhs), _p(rhs), _p(np.array([0, 1, 0], F)), 2.0, 3.0, _p(np.array([1, 0, 0], F)),
                            sign)
        el = np.zeros((6, 6), F)
        el[1, 1] = el[5, 5] = el[1, 5] = el[5, 1] = 2.0
        er = np.zeros(6, F)
        er[1] = er[5] = 3.0 * sign
        assert np.array_equal(lhs, el)
        assert np.array_equal(rhs, er)


def test_contact_force_known_answers():
    """Audit known answers for the fixed rules
    (docs/quality/avbd-audit-2026-09-15.md:135-159, :204-212)."""
    L = po.lib()
    f = np.zeros(3, F)
    # A3: K=100, C=(-0.01, 0.1, 0), lambda=(-9, 4, 0), mu=0.5 -> (-10, 5, 0)
    L.orc_contact_force_f32(F(-0.01), _p(np.full(3, 100, F)), _p(np.array([-9, 4, 0], F)), F(0.1), F(0.0),
                            _p(np.array([0.5, 0.5], F)), _p(f))
    np.testing.assert_allclose(f, [-10.0, 5.0, 0.0], rtol=1e-6)
    # A6: mu=(0, 0.5), raw tangent (0, 10), Fn=-10 -> (0, 5)
    L.orc_contact_force_f32(F(-10.0), _p(np.ones(3, F)), _p(np.zeros(3, F)), F(0.0), F(10.0),
                            _p(np.array([0.0, 0.5], F)), _p(f))
    np.testing.assert_allclose(f, [-10.0, 0.0, 5.0], rtol=1e-6)
    # push-only normal: a separating violation gives no force
    L.orc_contact_force_f32(F(0.5), _p(np.full(3, 100, F)), _p(np.zeros(3, F)), F(0.0), F(0.0),
                            _p(np.array([0.5, 0.5], F)), _p(f))
    assert np.array_equal(f, np.zeros(3, F))


def test_tangent_basis_orthonormal():
    """contact_tests.rs:172-186."""
    L = po.lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(F)
        t1, t2 = np.zeros(3, F), np.zeros(3, F)
        L.orc_tangent_basis_f32(_p(n), _p(t1), _p(t2))
        for a, b in ((t1, n), (t2, n), (t1, t2)):
            assert abs(np.dot(a, b)) < 1e-6
        assert abs(np.linalg.norm(t1) - 1) < 1e-6 and abs(np.linalg.norm(t2) - 1) < 1e-6


@pytest.mark.parametrize("math", [po.MATH_LIBM, po.MATH_PORTABLE])
def test_quat_integrate_and_diff_are_inverse(math):
    """exp-map integration and angle-axis recovery invert each other
    (rows.rs:110-146), and the rotation is exact (axis-angle closed form)."""
    L = po.lib()
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rng.normal(size=4)
        q = (q / np.linalg.norm(q)).astype(F)
        u = rng.normal(size=3)
        v = (u / np.linalg.norm(u) * rng.uniform(1e-4, 2.5)).astype(F)
        q2, back = np.zeros(4, F), np.zeros(3, F)
        L.orc_quat_integrate_f32(_p(q), _p(v), math, _p(q2))
        L.orc_quat_diff_vec_f32(_p(q2), _p(q), math, _p(back))
        np.testing.assert_allclose(back, v, atol=2e-5 * max(1.0, np.linalg.norm(v)))
        th = np.linalg.norm(v.astype(np.float64))
        ax = v / th
        dq = np.concatenate([ax * np.sin(th / 2), [np.cos(th / 2)]])
        x1, y1, z1, w1 = dq
        x2, y2, z2, w2 = q.astype(np.float64)
        ref = np.array([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                        w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])
        np.testing.assert_allclose(q2, ref, atol=3e-6)


def _ulp_err(got, want):
    got = got.astype(np.float64)
    sp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    return np.abs(got - want) / sp


def test_portable_sincos_asin_accuracy():
    """The portable float transcendentals stay within 2 ulp of the
    correctly rounded values (x in [-50, 50] for sin/cos, [-1, 1] for asin)."""
    L = po.lib()
    xs = np.concatenate([np.linspace(-50, 50, 40001), np.geomspace(1e-8, 1.6, 4000)]).astype(F)
    s, c = np.zeros(1, F), np.zeros(1, F)
    es, ec = [], []
    for x in xs[::7]:
        L.orc_fmath_sincos(F(x), _p(s), _p(c))
        es.append(_ulp_err(s, np.sin(np.float64(x)))[0])
        ec.append(_ulp_err(c, np.cos(np.float64(x)))[0])
    assert max(es) <= 2.0 and max(ec) <= 2.0
    ea = []
    for x in np.concatenate([np.linspace(-1, 1, 20001), np.geomspace(1e-8, 1.0, 3000)]).astype(F):
        a = L.orc_fmath_asin(F(x))
        ea.append(_ulp_err(np.array([a], F), np.arcsin(np.float64(x)))[0])
    assert max(ea) <= 2.0


def test_portable_atan2_accuracy():
    """oavbd_atan2f within 2 ulp of the correctly rounded atan2 over every
    octant and magnitude ratio, with C's signed-zero, infinity and NaN
    conventions."""
    L = po.lib()
    rng = np.random.default_rng(3)
    ang = rng.uniform(-np.pi, np.pi, 20000)
    mag = np.exp(rng.uniform(-12, 4, 20000))
    ys = (mag * np.sin(ang)).astype(F)
    xs = (mag * np.cos(ang)).astype(F)
    extra = np.array([(1.0, 1.0), (1.0, -1.0), (-1.0, -1.0), (1e-30, 1.0), (1.0, 1e-30), (0.41421356, 1.0),
                      (0.4142136, 1.0), (3.0, -1e-7)], F)
    ys, xs = np.concatenate([ys, extra[:, 0]]), np.concatenate([xs, extra[:, 1]])
    err = max(_ulp_err(np.array([L.orc_fmath_atan2(F(y), F(x))], F),
                       np.arctan2(np.float64(y), np.float64(x)))[0] for y, x in zip(ys, xs))
    assert err <= 2.0, err
    for y, x in ((0.0, 0.0), (-0.0, 0.0), (0.0, -0.0), (-0.0, -0.0), (0.0, -1.0), (-0.0, -1.0), (2.0, 0.0),
                 (-2.0, -0.0), (np.inf, 1.0), (1.0, -np.inf), (np.inf, np.inf), (-np.inf, -np.inf)):
        got, want = L.orc_fmath_atan2(F(y), F(x)), np.float32(np.arctan2(F(y), F(x)))
        assert got == want and np.signbit(got) == np.signbit(want), (y, x, got, want)
    assert np.isnan(L.orc_fmath_atan2(F(np.nan), F(1.0)))


def test_fp32_tracks_fp64_on_ordered_scenes():
    """FP32-vs-FP64 drift of the step (the GPU step equals the FP32 oracle
    bit for bit, so this is the GPU path's drift too): within 1e-3 m on the
    reference snapshot scene after 120 steps and 1e-4 m on a resting stack
    after 300.  Chaotic piles diverge by design; tests/fp64_drift.py reports
    them."""
    from fp64_drift import drift
    assert drift(snapshot_scene(), 120)["pos_max"] < 1e-3
    assert drift(stack_scene(4), 300)["pos_max"] < 1e-4
