"""The derivatives of the differentiable step, one piece at a time: every
function of ornis_amd/csrc/avbd_adjoint.hpp, built for the host in double
precision (tests/adjoint_prims.cpp), against central finite differences of
the forward function it differentiates, on seeded random inputs covering its
branches (touching and separated contacts, either side, clamped and free
forces, active and inactive penalty updates)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
D = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def L(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("adjoint") / "libadjoint_prims.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", so, os.path.join(HERE, "adjoint_prims.cpp")],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.ap_iform.restype = ctypes.c_double
    lib.ap_side_S.restype = ctypes.c_double
    lib.ap_inertial_S.restype = ctypes.c_double
    lib.ap_inertial_S.argtypes = [D, D, D, ctypes.c_double]
    lib.ap_inertial_vjp.argtypes = [D, D, D, ctypes.c_double, D]
    return lib


def p(a):
    return a.ctypes.data_as(D)


def f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def fd_vjp(fn, x, o, h=1e-6):
    """sum_j o_j d f_j / d x_i by central differences."""
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        step = h * max(1.0, abs(x[i]))
        xp[i] += step
        xm[i] -= step
        g[i] = np.dot(np.atleast_1d(fn(xp)) - np.atleast_1d(fn(xm)), o) / (2 * step)
    return g


def close(a, b, rtol=2e-6, atol=1e-7):
    scale = max(np.abs(b).max(), 1.0)
    assert np.abs(a - b).max() <= rtol * scale + atol, (a, b)


def unit(rng, n=4):
    v = rng.normal(size=n)
    return v / np.linalg.norm(v)


def test_quaternion_pieces(L):
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = f64(unit(rng) * rng.uniform(0.8, 1.2))
        v = f64(rng.normal(size=3))
        o3 = f64(rng.normal(size=3))
        o4 = f64(rng.normal(size=4))

        def qrot(x):
            out = np.zeros(3)
            L.ap_qrot(p(f64(x[:4])), p(f64(x[4:])), p(out))
            return out
        qb, vb = np.zeros(4), np.zeros(3)
        L.ap_qrot_vjp(p(q), p(v), p(o3), p(qb), p(vb))
        close(np.concatenate([qb, vb]), fd_vjp(qrot, np.concatenate([q, v]), o3))

        w = f64(rng.normal(size=3) * rng.uniform(0.01, 1.0))

        def qint(x):
            out = np.zeros(4)
            L.ap_qint(p(f64(x[:4])), p(f64(x[4:])), p(out))
            return out
        qb, vb = np.zeros(4), np.zeros(3)
        L.ap_qint_vjp(p(q), p(w), p(o4), p(qb), p(vb))
        close(np.concatenate([qb, vb]), fd_vjp(qint, np.concatenate([q, w]), o4))

        # relative rotations below about 115 degrees (asin's derivative is
        # singular at 180; the step compares nearby rotations)
        a = unit(rng)
        b = a + rng.normal(size=4) * 0.4
        a, b = f64(a), f64(b / np.linalg.norm(b))

        def qdiff(x):
            out = np.zeros(3)
            L.ap_qdiff(p(f64(x[:4])), p(f64(x[4:])), p(out))
            return out
        ab, bb = np.zeros(4), np.zeros(4)
        L.ap_qdiff_vjp(p(a), p(b), p(o3), p(ab), p(bb))
        close(np.concatenate([ab, bb]), fd_vjp(qdiff, np.concatenate([a, b]), o3), rtol=1e-5)

        def qnorm(x):
            out = np.zeros(4)
            L.ap_qnorm(p(f64(x)), p(out))
            return out
        qb = np.zeros(4)
        L.ap_qnorm_vjp(p(q), p(o4), p(qb))
        close(qb, fd_vjp(qnorm, q, o4))

        Dg, wv, zv = f64(rng.uniform(0.1, 2.0, 3)), f64(rng.normal(size=3)), f64(rng.normal(size=3))

        def iform(x):
            return L.ap_iform(p(f64(x[:4])), p(f64(x[4:7])), p(f64(x[7:10])), p(f64(x[10:13])))
        qb, Db, wb, zb = np.zeros(4), np.zeros(3), np.zeros(3), np.zeros(3)
        L.ap_iform_vjp(p(q), p(Dg), p(wv), p(zv), p(qb), p(Db), p(wb), p(zb))
        close(np.concatenate([qb, Db, wb, zb]), fd_vjp(iform, np.concatenate([q, Dg, wv, zv]), np.ones(1)))


def test_contact_force_branches(L):
    rng = np.random.default_rng(2)
    seen = set()
    for i in range(400):
        cn = rng.uniform(-0.05, 0.02)
        ct = rng.normal(size=2) * rng.choice([1e-4, 1e-2, 1.0])
        lam = np.array([rng.uniform(-20, 5), rng.normal() * 5, rng.normal() * 5])
        pen = rng.uniform(1.0, 1e4, 3)
        mu = rng.uniform(0.0, 0.9, 2) if i % 7 else np.array([0.0, rng.uniform(0.1, 0.9)])
        x = f64(np.concatenate([[cn], ct, lam, pen, mu]))

        def force(z):
            out = np.zeros(3)
            L.ap_force(p(f64(z)), p(out))
            return out
        f = force(x)
        a = x[6] * x[0] + x[3]
        b1, b2 = abs(f[0]) * x[9], abs(f[0]) * x[10]
        if b1 > 0 and b2 > 0:
            s = ((x[7] * x[1] + x[4]) / b1) ** 2 + ((x[8] * x[2] + x[5]) / b2) ** 2
            if abs(s - 1.0) < 1e-3:
                continue  # on the cone: the branch of a finite difference may differ
            seen.add("cone" if s > 1 else "inside")
        else:
            seen.add("degenerate")
        if abs(a) < 1e-6:
            continue
        o = f64(rng.normal(size=3))
        g = np.zeros(11)
        L.ap_force_vjp(p(x), p(o), p(g))
        want = fd_vjp(force, x, o, h=1e-7)
        keep = [k for k in range(11) if not (k >= 9 and x[k] == 0.0)]  # mu = 0 is a branch boundary
        close(g[keep], want[keep], rtol=1e-5, atol=1e-6)
    assert {"cone", "inside", "degenerate"} <= seen


def random_contact(rng, touching, sphere_a=False, sphere_b=False):
    n = unit(rng, 3)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t1 = np.cross(n, a)
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(t1, n)
    mu = rng.uniform(0.2, 0.9)
    damp = rng.uniform(10.0, 1e4)
    ra, rb = rng.normal(size=3) * 0.3, rng.normal(size=3) * 0.3
    rad_a = rng.uniform(0.2, 0.5) if sphere_a else -1.0
    rad_b = rng.uniform(0.2, 0.5) if sphere_b else -1.0
    xa0, xb0 = rng.normal(size=3), rng.normal(size=3)
    xa, xb = xa0 + rng.normal(size=3) * 0.02, xb0 + rng.normal(size=3) * 0.02
    qa0, qb0 = unit(rng), unit(rng)
    qa = qa0 + rng.normal(size=4) * 0.02
    qb = qb0 + rng.normal(size=4) * 0.02
    qa, qb = qa / np.linalg.norm(qa), qb / np.linalg.norm(qb)
    lam = np.array([rng.uniform(-30, 0), rng.normal() * 3, rng.normal() * 3])
    pen = rng.uniform(1.0, 5e4, 3)
    return f64(np.concatenate([n, t1, t2, [mu, mu, damp, 1.0 if touching else 0.0], ra, rb, [rad_a, rad_b], xa, xb,
                               xa0, xb0, qa, qb, qa0, qb0, lam, pen]))


# the differentiable entries of a contact, in the adjoint's layout:
# contact[21:55] (poses, duals) then mu1, mu2, damp
LIVE = list(range(21, 55)) + [9, 10, 11]


def test_point_side_rows(L):
    rng = np.random.default_rng(3)
    n_ok = 0
    for i in range(300):
        touching = i % 4 != 0
        c = random_contact(rng, touching, sphere_a=(i % 9 == 3), sphere_b=(i % 11 == 5))
        w, dx = f64(rng.normal(size=6)), f64(rng.normal(size=6) * 0.01)
        is_a = i % 2

        def S(z):
            cc = c.copy()
            cc[LIVE] = z
            return L.ap_side_S(p(f64(cc)), is_a, p(w), p(dx))
        x = c[LIVE].copy()
        # skip configurations on a branch boundary (a clamp or the cone)
        v0, vp, vm = S(x), S(x + 1e-5), S(x - 1e-5)
        bar = np.zeros(37)
        L.ap_side_vjp(p(c), is_a, p(w), p(dx), p(bar))
        g = fd_vjp(S, x, np.ones(1), h=1e-7)
        if abs(vp - 2 * v0 + vm) > 1e-4 * (abs(v0) + 1.0):
            continue
        close(-bar, g, rtol=1e-5, atol=1e-6)
        n_ok += 1
    assert n_ok > 200


def test_dual_update(L):
    rng = np.random.default_rng(4)
    n_ok = 0
    for i in range(300):
        c = random_contact(rng, touching=i % 5 != 0)
        lb, pb = f64(rng.normal(size=3)), f64(rng.normal(size=3) * 1e-3)

        def dual(z):
            cc = c.copy()
            cc[LIVE] = z
            lo, po = np.zeros(3), np.zeros(3)
            L.ap_dual(p(f64(cc)), p(lo), p(po))
            return np.concatenate([lo, po])
        x = c[LIVE].copy()
        o = np.concatenate([lb, pb])
        vp, vm, v0 = dual(x + 1e-5), dual(x - 1e-5), dual(x)
        if np.abs(vp - 2 * v0 + vm).max() > 1e-4 * (np.abs(v0).max() + 1.0):
            continue
        bar = np.zeros(37)
        L.ap_dual_vjp(p(c), p(lb), p(pb), p(bar))
        close(bar, fd_vjp(dual, x, o, h=1e-7), rtol=1e-5, atol=1e-6)
        n_ok += 1
    assert n_ok > 200


def test_inertial_terms(L):
    rng = np.random.default_rng(5)
    for _ in range(100):
        q, qy = unit(rng), unit(rng)
        qy = q + 0.1 * (qy - q)
        x = f64(np.concatenate([rng.normal(size=3), q, rng.normal(size=3), qy / np.linalg.norm(qy),
                                [rng.uniform(0.5, 3.0)], rng.uniform(0.01, 0.2, 3)]))
        w, dx = f64(rng.normal(size=6)), f64(rng.normal(size=6) * 0.01)
        dt = 1.0 / 60.0

        def S(z):
            return L.ap_inertial_S(p(f64(z)), p(w), p(dx), dt)
        bar = np.zeros(15)
        L.ap_inertial_vjp(p(x), p(w), p(dx), dt, p(bar))
        close(-bar, fd_vjp(S, x, np.ones(1), h=1e-7)[:15], rtol=1e-5, atol=1e-4)
