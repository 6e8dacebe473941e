"""ctypes front end of the CPU oracle (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It drives oracle/liboavbd_oracle.so, the C++ restatement of
the reference AVBD engine (crates/physics/src/avbd/*.rs).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboavbd_oracle.so")
_lib = None

ORDER_REFERENCE = 0
ORDER_COLOUR = 1
MATH_LIBM = 0
MATH_PORTABLE = 1

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_add_bodies.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                     _f32p, _f32p, _f32p, _u32p, _u32p, _u8p, _i32p, _f32p, _f32p]
        L.orc_add_joint.restype = ctypes.c_int
        L.orc_add_joint.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p,
                                    _f32p]
        L.orc_get_joint_limit_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p]
        L.orc_set_joint_limit_state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.orc_set_gravity.argtypes = [ctypes.c_void_p, _f32p]
        L.orc_wake_body.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_joint_sacc.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p]
        L.orc_set_vel.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, _f32p]
        L.orc_add_joint_prm.restype = ctypes.c_int
        L.orc_add_joint_prm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f32p, _f32p,
                                        _f32p, _f32p, _f32p]
        L.orc_add_ball_joint.restype = ctypes.c_int
        L.orc_add_ball_joint.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_int]
        L.orc_step_inner.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_body_count.restype = ctypes.c_int
        L.orc_body_count.argtypes = [ctypes.c_void_p]
        L.orc_get_state.argtypes = [ctypes.c_void_p, _f32p]
        L.orc_get_state_f64.argtypes = [ctypes.c_void_p, _f64p]
        L.orc_pair_count.restype = ctypes.c_int
        L.orc_pair_count.argtypes = [ctypes.c_void_p]
        L.orc_get_pairs.argtypes = [ctypes.c_void_p, _u32p, _u32p, _f32p, _f32p]
        L.orc_get_points.argtypes = [ctypes.c_void_p, _f32p]
        L.orc_get_colours.argtypes = [ctypes.c_void_p, _i32p]
        L.orc_get_sleep.argtypes = [ctypes.c_void_p, _u8p, _f32p]
        L.orc_get_joints.argtypes = [ctypes.c_void_p, _f32p]
        L.orc_get_stats.argtypes = [ctypes.c_void_p, _i32p]
        L.orc_get_toi.argtypes = [ctypes.c_void_p, _f32p, _i32p]
        L.orc_stage_diag_solve_f32.argtypes = [_f32p, _f32p, ctypes.c_uint32, ctypes.c_float, _f32p]
        L.orc_stage_contact_solve_f32.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_float, _f32p, _f32p]
        L.orc_solve_6x6_f32.restype = ctypes.c_int
        L.orc_solve_6x6_f32.argtypes = [_f32p, _f32p, _f32p]
        L.orc_stamp_row_f32.argtypes = [_f32p, _f32p, _f32p, ctypes.c_float, ctypes.c_float, _f32p, ctypes.c_float]
        L.orc_contact_force_f32.argtypes = [ctypes.c_float, _f32p, _f32p, ctypes.c_float, ctypes.c_float, _f32p, _f32p]
        L.orc_quat_integrate_f32.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p]
        L.orc_quat_diff_vec_f32.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p]
        L.orc_tangent_basis_f32.argtypes = [_f32p, _f32p, _f32p]
        L.orc_shape_distance_f32.argtypes = [ctypes.c_int, _f32p, _f32p, _f32p, ctypes.c_int, _f32p, _f32p, _f32p, _f32p]
        L.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_friction_axes.argtypes = [ctypes.c_void_p, _f32p, _f32p, ctypes.c_int]
        L.orc_drain_events.restype = ctypes.c_int
        L.orc_drain_events.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p, _f32p]
        L.orc_get_event_state.restype = ctypes.c_int
        L.orc_get_event_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p]
        L.orc_set_event_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_int]
        L.orc_set_freeze.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_frozen_steps.restype = ctypes.c_int
        L.orc_frozen_steps.argtypes = [ctypes.c_void_p]
        L.orc_frozen_toi.restype = ctypes.c_int
        L.orc_frozen_toi.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_frozen_pairs.restype = ctypes.c_int
        L.orc_frozen_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_int]
        L.orc_copy_frozen.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_set_pose.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, _f32p]
        L.orc_set_step_start.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, _f32p]
        L.orc_solve_body.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_dual_update.argtypes = [ctypes.c_void_p]
        L.orc_set_joint_duals.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p]
        L.orc_set_inertial.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, _f32p]
        L.orc_get_inertia.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p]
        L.orc_remove_body.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_contacts.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p, _f32p, _f32p, _u32p, _f32p]
        L.orc_set_sleep.argtypes = [ctypes.c_void_p, _u8p, _f32p]
        L.orc_set_params_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p, ctypes.c_double, ctypes.c_double]
        L.orc_fmath_sincos.argtypes = [ctypes.c_float, _f32p, _f32p]
        L.orc_fmath_asin.restype = ctypes.c_float
        L.orc_fmath_asin.argtypes = [ctypes.c_float]
        L.orc_fmath_atan2.restype = ctypes.c_float
        L.orc_fmath_atan2.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def quat_integrate(q, v, math=MATH_LIBM):
    """quat_integrate (rows.rs:110-119) in float32."""
    out = np.zeros(4, np.float32)
    lib().orc_quat_integrate_f32(_p(f32(q), _f32p), _p(f32(v), _f32p), int(math), _p(out, _f32p))
    return out


def stage_diag_solve(masses, rhs, dt=1.0 / 60.0):
    """Rung-1 staged solve over (n, 4) mass records and (n, 8) residual
    records; returns the (n, 8) deltas."""
    m = f32(masses).reshape(-1, 4)
    r = f32(rhs).reshape(-1, 8)
    assert m.shape[0] == r.shape[0]
    delta = np.zeros((m.shape[0], 8), np.float32)
    lib().orc_stage_diag_solve_f32(_p(m, _f32p), _p(r, _f32p), m.shape[0], ctypes.c_float(dt), _p(delta, _f32p))
    return delta


def stage_contact_solve(systems, dt=1.0 / 60.0):
    """Rung-2 staged solve over (n, 24) packed system records; returns
    ((n, 8) deltas as body-state records, (n,) ok flags)."""
    sysa = f32(systems).reshape(-1, 24)
    n = sysa.shape[0]
    delta = np.zeros((n, 8), np.float32)
    ok = np.zeros(n, np.float32)
    lib().orc_stage_contact_solve_f32(_p(sysa, _f32p), n, ctypes.c_float(dt), _p(delta, _f32p), _p(ok, _f32p))
    return delta, ok


class OracleEngine:
    """CPU AVBD engine: ORDER_REFERENCE (the reference's sweep) or
    ORDER_COLOUR (the GPU step's order), float32 or float64."""

    def __init__(self, scene, order=ORDER_REFERENCE, math=MATH_LIBM, precision=32, iters=10, brute=False):
        self._L = lib()
        g = f32(scene.get("gravity", (0.0, -9.81, 0.0)))
        self.h = self._L.orc_create(int(precision), _p(g, _f32p), int(order), int(math), int(iters), int(brute))
        self.n = 0
        self.add_bodies(scene)
        for jt in scene.get("joints", []):
            self.add_joint(*jt)

    def add_bodies(self, s):
        n = len(s["shape"])
        shape = np.ascontiguousarray(s["shape"], dtype=np.int32)
        arrs = [f32(s[k]) for k in ("shape_param", "pos", "quat", "vel", "angvel", "mass", "friction", "restitution")]
        layer = np.ascontiguousarray(s.get("layer", np.ones(n)), dtype=np.uint32)
        mask = np.ascontiguousarray(s.get("mask", np.full(n, 0xFFFFFFFF)), dtype=np.uint32)
        trig = np.ascontiguousarray(s.get("trigger", np.zeros(n)), dtype=np.uint8)
        btype = np.ascontiguousarray(s["body_type"], dtype=np.int32)
        roll = f32(s.get("rolling", np.zeros(n)))
        tors = f32(s.get("torsion", np.zeros(n)))
        self._L.orc_add_bodies(self.h, n, _p(shape, _i32p), *[_p(a, _f32p) for a in arrs], _p(layer, _u32p),
                               _p(mask, _u32p), _p(trig, _u8p), _p(btype, _i32p), _p(roll, _f32p), _p(tors, _f32p))
        self.n += n

    def set_pose(self, i, pos, quat):
        """Overwrite body i's pose (a driver teleport)."""
        self._L.orc_set_pose(self.h, int(i), _p(f32(pos), _f32p), _p(f32(quat), _f32p))

    # unit-pin hooks: single body solves and dual updates outside a step
    def set_step_start(self, i, pos, quat):
        self._L.orc_set_step_start(self.h, int(i), _p(f32(pos), _f32p), _p(f32(quat), _f32p))

    def set_inertial(self, i, pos, quat):
        self._L.orc_set_inertial(self.h, int(i), _p(f32(pos), _f32p), _p(f32(quat), _f32p))

    def set_joint_duals(self, j, d12):
        self._L.orc_set_joint_duals(self.h, int(j), _p(f32(d12), _f32p))

    def joint_limit_state(self, j):
        """(acc_lim, lim_dual) of a Revolute / Prismatic joint."""
        out = np.zeros(2, np.float32)
        self._L.orc_get_joint_limit_state(self.h, int(j), _p(out, _f32p))
        return out

    def set_joint_limit_state(self, j, acc_lim, lim_dual):
        self._L.orc_set_joint_limit_state(self.h, int(j), ctypes.c_float(acc_lim), ctypes.c_float(lim_dual))

    def joint_sacc(self, j):
        """A SixDof joint's limited-axis accumulators (linear xyz, angular xyz)."""
        out = np.zeros(6, np.float32)
        self._L.orc_get_joint_sacc(self.h, int(j), _p(out, _f32p))
        return out

    def wake_body(self, i):
        self._L.orc_wake_body(self.h, int(i))

    def set_gravity(self, g):
        self._L.orc_set_gravity(self.h, _p(f32(g), _f32p))

    def set_vel(self, i, vel, angvel=(0.0, 0.0, 0.0)):
        self._L.orc_set_vel(self.h, int(i), _p(f32(vel), _f32p), _p(f32(angvel), _f32p))

    def solve_body(self, i):
        self._L.orc_solve_body(self.h, int(i))

    def dual_update(self):
        self._L.orc_dual_update(self.h)

    def inertia(self, i):
        out = np.zeros(3, np.float32)
        self._L.orc_get_inertia(self.h, int(i), _p(out, _f32p))
        return out

    def remove_body(self, h):
        """Swap-remove body h (the reference's remove_body)."""
        if 0 <= int(h) < self.n:
            self._L.orc_remove_body(self.h, int(h))
            self.n -= 1

    def set_contacts(self, pairs, points):
        """Replace the pair table by the library's export form
        (oavbd_get_contacts): structured arrays with fields a, b,
        first_point, npoints, normal (xyz + gap), friction, and ra, rb, lam,
        pen (w: stuck), roll."""
        ab = np.ascontiguousarray(np.stack([pairs["a"], pairs["b"]], axis=1), dtype=np.uint32)
        n4 = np.ascontiguousarray(pairs["normal"], dtype=np.float32)
        mu2 = np.ascontiguousarray(pairs["friction"][:, 0:2], dtype=np.float32)
        pts = np.ascontiguousarray(np.stack([pairs["first_point"], pairs["npoints"]], axis=1), dtype=np.uint32)
        pt = np.ascontiguousarray(np.concatenate([points[k] for k in ("ra", "rb", "lam", "pen", "roll")], axis=1),
                                  dtype=np.float32)
        self._L.orc_set_contacts(self.h, len(ab), _p(ab, _u32p), _p(n4, _f32p), _p(mu2, _f32p), _p(pts, _u32p),
                                 _p(pt, _f32p))

    def set_sleep(self, asleep, timer):
        """Sleep flags and timers of every body (sleeping bodies take the
        state sleep_body leaves)."""
        a = np.ascontiguousarray(asleep, dtype=np.uint8)
        t = f32(timer)
        self._L.orc_set_sleep(self.h, _p(a, _u8p), _p(t, _f32p))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self._L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_threads(self, threads):
        """Pair discovery on `threads` threads (bit-identical results)."""
        self._L.orc_set_threads(self.h, int(threads))

    def record_active_sets(self):
        """Record every following step's active set (colour order only)."""
        self._L.orc_set_freeze(self.h, 1)

    def replay_active_sets(self, source=None):
        """Step through a recorded active set (this engine's own, or
        `source`'s) from its first step; step_inner then replays."""
        if source is not None:
            self._L.orc_copy_frozen(self.h, source.h)
        self._L.orc_set_freeze(self.h, 2)

    def set_params(self, vel=None, mass=None, friction=None):
        """Exact (float64) initial velocities (n x 3), masses and friction
        coefficients (n) of every body."""
        for i in range(self.n):
            v = None if vel is None else np.ascontiguousarray(vel[i], dtype=np.float64)
            self._L.orc_set_params_f64(self.h, i, _p(v, _f64p), float(mass[i]) if mass is not None else -1.0,
                                       float(friction[i]) if friction is not None else float("nan"))

    def recorded_steps(self):
        return self._L.orc_frozen_steps(self.h)

    def recorded_toi_hits(self, k):
        return self._L.orc_frozen_toi(self.h, int(k))

    def recorded_pairs(self, k):
        m = self._L.orc_frozen_pairs(self.h, int(k), None, 0)
        ab = np.zeros((m, 2), np.uint32)
        self._L.orc_frozen_pairs(self.h, int(k), _p(ab, _u32p), m)
        return ab

    def step(self, dt=1.0 / 60.0, nsteps=1):
        self._L.orc_step(self.h, ctypes.c_float(dt), int(nsteps))

    def step_inner(self, nsteps=1):
        self._L.orc_step_inner(self.h, int(nsteps))

    def state(self):
        out = np.zeros((self.n, 13), np.float32)
        self._L.orc_get_state(self.h, _p(out, _f32p))
        return out

    def state_f64(self):
        out = np.zeros((self.n, 13), np.float64)
        self._L.orc_get_state_f64(self.h, _p(out, _f64p))
        return out

    def pairs(self):
        m = self._L.orc_pair_count(self.h)
        ab = np.zeros((m, 2), np.uint32)
        npts = np.zeros(m, np.uint32)
        gap = np.zeros(m, np.float32)
        nrm = np.zeros((m, 3), np.float32)
        self._L.orc_get_pairs(self.h, _p(ab, _u32p), _p(npts, _u32p), _p(gap, _f32p), _p(nrm, _f32p))
        return ab, npts, gap, nrm

    def points(self):
        _, npts, _, _ = self.pairs()
        out = np.zeros((int(npts.sum()), 16), np.float32)
        self._L.orc_get_points(self.h, _p(out, _f32p))
        return out

    def colours(self):
        out = np.zeros(self.n, np.int32)
        self._L.orc_get_colours(self.h, _p(out, _i32p))
        return out

    def sleep(self):
        a = np.zeros(self.n, np.uint8)
        t = np.zeros(self.n, np.float32)
        self._L.orc_get_sleep(self.h, _p(a, _u8p), _p(t, _f32p))
        return a, t

    def add_joint(self, a, b, la, lb, kind=0, axis_a=(1.0, 0.0, 0.0), axis_b=(1.0, 0.0, 0.0), prm=None):
        """add_joint: kind 0 Ball, 1 Revolute, 2 Prismatic, 3 Fixed, 4 Distance,
        5 Rope, 6 Spring (local anchors la, lb; local hinge / slide axes for
        kinds 1 and 2; prm for Revolute / Prismatic (limit lo, hi, motor
        speed, budget, has limit, has motor), Rope (max distance) and
        Spring (target position, stiffness, damping, max force, target
        velocity, motor kind, motor model))."""
        if prm is None:
            r = self._L.orc_add_joint(self.h, int(kind), int(a), int(b), _p(f32(la), _f32p), _p(f32(lb), _f32p),
                                      _p(f32(axis_a), _f32p), _p(f32(axis_b), _f32p))
        else:
            pr = np.zeros(24, np.float32)
            pr[:len(prm)] = prm
            r = self._L.orc_add_joint_prm(self.h, int(kind), int(a), int(b), _p(f32(la), _f32p),
                                          _p(f32(lb), _f32p), _p(f32(axis_a), _f32p), _p(f32(axis_b), _f32p),
                                          _p(pr, _f32p))
        if r < 0:
            raise ValueError("bad joint")
        return r

    def joints(self, nj):
        """Joint duals, nj x 12: lam_l3 pen_l3 lam_a3 pen_a3."""
        out = np.zeros((nj, 12), np.float32)
        self._L.orc_get_joints(self.h, _p(out, _f32p))
        return out

    def toi(self):
        t = np.zeros(self.n, np.float32)
        tg = np.zeros(self.n, np.int32)
        self._L.orc_get_toi(self.h, _p(t, _f32p), _p(tg, _i32p))
        return t, tg

    def set_friction_axes(self, axis, transverse):
        """Anisotropic friction of the bodies: (n, 3) body-frame unit axes
        (zero rows: isotropic) and (n,) frictions across them."""
        ax = np.ascontiguousarray(axis, dtype=np.float32).reshape(-1, 3)
        tr = np.ascontiguousarray(transverse, dtype=np.float32).reshape(-1)
        self._L.orc_set_friction_axes(self.h, _p(ax, _f32p), _p(tr, _f32p), len(tr))

    def drain_events(self, trigger=False):
        """Queued contact (or trigger) events, oldest first: (kinds, ab
        (k, 2) int32, point (k, 3), normal (k, 3), approach (k,)) float32;
        the queue is emptied."""
        w = 1 if trigger else 0
        k = self._L.orc_drain_events(self.h, w, None, None)
        ii = np.zeros((k, 3), np.int32)
        ff = np.zeros((k, 7), np.float32)
        if k:
            self._L.orc_drain_events(self.h, w, _p(ii, _i32p), _p(ff, _f32p))
        return ii[:, 0].copy(), ii[:, 1:3].copy(), ff[:, 0:3].copy(), ff[:, 3:6].copy(), ff[:, 6].copy()

    def event_state(self, trigger=False):
        w = 1 if trigger else 0
        k = self._L.orc_get_event_state(self.h, w, None)
        ab = np.zeros((k, 2), np.uint32)
        if k:
            self._L.orc_get_event_state(self.h, w, _p(ab, _u32p))
        return ab

    def set_event_state(self, ab, trigger=False):
        ab = np.ascontiguousarray(ab, dtype=np.uint32).reshape(-1, 2)
        self._L.orc_set_event_state(self.h, 1 if trigger else 0, _p(ab, _u32p), len(ab))

    def stats(self):
        out = np.zeros(7, np.int32)
        self._L.orc_get_stats(self.h, _p(out, _i32p))
        return dict(zip(["pairs", "points", "colours", "max_colour", "sleeping", "breakdowns", "toi_hits"],
                        out.tolist()))
