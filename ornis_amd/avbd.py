"""Python mirror of the AVBD engine interface over the MI355X C ABI.

`AvbdEngine` follows the reference's `AvbdEngine` / `PhysicsEngine` seam
(crates/physics/src/avbd/mod.rs:424-464, crates/physics/src/engine.rs:24-133):
bodies are dense handles, `step(dt)` runs the fixed-step accumulator
(mod.rs:1212-1235), ball joints are added between handles
(mod.rs:1308-1510), and body state reads back as position, orientation,
velocity and angular velocity.  Every step runs on the GPU through
liboavbd.so; there is no CPU path — a missing library or device raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import scenes

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboavbd.so")

OK, EINVAL, ECAPACITY, EDEVICE, ENOMEM = 0, -1, -2, -3, -4

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class OavbdParams(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float * 3), ("iters", ctypes.c_uint32)]


class OavbdStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in (
        "pairs", "points", "candidates", "colours", "max_colour", "sleeping", "breakdowns", "toi_hits",
        "colour_rounds", "solved", "point_sides", "pair_sides", "joint_sides", "toi_casts", "joints",
        "tail_colours", "replays", "unwaited")]


class OavbdMotor(ctypes.Structure):
    """oavbd_motor: a Spring joint's motor (JointMotor, joint.rs:962-1126)."""
    _fields_ = [("kind", ctypes.c_uint32), ("model", ctypes.c_uint32), ("target_position", ctypes.c_float),
                ("target_velocity", ctypes.c_float), ("stiffness", ctypes.c_float), ("damping", ctypes.c_float),
                ("max_force", ctypes.c_float)]


class OavbdEvent(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("a", ctypes.c_uint32), ("b", ctypes.c_uint32),
                ("point", ctypes.c_float * 3), ("normal", ctypes.c_float * 3), ("approach", ctypes.c_float)]


EVENT_BEGIN, EVENT_END, EVENT_HIT, EVENT_ENTERED, EVENT_EXITED = 0, 1, 2, 3, 4


class OavbdKernelTimes(ctypes.Structure):
    _fields_ = [("primal_ms", ctypes.c_double), ("dual_ms", ctypes.c_double), ("primal_bytes", ctypes.c_double),
                ("dual_bytes", ctypes.c_double), ("primal_launches", ctypes.c_uint64),
                ("dual_launches", ctypes.c_uint64), ("steps", ctypes.c_uint64)]


PRIMAL_VARIANTS = 10  # oavbd_autotune's launch variants (include/oavbd.h)

# oavbd_contact_pair / oavbd_contact_point (include/oavbd.h), 80 bytes each.
CONTACT_PAIR = np.dtype([("a", "<u4"), ("b", "<u4"), ("first_point", "<u4"), ("npoints", "<u4"),
                         ("normal", "<f4", 4), ("tangent1", "<f4", 4), ("tangent2", "<f4", 4),
                         ("friction", "<f4", 4)])
CONTACT_POINT = np.dtype([("ra", "<f4", 4), ("rb", "<f4", 4), ("lam", "<f4", 4), ("pen", "<f4", 4),
                          ("roll", "<f4", 4)])
assert CONTACT_PAIR.itemsize == 80 and CONTACT_POINT.itemsize == 80

# Every entry point of include/oavbd.h with its ctypes signature.
SIGNATURES = {
    "oavbd_default_params": (None, [ctypes.POINTER(OavbdParams)]),
    "oavbd_create": (ctypes.c_int, [ctypes.POINTER(OavbdParams), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_void_p)]),
    "oavbd_destroy": (None, [ctypes.c_void_p]),
    "oavbd_set_bodies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                        _f32p, _f32p, _f32p, _i32p, _u32p, _u32p, _u8p, _f32p, _f32p, _u32p]),
    "oavbd_add_bodies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _i32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                        _f32p, _f32p, _f32p, _i32p, _u32p, _u32p, _u8p, _f32p, _f32p, _u32p, _u32p]),
    "oavbd_remove_body": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "oavbd_get_contacts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_uint32, _u32p, _u32p]),
    "oavbd_set_contacts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_uint32]),
    "oavbd_set_sleep": (ctypes.c_int, [ctypes.c_void_p, _u8p, _f32p, ctypes.c_uint32]),
    "oavbd_add_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _f32p,
                                       _f32p, _f32p, _f32p, _u32p]),
    "oavbd_add_ball_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                            _u32p]),
    "oavbd_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    "oavbd_step_n": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "oavbd_get_state": (ctypes.c_int, [ctypes.c_void_p, _f32p, ctypes.c_uint32]),
    "oavbd_set_state": (ctypes.c_int, [ctypes.c_void_p, _f32p, ctypes.c_uint32]),
    "oavbd_get_pairs": (ctypes.c_int, [ctypes.c_void_p, _u32p, ctypes.c_uint32, _u32p]),
    "oavbd_get_points": (ctypes.c_int, [ctypes.c_void_p, _f32p, ctypes.c_uint32, _u32p]),
    "oavbd_get_colours": (ctypes.c_int, [ctypes.c_void_p, _i32p, ctypes.c_uint32]),
    "oavbd_get_sleep": (ctypes.c_int, [ctypes.c_void_p, _u8p, _f32p, ctypes.c_uint32]),
    "oavbd_get_toi": (ctypes.c_int, [ctypes.c_void_p, _f32p, _i32p, ctypes.c_uint32]),
    "oavbd_get_joints": (ctypes.c_int, [ctypes.c_void_p, _f32p, ctypes.c_uint32]),
    "oavbd_add_joint_limited": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                               _f32p, _f32p, _f32p, _f32p, _f32p, _f32p,
                                               ctypes.POINTER(ctypes.c_uint32)]),
    "oavbd_add_wheel_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p, _f32p,
                                             _f32p, _f32p, _f32p, ctypes.c_float, ctypes.c_float, _f32p,
                                             ctypes.POINTER(ctypes.c_uint32)]),
    "oavbd_add_sixdof_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                              ctypes.POINTER(ctypes.c_uint32), _f32p, _f32p,
                                              ctypes.POINTER(ctypes.c_uint32)]),
    "oavbd_add_gear_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_uint32, ctypes.c_float, ctypes.POINTER(ctypes.c_uint32)]),
    "oavbd_add_motor_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                             ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_uint32)]),
    "oavbd_add_rope_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                            ctypes.c_float, ctypes.POINTER(ctypes.c_uint32)]),
    "oavbd_add_spring_joint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p,
                                              ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "oavbd_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OavbdStats)]),
    "oavbd_set_timing": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "oavbd_get_kernel_times": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OavbdKernelTimes), ctypes.c_int]),
    "oavbd_get_timing": (ctypes.c_int, [ctypes.c_void_p, _f32p]),
    "oavbd_set_primal_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "oavbd_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "oavbd_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _i32p, _f32p]),
    "oavbd_set_primal_table": (ctypes.c_int, [ctypes.c_void_p, _i32p, ctypes.c_uint32]),
    "oavbd_get_primal_table": (ctypes.c_int, [ctypes.c_void_p, _i32p, ctypes.c_uint32]),
    "oavbd_set_primal_tail": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "oavbd_set_unwaited": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]),
    "oavbd_set_events": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "oavbd_set_friction_axes": (ctypes.c_int, [ctypes.c_void_p, _f32p, _f32p, ctypes.c_uint32]),
    "oavbd_drain_events": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32, _u32p]),
    "oavbd_get_event_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_uint32, _u32p]),
    "oavbd_set_event_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_uint32]),
    "oavbd_solve_diag": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                        ctypes.c_float, ctypes.c_void_p]),
    "oavbd_solve_systems": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float,
                                           ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "oavbd_solve_systems_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                   ctypes.c_float, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "oavbd_checkpoint_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "oavbd_checkpoint_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "oavbd_diff_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "oavbd_diff_backward": (ctypes.c_int, [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p]),
    "oavbd_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load liboavbd.so and bind every C entry point (no device work)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"liboavbd.so not built ({path}); run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def pack_states(rhs):
    """(n, 6) residuals as GpuBodyState records {lin[3], 0, ang[3], 0}
    (crates/physics/src/gpu/mod.rs:200-207, from_residual :220-228)."""
    r = np.atleast_2d(np.asarray(rhs, np.float32))
    out = np.zeros((r.shape[0], 8), np.float32)
    out[:, 0:3] = r[:, 0:3]
    out[:, 4:7] = r[:, 3:6]
    return out


def pack_systems(inv_mass, inertia, rhs, axis, pen, lever, force, sign):
    """Packed 96-byte staged systems, GpuAvbdSystem::new
    (crates/physics/src/gpu/avbd.rs:329-337): GpuAvbdMass {inertia[3],
    inv_mass} (:55-60), the base residual as a GpuBodyState {lin[3], 0,
    ang[3], 0} (gpu/mod.rs:200-207) and one GpuAvbdRow {axis[3], pen,
    lever[3], force, sign, 0, 0, 0} (:264-281).  Per-system arguments are
    broadcast over the leading dimension; statics carry inv_mass 0 and a
    zero inertia, as GpuAvbdMass::from_body (:62-75) writes them."""
    inv_mass = np.atleast_1d(np.asarray(inv_mass, np.float32))
    n = max(inv_mass.shape[0], np.atleast_2d(np.asarray(rhs)).shape[0])
    out = np.zeros((n, 24), np.float32)
    out[:, 0:3] = np.broadcast_to(np.asarray(inertia, np.float32), (n, 3))
    out[:, 3] = np.broadcast_to(inv_mass, (n,))
    r = np.broadcast_to(np.asarray(rhs, np.float32), (n, 6))
    out[:, 4:7] = r[:, 0:3]
    out[:, 8:11] = r[:, 3:6]
    out[:, 12:15] = np.broadcast_to(np.asarray(axis, np.float32), (n, 3))
    out[:, 15] = np.broadcast_to(np.asarray(pen, np.float32), (n,))
    out[:, 16:19] = np.broadcast_to(np.asarray(lever, np.float32), (n, 3))
    out[:, 19] = np.broadcast_to(np.asarray(force, np.float32), (n,))
    out[:, 20] = np.broadcast_to(np.asarray(sign, np.float32), (n,))
    return out


class AvbdError(RuntimeError):
    pass


class AvbdEngine:
    """GPU AVBD engine (one context on one HIP device)."""

    def __init__(self, gravity=(0.0, -9.81, 0.0), iters=10, max_bodies=None, max_joints=None, device=0,
                 scene=None):
        self._L = load_library()
        self._gravity = tuple(float(g) for g in gravity)
        self._iters = int(iters)
        self._device = int(device)
        self._ctx = ctypes.c_void_p()
        self._max_bodies = max_bodies
        self._max_joints = max_joints
        self.n = 0
        self.njoints = 0
        self._tail = None
        self._unwaited = None
        if scene is not None:
            self.load(scene)

    def _check(self, rc):
        if rc != OK:
            msg = self._L.oavbd_last_error(self._ctx) if self._ctx else b""
            raise AvbdError(f"oavbd error {rc}: {msg.decode(errors='replace') if msg else ''}")

    def _ensure_ctx(self, n, nj):
        if self._ctx:
            return
        prm = OavbdParams()
        self._L.oavbd_default_params(ctypes.byref(prm))
        for k in range(3):
            prm.gravity[k] = self._gravity[k]
        prm.iters = self._iters
        mb = self._max_bodies or max(n, 1)
        mj = self._max_joints or max(nj, 64)
        self._check(self._L.oavbd_create(ctypes.byref(prm), mb, mj, self._device, ctypes.byref(self._ctx)))
        if self._tail is not None:
            self._check(self._L.oavbd_set_primal_tail(self._ctx, *self._tail))
        if self._unwaited is not None:
            self._check(self._L.oavbd_set_unwaited(self._ctx, *self._unwaited))

    @staticmethod
    def _body_arrays(s, scene_ids=None):
        """The oavbd_set_bodies / oavbd_add_bodies arrays of a scene dict,
        kept alive in the returned tuple, and the pointer arguments."""
        n = len(s["shape"])
        c = np.ascontiguousarray
        shape = c(s["shape"], dtype=np.int32)
        f = {k: c(s[k], dtype=np.float32) for k in ("shape_param", "pos", "quat", "vel", "angvel", "mass",
                                                    "friction", "restitution")}
        btype = c(s["body_type"], dtype=np.int32)
        layer = c(s.get("layer", np.ones(n)), dtype=np.uint32)
        mask = c(s.get("mask", np.full(n, 0xFFFFFFFF)), dtype=np.uint32)
        trig = c(s.get("trigger", np.zeros(n)), dtype=np.uint8)
        roll = c(s.get("rolling", np.zeros(n)), dtype=np.float32)
        tors = c(s.get("torsion", np.zeros(n)), dtype=np.float32)
        sid = s.get("scene_id") if scene_ids is None else scene_ids
        sid = None if sid is None else c(sid, dtype=np.uint32)
        keep = (shape, f, btype, layer, mask, trig, roll, tors, sid)
        args = (n, _p(shape, _i32p), _p(f["shape_param"], _f32p), _p(f["pos"], _f32p), _p(f["quat"], _f32p),
                _p(f["vel"], _f32p), _p(f["angvel"], _f32p), _p(f["mass"], _f32p), _p(f["friction"], _f32p),
                _p(f["restitution"], _f32p), _p(btype, _i32p), _p(layer, _u32p), _p(mask, _u32p), _p(trig, _u8p),
                _p(roll, _f32p), _p(tors, _f32p), _p(sid, _u32p))
        return keep, args

    def load(self, s, scene_ids=None):
        """Replace all bodies and joints with a scene (see ornis_amd.scenes)."""
        n = len(s["shape"])
        joints = s.get("joints", [])
        if "gravity" in s:
            self._gravity = tuple(float(g) for g in s["gravity"])
        self._ensure_ctx(n, len(joints))
        self._keep, args = self._body_arrays(s, scene_ids)
        self._check(self._L.oavbd_set_bodies(self._ctx, *args))
        self.n = n
        self.njoints = 0
        for jt in joints:
            self.add_joint(*jt)

    # PhysicsEngine::add_body / remove_body
    def add_bodies(self, s):
        """Append the bodies of scene dict `s` (no joints); returns the
        first new index.  Existing bodies keep state and warm starts."""
        keep, args = self._body_arrays(s)
        first = ctypes.c_uint32()
        self._check(self._L.oavbd_add_bodies(self._ctx, *args, ctypes.byref(first)))
        self.n += args[0]
        return first.value

    def remove_body(self, h):
        """Swap-remove body h (the last body takes index h)."""
        if not 0 <= int(h) < self.n:
            return
        self._check(self._L.oavbd_remove_body(self._ctx, int(h)))
        self.n -= 1
        self.njoints = self.stats()["joints"]

    def contacts(self):
        """The contact state carried between steps: (pairs, points) as
        structured arrays of CONTACT_PAIR / CONTACT_POINT."""
        m, q = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._L.oavbd_get_contacts(self._ctx, None, 0, None, 0, ctypes.byref(m), ctypes.byref(q)))
        pairs = np.zeros(m.value, CONTACT_PAIR)
        points = np.zeros(q.value, CONTACT_POINT)
        self._check(self._L.oavbd_get_contacts(self._ctx, pairs.ctypes.data, m.value, points.ctypes.data, q.value,
                                               ctypes.byref(m), ctypes.byref(q)))
        return pairs, points

    def set_contacts(self, pairs, points):
        pairs = np.ascontiguousarray(pairs, dtype=CONTACT_PAIR)
        points = np.ascontiguousarray(points, dtype=CONTACT_POINT)
        self._check(self._L.oavbd_set_contacts(self._ctx, pairs.ctypes.data, len(pairs), points.ctypes.data,
                                               len(points)))

    def set_sleep(self, asleep, timer):
        a = np.ascontiguousarray(asleep, dtype=np.uint8)
        t = np.ascontiguousarray(timer, dtype=np.float32)
        self._check(self._L.oavbd_set_sleep(self._ctx, _p(a, _u8p), _p(t, _f32p), self.n))

    def add_joint(self, a, b, la, lb, kind=0, axis_a=(1.0, 0.0, 0.0), axis_b=(1.0, 0.0, 0.0), prm=None):
        """Joint of `kind` (0 Ball, 1 Revolute, 2 Prismatic, 3 Fixed, 4 Distance,
        5 Rope with prm = (max distance,), 6 Spring with prm = (rest length,
        stiffness, damping, max force, target velocity, motor kind, motor
        model))."""
        la = np.ascontiguousarray(la, dtype=np.float32)
        lb = np.ascontiguousarray(lb, dtype=np.float32)
        idx = ctypes.c_uint32()
        if int(kind) == 7 and prm is not None:  # Wheel: axes = suspensions; prm = axle a 3, axle b 3, Hz, zeta, speed, budget, has motor
            f = lambda v: np.ascontiguousarray(v, dtype=np.float32)
            sa, sb, xa_, xb_ = f(axis_a), f(axis_b), f(prm[0:3]), f(prm[3:6])
            mot = f(prm[8:10])
            self._check(self._L.oavbd_add_wheel_joint(
                self._ctx, int(a), int(b), _p(la, _f32p), _p(lb, _f32p), _p(sa, _f32p), _p(sb, _f32p),
                _p(xa_, _f32p), _p(xb_, _f32p), ctypes.c_float(float(prm[6])), ctypes.c_float(float(prm[7])),
                _p(mot, _f32p) if prm[10] else None, ctypes.byref(idx)))
            self.njoints += 1
            return idx.value
        if int(kind) == 9 and prm is not None:  # SixDof: prm = 6 axis configurations, lo 6, hi 6
            cfg = np.ascontiguousarray(np.asarray(prm[0:6]).astype(np.uint32))
            lo = np.ascontiguousarray(prm[6:12], dtype=np.float32)
            hi = np.ascontiguousarray(prm[12:18], dtype=np.float32)
            self._check(self._L.oavbd_add_sixdof_joint(
                self._ctx, int(a), int(b), _p(la, _f32p), _p(lb, _f32p),
                cfg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _p(lo, _f32p), _p(hi, _f32p), ctypes.byref(idx)))
            self.njoints += 1
            return idx.value
        if int(kind) == 8 and prm is not None:  # Gear: prm = joint a, joint b, ratio
            self._check(self._L.oavbd_add_gear_joint(self._ctx, int(a), int(b), int(prm[0]), int(prm[1]),
                                                     ctypes.c_float(float(prm[2])), ctypes.byref(idx)))
            self.njoints += 1
            return idx.value
        if int(kind) == 10 and prm is not None:  # Motor: prm = linear target 3, angular target 3, max force, max torque
            lt = np.ascontiguousarray(prm[0:3], dtype=np.float32)
            at = np.ascontiguousarray(prm[3:6], dtype=np.float32)
            self._check(self._L.oavbd_add_motor_joint(self._ctx, int(a), int(b), _p(lt, _f32p), _p(at, _f32p),
                                                      ctypes.c_float(float(prm[6])), ctypes.c_float(float(prm[7])),
                                                      ctypes.byref(idx)))
            self.njoints += 1
            return idx.value
        if int(kind) == 5 and prm is not None:
            self._check(self._L.oavbd_add_rope_joint(self._ctx, int(a), int(b), _p(la, _f32p), _p(lb, _f32p),
                                                     ctypes.c_float(float(prm[0])), ctypes.byref(idx)))
            self.njoints += 1
            return idx.value
        if int(kind) == 6 and prm is not None:
            m = OavbdMotor(kind=int(prm[5]), model=int(prm[6]), target_position=prm[0], target_velocity=prm[4],
                           stiffness=prm[1], damping=prm[2], max_force=prm[3])
            self._check(self._L.oavbd_add_spring_joint(self._ctx, int(a), int(b), _p(la, _f32p), _p(lb, _f32p),
                                                       ctypes.byref(m), ctypes.byref(idx)))
            self.njoints += 1
            return idx.value
        xa = np.ascontiguousarray(axis_a, dtype=np.float32)
        xb = np.ascontiguousarray(axis_b, dtype=np.float32)
        if prm is not None and int(kind) in (1, 2):  # (lo, hi, speed, budget, has limit, has motor)
            lim = np.ascontiguousarray(prm[0:2], dtype=np.float32)
            mot = np.ascontiguousarray(prm[2:4], dtype=np.float32)
            self._check(self._L.oavbd_add_joint_limited(
                self._ctx, int(kind), int(a), int(b), _p(la, _f32p), _p(lb, _f32p), _p(xa, _f32p), _p(xb, _f32p),
                _p(lim, _f32p) if prm[4] else None, _p(mot, _f32p) if prm[5] else None, ctypes.byref(idx)))
            self.njoints += 1
            return idx.value
        self._check(self._L.oavbd_add_joint(self._ctx, int(kind), int(a), int(b), _p(la, _f32p), _p(lb, _f32p),
                                            _p(xa, _f32p), _p(xb, _f32p), ctypes.byref(idx)))
        self.njoints += 1
        return idx.value

    def add_ball_joint(self, a, b, la, lb):
        la = np.ascontiguousarray(la, dtype=np.float32)
        lb = np.ascontiguousarray(lb, dtype=np.float32)
        idx = ctypes.c_uint32()
        self._check(self._L.oavbd_add_ball_joint(self._ctx, int(a), int(b), _p(la, _f32p), _p(lb, _f32p),
                                                 ctypes.byref(idx)))
        self.njoints += 1
        return idx.value

    # PhysicsEngine::step
    def step(self, dt=1.0 / 60.0):
        self._check(self._L.oavbd_step(self._ctx, ctypes.c_float(dt)))

    def step_n(self, nsteps=1):
        self._check(self._L.oavbd_step_n(self._ctx, int(nsteps)))

    def sync(self):
        self._check(self._L.oavbd_sync(self._ctx))

    def state(self):
        out = np.zeros((self.n, 13), np.float32)
        self._check(self._L.oavbd_get_state(self._ctx, _p(out, _f32p), self.n))
        return out

    def set_state(self, st):
        st = np.ascontiguousarray(st, dtype=np.float32)
        self._check(self._L.oavbd_set_state(self._ctx, _p(st, _f32p), st.shape[0]))

    def pairs(self):
        m = ctypes.c_uint32()
        self._check(self._L.oavbd_get_pairs(self._ctx, None, 0, ctypes.byref(m)))
        ab = np.zeros((m.value, 2), np.uint32)
        self._check(self._L.oavbd_get_pairs(self._ctx, _p(ab, _u32p), m.value, ctypes.byref(m)))
        return ab

    def points(self):
        m = ctypes.c_uint32()
        self._check(self._L.oavbd_get_points(self._ctx, None, 0, ctypes.byref(m)))
        out = np.zeros((m.value, 16), np.float32)
        self._check(self._L.oavbd_get_points(self._ctx, _p(out, _f32p), m.value, ctypes.byref(m)))
        return out

    def colours(self):
        out = np.zeros(self.n, np.int32)
        self._check(self._L.oavbd_get_colours(self._ctx, _p(out, _i32p), self.n))
        return out

    def sleep(self):
        a = np.zeros(self.n, np.uint8)
        t = np.zeros(self.n, np.float32)
        self._check(self._L.oavbd_get_sleep(self._ctx, _p(a, _u8p), _p(t, _f32p), self.n))
        return a, t

    def toi(self):
        """Last step's TOI clamps: (t per body, -1 = none; target body)."""
        t = np.zeros(self.n, np.float32)
        tg = np.zeros(self.n, np.int32)
        self._check(self._L.oavbd_get_toi(self._ctx, _p(t, _f32p), _p(tg, _i32p), self.n))
        return t, tg

    def joints(self):
        out = np.zeros((self.njoints, 12), np.float32)
        self._check(self._L.oavbd_get_joints(self._ctx, _p(out, _f32p), self.njoints))
        return out

    def stats(self):
        s = OavbdStats()
        self._check(self._L.oavbd_get_stats(self._ctx, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in OavbdStats._fields_}

    def solve_diag(self, masses, rhs, dt=1.0 / 60.0):
        """Rung-1 staged bodies on the device: `masses` is (n, 4) float32
        GpuAvbdMass records, `rhs` (n, 8) GpuBodyState residuals; returns the
        (n, 8) GpuBodyState deltas."""
        m = np.ascontiguousarray(masses, dtype=np.float32).reshape(-1, 4)
        r = np.ascontiguousarray(rhs, dtype=np.float32).reshape(-1, 8)
        if m.shape[0] != r.shape[0]:
            raise ValueError("masses and rhs differ in length")
        n = m.shape[0]
        self._ensure_ctx(max(self.n, 1), self.njoints)
        delta = np.zeros((n, 8), np.float32)
        self._check(self._L.oavbd_solve_diag(self._ctx, m.ctypes.data_as(ctypes.c_void_p),
                                             r.ctypes.data_as(ctypes.c_void_p), n, ctypes.c_float(dt),
                                             delta.ctypes.data_as(ctypes.c_void_p)))
        return delta

    def solve_systems(self, systems, dt=1.0 / 60.0):
        """Rung-2 staged systems on the device: `systems` is (n, 24) float32
        (packed 96-byte GpuAvbdSystem records); returns ((n, 8) deltas as
        GpuBodyState records, (n,) ok flags)."""
        sysa = np.ascontiguousarray(systems, dtype=np.float32).reshape(-1, 24)
        n = sysa.shape[0]
        self._ensure_ctx(max(self.n, 1), self.njoints)
        delta = np.zeros((n, 8), np.float32)
        ok = np.zeros(n, np.float32)
        self._check(self._L.oavbd_solve_systems(self._ctx, sysa.ctypes.data_as(ctypes.c_void_p), n,
                                                ctypes.c_float(dt), delta.ctypes.data_as(ctypes.c_void_p),
                                                _p(ok, _f32p)))
        return delta, ok

    def solve_systems_noalloc(self, systems, dt=1.0 / 60.0):
        """solve_systems through the library's fixed staging array (no device
        allocation; at most 65536 systems per call)."""
        sysa = np.ascontiguousarray(systems, dtype=np.float32).reshape(-1, 24)
        n = sysa.shape[0]
        self._ensure_ctx(max(self.n, 1), self.njoints)
        delta = np.zeros((n, 8), np.float32)
        ok = np.zeros(n, np.float32)
        self._check(self._L.oavbd_solve_systems_noalloc(self._ctx, sysa.ctypes.data_as(ctypes.c_void_p), n,
                                                        ctypes.c_float(dt), delta.ctypes.data_as(ctypes.c_void_p),
                                                        _p(ok, _f32p)))
        return delta, ok

    def checkpoint_save(self, slot=0):
        """Keep the whole solver state (bodies, contact pairs with their warm
        starts, joint duals) in device checkpoint `slot`."""
        self._check(self._L.oavbd_checkpoint_save(self._ctx, int(slot)))

    def checkpoint_load(self, slot=0):
        """Return to checkpoint `slot`; the steps that follow repeat bit for bit."""
        self._check(self._L.oavbd_checkpoint_load(self._ctx, int(slot)))

    def diff_forward(self, nsteps, every=10):
        """nsteps differentiable steps with a checkpoint every `every` steps."""
        self._check(self._L.oavbd_diff_forward(self._ctx, int(nsteps), int(every)))

    def diff_backward(self, dfinal):
        """Gradients through the last diff_forward: given dL/d(final state)
        (n x 13), returns (dL/d(initial state) n x 13, dL/d(mass) n,
        dL/d(friction) n)."""
        g = np.ascontiguousarray(dfinal, dtype=np.float32).reshape(self.n, 13)
        d0 = np.zeros((self.n, 13), np.float32)
        dm = np.zeros(self.n, np.float32)
        df = np.zeros(self.n, np.float32)
        self._check(self._L.oavbd_diff_backward(self._ctx, _p(g, _f32p), _p(d0, _f32p), _p(dm, _f32p), _p(df, _f32p)))
        return d0, dm, df

    def set_primal_shape(self, lanes=16, waves=2):
        """Primal launch shape: lanes per body (16) and wavefronts per
        workgroup (bitwise neutral)."""
        self._check(self._L.oavbd_set_primal_shape(self._ctx, int(lanes), int(waves)))

    def autotune(self, nsteps=3):
        """Time the next nsteps steps under every primal launch variant
        (state restored afterwards) and keep the fastest per class size;
        returns the 24-entry table."""
        t = np.zeros(24, np.int32)
        self.tune_us = np.zeros((PRIMAL_VARIANTS, 24), np.float32)
        self._check(self._L.oavbd_autotune(self._ctx, int(nsteps), _p(t, _i32p), _p(self.tune_us, _f32p)))
        return t

    def set_primal_table(self, table):
        t = np.ascontiguousarray(table, dtype=np.int32)
        self._check(self._L.oavbd_set_primal_table(self._ctx, _p(t, _i32p), len(t)))

    def primal_table(self):
        t = np.zeros(24, np.int32)
        self._check(self._L.oavbd_get_primal_table(self._ctx, _p(t, _i32p), 24))
        return t

    def set_primal_tail(self, max_bodies, max_waves=0):
        """Bodies of the last colour classes swept by one fused launch (0: a
        launch per class), and its wavefront limit (0: device capacity).  Kept
        for the context a later load creates."""
        self._tail = (int(max_bodies), int(max_waves))
        if self._ctx:
            self._check(self._L.oavbd_set_primal_tail(self._ctx, *self._tail))

    def set_unwaited(self, on=True, spare_colours=1):
        """Unwaited stepping (default on): once a step's counts have been read
        back, step_n runs its steps without the host reading any count,
        repeating a batch whose counts outgrew their buffers; the same bits
        either way.  Kept for the context a later load creates."""
        self._unwaited = (1 if on else 0, int(spare_colours))
        if self._ctx:
            self._check(self._L.oavbd_set_unwaited(self._ctx, *self._unwaited))

    def set_friction_axes(self, axis, transverse):
        """Anisotropic friction: (n, 3) body-frame axes (zero rows:
        isotropic) and (n,) frictions across them, for the first n bodies."""
        ax = np.ascontiguousarray(axis, dtype=np.float32).reshape(-1, 3)
        tr = np.ascontiguousarray(transverse, dtype=np.float32).reshape(-1)
        self._check(self._L.oavbd_set_friction_axes(self._ctx, _p(ax, _f32p), _p(tr, _f32p), len(tr)))

    def set_events(self, on=True):
        """Contact and trigger events (off by default)."""
        self._check(self._L.oavbd_set_events(self._ctx, 1 if on else 0))

    def drain_events(self, trigger=False):
        """Queued contact (or trigger) events, oldest first, as arrays:
        (kinds, ab (k, 2), point (k, 3), normal (k, 3), approach (k,)); the
        queue is emptied."""
        w = 1 if trigger else 0
        n = ctypes.c_uint32(0)
        self._check(self._L.oavbd_drain_events(self._ctx, w, None, 0, ctypes.byref(n)))
        k = n.value
        buf = (OavbdEvent * max(k, 1))()
        if k:
            self._check(self._L.oavbd_drain_events(self._ctx, w, ctypes.cast(buf, ctypes.c_void_p), k,
                                                   ctypes.byref(n)))
        ev = buf[:k]
        kinds = np.array([e.kind for e in ev], np.uint32)
        ab = np.array([(e.a, e.b) for e in ev], np.uint32).reshape(-1, 2)
        pt = np.array([list(e.point) for e in ev], np.float32).reshape(-1, 3)
        nrm = np.array([list(e.normal) for e in ev], np.float32).reshape(-1, 3)
        app = np.array([e.approach for e in ev], np.float32)
        return kinds, ab, pt, nrm, app

    def event_state(self, trigger=False):
        w = 1 if trigger else 0
        n = ctypes.c_uint32(0)
        self._check(self._L.oavbd_get_event_state(self._ctx, w, None, 0, ctypes.byref(n)))
        ab = np.zeros((n.value, 2), np.uint32)
        if n.value:
            self._check(self._L.oavbd_get_event_state(self._ctx, w, _p(ab, _u32p), n.value, ctypes.byref(n)))
        return ab

    def set_event_state(self, ab, trigger=False):
        ab = np.ascontiguousarray(ab, dtype=np.uint32).reshape(-1, 2)
        self._check(self._L.oavbd_set_event_state(self._ctx, 1 if trigger else 0, _p(ab, _u32p), len(ab)))

    def set_timing(self, on=True):
        self._check(self._L.oavbd_set_timing(self._ctx, 1 if on else 0))

    def timing(self):
        out = np.zeros(5, np.float32)
        self._check(self._L.oavbd_get_timing(self._ctx, _p(out, _f32p)))
        return dict(zip(["setup", "primal", "dual", "finish", "step"], out.tolist()))

    def kernel_times(self, reset=False):
        k = OavbdKernelTimes()
        self._check(self._L.oavbd_get_kernel_times(self._ctx, ctypes.byref(k), 1 if reset else 0))
        return {f: getattr(k, f) for f, _ in OavbdKernelTimes._fields_}

    @property
    def sleeping_count(self):
        return int(self.sleep()[0].sum())

    @property
    def body_count(self):
        return self.n

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.oavbd_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["AvbdEngine", "AvbdError", "load_library", "SIGNATURES", "scenes"]
