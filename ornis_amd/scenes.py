"""Scene construction for the AVBD step: body arrays in the layout the C ABI
(`include/oavbd.h`) takes, plus seeded generators for the benchmark
workloads.

Material defaults mirror the reference constructors
(crates/physics/src/body.rs:20-32): boxes restitution 0.3 / friction 0.5,
spheres 0.5 / 0.3, half-space planes 0.3 / 0.6
(`RigidBody::try_new_halfspace`, body.rs:873-889).  Mass 0 means static
(`RigidBody::build`, body.rs:346-376).
"""
from __future__ import annotations

import math

import numpy as np

SHAPE_SPHERE = 0
SHAPE_BOX = 1
SHAPE_HALFSPACE = 2
SHAPE_CAPSULE = 3  # shape_param (radius, half height, 0), along the body's y axis

BODY_STATIC = 0
BODY_DYNAMIC = 1
BODY_KINEMATIC = 2

_FIELDS3 = ("shape_param", "pos", "vel", "angvel")


# joint kinds (include/oavbd.h)
JOINT_BALL, JOINT_REVOLUTE, JOINT_PRISMATIC, JOINT_FIXED = 0, 1, 2, 3
JOINT_DISTANCE, JOINT_ROPE, JOINT_SPRING, JOINT_WHEEL = 4, 5, 6, 7
JOINT_GEAR, JOINT_SIXDOF, JOINT_MOTOR = 8, 9, 10
# SixDof axis configurations
AXIS_FREE, AXIS_LOCKED, AXIS_LIMITED = 0, 1, 2
# motor kinds and models of a Spring joint (oavbd_motor)
MOTOR_VELOCITY, MOTOR_POSITION, MOTOR_SERVO = 0, 1, 2
MOTOR_ACCELERATION_BASED, MOTOR_FORCE_BASED = 0, 1


class SceneBuilder:
    """Accumulates bodies and ball joints; `build()` returns numpy arrays."""

    def __init__(self, gravity=(0.0, -9.81, 0.0)):
        self.gravity = tuple(float(g) for g in gravity)
        self.rows = []
        self.joints = []

    def _add(self, shape, sp, pos, mass, restitution, friction, quat=(0.0, 0.0, 0.0, 1.0),
             vel=(0.0, 0.0, 0.0), angvel=(0.0, 0.0, 0.0), body_type=None):
        if body_type is None:
            body_type = BODY_DYNAMIC if mass > 0 else BODY_STATIC
        self.rows.append(dict(shape=shape, shape_param=sp, pos=pos, quat=quat, vel=vel, angvel=angvel,
                              mass=mass, friction=friction, restitution=restitution, body_type=body_type))
        return len(self.rows) - 1

    def box(self, pos, half, mass, **kw):
        kw.setdefault("restitution", 0.3)
        kw.setdefault("friction", 0.5)
        return self._add(SHAPE_BOX, tuple(half), tuple(pos), float(mass), **kw)

    def sphere(self, pos, radius, mass, **kw):
        kw.setdefault("restitution", 0.5)
        kw.setdefault("friction", 0.3)
        return self._add(SHAPE_SPHERE, (float(radius), 0.0, 0.0), tuple(pos), float(mass), **kw)

    def capsule(self, pos, radius, half_height, mass, **kw):
        """RigidBody::new_capsule (crates/physics/src/body.rs:562-575): the
        core segment runs along the body's y axis, capsule material."""
        kw.setdefault("restitution", 0.4)
        kw.setdefault("friction", 0.4)
        return self._add(SHAPE_CAPSULE, (float(radius), float(half_height), 0.0), tuple(pos), float(mass), **kw)

    def halfspace(self, pos, normal, **kw):
        kw.setdefault("restitution", 0.3)
        kw.setdefault("friction", 0.6)
        n = np.asarray(normal, np.float32)
        n = n * (np.float32(1.0) / np.float32(np.sqrt(np.float32(np.dot(n, n)))))
        return self._add(SHAPE_HALFSPACE, tuple(float(x) for x in n), tuple(pos), 0.0, **kw)

    def ball_joint(self, a, b, la, lb):
        self.joints.append((int(a), int(b), tuple(la), tuple(lb)))
        return len(self.joints) - 1

    def joint(self, kind, a, b, la, lb, axis_a=(1.0, 0.0, 0.0), axis_b=(1.0, 0.0, 0.0), limit=None, motor=None):
        """A joint of kind JOINT_BALL / JOINT_REVOLUTE / JOINT_PRISMATIC /
        JOINT_FIXED / JOINT_DISTANCE between bodies a and b: local anchors
        la, lb and (hinge / slide kinds) local axes, with an optional travel
        `limit` (lo, hi) about the assembly angle / length and an optional
        velocity `motor` (target speed, force or torque budget)."""
        row = (int(a), int(b), tuple(la), tuple(lb), int(kind), tuple(axis_a), tuple(axis_b))
        if limit is not None or motor is not None:
            lo, hi = limit if limit is not None else (0.0, 0.0)
            sp, mx = motor if motor is not None else (0.0, 0.0)
            row += ((float(lo), float(hi), float(sp), float(mx), float(limit is not None), float(motor is not None)),)
        self.joints.append(row)
        return len(self.joints) - 1

    def rope(self, a, b, la, lb, max_distance):
        """Rope joint: the anchors part at most max_distance."""
        self.joints.append((int(a), int(b), tuple(la), tuple(lb), JOINT_ROPE, (1.0, 0.0, 0.0), (1.0, 0.0, 0.0),
                            (float(max_distance),)))
        return len(self.joints) - 1

    def spring(self, a, b, la, lb, rest, stiffness, damping, max_force=1e6, target_velocity=0.0,
               motor_kind=MOTOR_POSITION, model=MOTOR_ACCELERATION_BASED):
        """Spring joint about `rest` (a JointMotor of `motor_kind` / `model`)."""
        self.joints.append((int(a), int(b), tuple(la), tuple(lb), JOINT_SPRING, (1.0, 0.0, 0.0), (1.0, 0.0, 0.0),
                            (float(rest), float(stiffness), float(damping), float(max_force), float(target_velocity),
                             float(motor_kind), float(model))))
        return len(self.joints) - 1

    def wheel(self, a, b, la, lb, suspension_a, suspension_b, axle_a, axle_b, hz, zeta, motor=None):
        """Wheel joint: a slide along the suspension axis held by a spring of
        frequency `hz` and damping ratio `zeta` (rigid when hz <= 0), free
        spin about the axle, an optional axle motor (speed, torque budget)."""
        sp, mx = motor if motor is not None else (0.0, 0.0)
        self.joints.append((int(a), int(b), tuple(la), tuple(lb), JOINT_WHEEL, tuple(suspension_a),
                            tuple(suspension_b), tuple(float(x) for x in tuple(axle_a) + tuple(axle_b)) +
                            (float(hz), float(zeta), float(sp), float(mx), float(motor is not None))))
        return len(self.joints) - 1

    def sixdof(self, a, b, la, lb, kinds, lo=(0.0,) * 6, hi=(0.0,) * 6):
        """SixDof joint: per axis of A's frame (linear x, y, z then angular
        x, y, z) AXIS_FREE / AXIS_LOCKED / AXIS_LIMITED with window [lo, hi]
        from the assembly pose."""
        self.joints.append((int(a), int(b), tuple(la), tuple(lb), JOINT_SIXDOF, (1.0, 0.0, 0.0), (1.0, 0.0, 0.0),
                            tuple(float(k) for k in kinds) + tuple(float(x) for x in lo) +
                            tuple(float(x) for x in hi)))
        return len(self.joints) - 1

    def gear(self, a, b, joint_a, joint_b, ratio):
        """Gear joint: holds coord(joint_a) + ratio * coord(joint_b) at its
        value at creation (Revolute: hinge twist, Prismatic: slide)."""
        self.joints.append((int(a), int(b), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), JOINT_GEAR, (1.0, 0.0, 0.0),
                            (1.0, 0.0, 0.0), (float(joint_a), float(joint_b), float(ratio))))
        return len(self.joints) - 1

    def motor(self, a, b, linear_target, angular_target, max_force, max_torque):
        """Free Motor joint: drives b's velocity relative to a's (world frame)
        towards the targets within the force / torque budgets."""
        self.joints.append((int(a), int(b), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), JOINT_MOTOR, (1.0, 0.0, 0.0),
                            (1.0, 0.0, 0.0), tuple(float(x) for x in tuple(linear_target) + tuple(angular_target)) +
                            (float(max_force), float(max_torque))))
        return len(self.joints) - 1

    def build(self):
        n = len(self.rows)
        s = {"gravity": np.asarray(self.gravity, np.float32)}
        s["shape"] = np.array([r["shape"] for r in self.rows], np.int32)
        for k in _FIELDS3:
            s[k] = np.array([r[k] for r in self.rows], np.float32).reshape(n, 3)
        s["quat"] = np.array([r["quat"] for r in self.rows], np.float32).reshape(n, 4)
        for k in ("mass", "friction", "restitution"):
            s[k] = np.array([r[k] for r in self.rows], np.float32)
        s["body_type"] = np.array([r["body_type"] for r in self.rows], np.int32)
        s["layer"] = np.ones(n, np.uint32)
        s["mask"] = np.full(n, 0xFFFFFFFF, np.uint32)
        s["trigger"] = np.zeros(n, np.uint8)
        s["rolling"] = np.zeros(n, np.float32)
        s["torsion"] = np.zeros(n, np.float32)
        s["joints"] = list(self.joints)
        return s


def concat(scenes):
    """Concatenate scenes into one body set (joint indices shifted)."""
    out = {"gravity": scenes[0]["gravity"]}
    keys = [k for k in scenes[0] if k not in ("gravity", "joints")]
    for k in keys:
        out[k] = np.concatenate([s[k] for s in scenes])
    joints = []
    off = 0
    for s in scenes:
        for (a, b, la, lb) in s.get("joints", []):
            joints.append((a + off, b + off, la, lb))
        off += len(s["shape"])
    out["joints"] = joints
    return out


def _uniform_quats(rng, n):
    """Shoemake's uniform random unit quaternions (x, y, z, w)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * math.pi * u2), a * np.cos(2 * math.pi * u2),
                  b * np.sin(2 * math.pi * u3), b * np.cos(2 * math.pi * u3)], axis=1)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32)


def box_pyramid(base=20, rows=20, edge=1.0, density=1000.0, friction=0.5):
    """Box pyramid: row r holds base - r unit cubes resting on a static plane
    at y = 0 (the plain CPU-path workload: 210 cubes for base = rows = 20)."""
    sb = SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=friction, restitution=0.0)
    h = edge * 0.5
    mass = float(np.float32(density) * np.float32(edge) ** 3)
    for r in range(rows):
        cnt = base - r
        if cnt <= 0:
            break
        x0 = -0.5 * (cnt - 1) * edge
        for i in range(cnt):
            sb.box((x0 + i * edge, h + r * edge, 0.0), (h, h, h), mass, friction=friction, restitution=0.0)
    return sb.build()


def box_drop_lattice(nx=50, nz=50, ny=40, spacing=2.0, jitter=0.1, hmin=0.25, hmax=0.75,
                     density=1000.0, mu=(0.3, 0.8), container=110.0, seed=0xA5BD2026):
    """Random boxes in an nx x nz column lattice, ny layers high, dropped into
    a walled container on a static plane (100,000 boxes at the defaults)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = nx * nz * ny
    ix, iz, iy = np.meshgrid(np.arange(nx), np.arange(nz), np.arange(ny), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    pos = np.stack([-(nx - 1) * spacing * 0.5 + spacing * ix,
                    2.0 + spacing * iy,
                    -(nz - 1) * spacing * 0.5 + spacing * iz], axis=1)
    pos = pos + rng.uniform(-jitter, jitter, size=(n, 3))
    half = rng.uniform(hmin, hmax, size=(n, 3)).astype(np.float32)
    fr = rng.uniform(mu[0], mu[1], size=n).astype(np.float32)
    quat = _uniform_quats(rng, n)
    vol = np.float32(8.0) * half[:, 0] * half[:, 1] * half[:, 2]
    mass = (np.float32(density) * vol).astype(np.float32)
    sb = SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=float(mu[1]), restitution=0.0)
    c = container * 0.5
    wall_h, wall_t = 20.0, 1.0
    for p, hh in [((c + wall_t, wall_h, 0.0), (wall_t, wall_h, c + 2 * wall_t)),
                  ((-c - wall_t, wall_h, 0.0), (wall_t, wall_h, c + 2 * wall_t)),
                  ((0.0, wall_h, c + wall_t), (c, wall_h, wall_t)),
                  ((0.0, wall_h, -c - wall_t), (c, wall_h, wall_t))]:
        sb.box(p, hh, 0.0, friction=float(mu[1]), restitution=0.0)
    statics = sb.build()
    dyn = {
        "gravity": statics["gravity"],
        "shape": np.full(n, SHAPE_BOX, np.int32),
        "shape_param": half,
        "pos": pos.astype(np.float32),
        "quat": quat,
        "vel": np.zeros((n, 3), np.float32),
        "angvel": np.zeros((n, 3), np.float32),
        "mass": mass,
        "friction": fr,
        "restitution": np.zeros(n, np.float32),
        "body_type": np.full(n, BODY_DYNAMIC, np.int32),
        "layer": np.ones(n, np.uint32),
        "mask": np.full(n, 0xFFFFFFFF, np.uint32),
        "trigger": np.zeros(n, np.uint8),
        "rolling": np.zeros(n, np.float32),
        "torsion": np.zeros(n, np.float32),
        "joints": [],
    }
    return concat([statics, dyn])


def chains(n_chains=2000, links=32, grid=45, spacing=None, height=20.0, half=(0.05, 0.2, 0.05),
           n_spheres=5000, rmin=0.2, rmax=1.0, density=1000.0, friction=0.5, seed=0xA5BD2026):
    """Chains of 0.1 x 0.4 x 0.1 m boxes hanging along -y, consecutive links
    joined by ball joints at their shared end, roots on a grid at `height`,
    falling onto a plane scattered with static spheres (64,000 bodies and
    62,000 joints at the defaults)."""
    rng = np.random.Generator(np.random.PCG64(seed + 3))
    hx, hy, hz = half
    link_len = 2.0 * hy
    if spacing is None:
        spacing = 2.0 * link_len
    extent = spacing * (grid - 1)
    sb = SceneBuilder()
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=friction, restitution=0.0)
    mass = float(np.float32(density) * np.float32(8.0 * hx * hy * hz))
    for c in range(n_chains):
        gx, gz = c % grid, c // grid
        x0 = -0.5 * extent + spacing * gx
        z0 = -0.5 * extent + spacing * gz
        prev = None
        for k in range(links):
            y = height - k * link_len
            h = sb.box((x0, y, z0), half, mass, friction=friction, restitution=0.0)
            if prev is not None:
                sb.ball_joint(prev, h, (0.0, -hy, 0.0), (0.0, hy, 0.0))
            prev = h
    base = sb.build()
    if n_spheres:
        r = rng.uniform(rmin, rmax, size=n_spheres).astype(np.float32)
        span = 0.5 * extent + 2.0
        px = rng.uniform(-span, span, size=n_spheres)
        pz = rng.uniform(-span, span, size=n_spheres)
        sp = SceneBuilder()
        for i in range(n_spheres):
            sp.sphere((float(px[i]), float(r[i]), float(pz[i])), float(r[i]), 0.0, friction=friction,
                      restitution=0.0)
        return concat([base, sp.build()])
    return base


def box_drop_scene(n, rng, half_lo=0.2, half_hi=0.6, mu=(0.2, 0.9), floor=10.0, density=1000.0, height=1.0):
    """One small scene: n random boxes dropped onto a static floor box
    `floor` metres wide (used in batches)."""
    sb = SceneBuilder()
    sb.box((0.0, -0.5, 0.0), (floor * 0.5, 0.5, floor * 0.5), 0.0, friction=float(mu[1]), restitution=0.0)
    statics = sb.build()
    side = max(1, int(math.ceil(n ** (1.0 / 3.0))))
    pitch = 2.0 * half_hi + 0.1
    idx = np.arange(n)
    ix, iz, iy = idx % side, (idx // side) % side, idx // (side * side)
    pos = np.stack([(ix - (side - 1) * 0.5) * pitch, height + half_hi + iy * pitch,
                    (iz - (side - 1) * 0.5) * pitch], axis=1)
    half = rng.uniform(half_lo, half_hi, size=(n, 3)).astype(np.float32)
    fr = rng.uniform(mu[0], mu[1], size=n).astype(np.float32)
    quat = _uniform_quats(rng, n)
    mass = (np.float32(density) * np.float32(8.0) * half[:, 0] * half[:, 1] * half[:, 2]).astype(np.float32)
    dyn = {
        "gravity": statics["gravity"], "shape": np.full(n, SHAPE_BOX, np.int32), "shape_param": half,
        "pos": pos.astype(np.float32), "quat": quat, "vel": np.zeros((n, 3), np.float32),
        "angvel": np.zeros((n, 3), np.float32), "mass": mass, "friction": fr,
        "restitution": np.zeros(n, np.float32), "body_type": np.full(n, BODY_DYNAMIC, np.int32),
        "layer": np.ones(n, np.uint32), "mask": np.full(n, 0xFFFFFFFF, np.uint32),
        "trigger": np.zeros(n, np.uint8), "rolling": np.zeros(n, np.float32),
        "torsion": np.zeros(n, np.float32), "joints": []}
    return concat([statics, dyn])


def batched_drops(n_scenes=4096, nmin=8, nmax=512, seed=0xA5BD2026):
    """Independent box-drop scenes with log-uniform body counts in
    [nmin, nmax]; returns (list of scenes, concatenated scene with scene_id)."""
    rng = np.random.Generator(np.random.PCG64(seed + 4))
    counts = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), size=n_scenes))).astype(int)
    parts = [box_drop_scene(int(k), rng) for k in counts]
    allsc = concat(parts)
    allsc["scene_id"] = np.concatenate([np.full(len(p["shape"]), i, np.uint32) for i, p in enumerate(parts)])
    return parts, allsc


def throw_scene(rng, n_boxes=64, half=0.25, spacing=1.6, mu=None):
    """One differentiable-workload scene: n_boxes boxes (half-extents
    `half`) on an 8-wide grid `spacing` apart, centres 0.3-0.6 m above a
    ground plane at y = 0, thrown with initial velocity uniform [-3, 3] m/s
    per axis; one friction coefficient (uniform [0.2, 0.9] unless given) for
    every body of the scene, the plane included; masses lognormal(0, 0.3);
    and a target point for the scene's final centroid.  Speeds stay below
    the TOI clamp's reach (7.5 m/s for these boxes)."""
    sb = SceneBuilder()
    mu = float(rng.uniform(0.2, 0.9)) if mu is None else float(mu)
    sb.halfspace((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), friction=mu, restitution=0.0)
    side = 8
    for i in range(n_boxes):
        gx, gz = i % side, i // side
        pos = ((gx - (side - 1) / 2.0) * spacing, rng.uniform(0.3, 0.6), (gz - (side - 1) / 2.0) * spacing)
        vel = tuple(rng.uniform(-3.0, 3.0, 3))
        m = float(rng.lognormal(0.0, 0.3))
        sb.box(pos, (half, half, half), m, vel=vel, friction=mu, restitution=0.0)
    s = sb.build()
    s["target"] = np.array([rng.uniform(-2.0, 2.0), 0.25, rng.uniform(-2.0, 2.0)], np.float32)
    return s


def diff_throws(n_scenes=1024, n_boxes=64, seed=0xA5BD2026):
    """The differentiable workload: n_scenes independent throw scenes;
    returns (list of scenes, concatenated scene with scene_id)."""
    rng = np.random.Generator(np.random.PCG64(seed + 5))
    parts = [throw_scene(rng, n_boxes) for _ in range(n_scenes)]
    allsc = concat([{k: v for k, v in p.items() if k != "target"} for p in parts])
    allsc["scene_id"] = np.concatenate([np.full(len(p["shape"]), i, np.uint32) for i, p in enumerate(parts)])
    return parts, allsc
