/*
 * oavbd.h — C ABI of the MI355X AVBD rigid-body step (liboavbd.so).
 *
 * This is the device boundary behind the reference's physics-solver seam:
 * a Rust `GpuAvbdDispatch` implementation (crates/physics/src/gpu/avbd.rs:
 * 146-151, consulted by AvbdEngine::step, crates/physics/src/avbd/mod.rs:
 * 1220-1223) would hold one `oavbd_ctx*` and call these entry points
 * through `extern "C"`.  Plain pointers and sizes only; no C++ exceptions
 * cross this boundary; every entry point returns an oavbd_status.
 *
 * Threading: a context is used by one host thread at a time (the
 * reference's `step(&mut self)`, crates/physics/src/engine.rs:24);
 * separate contexts may run concurrently.  All device work of a context
 * goes to its own HIP stream.  Host arrays passed in are borrowed for the
 * duration of the call only.
 *
 * Bodies are dense indices 0..n-1 (the reference's BodyHandle,
 * crates/physics/src/body.rs).  Shapes: 0 sphere (param[0] = radius),
 * 1 box (param = half extents), 2 half-space (param = outward unit normal,
 * static only), 3 capsule (param = radius, half height of the core segment,
 * which runs along the body's y axis; Shape::Capsule,
 * crates/physics/src/collision/shape.rs:157).  Body types: 0 static,
 * 1 dynamic, 2 kinematic.  Mass 0 makes a dynamic body static
 * (RigidBody::build, body.rs:346-376).
 *
 * Scenes: every body carries a scene id; bodies of different scenes never
 * collide, so one context steps thousands of independent scenes in one
 * launch sequence.  Within a scene the result equals stepping that scene
 * alone.
 */
#ifndef OAVBD_H
#define OAVBD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum oavbd_status {
  OAVBD_OK = 0,
  OAVBD_EINVAL = -1,    /* bad argument (null pointer, index out of range)  */
  OAVBD_ECAPACITY = -2, /* more bodies / joints than the context holds      */
  OAVBD_EDEVICE = -3,   /* HIP runtime error; see oavbd_last_error           */
  OAVBD_ENOMEM = -4     /* device allocation failed                          */
} oavbd_status;

typedef struct oavbd_ctx oavbd_ctx;

/* Solver parameters.  Defaults (oavbd_default_params) equal the reference
 * constants (crates/physics/src/avbd/mod.rs:154-247). */
typedef struct oavbd_params {
  float gravity[3]; /* world acceleration, default (0, -9.81, 0)            */
  uint32_t iters;   /* primal/dual iterations per step, default 10 (ITERS)  */
} oavbd_params;

/* Per-step counters (observability; the reference's sleeping_count,
 * crates/physics/src/avbd/mod.rs:788-790, plus solver sizes). */
typedef struct oavbd_stats {
  uint32_t pairs;       /* contact pairs after the merge                     */
  uint32_t points;      /* contact points                                    */
  uint32_t candidates;  /* broadphase candidate pairs                        */
  uint32_t colours;     /* colour classes swept                              */
  uint32_t max_colour;  /* bodies in the largest colour                      */
  uint32_t sleeping;    /* dynamic bodies frozen by sleep                    */
  uint32_t breakdowns;  /* 6x6 LDL breakdowns (body kept its pose)           */
  uint32_t toi_hits;    /* TOI clamps applied                                */
  uint32_t colour_rounds; /* Jones-Plassmann rounds of the colouring         */
  uint32_t solved;      /* bodies swept by the primal solve                  */
  uint32_t point_sides; /* contact point-sides of swept bodies               */
  uint32_t pair_sides;  /* contact pair-sides (with points) of swept bodies  */
  uint32_t joint_sides; /* joint sides of swept bodies                       */
  uint32_t toi_casts;   /* (mover, target) casts tested by the TOI clamp     */
  uint32_t joints;      /* joints held by the context                        */
  uint32_t tail_colours; /* last colour classes swept by one fused launch    */
  uint32_t replays;     /* unwaited batches repeated after an overflow (total) */
  uint32_t unwaited;    /* 1 if the last oavbd_step_n ran unwaited            */
} oavbd_stats;

/* Kernel time and algorithmic bytes accumulated since the last reset
 * while timing is enabled (oavbd_set_timing).  Every primal-solve and
 * dual-update launch is bracketed by HIP events on the context stream.
 * Algorithmic bytes count each datum a launch must read or write once, at
 * its packed size (3 floats per vector, 4 per quaternion):
 *   primal, per iteration: 128 B per swept body (position, rotation, their
 *     step-start values, inertial targets, inverse mass, inertia: 100 B
 *     read; pose: 28 B written) + 96 B per pair-side (pair header 24 B,
 *     partner pose and start pose 56 B, shape 16 B) + 48 B per point-side
 *     (anchors, duals, penalties) + 168 B per joint-side;
 *   dual, per iteration: 136 B per pair (24 B header + four 28 B poses) +
 *     76 B per point (48 B read, 28 B written) + 264 B per joint. */
typedef struct oavbd_kernel_times {
  double primal_ms;      /* summed primal launch durations                 */
  double dual_ms;        /* summed dual launch durations                   */
  double primal_bytes;   /* summed algorithmic bytes of those launches     */
  double dual_bytes;
  uint64_t primal_launches;
  uint64_t dual_launches;
  uint64_t steps;
} oavbd_kernel_times;

int oavbd_get_kernel_times(oavbd_ctx* ctx, oavbd_kernel_times* out, int reset);

void oavbd_default_params(oavbd_params* p);

/* Create a context able to hold max_bodies bodies and max_joints joints on
 * HIP device `device`.  Replaces: AvbdEngine::new (mod.rs:1183-1204) plus
 * the device set-up of WgpuAvbdSolver::new (crates/physics/src/gpu/avbd.rs:
 * 502-632). */
int oavbd_create(const oavbd_params* params, uint32_t max_bodies, uint32_t max_joints, int device,
                 oavbd_ctx** out);
void oavbd_destroy(oavbd_ctx* ctx);

/* Replace all bodies (and drop joints, pairs and warm starts).  Arrays of n
 * entries; vectors are packed xyz (3 floats) or xyzw (4 floats).  Optional
 * arrays may be NULL: layer (default 1), mask (default all), trigger (0),
 * rolling/torsion friction (0), scene (0).  Replaces: repeated
 * PhysicsEngine::add_body (mod.rs:1236-1239) and WgpuAvbdSolver::upload
 * (crates/physics/src/gpu/avbd.rs:645-656). */
int oavbd_set_bodies(oavbd_ctx* ctx, uint32_t n, const int32_t* shape, const float* shape_param3,
                     const float* pos3, const float* quat4, const float* vel3, const float* angvel3,
                     const float* mass, const float* friction, const float* restitution,
                     const int32_t* body_type, const uint32_t* layer, const uint32_t* mask,
                     const uint8_t* trigger, const float* rolling, const float* torsion,
                     const uint32_t* scene);

/* Append n bodies (same arrays and defaults as oavbd_set_bodies) after the
 * existing ones, which keep their state, pairs, warm starts, joints and
 * sleep state; *first_index (may be NULL) receives the first new index.
 * Replaces: PhysicsEngine::add_body (mod.rs:1236-1239), whose handles are
 * table indices, with the scratch of ensure_scratch (mod.rs:1183-1199):
 * new bodies start awake with a zero sleep timer. */
int oavbd_add_bodies(oavbd_ctx* ctx, uint32_t n, const int32_t* shape, const float* shape_param3,
                     const float* pos3, const float* quat4, const float* vel3, const float* angvel3,
                     const float* mass, const float* friction, const float* restitution,
                     const int32_t* body_type, const uint32_t* layer, const uint32_t* mask,
                     const uint8_t* trigger, const float* rolling, const float* torsion,
                     const uint32_t* scene, uint32_t* first_index);

/* Remove body h by swap-remove: the last body takes index h.  Pairs and
 * joints on h are dropped; the others are renumbered in place (a pair keeps
 * its (a, b) orientation, so one whose first body now has the larger index
 * finds no match at the next step and restarts cold, as in the reference);
 * warm starts of the rest are kept; every sleeping body wakes.  An index out
 * of range is a no-op.  Replaces: PhysicsEngine::remove_body (engine.rs:
 * 40-43; mod.rs:1241-1300; trigger and contact events are not produced). */
int oavbd_remove_body(oavbd_ctx* ctx, uint32_t h);

/* The contact state carried from one step to the next (the warm starts of
 * the reference's AvbdPoint / AvbdPair, crates/physics/src/avbd/mod.rs:
 * 253-281): each pair with its body indices, its range of points, the last
 * step's normal and signed gap, friction tangents and coefficients; each
 * point with its body-frame anchors, stuck flag, duals and penalties. */
typedef struct oavbd_contact_pair {
  uint32_t a, b;             /* bodies (a < b as discovered) */
  uint32_t first_point;      /* index of the pair's first point */
  uint32_t npoints;          /* 0..8 (a pair may hold no points) */
  float normal[4];           /* xyz: unit normal from B to A; w: signed gap */
  float tangent1[4];         /* xyz: friction tangent; w: damper stiffness */
  float tangent2[4];         /* xyz: t1 x n */
  float friction[4];         /* mu1, mu2, rolling, torsion */
} oavbd_contact_pair;        /* 80 bytes */
typedef struct oavbd_contact_point {
  float ra[4];               /* xyz: anchor on A, body frame; w: step-start violation */
  float rb[4];               /* xyz: anchor on B, body frame */
  float lam[4];              /* xyz: normal and two friction duals */
  float pen[4];              /* xyz: their penalties; w: stuck (0 or 1) */
  float roll[4];             /* xyz: rolling and torsion duals */
} oavbd_contact_point;       /* 80 bytes */

/* Export the contact state: *npairs and *npoints receive the counts; with
 * non-NULL arrays, cap_pairs / cap_points must hold them (else ECAPACITY).
 * Pairs are in canonical (a, b) order with their points contiguous.
 * Replaces: the pair vector a checkpoint or solver migration carries
 * (mod.rs:469-636 hand over bodies, joints and event state; the warm starts
 * are this build's addition, for bit-exact continuation). */
int oavbd_get_contacts(oavbd_ctx* ctx, oavbd_contact_pair* pairs, uint32_t cap_pairs,
                       oavbd_contact_point* points, uint32_t cap_points, uint32_t* npairs, uint32_t* npoints);

/* Replace the contact state by an exported one (or one a host built):
 * bodies in range and distinct, (a, b) unique, 0..8 points per pair inside
 * the point array.  The pairs are stored sorted by (a, b). */
int oavbd_set_contacts(oavbd_ctx* ctx, const oavbd_contact_pair* pairs, uint32_t npairs,
                       const oavbd_contact_point* points, uint32_t npoints);

/* Set every body's sleep flag and timer (n = body count).  A dynamic body
 * set asleep takes the state sleep_body leaves (zero velocities, inverse
 * mass and inertia: mod.rs:822-832); one set awake from sleep gets its mass
 * properties back (wake_body, mod.rs:833-842).  With oavbd_set_state and
 * oavbd_set_contacts this restores a step boundary exported with
 * oavbd_get_state, oavbd_get_sleep and oavbd_get_contacts. */
int oavbd_set_sleep(oavbd_ctx* ctx, const uint8_t* asleep, const float* timer, uint32_t n);

/* Joint kinds (the reference's JointKind arms, numbered as its
 * AvbdJointKind; Gear 8 is not built).  Rope, Spring, Wheel, SixDof and
 * Motor are added with oavbd_add_rope_joint / oavbd_add_spring_joint /
 * oavbd_add_wheel_joint / oavbd_add_sixdof_joint / oavbd_add_motor_joint,
 * Revolute and Prismatic
 * with a limit or motor with oavbd_add_joint_limited, the others with
 * oavbd_add_joint. */
enum {
  OAVBD_JOINT_BALL = 0,
  OAVBD_JOINT_REVOLUTE = 1,
  OAVBD_JOINT_PRISMATIC = 2,
  OAVBD_JOINT_FIXED = 3,
  OAVBD_JOINT_DISTANCE = 4,
  OAVBD_JOINT_ROPE = 5,
  OAVBD_JOINT_SPRING = 6,
  OAVBD_JOINT_WHEEL = 7,
  OAVBD_JOINT_SIXDOF = 9,
  OAVBD_JOINT_MOTOR = 10
};

/* A joint motor as a Spring joint carries it (JointMotor,
 * crates/physics/src/joint.rs:962-1126): kind (MotorKind), model
 * (MotorModel: acceleration-based scales stiffness and damping by the
 * bodies' reduced mass), rest length (target_position), stiffness, damping,
 * force budget and target velocity.  The step reads rest length, stiffness,
 * damping and model; kind, budget and target velocity decide whether the
 * spring keeps its bodies awake (JointMotor::keeps_awake). */
enum { OAVBD_MOTOR_VELOCITY = 0, OAVBD_MOTOR_POSITION = 1, OAVBD_MOTOR_SERVO = 2 };
enum { OAVBD_MOTOR_ACCELERATION_BASED = 0, OAVBD_MOTOR_FORCE_BASED = 1 };
typedef struct oavbd_motor {
  uint32_t kind;  /* OAVBD_MOTOR_VELOCITY / _POSITION / _SERVO           */
  uint32_t model; /* OAVBD_MOTOR_ACCELERATION_BASED / _FORCE_BASED       */
  float target_position, target_velocity, stiffness, damping, max_force;
} oavbd_motor;

/* Joint of `kind` between bodies a and b: body-local anchors la, lb and,
 * for Revolute and Prismatic, body-local hinge / slide axes axis_a, axis_b
 * (normalised here; NULL for the other kinds).  A Fixed joint captures the
 * bodies' current relative rotation and anchor offset as its assembly pose,
 * a Distance joint the anchors' current distance as its rod length.
 * Both bodies wake first, as in the reference, even when the spec is then
 * rejected: EINVAL for a == b, an out-of-range body, an unknown kind or a
 * degenerate axis.  Jointed bodies do not collide, except through a Fixed
 * joint.  Returns the joint index through *out_index (may be NULL).
 * Replaces: PhysicsEngine::add_joint for JointKind::{Ball, Revolute,
 * Prismatic, Fixed, Distance} with limit and motor None (mod.rs:1308-1510;
 * frames of resolve_joint, crates/physics/src/joint.rs:1251-1332). */
int oavbd_add_joint(oavbd_ctx* ctx, uint32_t kind, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                    const float axis_a[3], const float axis_b[3], uint32_t* out_index);

/* oavbd_add_joint with OAVBD_JOINT_BALL.  Replaces: PhysicsEngine::add_joint
 * with JointKind::Ball (mod.rs:1308-1510, crates/physics/src/joint.rs:
 * 1273-1279). */
int oavbd_add_ball_joint(oavbd_ctx* ctx, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                         uint32_t* out_index);

/* A Revolute or Prismatic joint (as oavbd_add_joint) with an optional
 * travel window limit[0] <= limit[1] (radians of twist about the hinge
 * axis from the assembly angle, or metres along the slide axis from the
 * assembly length; a one-sided row whose force the dual accumulates) and
 * an optional velocity motor (motor[0] target speed in rad/s or m/s,
 * motor[1] torque or force budget: a budget-clamped velocity impulse each
 * step, paused while the limit is violated).  NULL for none.  EINVAL,
 * before any wake, for a bound that is not finite or lo > hi, or a motor
 * with a non-finite value or a negative budget.  Replaces:
 * PhysicsEngine::add_joint for JointKind::{Revolute, Prismatic} with
 * limit / motor Some (mod.rs:1437-1459; rows assembly.rs:809-874, dual
 * joints_ext.rs:820-890, motor joints_ext.rs:237-356). */
int oavbd_add_joint_limited(oavbd_ctx* ctx, uint32_t kind, uint32_t a, uint32_t b, const float la[3],
                            const float lb[3], const float axis_a[3], const float axis_b[3], const float limit[2],
                            const float motor[2], uint32_t* out_index);

/* Wheel between a and b (Box2D's wheel formulation): a slide along the
 * suspension axis (suspension_a / _b in each body's frame, normalised,
 * +y when degenerate) held by a spring of frequency_hz and damping_ratio
 * (a fixed penalty m w^2 and damping 2 m zeta w with m the bodies' reduced
 * mass; a rigid slide when frequency_hz is 0), the two directions across
 * the axle (axle_a / _b, orthogonalised against the suspension) locked,
 * spin about it free, and an optional axle velocity motor (motor[0]
 * rad/s, motor[1] torque budget; NULL for none).  EINVAL, before any wake,
 * for a non-finite axis or anchor, a negative frequency or damping ratio,
 * or a bad motor.  Replaces: PhysicsEngine::add_joint with JointKind::Wheel
 * (mod.rs:1494-1501; resolve_joint, crates/physics/src/joint.rs:1352-1380;
 * rows assembly.rs:521-580, :753-781; duals joints_ext.rs:738-805; motor
 * joints_ext.rs:337-369). */
int oavbd_add_wheel_joint(oavbd_ctx* ctx, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                          const float suspension_a[3], const float suspension_b[3], const float axle_a[3],
                          const float axle_b[3], float frequency_hz, float damping_ratio, const float motor[2],
                          uint32_t* out_index);

/* SixDof between anchors la, lb: per axis of body A's live frame (axes[0..2]
 * linear x y z, axes[3..5] angular x y z) OAVBD_AXIS_FREE, OAVBD_AXIS_LOCKED
 * (an equality row on the drift from the assembly pose) or
 * OAVBD_AXIS_LIMITED (a one-sided row keeping the travel from the assembly
 * pose inside [lo[k], hi[k]]: metres along the axis, or radians of twist
 * about it).  lo / hi may be NULL when no axis is limited.  Its bodies keep
 * colliding (weld-like, as Fixed).  EINVAL, before any wake, for an
 * unknown configuration or a limited axis without finite lo <= hi.
 * Replaces: PhysicsEngine::add_joint with JointKind::SixDof (mod.rs:
 * 1502-1511; rows assembly.rs:452-519, :682-740; duals joints_ext.rs:
 * 646-714). */
enum { OAVBD_AXIS_FREE = 0, OAVBD_AXIS_LOCKED = 1, OAVBD_AXIS_LIMITED = 2 };
int oavbd_add_sixdof_joint(oavbd_ctx* ctx, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                           const uint32_t axes[6], const float lo[6], const float hi[6], uint32_t* out_index);

/* Free motor between a and b: each step begins with budget-clamped
 * velocity impulses (per world axis, through the pair's effective mass at
 * the centres) driving b's velocity relative to a's towards linear_target
 * (m/s, within max_force) and angular_target (rad/s, within max_torque);
 * no position rows, no dual state.  It captures the bodies' relative pose
 * for its rest residual; a budgeted non-zero target keeps the bodies
 * awake.  EINVAL, before any wake, for a non-finite target or a negative
 * budget.  Replaces: PhysicsEngine::add_joint with JointKind::Motor
 * (mod.rs:1517-1525; free_motor_impulse, joints_ext.rs:368-421). */
int oavbd_add_motor_joint(oavbd_ctx* ctx, uint32_t a, uint32_t b, const float linear_target[3],
                          const float angular_target[3], float max_force, float max_torque, uint32_t* out_index);

/* Gear: holds coord(joint_a) + ratio * coord(joint_b) at its value when the
 * gear is made, where coord is a Revolute joint's hinge twist past its
 * assembly angle (unwrapped across +-pi step to step) or a Prismatic
 * joint's slide past its assembly length.  One shared row over the bodies
 * of both joints (up to four); a and b only name the gear's own pair (both
 * are woken; their contacts stay).  EINVAL, before any wake, for a
 * non-finite ratio, an unknown joint index or a joint that is neither
 * Revolute nor Prismatic.  Replaces: PhysicsEngine::add_joint with
 * JointKind::Gear (mod.rs:1329-1379; gear_sides / update_gear_mem,
 * joints_ext.rs:38-161; rows assembly.rs:426-449, :650-679; dual
 * joints_ext.rs:897-917). */
int oavbd_add_gear_joint(oavbd_ctx* ctx, uint32_t a, uint32_t b, uint32_t joint_a, uint32_t joint_b, float ratio,
                         uint32_t* out_index);

/* Rope between a and b: anchors la, lb may come closer freely but never
 * part beyond max_distance (one-sided row, a velocity catch each step).
 * EINVAL, before any wake, unless max_distance is finite and positive.
 * Replaces: PhysicsEngine::add_joint with JointKind::Rope (mod.rs:1483-1486;
 * JointKind::rope_checked, crates/physics/src/joint.rs:539-553; rows
 * assembly.rs:344-362, dual joints_ext.rs:618-641, catch joints_ext.rs:
 * 173-223). */
int oavbd_add_rope_joint(oavbd_ctx* ctx, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                         float max_distance, uint32_t* out_index);

/* Spring between anchors la, lb about the motor's rest length (a
 * position-level row with a fixed penalty, no dual state).  EINVAL, before
 * any wake, for a motor failing JointMotor::check (joint.rs:1101-1110).
 * Replaces: PhysicsEngine::add_joint with JointKind::Spring (mod.rs:
 * 1487-1493; rows assembly.rs:363-425). */
int oavbd_add_spring_joint(oavbd_ctx* ctx, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                           const oavbd_motor* motor, uint32_t* out_index);

/* Advance by dt seconds through the reference's fixed-step accumulator
 * (1/60 s substeps, at most 4 per call; non-finite or non-positive dt is
 * a no-op).  Replaces: PhysicsEngine::step for AVBD (mod.rs:1212-1235). */
int oavbd_step(oavbd_ctx* ctx, float dt);

/* Run exactly nsteps fixed 1/60 s substeps (step_inner, mod.rs:846-1143). */
int oavbd_step_n(oavbd_ctx* ctx, uint32_t nsteps);

/* Copy the body state to host: n x 13 floats (pos3, quat4, vel3, angvel3),
 * the layout of the reference's snapshot render (crates/physics/tests/
 * avbd_engine.rs:1605-1630).  Replaces: get_body reads / try_download
 * (crates/physics/src/gpu/avbd.rs:712-758). */
int oavbd_get_state(oavbd_ctx* ctx, float* out13, uint32_t n);

/* Overwrite the body state from host (same layout as oavbd_get_state);
 * the reference's get_body_mut pose/velocity edits. */
int oavbd_set_state(oavbd_ctx* ctx, const float* in13, uint32_t n);

/* Contact pairs in canonical (a, b) order: ab[2i], ab[2i+1]; *n receives
 * the pair count (up to cap entries are written; ab may be NULL to query). */
int oavbd_get_pairs(oavbd_ctx* ctx, uint32_t* ab, uint32_t cap, uint32_t* n);

/* Colour of every body in the last step's sweep (-1: not swept). */
int oavbd_get_colours(oavbd_ctx* ctx, int32_t* colour, uint32_t n);

/* Sleep flags and timers (reference sleep_body / sleep_timer). */
int oavbd_get_sleep(oavbd_ctx* ctx, uint8_t* asleep, float* timer, uint32_t n);

/* Contact points in pair order, 16 floats each: ra3 rb3 lam3 pen3 roll3
 * stuck; *n receives the point count. */
int oavbd_get_points(oavbd_ctx* ctx, float* out16, uint32_t cap, uint32_t* n);

/* TOI clamps of the last step per body: distance travelled to the hit
 * (t < 0: no clamp) and the body hit (the data of the reference's CCD
 * ContactEventKind::Hit, crates/physics/src/avbd/mod.rs:919-937). */
int oavbd_get_toi(oavbd_ctx* ctx, float* t, int32_t* target, uint32_t n);

/* Joint duals, 12 floats each: linear lam3, pen3, angular lam3, pen3. */
int oavbd_get_joints(oavbd_ctx* ctx, float* out12, uint32_t n);

int oavbd_get_stats(oavbd_ctx* ctx, oavbd_stats* out);

/* Per-kernel-class device time of the last step in milliseconds, measured
 * with HIP events on the context stream: [0] setup (TOI, broadphase,
 * narrowphase, merge, colouring), [1] primal sweep, [2] dual update,
 * [3] finish (BDF1, sleep), [4] whole step.  Timing is recorded only while
 * enabled with oavbd_set_timing(ctx, 1). */
int oavbd_set_timing(oavbd_ctx* ctx, int enable);
int oavbd_get_timing(oavbd_ctx* ctx, float* ms5);

/* Launch shape of the primal solve: `lanes` lanes cooperate on one body
 * (16; the lanes' row sums are combined by a fixed tree, so the lane
 * count is part of the arithmetic) and a workgroup holds `waves` wavefronts
 * (1, 2, 4 or 8).  Results are bitwise independent of `waves`; only speed
 * changes.  Default 16 lanes, 2 wavefronts. */
int oavbd_set_primal_shape(oavbd_ctx* ctx, uint32_t lanes, uint32_t waves);

/* Primal launch variants, chosen per colour class by its size (bucket b
 * holds classes of 2^b .. 2^(b+1) - 1 bodies, b < 24): 0-3 sixteen lanes
 * per body with 1, 2, 4 or 8 wavefronts per workgroup; 4-6 one body per
 * wavefront, its contact records staged in LDS first, 1, 2 or 4 wavefronts
 * per workgroup; 7-9 sixteen lanes per body with the next item's records
 * fetched while the current item is evaluated, 1, 2 or 4 wavefronts per
 * workgroup.  Every variant gives the same bits.  -1 leaves a bucket to
 * the default rule (variant 7 below 8,192 bodies, the oavbd_set_primal_shape
 * wavefront count above).  No reference counterpart: the reference sweeps
 * bodies serially (mod.rs:1052-1059).
 *
 * oavbd_autotune runs the next nsteps (1..64) steps once per variant from
 * one saved state, times every primal launch with HIP events, and gives
 * each bucket the variant with the least mean launch time (unmeasured
 * buckets take the nearest measured bucket's choice).  The context's state,
 * statistics and kernel times are restored afterwards; table_out (24
 * entries, may be NULL) receives the table and mean_us_out (10 x 24, may be
 * NULL) the mean launch time in microseconds of every variant in every
 * bucket (-1: no launch).  The table can be carried to other contexts with
 * oavbd_set_primal_table (n <= 24 entries). */
int oavbd_autotune(oavbd_ctx* ctx, uint32_t nsteps, int32_t* table_out, float* mean_us_out);
int oavbd_set_primal_table(oavbd_ctx* ctx, const int32_t* table, uint32_t n);
int oavbd_get_primal_table(oavbd_ctx* ctx, int32_t* table, uint32_t n);

/* Fused tail: the last colour classes of a sweep, when at least two of them
 * hold at most `max_bodies` bodies together, are solved by one launch in
 * which a body waits only for its own partners of lower classes instead of
 * for a whole class (the colour-by-colour order, the same bits).  0 turns
 * it off (one launch per class), the default: on the workloads measured the
 * fused launch is slower (its cross-workgroup waits pay for L2 write-backs
 * and invalidations).  The launch holds one body
 * per wavefront and at most `max_waves` wavefronts (0: as many as are
 * resident on the device at once; fewer make each wavefront solve several
 * bodies in list order).  The tail is also off while oavbd_autotune
 * measures the per-class variants.  No reference counterpart (the
 * reference sweeps bodies serially, mod.rs:1052-1059). */
int oavbd_set_primal_tail(oavbd_ctx* ctx, uint32_t max_bodies, uint32_t max_waves);

/* Anisotropic friction of the first n bodies.  Replaces:
 * RigidBody::set_friction_axis / clear_friction_axis and
 * friction_transverse (crates/physics/src/body.rs:276-286, :444-466) as
 * friction_frame reads them (crates/physics/src/avbd/rows.rs:32-52):
 * axis3 is a body-frame direction (normalised here; zero: isotropic, a
 * non-normalizable axis is rejected), transverse the friction across it.
 * A contact's first friction direction is body a's axis (else b's) turned
 * with the body's current orientation and projected on the contact plane,
 * re-evaluated at every solve and dual update; the coefficients along and
 * across it are the larger of the two bodies' (a body without an axis
 * counts its friction both ways).  Axes persist through body edits; new
 * bodies are isotropic; oavbd_set_bodies resets them.  Differentiable
 * stepping rejects a context with axes. */
int oavbd_set_friction_axes(oavbd_ctx* ctx, const float* axis3, const float* transverse, uint32_t n);

/* Contact and trigger events.  Replaces: PhysicsEngine::drain_contact_events
 * / drain_trigger_events (crates/physics/src/avbd/mod.rs:1591-1597) with
 * the events emit_events produces (mod.rs:1606-1669) and the TOI clamp's
 * impact events (mod.rs:943-960, :1126-1136).  Off by default (each
 * reported step then waits for its counts and copies the touching pairs
 * back); oavbd_set_events(ctx, 1) turns them on.  The last substep of each
 * oavbd_step / oavbd_step_n call reports: trigger pairs whose overlap
 * began or ended since the last report (Entered / Exited, pairs ascending),
 * contact pairs that started or stopped touching by more than 5 mm (Begin /
 * End, ascending), then Hit for newly touching pairs closing faster than
 * 1 m/s (pair order: ascending (a, b)); every substep adds the TOI clamp's
 * impacts above 1 m/s (movers ascending), those of a pair already reported
 * as a hit in the same substep dropped.  Hit carries the contact point,
 * the normal from body a towards body b and the closing speed. */
typedef struct oavbd_event {
  uint32_t kind;      /* OAVBD_EVENT_*                                      */
  uint32_t a, b;      /* body handles, a < b                                */
  float point[3];     /* Hit: contact point (world)                         */
  float normal[3];    /* Hit: unit normal from a towards b                  */
  float approach;     /* Hit: closing speed along the normal (m/s, > 1)    */
} oavbd_event;
enum {
  OAVBD_EVENT_BEGIN = 0,
  OAVBD_EVENT_END = 1,
  OAVBD_EVENT_HIT = 2,
  OAVBD_EVENT_ENTERED = 3, /* trigger events */
  OAVBD_EVENT_EXITED = 4
};
int oavbd_set_events(oavbd_ctx* ctx, int enable);
/* which = 0: contact events, 1: trigger events.  With out = NULL, *n is the
 * number queued; otherwise the oldest min(cap, queued) are copied to out,
 * removed from the queue, and *n is their number. */
int oavbd_drain_events(oavbd_ctx* ctx, int which, oavbd_event* out, uint32_t cap, uint32_t* n);
/* The last report's touching (which = 0) or overlapping trigger (1) pairs,
 * ascending (a, b) pairs: the solver-migration event baseline.  Replaces:
 * AvbdEngine::event_state / restore_event_state (mod.rs:516-556); a host
 * that rebuilds a context restores it so no Begin / Entered repeats.
 * oavbd_remove_body reports Exited for h's trigger pairs, renumbers the
 * rest and drops the touch baseline and the undrained contact events
 * (mod.rs:1240-1290). */
int oavbd_get_event_state(oavbd_ctx* ctx, int which, uint32_t* ab, uint32_t cap, uint32_t* n);
int oavbd_set_event_state(oavbd_ctx* ctx, int which, const uint32_t* ab, uint32_t n);

/* Unwaited stepping (off by default: one unwaited batch that outgrew its
 * buffers ended in a GPU memory fault whose cause has not been found, so
 * the path stays disabled until it is).  Once one step's counts have been
 * read back, the steps of an oavbd_step_n call run without the host
 * reading any count mid-step (TOI casts, candidates, pairs, points, colour
 * classes): launches are sized from buffers kept a quarter above the last
 * counts and from the last colour count plus `spare_colours` extra class
 * launches per sweep, and every kernel takes its count from the device.  A
 * batch in which a count outgrew its buffer is restored from a device
 * checkpoint taken at its start and repeated with the host reading every
 * count (oavbd_stats.replays), so the results are the same bits either
 * way.  Recorded (differentiable), timed and autotuning steps always wait.
 * No reference counterpart (the reference's CPU step has no host/device
 * split). */
int oavbd_set_unwaited(oavbd_ctx* ctx, int enable, uint32_t spare_colours);

/* Rung-1 staged bodies: n GpuAvbdMass records (16 bytes: inertia[3],
 * inv_mass) and n GpuBodyState residuals (32 bytes: lin[3], pad, ang[3],
 * pad) are solved on the device -- lumped inertial Hessian diagonal and its
 * diagonal solve, 0 on axes at or below the degenerate threshold -- into n
 * GpuBodyState deltas.  Replaces: the rung-1 kernel cs_main
 * (crates/gpu-avbd/src/lib.rs:113-148) with the semantics of
 * GpuAvbdStub::stage_diag_solve (crates/physics/src/gpu/avbd.rs:207-226),
 * avbd_inertial_hessian_diag (:88-107) and avbd_diag_solve_cpu (:117-128).
 * Independent of the context's bodies. */
int oavbd_solve_diag(oavbd_ctx* ctx, const void* masses16, const void* rhs32, uint32_t n, float dt, void* delta32);

/* Rung-2 staged systems: n packed GpuAvbdSystem records (96 bytes each:
 * inertia[3], inv_mass | residual lin[3], pad, ang[3], pad | row axis[3],
 * pen, lever[3], force, sign, pad[3]) are solved on the device -- lumped
 * inertial Hessian diagonal, one stamped row, dense 6x6 LDL -- into n
 * GpuBodyState deltas (32 bytes: lin[3], 0, ang[3], 0) and ok flags (1.0
 * solved, 0.0 breakdown with a zero delta).  Replaces: WgpuAvbdSolver::upload
 * / solve / try_download (crates/physics/src/gpu/avbd.rs:645-760) with the
 * semantics of avbd_stage_contact_solve (:415-449); the record layouts are
 * the reference's (:55-60, :264-281, :320-327; crates/physics/src/gpu/
 * mod.rs:200-207).  Independent of the context's bodies. */
int oavbd_solve_systems(oavbd_ctx* ctx, const void* systems96, uint32_t n, float dt, void* delta32, float* ok);

/* Rung-2 staged systems exactly as oavbd_solve_systems, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The call allocates nothing: the records are staged through a fixed device
 * array of the library, one per GPU (resolved when a context is created on
 * it), holding at most 65536 systems per call (8.65 MB, held until the
 * process ends); larger n returns OAVBD_ECAPACITY.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (from several threads
 * or contexts) race and must be serialised by the caller. */
int oavbd_solve_systems_noalloc(oavbd_ctx* ctx, const void* systems96, uint32_t n, float dt, void* delta32,
                                float* ok);

/* Device checkpoints of the solver state: bodies (pose, velocities, sleep
 * state), the contact pair table with its points' anchors, duals,
 * penalties and stuck flags, and the joint duals -- everything a step reads
 * from earlier steps, so stepping on from a loaded checkpoint repeats the
 * steps taken after the save bit for bit.  Replaces: the warm-start state
 * the reference keeps across steps in AvbdEngine's pairs / joints
 * (crates/physics/src/avbd/mod.rs:424-464), which its snapshot helpers
 * read (:469-636).  Slots are numbered from 0; a slot keeps its device
 * memory until the context is destroyed. */
int oavbd_checkpoint_save(oavbd_ctx* ctx, uint32_t slot);
int oavbd_checkpoint_load(oavbd_ctx* ctx, uint32_t slot);

/* Differentiable stepping (no counterpart in the reference).
 * oavbd_diff_forward runs nsteps fixed steps like oavbd_step_n and keeps a
 * checkpoint every `every` steps (they use checkpoint slots 0..).
 * oavbd_diff_backward then returns the gradient of a loss L given dL/d(final
 * state) (n x 13, the oavbd_get_state layout): dL/d(initial state) (n x 13;
 * pose, velocity, angular velocity), dL/d(mass) and dL/d(friction
 * coefficient) per body (any output may be NULL).  Each interval between
 * checkpoints is re-run (bit for bit) while its steps are recorded, and a
 * hand-written adjoint of the step runs backwards over them: the active set
 * of every step -- its pairs, contact anchors, normals and friction frames,
 * colouring, warm-start matches, sleep states and the branches of the
 * force clamps and penalty updates -- is held fixed and everything
 * continuous is differentiated.  A pair's friction is the larger of its
 * bodies' coefficients (rows.rs:32-52); its adjoint goes to that body (the
 * second on a tie).  Scenes with joints, rolling / torsion friction or
 * kinematic bodies, and steps where a TOI clamp fires, are rejected
 * (OAVBD_EINVAL / OAVBD_EDEVICE with a message).  Gradients are summed with
 * float atomics, so their last bits may vary from run to run. */
int oavbd_diff_forward(oavbd_ctx* ctx, uint32_t nsteps, uint32_t every);
int oavbd_diff_backward(oavbd_ctx* ctx, const float* dfinal13, float* dinit13, float* dmass, float* dfriction);

/* Block until all work of the context has finished. */
int oavbd_sync(oavbd_ctx* ctx);

/* Last error message of the context (never NULL). */
const char* oavbd_last_error(oavbd_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* OAVBD_H */
