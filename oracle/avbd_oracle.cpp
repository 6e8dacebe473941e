// avbd_oracle.cpp — CPU restatement of the reference AVBD rigid-body step.
//
// TEST INFRASTRUCTURE.  This is the checker the HIP step is compared with;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  The product library (ornis_amd/csrc) neither includes nor links
// anything under oracle/.
//
// Restated from the reference's Rust source, read as text (it cannot be
// built here: no Rust toolchain).  Scope: Box / Sphere / HalfSpace shapes,
// Ball joints, static / dynamic / kinematic bodies, sleep, TOI, contact
// manifolds with persistent duals, rolling/torsion friction.
//
//   step (accumulator)      crates/physics/src/avbd/mod.rs:1212-1235
//   step_inner              crates/physics/src/avbd/mod.rs:846-1143
//   generate_pairs / merge  crates/physics/src/avbd/rows.rs:767-943
//   discover_pair           crates/physics/src/avbd/rows.rs:432-756
//   solve_body / solve_6x6  crates/physics/src/avbd/assembly.rs:25-903
//   dual_update             crates/physics/src/avbd/joints_ext.rs:464-919
//   sleep / joint wake      crates/physics/src/avbd/sleep.rs:1-175
//   shape_distance et al.   crates/physics/src/collision/distance.rs:63-891
//   box SAT / features      crates/physics/src/collision/distance_box.rs:21-147
//   staged GPU rungs 1, 2   crates/physics/src/gpu/avbd.rs:88-128, :207-226, :415-449
//
// Two orders:
//   ORDER_REFERENCE  the reference exactly: reverse-handle Gauss-Seidel sweep,
//                    history-ordered pair vector (new pairs appended,
//                    unseen pairs swap-removed), sequential TOI clamp.
//   ORDER_COLOUR     the GPU step's order: the same arithmetic, with the pair
//                    vector sorted by (a, b) after every merge, the sweep run
//                    colour by colour of a deterministic greedy colouring,
//                    and every TOI cast taken against the step-top poses
//                    before any clamp is applied.
// Within one body solve both orders sum the 6x6 rows in pair-vector order,
// exactly as the reference does.
//
// Discovery and TOI candidate search use a uniform grid above a small body
// count; the exact reference prefilters run on every candidate, so the
// discovered set (and everything downstream) is identical to the O(n^2)
// scan, which stays selectable for tests.
//
// What pins this restatement (tests/test_oracle.py):
//   - the reference's golden file crates/physics/tests/data/avbd_snapshot.hex
//     (9 bodies x 120 steps, ORDER_REFERENCE, bit-exact with libm and with
//     the portable trig): discovery, damper, TOI, sleep, dual, Ball joint;
//   - closed forms: solve_6x6 vs a dense solve and its breakdown cases,
//     stamp_row vs the analytic J J^T, contact_force audit answers (A3, A6,
//     push-only), tangent basis, quat_integrate / quat_diff_vec inverse;
//   - invariants: stack stands, damper momentum, determinism, grid = brute
//     discovery, proper greedy colouring, FP64 tracks FP32;
//   - staged rungs (tests/test_stage_solve.py): the reference's known
//     answers for the lumped Hessian diagonal and the staged contact cases
//     (crates/physics/tests/gpu_batches.rs:537-572, :904-926, :968-1010),
//     and a float64 dense solve of the stamped system on random systems.
// Parity unpinned: glam Quat op order under non-identity rotations (the
// golden file only holds identity rotations), the reference's libm
// sin_cos / asin bits (MATH_PORTABLE is pinned to libm within 2 ulp, not to
// the reference), kinematic bodies and rolling / torsion rows.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <unordered_set>
#include <thread>
#include <type_traits>
#include <vector>

#include "oracle_math.hpp"

namespace orc {

enum Shape : int { SH_SPHERE = 0, SH_BOX = 1, SH_HALFSPACE = 2, SH_CAPSULE = 3 };  // capsule sp: (radius, half height)
enum BodyType : int { BT_STATIC = 0, BT_DYNAMIC = 1, BT_KINEMATIC = 2 };
enum Order : int { ORDER_REFERENCE = 0, ORDER_COLOUR = 1 };

// Reference constants, crates/physics/src/avbd/mod.rs:154-247, :1205-1210;
// crates/physics/src/constants.rs; crates/gpu-abi/src/lib.rs:52.
#define CF(x) T(x##f)

template <class T>
struct K {
  static T ALPHA() { return CF(0.99); }
  static T GAMMA() { return CF(0.999); }
  static T BETA() { return CF(10000.0); }
  static T BETA_ANG() { return CF(100.0); }
  static T MARGIN() { return CF(0.01); }
  static T PEN_MIN() { return CF(1.0); }
  static T PEN_MAX() { return CF(1.0e10); }
  static T PEN_INIT() { return CF(1.0); }
  static T JOINT_PEN_INIT() { return CF(1.0); }
  static T LIM_PEN_MAX() { return CF(1e4); }
  static T PI() { return CF(3.14159265358979323846); }
  static T TAU() { return CF(6.28318530717958647692); }
  static T LIMIT_SLOP_ANG() { return CF(0.005); }
  static T LIMIT_SLOP_LIN() { return CF(0.002); }
  static T MIN_EFF_MASS() { return CF(1e-9); }
  static T GEN_MARGIN() { return CF(0.05); }
  static T EXPAND_SLOP() { return CF(0.005); }
  static T POINT_MATCH() { return CF(0.03); }
  static T LAMBDA_MATCH() { return CF(0.05); }
  static T C_EPS() { return CF(1e-7); }
  static T STICK() { return CF(0.00001); }
  static T SLEEP_LIN() { return CF(0.15); }
  static T SLEEP_ANG() { return CF(0.15); }
  static T HALF() { return CF(0.5); }
  static T SLEEP_TIME() { return CF(0.5); }
  static T WAKE_IMPACT() { return CF(0.5); }
  static T WAKE_PEN() { return CF(0.01); }
  static T ROLL_PEN() { return CF(100.0); }
  static T MIN_TOI_LEN() { return CF(1e-9); }
  static T SPIN_REST() { return CF(1e-9); }
  static T CCD_BOUNCE() { return CF(1.0); }
  static T DT() { return T(1.0f / 60.0f); }
  static T NEAR_ZERO() { return CF(1e-9); }
  static T AXIS_REST_LEN2() { return CF(1e-16); }
  static T DEGENERATE() { return CF(1e-12); }
  static T SHAPE_TOUCH() { return CF(1e-3); }
  static T CCD_GATE() { return CF(0.5); }
  static T JOINT_REST() { return CF(0.005); }
  static T TANGENT_DOT() { return CF(0.9); }
  static T FACE_AXIS_EPS() { return CF(1e-6); }
  static T SPHERE_I() { return CF(0.4); }
  static T BOX_I_DIV() { return CF(12.0); }
};
static const int MAX_POINTS = 8;
static const int MAX_SUBSTEPS = 4;

template <class T>
struct Body {
  V3<T> pos;
  Q<T> rot;
  V3<T> vel, angvel, torque;
  T mass, inv_mass;
  V3<T> inertia;
  T restitution, friction, rolling, torsion;
  int shape;
  V3<T> sp;  // sphere: (r, 0, 0); box: half extents; half-space: unit normal
  uint32_t layer, mask;
  bool trigger;
  int type;
  // anisotropic friction (body.rs:276-286): unit axis in the body frame
  // (friction_dir; zero: isotropic) and the friction across it
  V3<T> fdir{T(0), T(0), T(0)};
  T ftrans = T(0);
  bool has_fdir() const { return fdir.x != T(0) || fdir.y != T(0) || fdir.z != T(0); }
};

template <class T>
struct Point {
  V3<T> ra, rb;
  T lam[3], pen[3];
  bool stuck;
  T roll[3];
  int src_pair = -1, src_pt = -1;  // point of the previous step's pair vector inherited from
};

// A contact or trigger event (ContactEvent / TriggerEvent, trigger.rs:29-81):
// kind EV_*, the pair, and for a hit the point, normal and approach speed.
enum EventKind : int { EV_BEGIN = 0, EV_END = 1, EV_HIT = 2, EV_ENTERED = 3, EV_EXITED = 4 };
template <class T>
struct Event {
  int kind, a, b;
  V3<T> point, normal;
  T approach;
};

template <class T>
struct Pair {
  int a, b;
  V3<T> n;
  T mu[2];
  T gap;
  std::vector<Point<T>> pts;
};

// Joint kinds of this restatement (AvbdJointKind, mod.rs:286-300, same
// numbering): Gear is out of its scope.
enum JointKindC : int {
  JK_BALL = 0, JK_REVOLUTE = 1, JK_PRISMATIC = 2, JK_FIXED = 3, JK_DISTANCE = 4, JK_ROPE = 5, JK_SPRING = 6,
  JK_WHEEL = 7, JK_GEAR = 8, JK_SIXDOF = 9, JK_MOTOR = 10
};
// AxisConfig (joint.rs:676-690) of a SixDof axis.
enum AxisCfg : int { AX_FREE = 0, AX_LOCKED = 1, AX_LIMITED = 2 };
inline bool is_tether(int kind) { return kind == JK_DISTANCE || kind == JK_ROPE || kind == JK_SPRING; }
// JointMotor (joint.rs:962-1126) as a Spring joint carries it: kind
// (Velocity, Position, Servo), model (AccelerationBased, ForceBased).
enum MotorKindC : int { MK_VELOCITY = 0, MK_POSITION = 1, MK_SERVO = 2 };
enum MotorModelC : int { MM_ACCELERATION = 0, MM_FORCE = 1 };

template <class T>
struct Joint {
  int a, b;
  int kind = JK_BALL;
  V3<T> la, lb;
  V3<T> ax_a{T(1), T(0), T(0)}, ax_b{T(1), T(0), T(0)};  // unit hinge / slide axes (body frames)
  V3<T> dref{T(0), T(0), T(0)};                          // Fixed: assembly anchor offset in A's frame
  Q<T> q_ref{T(0), T(0), T(0), T(1)};                    // Fixed: assembly relative rotation conj(qa) qb
  // Distance: assembly anchor distance; Rope: maximum; Spring: the motor's
  // target position (rest length)
  T ref_val = T(0);
  // Spring: the motor's stiffness, damping, force budget, target velocity
  T m_stiff = T(0), m_damp = T(0), m_max = T(0), m_tvel = T(0);
  int m_kind = MK_POSITION, m_model = MM_ACCELERATION;
  // Revolute / Prismatic: travel window about the assembly angle / length
  // (ref_val), a velocity motor [target speed, budget], and the limit's
  // one-sided force accumulators (acc_lim: revolute, lim_dual: prismatic)
  bool has_lim = false, has_mot = false;
  T lim_lo = T(0), lim_hi = T(0), mot_speed = T(0), mot_max = T(0);
  T acc_lim = T(0), lim_dual = T(0);
  // Motor (free drive): relative linear / angular velocity targets (world,
  // B minus A) and their force / torque budgets
  V3<T> lin_target{T(0), T(0), T(0)}, ang_target{T(0), T(0), T(0)};
  T max_force = T(0), max_torque = T(0);
  // Wheel: the axle in each body's frame (ax_a / ax_b hold the suspension
  // axes) and the suspension's frequency (Hz) and damping ratio
  V3<T> bx_a{T(1), T(0), T(0)}, bx_b{T(1), T(0), T(0)};
  T susp_hz = T(0), susp_zeta = T(0);
  // SixDof: per axis of A's frame (linear x y z, then angular x y z) its
  // AxisCfg and window, and the limited axes' force accumulators
  int six[6] = {AX_FREE, AX_FREE, AX_FREE, AX_FREE, AX_FREE, AX_FREE};
  T six_lo[6] = {}, six_hi[6] = {}, sacc[6] = {};
  // Gear: the two Revolute / Prismatic joints it couples (ca + gratio cb =
  // ref_val) and the step-start memory of their raw and continuous
  // coordinates (joints_ext.rs:53-115), set once gmem
  int gb[2] = {-1, -1};
  T gratio = T(0);
  bool gmem = false;
  T graw[2] = {}, gcont[2] = {};
  T lam_l[3], pen_l[3];
  T lam_a[3], pen_a[3];
};

template <class T>
struct Dist {
  T dist;
  V3<T> pa, pb;
};

template <class T>
struct SRef {
  int shape;
  V3<T> sp;
  V3<T> pos;
  Q<T> rot;
};

// ---------------------------------------------------------------------------
// Shape helpers (crates/physics/src/collision/shape.rs, body.rs, rows.rs)
// ---------------------------------------------------------------------------

// Shape::inertia for the dynamic arms (shape.rs:713-767): sphere
// 0.4*m*r^2, box m/12*(y^2+z^2) per axis over full side lengths.
template <class T>
V3<T> shape_inertia(int shape, V3<T> sp, T mass) {
  if (!(mass > T(0)) || !std::isfinite(mass)) return V3<T>{T(0), T(0), T(0)};
  if (shape == SH_SPHERE) {
    T v = K<T>::SPHERE_I() * mass * sp.x * sp.x;
    return V3<T>{v, v, v};
  }
  if (shape == SH_BOX) {
    V3<T> s = sp * T(2);
    T x = s.x * s.x, y = s.y * s.y, z = s.z * s.z;
    T k = mass / K<T>::BOX_I_DIV();
    return V3<T>{k * (y + z), k * (z + x), k * (x + y)};
  }
  if (shape == SH_CAPSULE) {  // shape.rs:769-781, along the body's y axis
    const T r = sp.x, h = sp.y;
    const T iy = T(0.5) * mass * r * r;
    const T ixz = T(0.25) * mass * (r * r) + (mass / T(3)) * h * r + T(0.25) * mass * h * h;
    return V3<T>{ixz, iy, ixz};
  }
  return V3<T>{T(0), T(0), T(0)};
}

// bound_radius (rows.rs:253-292).
template <class T>
T bound_radius(int shape, V3<T> sp) {
  if (shape == SH_SPHERE) return sp.x;
  if (shape == SH_BOX) return length(sp);
  if (shape == SH_CAPSULE) return sp.y + sp.x;
  return std::numeric_limits<T>::infinity();
}

// shape_min_dimension (rows.rs:373-403).
template <class T>
T shape_min_dim(int shape, V3<T> sp) {
  if (shape == SH_SPHERE) return sp.x;
  if (shape == SH_BOX) return min_element(sp);
  if (shape == SH_CAPSULE) return sp.x;
  return std::numeric_limits<T>::infinity();
}

// tangent_basis (crates/physics/src/math.rs:152-160).
template <class T>
void tangent_basis(V3<T> n, V3<T>* t1, V3<T>* t2) {
  V3<T> axis = std::fabs(n.x) < K<T>::TANGENT_DOT() ? V3<T>{T(1), T(0), T(0)} : V3<T>{T(0), T(1), T(0)};
  *t1 = normalize_or(cross(n, axis), V3<T>{T(0), T(0), T(1)});
  *t2 = cross(*t1, n);
}

// orthogonalize_axle (crates/physics/src/math.rs:179-186): the axle with
// its suspension component removed, normalised; the suspension's first
// tangent when that leaves (almost) nothing.
template <class T>
V3<T> orthogonalize_axle(V3<T> suspension, V3<T> axle) {
  const V3<T> a = axle - suspension * dot(axle, suspension);
  if (length_sq(a) < CF(1e-6)) {
    V3<T> t1, t2;
    tangent_basis(suspension, &t1, &t2);
    return t1;
  }
  return normalize(a);
}

// friction_frame (rows.rs:32-52): the first friction direction from body
// a's axis, else b's, projected on the contact plane and normalised
// (UnitVec3::normalize_checked, invariants.rs:55-60: none when not finite
// or shorter than DEGENERATE_LEN2), else tangent_basis(n); the friction
// coefficients along it and across it.
template <class T>
V3<T> friction_t1(const Body<T>& a, const Body<T>& b, V3<T> n) {
  auto pick = [&](const Body<T>& bd, V3<T>* out) {
    if (!bd.has_fdir()) return false;
    const V3<T> world = rotate(bd.rot, bd.fdir);
    const V3<T> proj = world - n * dot(world, n);
    if (!(std::isfinite((double)proj.x) && std::isfinite((double)proj.y) && std::isfinite((double)proj.z)) ||
        dot(proj, proj) < T(1e-12f))
      return false;
    *out = normalize(proj);
    return true;
  };
  V3<T> t1;
  if (pick(a, &t1) || pick(b, &t1)) return t1;
  V3<T> t2;
  tangent_basis(n, &t1, &t2);
  return t1;
}
template <class T>
void friction_mu(const Body<T>& a, const Body<T>& b, T* mu1, T* mu2) {
  const T a2 = a.has_fdir() ? a.ftrans : a.friction, b2 = b.has_fdir() ? b.ftrans : b.friction;
  *mu1 = fmax_(a.friction, b.friction);
  *mu2 = fmax_(a2, b2);
}

// quat_mul (rows.rs:91-98).
template <class T>
Q<T> quat_mul(Q<T> a, Q<T> b) {
  return Q<T>{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

// quat_integrate (rows.rs:110-119).
template <class T>
Q<T> quat_integrate(Q<T> q, V3<T> v, int mm) {
  T theta = length(v);
  if (theta < K<T>::NEAR_ZERO()) return q;
  T s, c;
  sincos_t(K<T>::HALF() * theta, &s, &c, mm);
  T k = s / theta;
  Q<T> dq = quat_mul(Q<T>{v.x * k, v.y * k, v.z * k, c}, q);
  return qnormalize(dq);
}

// quat_diff_vec (rows.rs:128-146).
template <class T>
V3<T> quat_diff_vec(Q<T> a, Q<T> b, int mm) {
  Q<T> r = quat_mul(a, Q<T>{-b.x, -b.y, -b.z, b.w});
  V3<T> xyz{r.x, r.y, r.z};
  if (r.w < T(0)) xyz = -xyz;
  T s = fmin_(length(xyz), T(1));
  if (s < K<T>::NEAR_ZERO()) return V3<T>{T(0), T(0), T(0)};
  return xyz * (T(2) * asin_t(s, mm) / s);
}

// world_inertia (rows.rs:194-207).
template <class T>
void world_inertia(V3<T> inertia, Q<T> rot, T m[3][3]) {
  M3c<T> r = mat3_from_quat(rot);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      T ca0 = r.c[0][a], ca1 = r.c[1][a], ca2 = r.c[2][a];
      T cb0 = r.c[0][b], cb1 = r.c[1][b], cb2 = r.c[2][b];
      m[a][b] = ca0 * inertia.x * cb0 + ca1 * inertia.y * cb1 + ca2 * inertia.z * cb2;
    }
}

// inverse_symmetric (rows.rs:211-238).
template <class T>
void inverse_symmetric(const T m[3][3], V3<T> diag, T out[3][3]) {
  T a = m[0][0], b = m[0][1], c = m[0][2];
  T d = m[1][1], e = m[1][2], f = m[2][2];
  T det = a * (d * f - e * e) - b * (b * f - c * e) + c * (b * e - c * d);
  if (std::fabs(det) < K<T>::DEGENERATE()) {
    T ix = diag.x > T(0) ? T(1) / diag.x : T(0);
    T iy = diag.y > T(0) ? T(1) / diag.y : T(0);
    T iz = diag.z > T(0) ? T(1) / diag.z : T(0);
    T z = T(0);
    T r[3][3] = {{ix, z, z}, {z, iy, z}, {z, z, iz}};
    std::memcpy(out, r, sizeof(r));
    return;
  }
  T id = T(1) / det;
  out[0][0] = (d * f - e * e) * id;
  out[0][1] = (c * e - b * f) * id;
  out[0][2] = (b * e - c * d) * id;
  out[1][0] = (c * e - b * f) * id;
  out[1][1] = (a * f - c * c) * id;
  out[1][2] = (b * c - a * e) * id;
  out[2][0] = (b * e - c * d) * id;
  out[2][1] = (b * c - a * e) * id;
  out[2][2] = (a * d - b * b) * id;
}

template <class T>
V3<T> mat3_vec(const T m[3][3], V3<T> v) {
  return V3<T>{m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
               m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}

// ---------------------------------------------------------------------------
// Distance queries (crates/physics/src/collision/distance.rs,
// distance_box.rs, gjk.rs:68-79)
// ---------------------------------------------------------------------------

template <class T>
V3<T> seg_pt_dummy(V3<T> a) {
  return a;
}

// seg_seg_closest (distance.rs:75-113).
template <class T>
void seg_seg_closest(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1, V3<T>* x, V3<T>* y) {
  const T EPS = CF(1e-12);
  V3<T> d1 = a1 - a0, d2 = b1 - b0, r = a0 - b0;
  T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  if (a < EPS && e < EPS) {
    *x = a0;
    *y = b0;
    return;
  }
  if (a < EPS) {
    T t = clampf(f / e, T(0), T(1));
    *x = a0;
    *y = b0 + d2 * t;
    return;
  }
  T c = dot(d1, r);
  if (e < EPS) {
    T t = clampf(-c / a, T(0), T(1));
    *x = a0 + d1 * t;
    *y = b0;
    return;
  }
  T b = dot(d1, d2);
  T denom = a * e - b * b;
  T s = std::fabs(denom) > EPS ? clampf((b * f - c * e) / denom, T(0), T(1)) : T(0);
  T t = e > EPS ? (b * s + f) / e : T(0);
  if (!(t >= T(0) && t <= T(1))) {
    t = clampf(t, T(0), T(1));
    s = clampf((b * t - c) / a, T(0), T(1));
  }
  *x = a0 + d1 * s;
  *y = b0 + d2 * t;
}

// point_obb_closest (distance.rs:117-138).
// ORC_PLANT_CONJ builds a deliberately wrong checker (the box's local frame
// taken with the rotation instead of its conjugate): the float64 geometry
// pins must reject that build (tests/test_oracle_rotation.py).
template <class T>
V3<T> point_obb_closest(V3<T> p, V3<T> pos, V3<T> half, Q<T> rot) {
#ifdef ORC_PLANT_CONJ
  V3<T> l = rotate(rot, p - pos);
#else
  V3<T> l = rotate(conj(rot), p - pos);
#endif
  V3<T> cl = vclamp(l, -half, half);
  V3<T> q;
  if (cl == l) {
    T dx = half.x - std::fabs(l.x), dy = half.y - std::fabs(l.y), dz = half.z - std::fabs(l.z);
    q = l;
    if (dx <= dy && dx <= dz)
      q.x = std::copysign(half.x, l.x);
    else if (dy <= dz)
      q.y = std::copysign(half.y, l.y);
    else
      q.z = std::copysign(half.z, l.z);
  } else {
    q = cl;
  }
  return pos + rotate(rot, q);
}

// obb_corners (distance.rs:142-156): bit2 = x, bit1 = y, bit0 = z, 0 = +.
template <class T>
void obb_corners(V3<T> pos, V3<T> half, Q<T> rot, V3<T> out[8]) {
  V3<T> x = rotate(rot, V3<T>{T(1) * half.x, T(0) * half.x, T(0) * half.x});
  V3<T> y = rotate(rot, V3<T>{T(0) * half.y, T(1) * half.y, T(0) * half.y});
  V3<T> z = rotate(rot, V3<T>{T(0) * half.z, T(0) * half.z, T(1) * half.z});
  out[0] = ((pos + x) + y) + z;
  out[1] = ((pos + x) + y) - z;
  out[2] = ((pos + x) - y) + z;
  out[3] = ((pos + x) - y) - z;
  out[4] = ((pos - x) + y) + z;
  out[5] = ((pos - x) + y) - z;
  out[6] = ((pos - x) - y) + z;
  out[7] = ((pos - x) - y) - z;
}

static const int OBB_EDGES[12][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 2}, {1, 3},
                                     {4, 6}, {5, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}};

// sphere_sphere (distance.rs:180-189).
template <class T>
Dist<T> sphere_sphere(const SRef<T>& a, T ra, const SRef<T>& b, T rb) {
  V3<T> d = b.pos - a.pos;
  T cd = length(d);
  V3<T> n = normalize_or(d, V3<T>{T(1), T(0), T(0)});
  return Dist<T>{cd - ra - rb, a.pos + n * ra, b.pos - n * rb};
}

// sphere_obb (distance.rs:192-216).
template <class T>
Dist<T> sphere_obb(const SRef<T>& a, T r, const SRef<T>& b, V3<T> half) {
  V3<T> pb = point_obb_closest(a.pos, b.pos, half, b.rot);
  V3<T> dv = a.pos - pb;
  T core = length(dv);
  V3<T> local = rotate(conj(b.rot), a.pos - b.pos);
  V3<T> al = vabs(local);
  bool inside = al.x <= half.x && al.y <= half.y && al.z <= half.z;
  V3<T> n;
  if (inside) {
    V3<T> cl = half - al;
    V3<T> axis;
    if (cl.x <= cl.y && cl.x <= cl.z) {
      T s = signum(local.x);
      axis = V3<T>{T(1) * s, T(0) * s, T(0) * s};
    } else if (cl.y <= cl.z) {
      T s = signum(local.y);
      axis = V3<T>{T(0) * s, T(1) * s, T(0) * s};
    } else {
      T s = signum(local.z);
      axis = V3<T>{T(0) * s, T(0) * s, T(1) * s};
    }
    n = normalize_or(pb - a.pos, rotate(b.rot, axis));
  } else {
    n = normalize_or(dv, V3<T>{T(1), T(0), T(0)});
  }
  return Dist<T>{inside ? -core - r : core - r, a.pos - n * r, pb};
}

// gjk::support, box/sphere/capsule arms (gjk.rs:68-79).
template <class T>
V3<T> support(const SRef<T>& s, V3<T> dir) {
  if (s.shape == SH_SPHERE) return s.pos + normalize_or(dir, V3<T>{T(1), T(0), T(0)}) * s.sp.x;
  if (s.shape == SH_CAPSULE) {
    const V3<T> axis = rotate(s.rot, V3<T>{T(0), T(1), T(0)});
    const V3<T> tip = (dot(dir, axis) >= T(0) ? axis : -axis) * s.sp.y;
    return s.pos + tip + normalize_or(dir, V3<T>{T(1), T(0), T(0)}) * s.sp.x;
  }
  V3<T> local = rotate(conj(s.rot), dir);
  V3<T> sg{local.x >= T(0) ? T(1) : T(-1), local.y >= T(0) ? T(1) : T(-1), local.z >= T(0) ? T(1) : T(-1)};
  return s.pos + rotate(s.rot, cmul(s.sp, sg));
}

// halfspace_distance (distance.rs:607-641).
template <class T>
Dist<T> halfspace_distance(const SRef<T>& hs, const SRef<T>& other, bool first) {
  const T INF = std::numeric_limits<T>::infinity();
  Dist<T> none{INF, first ? hs.pos : other.pos, first ? other.pos : hs.pos};
  if (other.shape == SH_HALFSPACE) return none;
  V3<T> n = rotate(hs.rot, hs.sp);
  if (!finite3(n)) return none;
  V3<T> s = support(other, -n);
  T dist = dot(s - hs.pos, n);
  Dist<T> d{dist, s - n * dist, s};
  if (first) return d;
  return Dist<T>{d.dist, d.pb, d.pa};
}

// separating_axis (distance_box.rs:21-52).
template <class T>
T separating_axis(const SRef<T>& a, V3<T> ha, const SRef<T>& b, V3<T> hb, V3<T>* normal_out) {
  M3c<T> ra = mat3_from_quat(a.rot), rb = mat3_from_quat(b.rot);
  V3<T> axes[15];
  const V3<T> Z{T(0), T(0), T(0)};
  for (int i = 0; i < 3; ++i) {
    axes[i] = ra.c[i];
    axes[3 + i] = rb.c[i];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      V3<T> c = cross(ra.c[i], rb.c[j]);
      axes[6 + 3 * i + j] = length_sq(c) > CF(1e-12) ? normalize(c) : Z;
    }
  V3<T> delta = b.pos - a.pos;
  T best = -std::numeric_limits<T>::infinity();
  V3<T> bn{T(0), T(1), T(0)};
  for (int k = 0; k < 15; ++k) {
    V3<T> n = axes[k];
    if (n == Z) continue;
    T ea = ha.x * std::fabs(dot(n, ra.c[0])) + ha.y * std::fabs(dot(n, ra.c[1])) + ha.z * std::fabs(dot(n, ra.c[2]));
    T eb = hb.x * std::fabs(dot(n, rb.c[0])) + hb.y * std::fabs(dot(n, rb.c[1])) + hb.z * std::fabs(dot(n, rb.c[2]));
    T proj = dot(delta, n);
    T gap = std::fabs(proj) - ea - eb;
    if (gap > best) {
      best = gap;
      bn = proj >= T(0) ? n : -n;
    }
  }
  if (normal_out) *normal_out = bn;
  return best;
}

// support_face (distance_box.rs:81-98).
template <class T>
void support_face(const SRef<T>& body, V3<T> half, V3<T> dir, SRef<T>* out, V3<T>* ext) {
  V3<T> local = rotate(conj(body.rot), dir);
  V3<T> off{T(0), T(0), T(0)};
  V3<T> e = half;
  for (int i = 0; i < 3; ++i) {
    if (std::fabs(local[i]) > K<T>::FACE_AXIS_EPS()) {
      off[i] = signum(local[i]) * half[i];
      e[i] = T(0);
    }
  }
  *out = body;
  out->pos = body.pos + rotate(body.rot, off);
  *ext = e;
}

// feature_distance (distance_box.rs:100-133).
template <class T>
Dist<T> feature_distance(const SRef<T>& a, V3<T> ha, const SRef<T>& b, V3<T> hb) {
  V3<T> ca[8], cb[8];
  obb_corners(a.pos, ha, a.rot, ca);
  obb_corners(b.pos, hb, b.rot, cb);
  Dist<T> res{std::numeric_limits<T>::infinity(), a.pos, b.pos};
  T best = std::numeric_limits<T>::infinity();
  auto consider = [&](V3<T> x, V3<T> y) {
    T sq = length_sq(y - x);
    if (sq < best) {
      best = sq;
      res = Dist<T>{std::sqrt(sq), x, y};
    }
  };
  for (int i = 0; i < 8; ++i) consider(ca[i], point_obb_closest(ca[i], b.pos, hb, b.rot));
  for (int i = 0; i < 8; ++i) consider(point_obb_closest(cb[i], a.pos, ha, a.rot), cb[i]);
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      V3<T> x, y;
      seg_seg_closest(ca[OBB_EDGES[i][0]], ca[OBB_EDGES[i][1]], cb[OBB_EDGES[j][0]], cb[OBB_EDGES[j][1]], &x, &y);
      consider(x, y);
    }
  return res;
}

// box_box (distance_box.rs:135-147); returns the SAT gap through *sat.
template <class T>
Dist<T> box_box(const SRef<T>& a, V3<T> ha, const SRef<T>& b, V3<T> hb, T* sat) {
  V3<T> normal;
  T gap = separating_axis(a, ha, b, hb, &normal);
  *sat = gap;
  if (gap > T(0)) return feature_distance(a, ha, b, hb);
  SRef<T> fa, fb;
  V3<T> ea, eb;
  support_face(a, ha, normal, &fa, &ea);
  support_face(b, hb, -normal, &fb, &eb);
  Dist<T> r = feature_distance(fa, ea, fb, eb);
  r.dist = gap;
  return r;
}

// Capsule core segment (distance.rs:174-178): the capsule's y axis.
template <class T>
void capsule_segment(V3<T> pos, T hh, Q<T> rot, V3<T>* s0, V3<T>* s1) {
  const V3<T> axis = rotate(rot, V3<T>{T(0), T(1), T(0)});
  *s0 = pos - axis * hh;
  *s1 = pos + axis * hh;
}
// point_segment_closest (distance.rs:63-71).
template <class T>
V3<T> point_segment_closest(V3<T> p, V3<T> a, V3<T> b) {
  const V3<T> ab = b - a;
  const T len_sq = length_sq(ab);
  if (len_sq < CF(1e-12)) return a;
  const T t = clampf(dot(p - a, ab) / len_sq, T(0), T(1));
  return a + ab * t;
}
// sphere_capsule (distance.rs:218-230).
template <class T>
Dist<T> sphere_capsule(const SRef<T>& a, T r, const SRef<T>& b, T cr, T hh) {
  V3<T> s0, s1;
  capsule_segment(b.pos, hh, b.rot, &s0, &s1);
  const V3<T> core_b = point_segment_closest(a.pos, s0, s1);
  const V3<T> d_vec = a.pos - core_b;
  const T core = length(d_vec);
  const V3<T> n = normalize_or(d_vec, V3<T>{T(1), T(0), T(0)});
  return Dist<T>{core - r - cr, a.pos - n * r, core_b + n * cr};
}
// capsule_capsule (distance.rs:232-245).
template <class T>
Dist<T> capsule_capsule(const SRef<T>& a, T ra, T ha, const SRef<T>& b, T rb, T hb) {
  V3<T> a0, a1, b0, b1, ca, cb;
  capsule_segment(a.pos, ha, a.rot, &a0, &a1);
  capsule_segment(b.pos, hb, b.rot, &b0, &b1);
  seg_seg_closest(a0, a1, b0, b1, &ca, &cb);
  const V3<T> d_vec = cb - ca;
  const T core = length(d_vec);
  const V3<T> n = normalize_or(d_vec, V3<T>{T(1), T(0), T(0)});
  return Dist<T>{core - ra - rb, ca + n * ra, cb - n * rb};
}
// segment_box (distance_box.rs:148-197): closest points between the segment
// [start, end] and the solid box [lo, hi], and their squared distance:
// the segment is cut at the box's slab boundaries, and each piece's
// stationary point (and ends) tested.
template <class T>
void segment_box(V3<T> start, V3<T> end, V3<T> lo, V3<T> hi, V3<T>* core, V3<T>* surface, T* squared) {
  const V3<T> delta = end - start;
  T cuts[8];
  int nc = 0;
  cuts[nc++] = T(0);
  cuts[nc++] = T(1);
  for (int i = 0; i < 3; ++i) {
    if (std::fabs(delta[i]) > CF(1e-12)) {
      const T bounds[2] = {lo[i], hi[i]};
      for (T boundary : bounds) {
        const T t = (boundary - start[i]) / delta[i];
        if (t > T(0) && t < T(1)) cuts[nc++] = t;
      }
    }
  }
  std::sort(cuts, cuts + nc);
  *core = start;
  *surface = vclamp(start, lo, hi);
  *squared = std::numeric_limits<T>::infinity();
  for (int w = 0; w + 1 < nc; ++w) {
    const T r0 = cuts[w], r1 = cuts[w + 1];
    const T middle = (r0 + r1) * T(0.5);
    const V3<T> p = start + delta * middle;
    T num = T(0), den = T(0);
    for (int i = 0; i < 3; ++i) {
      T boundary;
      if (p[i] < lo[i])
        boundary = lo[i];
      else if (p[i] > hi[i])
        boundary = hi[i];
      else
        continue;
      num += delta[i] * (start[i] - boundary);
      den += delta[i] * delta[i];
    }
    const T stationary = den > T(0) ? clampf(-num / den, r0, r1) : middle;
    const T ts[3] = {stationary, r0, r1};
    for (T t : ts) {
      const V3<T> point = start + delta * t;
      const V3<T> closest = vclamp(point, lo, hi);
      const T sq = length_sq(point - closest);
      if (sq < *squared) {
        *core = point;
        *surface = closest;
        *squared = sq;
      }
    }
  }
}
// spine_penetration (distance_box.rs:244-265): the deepest separating axis
// between a box and a capsule whose spine meets it (box frame).
template <class T>
T spine_penetration(V3<T> center, V3<T> axis, V3<T> half, T radius, T height, V3<T>* normal) {
  T best = -std::numeric_limits<T>::infinity();
  *normal = V3<T>{T(1), T(0), T(0)};
  const V3<T> E[3] = {{T(1), T(0), T(0)}, {T(0), T(1), T(0)}, {T(0), T(0), T(1)}};
  for (const V3<T>& e : E) {
    const V3<T> cand[2] = {e, normalize_or(cross(axis, e), V3<T>{T(0), T(0), T(0)})};
    for (const V3<T>& n : cand) {
      if (n == V3<T>{T(0), T(0), T(0)}) continue;
      const T projection = dot(center, n);
#ifdef ORC_PLANT_CAPSULE  // a planted mistake for the pin tests: the spine's extent dropped
      const T gap = std::fabs(projection) - dot(half, vabs(n)) - radius;
#else
      const T gap = std::fabs(projection) - dot(half, vabs(n)) - height * std::fabs(dot(axis, n)) - radius;
#endif
      if (gap > best) {
        best = gap;
        *normal = projection >= T(0) ? n : -n;
      }
    }
  }
  return best;
}
// box_capsule (distance_box.rs:200-240): box a, capsule b.
template <class T>
Dist<T> box_capsule(const SRef<T>& a, V3<T> half, const SRef<T>& b, T radius, T height) {
  const Q<T> inv = conj(a.rot);
  const V3<T> center = rotate(inv, b.pos - a.pos);
  const V3<T> axis = rotate(inv, rotate(b.rot, V3<T>{T(0), T(1), T(0)}));
  const V3<T> start = center - axis * height, end = center + axis * height;
  V3<T> core, surface, normal;
  T squared, distance;
  segment_box(start, end, -half, half, &core, &surface, &squared);
  if (squared > CF(1e-12)) {
    const T len = std::sqrt(squared);
    distance = len - radius;
    normal = (core - surface) / len;
  } else {
    distance = spine_penetration(center, axis, half, radius, height, &normal);
    V3<T> lo = -half, hi = half;
    for (int i = 0; i < 3; ++i)
      if (std::fabs(normal[i]) > CF(1e-6)) {
        lo[i] = signum(normal[i]) * half[i];
        hi[i] = lo[i];
      }
    const T projection = dot(axis, normal);
    V3<T> e0 = start, e1 = end;
    if (projection > CF(1e-6))
      e1 = start;
    else if (projection < -CF(1e-6))
      e0 = end;
    T sq;
    segment_box(e0, e1, lo, hi, &core, &surface, &sq);
  }
  return Dist<T>{distance, a.pos + rotate(a.rot, surface), a.pos + rotate(a.rot, core - normal * radius)};
}

// shape_distance (distance.rs:654-823) for box/sphere/half-space; *sat
// receives the box-box SAT gap (box_box_signed_gap, distance_box.rs:54-79).
template <class T>
Dist<T> shape_distance(const SRef<T>& a, const SRef<T>& b, T* sat) {
  *sat = T(0);
  if (a.shape == SH_HALFSPACE) return halfspace_distance(a, b, true);
  if (b.shape == SH_HALFSPACE) return halfspace_distance(b, a, false);
  if (a.shape == SH_SPHERE && b.shape == SH_SPHERE) return sphere_sphere(a, a.sp.x, b, b.sp.x);
  if (a.shape == SH_SPHERE && b.shape == SH_BOX) return sphere_obb(a, a.sp.x, b, b.sp);
  if (a.shape == SH_BOX && b.shape == SH_SPHERE) {
    Dist<T> d = sphere_obb(b, b.sp.x, a, a.sp);
    return Dist<T>{d.dist, d.pb, d.pa};
  }
  if (a.shape == SH_SPHERE && b.shape == SH_CAPSULE) return sphere_capsule(a, a.sp.x, b, b.sp.x, b.sp.y);
  if (a.shape == SH_CAPSULE && b.shape == SH_SPHERE) {
    Dist<T> d = sphere_capsule(b, b.sp.x, a, a.sp.x, a.sp.y);
    return Dist<T>{d.dist, d.pb, d.pa};
  }
  if (a.shape == SH_BOX && b.shape == SH_CAPSULE) return box_capsule(a, a.sp, b, b.sp.x, b.sp.y);
  if (a.shape == SH_CAPSULE && b.shape == SH_BOX) {
    Dist<T> d = box_capsule(b, b.sp, a, a.sp.x, a.sp.y);
    return Dist<T>{d.dist, d.pb, d.pa};
  }
  if (a.shape == SH_CAPSULE && b.shape == SH_CAPSULE)
    return capsule_capsule(a, a.sp.x, a.sp.y, b, b.sp.x, b.sp.y);
  return box_box(a, a.sp, b, b.sp, sat);
}

// ---------------------------------------------------------------------------
// Uniform grid for candidate search (acceleration only; exact tests follow).
// ---------------------------------------------------------------------------
struct Grid {
  double cell = 1.0;
  std::unordered_map<uint64_t, std::vector<int>> cells;
  static uint64_t key(int64_t x, int64_t y, int64_t z) {
    return (uint64_t(x + (1 << 20)) << 42) | (uint64_t(y + (1 << 20)) << 21) | uint64_t(z + (1 << 20));
  }
  int64_t ci(double v) const { return (int64_t)std::floor(v / cell); }
};

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
struct Stats {
  int pairs, points, colours, max_colour, sleeping, breakdowns, toi_hits;
};

template <class T>
struct Engine {
  V3<T> gravity;
  std::vector<Body<T>> bodies;
  std::vector<Pair<T>> pairs;
  std::vector<Joint<T>> joints;
  std::vector<V3<T>> pos0, inertial, prev_pos;
  std::vector<Q<T>> rot0, inertial_rot, prev_rot;
  std::vector<T> sleep_timer;
  std::vector<uint8_t> asleep;
  T time_debt = T(0);
  int iters = 10;
  int order = ORDER_REFERENCE;
  int math = MATH_LIBM;
  bool brute = false;
  int threads = 1;  // discovery threads (the reference parallelises discovery)
  std::vector<int> colour;  // per body, -1 = not swept
  int ncolours = 0;
  Stats stats{};
  // per-step incidence (pair indices in vector order; joint indices)
  std::vector<std::vector<int>> inc_pairs, inc_joints;
  // last step's TOI clamps: distance travelled to the hit (-1: none), target
  std::vector<T> toi_t;
  std::vector<int> toi_tgt;
  // events (mod.rs:424-464 prev_touch / prev_trigger / pre_vel, the drained
  // queues; emit_events mod.rs:1606-1669)
  std::vector<V3<T>> pre_vel;
  std::vector<Event<T>> contact_events, trigger_events, ccd_hits;
  std::vector<std::pair<int, int>> prev_touch, prev_trigger, trigger_now;

  // Frozen active set (test support for the differentiable GPU step): a
  // recording run keeps, per step, the pair vector (normals, gaps, anchors,
  // stuck flags and warm-start sources of its points, the damper's
  // imminence decision), the colouring and which bodies were awake and which
  // fell asleep; a replay run steps through exactly that active set with
  // its own (possibly perturbed) bodies, all continuous quantities live.
  struct FrozenPoint {
    V3<T> ra, rb;
    bool stuck;
    int src_pair, src_pt;
  };
  struct FrozenPair {
    int a, b;
    V3<T> n;
    T gap;
    bool damped;
    std::vector<FrozenPoint> pts;
  };
  struct FrozenStep {
    std::vector<FrozenPair> pairs;
    std::vector<int> colour;
    std::vector<uint8_t> awake, slept;
    int toi_hits = 0;
  };
  std::vector<FrozenStep> frozen;
  int freeze_mode = 0;  // 0 off, 1 record, 2 replay
  size_t replay_at = 0;
  std::vector<uint8_t> frozen_damped;  // per pair of the replayed step
  // The continuous branches of a recorded run (force clamps, the friction
  // cone, live rows, penalty updates and caps, the spin deadband), in
  // evaluation order; a replay takes every such branch from here, so it
  // evaluates the same piece of the step's piecewise-smooth map.
  std::vector<uint8_t> btape;
  size_t bpos = 0;
  bool br(bool v) {
    if (freeze_mode == 1)
      btape.push_back(v ? 1 : 0);
    else if (freeze_mode == 2)
      v = btape[bpos++] != 0;
    return v;
  }
  T clamp_b(T x, T lo, T hi) {
    if (br(x < lo)) x = lo;
    if (br(x > hi)) x = hi;
    return x;
  }
  // contact_force with its branches through br (identical results when no
  // active set is recorded or replayed).
  void contact_force_b(T cn, const T pens[3], const T lam[3], T ct0, T ct1, const T mu[2], T f[3]) {
    T a = pens[0] * cn + lam[0];
    T fn_c = br(a < T(0)) ? a : T(0);
    T r0 = pens[1] * ct0 + lam[1], r1 = pens[2] * ct1 + lam[2];
    T b1 = std::fabs(fn_c) * mu[0];
    T b2 = std::fabs(fn_c) * mu[1];
    if (br(b1 <= T(0) || b2 <= T(0))) {
      f[0] = fn_c;
      f[1] = br(b1 <= T(0)) ? T(0) : clamp_b(r0, -b1, b1);
      f[2] = br(b2 <= T(0)) ? T(0) : clamp_b(r1, -b2, b2);
      return;
    }
    T s = (r0 / b1) * (r0 / b1) + (r1 / b2) * (r1 / b2);
    T k = br(s > T(1)) ? T(1) / std::sqrt(s) : T(1);
    f[0] = fn_c;
    f[1] = r0 * k;
    f[2] = r1 * k;
  }

  bool solvable(int h) const { return bodies[h].type == BT_DYNAMIC && bodies[h].inv_mass > T(0); }
  T eff_inv_mass(int h) const { return bodies[h].type == BT_DYNAMIC ? bodies[h].inv_mass : T(0); }
  SRef<T> sref(int h) const { return SRef<T>{bodies[h].shape, bodies[h].sp, bodies[h].pos, bodies[h].rot}; }

  int add_body(const Body<T>& b) {
    bodies.push_back(b);
    pos0.push_back(V3<T>{T(0), T(0), T(0)});
    rot0.push_back(qident<T>());
    inertial.push_back(V3<T>{T(0), T(0), T(0)});
    inertial_rot.push_back(qident<T>());
    prev_pos.push_back(b.pos);
    prev_rot.push_back(b.rot);
    sleep_timer.push_back(T(0));
    asleep.push_back(0);
    colour.push_back(-1);
    return (int)bodies.size() - 1;
  }

  // remove_body (mod.rs:1241-1300): the last body moves into slot hi; pairs
  // and joints on hi go, the rest are renumbered in place (a pair keeps its
  // (a, b) orientation, so a renumbered pair whose first body now sorts
  // after its second finds no match at the next merge and restarts cold);
  // the per-body vectors swap-remove alongside, and every sleeping body
  // wakes.  (Trigger and contact events are not modelled.)
  void remove_body(int hi) {
    const int n = (int)bodies.size();
    if (hi < 0 || hi >= n) return;
    const int last = n - 1;
    // exiting trigger pairs report Exited (mod.rs:1246-1262); touch state
    // and undrained contact events are dropped, trigger state remapped
    // (mod.rs:1278-1286)
    for (const auto& ab : prev_trigger)
      if (ab.first == hi || ab.second == hi)
        trigger_events.push_back(Event<T>{EV_EXITED, ab.first, ab.second, {}, {}, T(0)});
    {
      auto mp = [&](int h) { return h == last ? hi : h; };
      std::vector<std::pair<int, int>> kept;
      for (const auto& ab : prev_trigger)
        if (ab.first != hi && ab.second != hi)
          kept.push_back({std::min(mp(ab.first), mp(ab.second)), std::max(mp(ab.first), mp(ab.second))});
      std::sort(kept.begin(), kept.end());
      kept.erase(std::unique(kept.begin(), kept.end()), kept.end());
      prev_trigger = std::move(kept);
      prev_touch.clear();
      contact_events.clear();
    }
    auto swap_rm = [&](auto& v) {
      if (hi < (int)v.size()) {
        v[hi] = v.back();
        v.pop_back();
      }
    };
    swap_rm(bodies);
    auto map = [&](int h) { return h == last ? hi : h; };
    std::vector<Pair<T>> kept;
    kept.reserve(pairs.size());
    for (auto& p : pairs) {
      if (p.a == hi || p.b == hi) continue;
      p.a = map(p.a);
      p.b = map(p.b);
      kept.push_back(std::move(p));
    }
    pairs = std::move(kept);
    std::vector<Joint<T>> kj;
    for (auto& j : joints) {
      if (j.a == hi || j.b == hi) continue;
      j.a = map(j.a);
      j.b = map(j.b);
      kj.push_back(j);
    }
    joints = std::move(kj);
    swap_rm(sleep_timer);
    swap_rm(asleep);
    swap_rm(prev_pos);
    swap_rm(prev_rot);
    // step scratch of this restatement, kept parallel to the bodies
    swap_rm(pos0);
    swap_rm(rot0);
    swap_rm(inertial);
    swap_rm(inertial_rot);
    swap_rm(colour);
    swap_rm(toi_t);
    swap_rm(toi_tgt);
    for (int h = 0; h < (int)bodies.size(); ++h) wake_body(h);
  }

  // add_joint for Ball, Revolute, Prismatic and Fixed without limits or
  // motors (mod.rs:1308-1510) with the frames of resolve_joint
  // (joint.rs:1251-1323): hinge / slide axes normalised (a degenerate axis
  // rejects), Fixed captures the assembly rotation and anchor offset.
  //
  // Distance, Rope and Spring (mod.rs:1479-1493; resolve_joint,
  // joint.rs:1325-1351): Distance captures the anchors' assembly distance,
  // Rope takes its maximum, Spring its motor (prm: max distance | target
  // position, stiffness, damping, max force, target velocity, motor kind,
  // motor model).  A rope maximum that is not finite and positive, or a
  // motor failing JointMotor::check (joint.rs:1101-1110), rejects before
  // the bodies wake (validate_joint runs first, mod.rs:1315).
  int add_joint(int kind, int a, int b, V3<T> la, V3<T> lb, V3<T> axa, V3<T> axb, const float* prm = nullptr) {
    if (kind < JK_BALL || kind > JK_MOTOR) return -1;
    if (kind == JK_GEAR) {  // mod.rs:1329-1365: a finite ratio over two existing Revolute / Prismatic joints
      if (!prm || !std::isfinite(prm[2])) return -1;
      for (int k = 0; k < 2; ++k) {
        const int r = (int)prm[k];
        if (!(prm[k] >= 0.0f) || r >= (int)joints.size()) return -1;
        if (joints[r].kind != JK_REVOLUTE && joints[r].kind != JK_PRISMATIC) return -1;
      }
    }
    if (kind == JK_SIXDOF) {  // validate_joint (migration.rs:188-191): limited windows finite, lo <= hi
      if (!prm) return -1;
      for (int k = 0; k < 6; ++k) {
        if (!(prm[k] == 0.0f || prm[k] == 1.0f || prm[k] == 2.0f)) return -1;
        if (prm[k] == 2.0f && !(std::isfinite(prm[6 + k]) && std::isfinite(prm[12 + k]) && prm[6 + k] <= prm[12 + k]))
          return -1;
      }
    }
    if (kind == JK_WHEEL) {  // validate_joint (migration.rs:288-315); prm = axle a 3, axle b 3, Hz, zeta,
                             // motor speed, budget, has motor
      if (!prm) return -1;
      for (int k = 0; k < 10; ++k)
        if (!std::isfinite(prm[k])) return -1;
      if (!(prm[6] >= 0.0f && prm[7] >= 0.0f)) return -1;
      if (prm[10] != 0.0f && !(prm[9] >= 0.0f)) return -1;
    }
    if (kind == JK_MOTOR) {  // validate_joint (migration.rs:335-350): finite targets, budgets >= 0
      if (!prm) return -1;
      for (int k = 0; k < 8; ++k)
        if (!std::isfinite(prm[k])) return -1;
      if (!(prm[6] >= 0.0f && prm[7] >= 0.0f)) return -1;
    }
    if (kind == JK_ROPE && !(prm && std::isfinite(prm[0]) && prm[0] > 0.0f)) return -1;
    if ((kind == JK_REVOLUTE || kind == JK_PRISMATIC) && prm) {  // validate_joint (migration.rs:160-190)
      if (prm[4] != 0.0f && !(std::isfinite(prm[0]) && std::isfinite(prm[1]) && prm[0] <= prm[1])) return -1;
      if (prm[5] != 0.0f && !(std::isfinite(prm[2]) && std::isfinite(prm[3]) && prm[3] >= 0.0f)) return -1;
    }
    if (kind == JK_SPRING) {
      if (!prm) return -1;
      for (int k = 0; k < 5; ++k)
        if (!std::isfinite(prm[k])) return -1;
      if (!(prm[1] >= 0.0f && prm[2] >= 0.0f && prm[3] >= 0.0f)) return -1;
      if (!(prm[5] >= 0.0f && prm[5] <= 2.0f && prm[6] >= 0.0f && prm[6] <= 1.0f)) return -1;
    }
    wake_body(a);
    wake_body(b);
    if (a == b || a < 0 || b < 0 || a >= (int)bodies.size() || b >= (int)bodies.size()) return -1;
    Joint<T> j;
    j.a = a;
    j.b = b;
    j.kind = kind;
    j.la = la;
    j.lb = lb;
    const V3<T> Z{T(0), T(0), T(1)};
    if (kind == JK_REVOLUTE || kind == JK_PRISMATIC) {
      if (dot(axa, axa) < K<T>::DEGENERATE() || dot(axb, axb) < K<T>::DEGENERATE()) return -1;
      j.ax_a = normalize_or(axa, Z);
      j.ax_b = normalize_or(axb, Z);
      // resolve_joint's assembly twist / slide (joint.rs:1280-1312)
      const Body<T>& A = bodies[a];
      const Body<T>& B = bodies[b];
      if (kind == JK_REVOLUTE) {
        j.ref_val = hinge_twist(A.rot, B.rot, j.ax_a);
      } else {
        const V3<T> wa = normalize_or(rotate(A.rot, j.ax_a), Z);
        j.ref_val = dot((B.pos + rotate(B.rot, lb)) - (A.pos + rotate(A.rot, la)), wa);
      }
      if (prm) {  // [lo, hi, target speed, budget, has limit, has motor]
        j.has_lim = prm[4] != 0.0f;
        j.has_mot = prm[5] != 0.0f;
        j.lim_lo = T(prm[0]);
        j.lim_hi = T(prm[1]);
        j.mot_speed = T(prm[2]);
        j.mot_max = T(prm[3]);
      }
    }
    if (kind == JK_FIXED) {
      const Body<T>& A = bodies[a];
      const Body<T>& B = bodies[b];
      j.q_ref = quat_mul(conj(A.rot), B.rot);
      V3<T> ra = rotate(A.rot, la), rb = rotate(B.rot, lb);
      j.dref = rotate(conj(A.rot), (B.pos + rb) - (A.pos + ra));
    }
    if (kind == JK_WHEEL) {  // resolve_joint (joint.rs:1352-1380)
      const Body<T>& A = bodies[a];
      const Body<T>& B = bodies[b];
      const V3<T> Y{T(0), T(1), T(0)};
      j.ax_a = normalize_or(axa, Y);
      j.ax_b = normalize_or(axb, Y);
      j.bx_a = orthogonalize_axle(j.ax_a, V3<T>{T(prm[0]), T(prm[1]), T(prm[2])});
      j.bx_b = orthogonalize_axle(j.ax_b, V3<T>{T(prm[3]), T(prm[4]), T(prm[5])});
      const V3<T> wa = normalize_or(rotate(A.rot, j.ax_a), Z);
      j.ref_val = dot((B.pos + rotate(B.rot, lb)) - (A.pos + rotate(A.rot, la)), wa);
      j.susp_hz = T(prm[6]);
      j.susp_zeta = T(prm[7]);
      j.has_mot = prm[10] != 0.0f;
      j.mot_speed = T(prm[8]);
      j.mot_max = T(prm[9]);
    }
    if (kind == JK_SIXDOF) {  // resolve_joint (joint.rs:1381-1391): the assembly relative pose
      const Body<T>& A = bodies[a];
      const Body<T>& B = bodies[b];
      j.q_ref = quat_mul(conj(A.rot), B.rot);
      j.dref = rotate(conj(A.rot), (B.pos + rotate(B.rot, lb)) - (A.pos + rotate(A.rot, la)));
      for (int k = 0; k < 6; ++k) {
        j.six[k] = (int)prm[k];
        j.six_lo[k] = T(prm[6 + k]);
        j.six_hi[k] = T(prm[12 + k]);
      }
    }
    if (kind == JK_MOTOR) {  // resolve_joint (joint.rs:1393-1399): the assembly relative pose
      const Body<T>& A = bodies[a];
      const Body<T>& B = bodies[b];
      j.la = j.lb = V3<T>{T(0), T(0), T(0)};
      j.q_ref = quat_mul(conj(A.rot), B.rot);
      j.dref = rotate(conj(A.rot), B.pos - A.pos);
      j.lin_target = V3<T>{T(prm[0]), T(prm[1]), T(prm[2])};
      j.ang_target = V3<T>{T(prm[3]), T(prm[4]), T(prm[5])};
      j.max_force = T(prm[6]);
      j.max_torque = T(prm[7]);
    }
    if (kind == JK_DISTANCE) {
      const Body<T>& A = bodies[a];
      const Body<T>& B = bodies[b];
      j.ref_val = length((B.pos + rotate(B.rot, lb)) - (A.pos + rotate(A.rot, la)));
    } else if (kind == JK_GEAR) {  // the assembly constant from the current coordinates (mod.rs:1354-1378)
      j.gb[0] = (int)prm[0];
      j.gb[1] = (int)prm[1];
      j.gratio = T(prm[2]);
      j.ref_val = joint_coordinate(joints[j.gb[0]]) + j.gratio * joint_coordinate(joints[j.gb[1]]);
    } else if (kind == JK_ROPE) {
      j.ref_val = T(prm[0]);
    } else if (kind == JK_SPRING) {
      j.ref_val = T(prm[0]);
      j.m_stiff = T(prm[1]);
      j.m_damp = T(prm[2]);
      j.m_max = T(prm[3]);
      j.m_tvel = T(prm[4]);
      j.m_kind = (int)prm[5];
      j.m_model = (int)prm[6];
    }
    for (int k = 0; k < 3; ++k) {
      j.lam_l[k] = T(0);
      j.pen_l[k] = K<T>::JOINT_PEN_INIT();
      j.lam_a[k] = T(0);
      j.pen_a[k] = K<T>::JOINT_PEN_INIT();
    }
    joints.push_back(j);
    return (int)joints.size() - 1;
  }
  int add_ball_joint(int a, int b, V3<T> la, V3<T> lb) {
    const V3<T> X{T(1), T(0), T(0)};
    return add_joint(JK_BALL, a, b, la, lb, X, X);
  }

  // sleep_body / wake_body (mod.rs:822-842; body.rs:403-418).
  void sleep_body(int h) {
    if (bodies[h].type != BT_DYNAMIC) return;
    Body<T>& b = bodies[h];
    b.vel = V3<T>{T(0), T(0), T(0)};
    b.angvel = V3<T>{T(0), T(0), T(0)};
    b.inv_mass = T(0);
    b.inertia = V3<T>{T(0), T(0), T(0)};
    asleep[h] = 1;
  }
  void wake_body(int h) {
    if (h < 0 || h >= (int)asleep.size() || !asleep[h]) return;
    asleep[h] = 0;
    sleep_timer[h] = T(0);
    Body<T>& b = bodies[h];
    if (b.type == BT_DYNAMIC) {
      b.inv_mass = T(1) / b.mass;
      b.inertia = shape_inertia(b.shape, b.sp, b.mass);
    }
  }

  static bool pair_allowed(const Body<T>& a, const Body<T>& b) {
    return (a.layer & b.mask) != 0 && (b.layer & a.mask) != 0;
  }
  // Every joint kind here but Fixed disables contact between its bodies.
  // The pairs are looked up in `jointed`, which generate_pairs fills from
  // the joint list before discovery (a scan of every joint per candidate
  // pair is quadratic on chain scenes).
  std::unordered_set<uint64_t> jointed;
  static uint64_t jkey(int ia, int ib) {
    const uint32_t lo = (uint32_t)std::min(ia, ib), hi = (uint32_t)std::max(ia, ib);
    return (uint64_t(lo) << 32) | hi;
  }
  void index_joints() {
    jointed.clear();
    for (const auto& j : joints)
      if (j.kind != JK_FIXED && j.kind != JK_SIXDOF && j.kind != JK_GEAR) jointed.insert(jkey(j.a, j.b));
  }
  bool ball_jointed(int ia, int ib) const { return jointed.count(jkey(ia, ib)) != 0; }

  // ---- joints helpers (sleep.rs) ----
  // quat_twist / hinge_twist (crates/physics/src/sequential_impulse/
  // joints.rs:38-48): the twist of a relative rotation about an axis,
  // wrapped to [-pi, pi]; wrap_pi (rows.rs:149-151).
  T wrap_pi(T x) const { return rem_euclid(x + K<T>::PI(), K<T>::TAU()) - K<T>::PI(); }
  T quat_twist(Q<T> q, V3<T> axis) const {
    return wrap_pi(T(2) * atan2_t(dot(V3<T>{q.x, q.y, q.z}, axis), q.w, math));
  }
  T hinge_twist(Q<T> qa, Q<T> qb, V3<T> axis) const { return quat_twist(quat_mul(conj(qa), qb), axis); }
  // joint_coordinate (joints_ext.rs:38-51): a Revolute's twist past its
  // assembly angle, a Prismatic's slide past its assembly length
  T joint_coordinate(const Joint<T>& r) const {
    const Body<T>& x = bodies[r.a];
    const Body<T>& y = bodies[r.b];
    if (r.kind == JK_REVOLUTE) return hinge_twist(x.rot, y.rot, r.ax_a) - r.ref_val;
    const V3<T> wa = normalize_or(rotate(x.rot, r.ax_a), V3<T>{T(0), T(0), T(1)});
    const V3<T> pa = x.pos + rotate(x.rot, r.la);
    const V3<T> pb = y.pos + rotate(y.rot, r.lb);
    return dot(pb - pa, wa) - r.ref_val;
  }
  // gear_sides (joints_ext.rs:120-161): side k of a gear, its coordinate
  // unwrapped against the step-start memory on a hinge side
  struct GearSide {
    int a, b;
    bool ang;
    V3<T> axis, ra, rb;
    T coord;
  };
  GearSide gear_side(const Joint<T>& g, int k) const {
    const Joint<T>& r = joints[g.gb[k]];
    const Body<T>& x = bodies[r.a];
    const Body<T>& y = bodies[r.b];
    GearSide s;
    s.a = r.a;
    s.b = r.b;
    s.ang = r.kind == JK_REVOLUTE;
    s.axis = normalize_or(rotate(x.rot, r.ax_a), V3<T>{T(0), T(0), T(1)});
    s.ra = s.ang ? V3<T>{T(0), T(0), T(0)} : rotate(x.rot, r.la);
    s.rb = s.ang ? V3<T>{T(0), T(0), T(0)} : rotate(y.rot, r.lb);
    const T raw = joint_coordinate(r);
    s.coord = (g.gmem && s.ang) ? g.gcont[k] + wrap_pi(raw - g.graw[k]) : raw;
    return s;
  }
  T gear_error(const Joint<T>& g) const {
    const GearSide sa = gear_side(g, 0), sb = gear_side(g, 1);
    return sa.coord + g.gratio * sb.coord - g.ref_val;
  }
  // update_gear_mem (joints_ext.rs:56-115), once per step
  void update_gear_mem() {
    for (auto& g : joints) {
      if (g.kind != JK_GEAR) continue;
      for (int k = 0; k < 2; ++k) {
        const Joint<T>& r = joints[g.gb[k]];
        const T raw = joint_coordinate(r);
        g.gcont[k] = (g.gmem && r.kind == JK_REVOLUTE) ? g.gcont[k] + wrap_pi(raw - g.graw[k]) : raw;
        g.graw[k] = raw;
      }
      g.gmem = true;
    }
  }
  // joint_members (sleep.rs:9-16): a gear's are its two joints' bodies
  void joint_members(const Joint<T>& j, int m[4]) const {
    if (j.kind == JK_GEAR) {
      m[0] = joints[j.gb[0]].a;
      m[1] = joints[j.gb[0]].b;
      m[2] = joints[j.gb[1]].a;
      m[3] = joints[j.gb[1]].b;
    } else {
      m[0] = m[2] = j.a;
      m[1] = m[3] = j.b;
    }
  }
  // limit_state / warm_limit_state / regularized_limit (rows.rs:156-191):
  // 1 = lower bound, 0 = upper bound, -1 = free inside the window.
  static int limit_state(T v, T lo, T hi, T slop) {
    if (v <= lo + slop) return 1;
    if (v >= hi - slop) return 0;
    return -1;
  }
  static int warm_limit_state(T v, T lo, T hi, T slop, T force) {
    const int st = limit_state(v, lo, hi, slop);
    if (st >= 0) return st;
    if (force < T(0)) return 1;
    if (force > T(0)) return 0;
    return -1;
  }
  static T regularized_limit(T value, T initial, T bound, bool lower) {
    const T old = initial - bound;
    const T violation = lower ? fmin_(old, T(0)) : fmax_(old, T(0));
    return value - bound - K<T>::ALPHA() * violation;
  }
  // The joint's live angle (Revolute) or slide (Prismatic) from the
  // assembly value, for its limit.
  T limit_value(const Joint<T>& j, Q<T> ra, Q<T> rb, V3<T> pa_w, V3<T> pb_w, V3<T> wa) const {
    if (j.kind == JK_REVOLUTE) return wrap_pi(hinge_twist(ra, rb, j.ax_a) - j.ref_val);
    return dot(pb_w - pa_w, wa) - j.ref_val;
  }
  // joint_limit_violated (joints_ext.rs:440-462), the motor pause rule.
  bool joint_limit_violated(const Joint<T>& j) const {
    if (!j.has_lim || (j.kind != JK_REVOLUTE && j.kind != JK_PRISMATIC)) return false;
    const Body<T>& a = bodies[j.a];
    const Body<T>& b = bodies[j.b];
    const V3<T> wa = normalize_or(rotate(a.rot, j.ax_a), V3<T>{T(0), T(0), T(1)});
    const V3<T> pa = a.pos + rotate(a.rot, j.la), pb = b.pos + rotate(b.rot, j.lb);
    const T v = limit_value(j, a.rot, b.rot, pa, pb, wa);
    return limit_state(v, j.lim_lo, j.lim_hi, j.kind == JK_REVOLUTE ? K<T>::LIMIT_SLOP_ANG() : K<T>::LIMIT_SLOP_LIN()) >= 0;
  }
  // free_motor_impulse (joints_ext.rs:368-421): per world axis, the
  // budget-clamped deadbeat linear then angular impulse through the pair's
  // effective mass at the centres (MotorDrive::linear_impulse /
  // angular_impulse, joint.rs:485-499).
  void free_motor_impulse(const Joint<T>& j) {
    const int ia = j.a, ib = j.b;
    auto clamp_imp = [](T err, T k, T budget) -> T {
      if (budget <= T(0) || k <= T(0)) return T(0);
      return clampf(err / k, -budget * K<T>::DT(), budget * K<T>::DT());
    };
    for (int e = 0; e < 3; ++e) {
      V3<T> axis{T(0), T(0), T(0)};
      axis[e] = T(1);
      const T v = dot(bodies[ib].vel - bodies[ia].vel, axis);
      const T ka_lin = eff_inv_mass(ia), kb_lin = eff_inv_mass(ib);
      const T dj = clamp_imp(dot(j.lin_target, axis) - v, ka_lin + kb_lin, j.max_force);
      if (dj != T(0)) {
        if (solvable(ia)) bodies[ia].vel = bodies[ia].vel + ((-dj) * ka_lin) * axis;
        if (solvable(ib)) bodies[ib].vel = bodies[ib].vel + (dj * kb_lin) * axis;
      }
      const T w = dot(bodies[ib].angvel - bodies[ia].angvel, axis);
      const V3<T> wia = ang_inv_wa(ia, axis), wib = ang_inv_wa(ib, axis);
      const T ka = dot(wia, axis) + dot(wib, axis);
      const T dw = clamp_imp(dot(j.ang_target, axis) - w, ka, j.max_torque);
      if (dw != T(0)) {
        if (solvable(ia)) bodies[ia].angvel = bodies[ia].angvel + (-dw) * wia;
        if (solvable(ib)) bodies[ib].angvel = bodies[ib].angvel + dw * wib;
      }
    }
  }

  // motor_impulse (joints_ext.rs:237-356) for Revolute / Prismatic velocity
  // motors: the budget-clamped deadbeat impulse driving the axis rate to
  // the target (JointMotor::velocity_impulse, joint.rs:1172-1177), paused
  // while the joint's limit is violated; joint order.
  void motor_impulse() {
    const V3<T> Z{T(0), T(0), T(1)};
    for (const auto& j : joints) {
      if (j.kind == JK_MOTOR) {
        free_motor_impulse(j);
        continue;
      }
      if (!j.has_mot || (j.kind != JK_REVOLUTE && j.kind != JK_PRISMATIC && j.kind != JK_WHEEL)) continue;
      if (j.mot_max <= T(0) || joint_limit_violated(j)) continue;
      const int ia = j.a, ib = j.b;
      // a wheel's motor turns about its axle (joints_ext.rs:337-369)
      const V3<T> wa = normalize_or(rotate(bodies[ia].rot, j.kind == JK_WHEEL ? j.bx_a : j.ax_a), Z);
      auto impulse = [&](T err, T k) -> T {
        if (k <= T(0)) return T(0);
        return clampf(err / k, -j.mot_max * K<T>::DT(), j.mot_max * K<T>::DT());
      };
      if (j.kind == JK_REVOLUTE || j.kind == JK_WHEEL) {
        const T w = dot(bodies[ib].angvel - bodies[ia].angvel, wa);
        const V3<T> wia = ang_inv_wa(ia, wa), wib = ang_inv_wa(ib, wa);
        const T k = dot(wia, wa) + dot(wib, wa);
        if (k < K<T>::MIN_EFF_MASS()) continue;
        const T dj = impulse(j.mot_speed - w, k);
        if (solvable(ia)) bodies[ia].angvel = bodies[ia].angvel + (-dj) * wia;
        if (solvable(ib)) bodies[ib].angvel = bodies[ib].angvel + dj * wib;
      } else {
        const T v = dot(bodies[ib].vel - bodies[ia].vel, wa);
        const T ka = eff_inv_mass(ia), kb = eff_inv_mass(ib);
        const T k = ka + kb;
        if (k < K<T>::MIN_EFF_MASS()) continue;
        const T dj = impulse(j.mot_speed - v, k);
        if (solvable(ia)) bodies[ia].vel = bodies[ia].vel + ((-dj) * ka) * wa;
        if (solvable(ib)) bodies[ib].vel = bodies[ib].vel + (dj * kb) * wa;
      }
    }
  }

  T joint_rest_error(const Joint<T>& j) const {
    if (j.kind == JK_GEAR) return std::fabs(gear_error(j));
    const Body<T>& a = bodies[j.a];
    const Body<T>& b = bodies[j.b];
    V3<T> delta = (b.pos + rotate(b.rot, j.lb)) - (a.pos + rotate(a.rot, j.la));
    V3<T> axis = rotate(a.rot, j.ax_a);
    switch (j.kind) {
      case JK_REVOLUTE:
        return fmax_(length(delta), length(axis - rotate(b.rot, j.ax_b)));
      case JK_FIXED:
        return fmax_(length(delta - rotate(a.rot, j.dref)),
                     length(quat_diff_vec(quat_mul(conj(a.rot), b.rot), j.q_ref, math)));
      case JK_PRISMATIC: {
        V3<T> perp = delta - axis * dot(delta, axis);
        T error = length(perp);
        if (j.has_lim) {
          const T value = dot(delta, axis) - j.ref_val;
          error = fmax_(error, fmax_(j.lim_lo - value, value - j.lim_hi));
        }
        return fmax_(error, length(axis - rotate(b.rot, j.ax_b)));
      }
      case JK_DISTANCE:
        return std::fabs(length(delta) - j.ref_val);
      case JK_ROPE:  // slack is rest
        return fmax_(length(delta) - j.ref_val, T(0));
      case JK_SPRING:
        return std::fabs(length(delta) - j.ref_val);
      case JK_WHEEL: {
        V3<T> perp = delta - axis * dot(delta, axis);
        T error = length(perp);
        error = fmax_(error, length(rotate(a.rot, j.bx_a) - rotate(b.rot, j.bx_b)));
        if (j.susp_hz <= T(0)) error = fmax_(error, std::fabs(dot(delta, axis) - j.ref_val));
        return error;
      }
      case JK_SIXDOF: {  // sixdof_rest_error (sleep.rs:86-103)
        const V3<T> lin = rotate(conj(a.rot), delta) - j.dref;
        const V3<T> ang = quat_diff_vec(quat_mul(conj(a.rot), b.rot), j.q_ref, math);
        T error = T(0);
        for (int i = 0; i < 3; ++i) {
          const T vals[2] = {lin[i], ang[i]};
          for (int s2 = 0; s2 < 2; ++s2) {
            const int k = 3 * s2 + i;
            const T v = vals[s2];
            T r = T(0);
            if (j.six[k] == AX_LOCKED) r = std::fabs(v);
            if (j.six[k] == AX_LIMITED) r = fmax_(fmax_(j.six_lo[k] - v, v - j.six_hi[k]), T(0));
            error = fmax_(error, r);
          }
        }
        return error;
      }
      case JK_MOTOR:
        return fmax_(length(delta - rotate(a.rot, j.dref)),
                     length(quat_diff_vec(quat_mul(conj(a.rot), b.rot), j.q_ref, math)));
      default:
        return length(delta);
    }
  }
  // JointMotor::keeps_awake (joint.rs:1116-1126) of a Spring's motor.
  static bool spring_keeps_awake(const Joint<T>& j) {
    if (j.kind != JK_SPRING || !(j.m_max > T(0))) return false;
    if (j.m_kind == MK_VELOCITY) return j.m_tvel != T(0);
    return j.m_stiff > T(0) || j.m_tvel != T(0);
  }
  bool joint_driven(const Joint<T>& j) const {
    if (spring_keeps_awake(j)) return true;
    if (j.has_mot && j.mot_speed != T(0) && j.mot_max > T(0)) return true;
    if (j.kind == JK_MOTOR && ((j.max_force > T(0) && !(j.lin_target == V3<T>{T(0), T(0), T(0)})) ||
                               (j.max_torque > T(0) && !(j.ang_target == V3<T>{T(0), T(0), T(0)}))))
      return true;  // MotorDrive::keeps_awake (joint.rs:474-477)
    int m[4];
    joint_members(j, m);
    for (int h : m) {
      const Body<T>& b = bodies[h];
      if (b.type != BT_DYNAMIC && (b.pos != prev_pos[h] || b.rot != prev_rot[h])) return true;
    }
    return false;
  }
  void propagate_joint_flags(std::vector<uint8_t>& flags, uint8_t value) const {
    for (;;) {
      bool changed = false;
      for (const auto& j : joints) {
        int m[4];
        joint_members(j, m);
        bool any = false;
        for (int h : m)
          if (bodies[h].type == BT_DYNAMIC && flags[h] == value) any = true;
        if (any)
          for (int h : m)
            if (bodies[h].type == BT_DYNAMIC && flags[h] != value) {
              flags[h] = value;
              changed = true;
            }
      }
      if (!changed) break;
    }
  }
  void wake_joint_motion() {
    size_t n = bodies.size();
    std::vector<uint8_t> awake(n);
    const V3<T> Z{T(0), T(0), T(0)};
    for (size_t h = 0; h < n; ++h) {
      const Body<T>& b = bodies[h];
      awake[h] = b.type == BT_DYNAMIC && (!asleep[h] || b.vel != Z || b.angvel != Z || b.torque != Z ||
                                          b.pos != prev_pos[h] || b.rot != prev_rot[h]);
    }
    if (joints.empty()) {
      // propagate is a no-op without joints
    }
    for (const auto& j : joints) {
      if (joint_driven(j) || joint_rest_error(j) > K<T>::JOINT_REST()) {
        int m[4];
        joint_members(j, m);
        for (int h : m)
          if (bodies[h].type == BT_DYNAMIC) awake[h] = 1;
      }
    }
    propagate_joint_flags(awake, 1);
    for (size_t h = 0; h < n; ++h)
      if (awake[h] && asleep[h]) wake_body((int)h);
  }
  std::vector<uint8_t> joint_sleep_ready() const {
    std::vector<uint8_t> ready(bodies.size(), 1);
    for (const auto& j : joints) {
      if (joint_driven(j) || joint_rest_error(j) > K<T>::JOINT_REST()) {
        ready[j.a] = 0;
        ready[j.b] = 0;
      }
    }
    propagate_joint_flags(ready, 0);
    return ready;
  }

  // ---- TOI clamp: cast_shape (distance.rs:841-891) ----
  struct Hit {
    T t;
    V3<T> point, normal;
    int handle;
    bool valid;
  };
  static Hit cast_one(const SRef<T>& mover, V3<T> dir, T len, int handle, const SRef<T>& target) {
    Hit hit{T(0), {}, {}, handle, false};
    T t = T(0);
    for (int it = 0; it < 24; ++it) {
      SRef<T> m = mover;
      m.pos = mover.pos + dir * t;
      T sat;
      Dist<T> d = shape_distance(m, target, &sat);
      if (d.dist <= K<T>::SHAPE_TOUCH()) {
        if (t > T(0)) {
          V3<T> nn = normalize_or(d.pa - d.pb, -dir);
          hit = Hit{t, d.pb, nn, handle, true};
        }
        break;
      }
      t += d.dist - K<T>::SHAPE_TOUCH() * K<T>::HALF();
      if (t >= len) break;
    }
    return hit;
  }
  // cast against candidate targets (ascending index); first minimum wins.
  Hit cast_shape(int h, V3<T> delta, const std::vector<int>& targets) const {
    Hit best{T(0), {}, {}, -1, false};
    T len = length(delta);
    if (len < K<T>::NEAR_ZERO()) return best;
    V3<T> dir = delta / len;
    SRef<T> mover = sref(h);
    for (int o : targets) {
      Hit hh = cast_one(mover, dir, len, o, sref(o));
      if (hh.valid && (!best.valid || hh.t < best.t)) best = hh;
    }
    return best;
  }
  bool toi_target_ok(int h, int o, const std::vector<int>& joined) const {
    if (o == h || bodies[o].trigger || !pair_allowed(bodies[h], bodies[o])) return false;
    for (int jn : joined)
      if (jn == o) return false;
    return true;
  }
  std::vector<int> joined_of(int h) const {
    std::vector<int> out;
    for (const auto& j : joints) {
      if (j.kind == JK_GEAR) continue;  // mod.rs:903-905
      if (j.a == h)
        out.push_back(j.b);
      else if (j.b == h)
        out.push_back(j.a);
    }
    return out;
  }
  bool toi_mover(int h) const {
    if (!solvable(h) || asleep[h] || bodies[h].trigger) return false;
    V3<T> disp = bodies[h].vel * K<T>::DT();
    return !(length(disp) <= K<T>::CCD_GATE() * shape_min_dim(bodies[h].shape, bodies[h].sp));
  }
  // Candidate TOI targets of mover h from a grid of the current poses.
  void toi_candidates(int h, const Grid* g, const std::vector<int>& large, std::vector<int>& out) const {
    out.clear();
    size_t n = bodies.size();
    if (!g) {
      for (size_t o = 0; o < n; ++o) out.push_back((int)o);
      return;
    }
    const Body<T>& b = bodies[h];
    V3<T> disp = b.vel * K<T>::DT();
    double r = (double)bound_radius(b.shape, b.sp) + g->cell * 0.5 + 0.01;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      double p0 = (double)b.pos[k], p1 = (double)(b.pos[k] + disp[k]);
      lo[k] = std::min(p0, p1) - r;
      hi[k] = std::max(p0, p1) + r;
    }
    for (int64_t x = g->ci(lo[0]); x <= g->ci(hi[0]); ++x)
      for (int64_t y = g->ci(lo[1]); y <= g->ci(hi[1]); ++y)
        for (int64_t z = g->ci(lo[2]); z <= g->ci(hi[2]); ++z) {
          auto it = g->cells.find(Grid::key(x, y, z));
          if (it == g->cells.end()) continue;
          for (int o : it->second) out.push_back(o);
        }
    for (int o : large) out.push_back(o);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }
  // Bodies split into grid members (finite bound radius below a cutoff)
  // and "large" ones tested against everything.
  void build_grid(Grid& g, std::vector<int>& large, double extra) const {
    size_t n = bodies.size();
    std::vector<double> rs;
    for (size_t h = 0; h < n; ++h) {
      double r = (double)bound_radius(bodies[h].shape, bodies[h].sp);
      if (std::isfinite(r)) rs.push_back(r);
    }
    double cut = 1.0;
    if (!rs.empty()) {
      std::vector<double> s = rs;
      std::nth_element(s.begin(), s.begin() + s.size() / 2, s.end());
      cut = std::max(4.0 * s[s.size() / 2], 1e-3);
    }
    double rmax = 0.0;
    for (double r : rs)
      if (r <= cut) rmax = std::max(rmax, r);
    g.cell = (2.0 * rmax + 0.05 + extra) * 1.01 + 1e-6;
    g.cells.clear();
    large.clear();
    for (size_t h = 0; h < n; ++h) {
      double r = (double)bound_radius(bodies[h].shape, bodies[h].sp);
      if (!std::isfinite(r) || r > cut) {
        large.push_back((int)h);
        continue;
      }
      const Body<T>& b = bodies[h];
      g.cells[Grid::key(g.ci(b.pos.x), g.ci(b.pos.y), g.ci(b.pos.z))].push_back((int)h);
    }
  }
  void grid_move(Grid& g, int h, V3<T> oldp) const {
    uint64_t k0 = Grid::key(g.ci(oldp.x), g.ci(oldp.y), g.ci(oldp.z));
    const Body<T>& b = bodies[h];
    uint64_t k1 = Grid::key(g.ci(b.pos.x), g.ci(b.pos.y), g.ci(b.pos.z));
    if (k0 == k1) return;
    auto it = g.cells.find(k0);
    if (it != g.cells.end()) {
      auto& v = it->second;
      v.erase(std::remove(v.begin(), v.end(), h), v.end());
    }
    g.cells[k1].push_back(h);
  }

  void apply_toi_hit(int h, const Hit& hit) {
    {  // the clamp's impact event (mod.rs:947-960)
      const T approach = -dot(pre_vel[h] - pre_vel[hit.handle], hit.normal);
      if (approach > T(1)) {
        const bool lo = h < hit.handle;
        ccd_hits.push_back(Event<T>{EV_HIT, lo ? h : hit.handle, lo ? hit.handle : h, hit.point,
                                    lo ? -hit.normal : hit.normal, approach});
      }
    }
    Body<T>& b = bodies[h];
    V3<T> disp = b.vel * K<T>::DT();
    T len = fmax_(length(disp), K<T>::MIN_TOI_LEN());
    T frac = clampf(hit.t / len, T(0), T(1));
    T e = fmin_(bodies[h].restitution, bodies[hit.handle].restitution);
    b.pos = b.pos + disp * frac;
    T vn = dot(b.vel, hit.normal);
    if (vn < T(0)) {
      T bounce = vn < -K<T>::CCD_BOUNCE() ? T(1) + e : T(1);
      b.vel = b.vel - hit.normal * (bounce * vn);
    }
    wake_body(hit.handle);
    sleep_timer[h] = T(0);
    stats.toi_hits++;
    toi_t[h] = hit.t;
    toi_tgt[h] = hit.handle;
  }

  // mod.rs:882-978.
  // The inverse world inertia applied to w (ang_inv_wa, joints_ext.rs:423-438).
  V3<T> ang_inv_wa(int h, V3<T> w) const {
    if (!solvable(h)) return V3<T>{T(0), T(0), T(0)};
    const Body<T>& b = bodies[h];
    T iw[3][3], iwi[3][3];
    world_inertia(b.inertia, b.rot, iw);
    inverse_symmetric(iw, V3<T>{fmax_(b.inertia.x, T(0)), fmax_(b.inertia.y, T(0)), fmax_(b.inertia.z, T(0))}, iwi);
    return mat3_vec(iwi, w);
  }
  // rope_impulse (joints_ext.rs:173-223): once per step, in joint order, a
  // stretched rope whose ends separate has the separating anchor velocity
  // removed through the pair's effective mass.
  void rope_impulse() {
    for (const auto& j : joints) {
      if (j.kind != JK_ROPE) continue;
      const int ia = j.a, ib = j.b;
      const V3<T> pa = bodies[ia].pos + rotate(bodies[ia].rot, j.la);
      const V3<T> pb = bodies[ib].pos + rotate(bodies[ib].rot, j.lb);
      const V3<T> delta = pb - pa;
      const T len = length(delta);
      if (len <= j.ref_val) continue;
      const V3<T> nrm = delta / len;
      const V3<T> ra = pa - bodies[ia].pos, rb = pb - bodies[ib].pos;
      const T ka = eff_inv_mass(ia), kb = eff_inv_mass(ib);
      const V3<T> ta = cross(ra, nrm), tb = cross(rb, nrm);
      const T k = ka + kb + dot(ang_inv_wa(ia, ta), ta) + dot(ang_inv_wa(ib, tb), tb);
      if (k < K<T>::MIN_EFF_MASS()) continue;
      const V3<T> va = bodies[ia].vel + cross(bodies[ia].angvel, ra);
      const V3<T> vb = bodies[ib].vel + cross(bodies[ib].angvel, rb);
      const T vsep = dot(vb - va, nrm);
      if (vsep <= T(0)) continue;
      const T dj = -vsep / k;
      if (solvable(ia)) {
        const V3<T> wa = ang_inv_wa(ia, ta);
        bodies[ia].vel = bodies[ia].vel - nrm * (dj * ka);
        bodies[ia].angvel = bodies[ia].angvel - wa * dj;
      }
      if (solvable(ib)) {
        const V3<T> wb = ang_inv_wa(ib, tb);
        bodies[ib].vel = bodies[ib].vel + nrm * (dj * kb);
        bodies[ib].angvel = bodies[ib].angvel + wb * dj;
      }
    }
  }

  void toi_clamp() {
    int n = (int)bodies.size();
    toi_t.assign(n, T(-1));
    toi_tgt.assign(n, -1);
    bool use_grid = !brute && n > 256;
    Grid g;
    std::vector<int> large, cand;
    if (use_grid) build_grid(g, large, 0.0);
    if (order == ORDER_REFERENCE) {
      for (int h = 0; h < n; ++h) {
        if (!toi_mover(h)) continue;
        std::vector<int> joined = joined_of(h);
        toi_candidates(h, use_grid ? &g : nullptr, large, cand);
        std::vector<int> tg;
        for (int o : cand)
          if (toi_target_ok(h, o, joined)) tg.push_back(o);
        Hit hit = cast_shape(h, bodies[h].vel * K<T>::DT(), tg);
        if (hit.valid) {
          V3<T> oldp = bodies[h].pos;
          apply_toi_hit(h, hit);
          if (use_grid && std::isfinite((double)bound_radius(bodies[h].shape, bodies[h].sp)) &&
              std::find(large.begin(), large.end(), h) == large.end())
            grid_move(g, h, oldp);
        }
      }
    } else {
      std::vector<Hit> hits(n, Hit{T(0), {}, {}, -1, false});
      for (int h = 0; h < n; ++h) {
        if (!toi_mover(h)) continue;
        std::vector<int> joined = joined_of(h);
        toi_candidates(h, use_grid ? &g : nullptr, large, cand);
        std::vector<int> tg;
        for (int o : cand)
          if (toi_target_ok(h, o, joined)) tg.push_back(o);
        hits[h] = cast_shape(h, bodies[h].vel * K<T>::DT(), tg);
      }
      for (int h = 0; h < n; ++h)
        if (hits[h].valid) apply_toi_hit(h, hits[h]);
    }
  }

  // ---- discovery (rows.rs:432-756) ----
  struct Disc {
    int ia, ib;
    bool trigger, exists, wake;
    V3<T> normal;
    T mu[2];
    T signed_;
    std::vector<std::pair<V3<T>, V3<T>>> fresh;
  };

  const Pair<T>* find_pair(int ia, int ib, const std::unordered_map<uint64_t, int>* index) const {
    if (index) {
      auto it = index->find((uint64_t(ia) << 32) | uint32_t(ib));
      return it == index->end() ? nullptr : &pairs[it->second];
    }
    for (const auto& p : pairs)
      if (p.a == ia && p.b == ib) return &p;
    return nullptr;
  }

  bool discover_pair(int ia, int ib, const std::unordered_map<uint64_t, int>* index, Disc& out) const {
    const Body<T>& a = bodies[ia];
    const Body<T>& b = bodies[ib];
    if (!pair_allowed(a, b)) return false;
    if (!a.trigger && !b.trigger && !joints.empty() && ball_jointed(ia, ib)) return false;
    T ra = bound_radius(a.shape, a.sp);
    T rb = bound_radius(b.shape, b.sp);
    if (length(a.pos - b.pos) > ra + rb + K<T>::GEN_MARGIN()) return false;
    T sat;
    SRef<T> sa = sref(ia), sb = sref(ib);
    Dist<T> d = shape_distance(sa, sb, &sat);
    T sgn = (a.shape == SH_BOX && b.shape == SH_BOX) ? sat : d.dist;
    out.ia = ia;
    out.ib = ib;
    out.fresh.clear();
    out.signed_ = sgn;
    if (a.trigger || b.trigger) {
      out.trigger = true;
      out.exists = false;
      out.wake = false;
      out.normal = V3<T>{T(0), T(1), T(0)};
      out.mu[0] = out.mu[1] = T(0);
      return true;
    }
    out.trigger = false;
    const Pair<T>* old = find_pair(ia, ib, index);
    // orient_contact_normal (rows.rs:303-333)
    V3<T> normal;
    if (a.shape == SH_HALFSPACE) {
      normal = -rotate(a.rot, a.sp);
    } else if (b.shape == SH_HALFSPACE) {
      normal = rotate(b.rot, b.sp);
    } else {
      normal = d.pa - d.pb;
      if (d.dist < T(0)) normal = -normal;
      if (length_sq(normal) < K<T>::AXIS_REST_LEN2()) normal = a.pos - b.pos;
      normal = normalize_or(normal, V3<T>{T(0), T(1), T(0)});
      if (old) {
        if (dot(normal, old->n) < T(0)) normal = -normal;
      } else if (dot(normal, a.pos - b.pos) < T(0)) {
        normal = -normal;
      }
    }
    bool exists = old != nullptr;
    T approach = fmax_(dot(b.vel - a.vel, normal), T(0));
    T reach = K<T>::GEN_MARGIN() + approach * K<T>::DT();
    if (sgn > reach && !exists) return false;
    bool separated = sgn > reach;
    // friction_frame's coefficients (rows.rs:32-52, :567-570)
    T mu1, mu2;
    friction_mu(a, b, &mu1, &mu2);
    out.mu[0] = fmax_(mu1, T(0));
    out.mu[1] = fmax_(mu2, T(0));
    T patch = fmin_(ra, rb) + K<T>::MARGIN();
    V3<T> pp = (d.pa + d.pb) * K<T>::HALF();
    V3<T> pa_c = a.pos - normal * dot(a.pos - pp, normal);
    V3<T> pb_c = b.pos - normal * dot(b.pos - pp, normal);
    auto plane = [&](const Body<T>& body, V3<T> witness, T sign) -> T {
      if (body.shape == SH_BOX) {
        T extent = dot(vabs(rotate(conj(body.rot), normal)), body.sp);
        return dot(normal, body.pos) - sign * extent;
      }
      return dot(normal, witness);
    };
    T plane_a = plane(a, d.pa, T(1));
    T plane_b = plane(b, d.pb, T(-1));
    auto& fresh = out.fresh;
    Q<T> inv_a = conj(a.rot), inv_b = conj(b.rot);
    auto dup = [&](V3<T> ra_l, V3<T> rb_l) {
      for (auto& e : fresh)
        if (length(e.first - ra_l) < K<T>::POINT_MATCH() && length(e.second - rb_l) < K<T>::POINT_MATCH())
          return true;
      return false;
    };
    if (!separated) {
      for (int side = 0; side < 2; ++side) {
        int h = side == 0 ? ia : ib;
        V3<T> witness = side == 0 ? d.pa : d.pb;
        T sign = side == 0 ? T(1) : T(-1);
        V3<T> pc = side == 0 ? pa_c : pb_c;
        const Body<T>& body = bodies[h];
        if (body.shape != SH_BOX) continue;
        V3<T> hh = body.sp;
        for (int cx = 0; cx < 2; ++cx)
          for (int cy = 0; cy < 2; ++cy)
            for (int cz = 0; cz < 2; ++cz) {
              if ((int)fresh.size() >= MAX_POINTS) goto next_side;
              {
                T sx = cx ? T(1) : T(-1), sy = cy ? T(1) : T(-1), sz = cz ? T(1) : T(-1);
                V3<T> corner{sx * hh.x, sy * hh.y, sz * hh.z};
                V3<T> world = body.pos + rotate(body.rot, corner);
                T along = dot(world - witness, sign * normal);
                if (std::fabs(along) > K<T>::EXPAND_SLOP() + fmax_(-sgn, T(0))) continue;
                V3<T> rel_c = world - pc;
                V3<T> tang = rel_c - normal * dot(rel_c, normal);
                if (length(tang) > patch) continue;
                V3<T> pa, pb;
                if (side == 0) {
                  pa = world;
                  pb = world + normal * (plane_b - dot(normal, world));
                } else {
                  pa = world + normal * (plane_a - dot(normal, world));
                  pb = world;
                }
                V3<T> ra_l = rotate(inv_a, pa - a.pos);
                V3<T> rb_l = rotate(inv_b, pb - b.pos);
                if (dup(ra_l, rb_l)) continue;
                fresh.push_back({ra_l, rb_l});
              }
            }
      next_side:;
      }
    }
    if (!separated && fresh.size() >= 2 && (int)fresh.size() < MAX_POINTS) {
      V3<T> mid = ra < rb ? pa_c : (rb < ra ? pb_c : (pa_c + pb_c) * K<T>::HALF());
      V3<T> pa = mid + normal * (plane_a - dot(normal, mid));
      V3<T> pb = mid + normal * (plane_b - dot(normal, mid));
      V3<T> ra_l = rotate(inv_a, pa - a.pos);
      V3<T> rb_l = rotate(inv_b, pb - b.pos);
      if (!dup(ra_l, rb_l)) fresh.push_back({ra_l, rb_l});
    }
    if (!separated && fresh.size() < 3) {
      V3<T> ra_l = rotate(inv_a, d.pa - a.pos);
      V3<T> rb_l = rotate(inv_b, d.pb - b.pos);
      if (!dup(ra_l, rb_l)) fresh.push_back({ra_l, rb_l});
    }
    bool wake;
    if (exists) {
      wake = (asleep[ia] || asleep[ib]) && sgn < -K<T>::WAKE_PEN();
    } else {
      bool impact = sgn < -K<T>::WAKE_PEN();
      if (!impact) {
        int ss[2] = {ia, ib}, oo[2] = {ib, ia};
        T sg[2] = {T(1), T(-1)};
        for (int k = 0; k < 2; ++k) {
          if (asleep[ss[k]]) {
            T app = dot(bodies[oo[k]].vel - bodies[ss[k]].vel, sg[k] * normal);
            if (app > K<T>::WAKE_IMPACT()) {
              impact = true;
              break;
            }
          }
        }
      }
      wake = impact;
    }
    out.exists = exists;
    out.normal = normal;
    out.wake = wake;
    return true;
  }

  // candidate (ia < ib) pairs: every pair (brute) or grid neighbours + large.
  void candidate_pairs(std::vector<std::pair<int, int>>& out) const {
    int n = (int)bodies.size();
    out.clear();
    if (brute || n <= 256) {
      for (int ia = 0; ia < n; ++ia)
        for (int ib = ia + 1; ib < n; ++ib) out.push_back({ia, ib});
      return;
    }
    Grid g;
    std::vector<int> large;
    build_grid(g, large, 0.0);
    std::vector<uint8_t> is_large(n, 0);
    for (int h : large) is_large[h] = 1;
    for (int ia = 0; ia < n; ++ia) {
      if (is_large[ia]) {
        for (int ib = ia + 1; ib < n; ++ib) out.push_back({ia, ib});
        continue;
      }
      const Body<T>& b = bodies[ia];
      int64_t cx = g.ci(b.pos.x), cy = g.ci(b.pos.y), cz = g.ci(b.pos.z);
      std::vector<int> nb;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            auto it = g.cells.find(Grid::key(cx + dx, cy + dy, cz + dz));
            if (it == g.cells.end()) continue;
            for (int o : it->second)
              if (o > ia) nb.push_back(o);
          }
      for (int o : large)
        if (o > ia) nb.push_back(o);
      std::sort(nb.begin(), nb.end());
      for (int o : nb) out.push_back({ia, o});
    }
  }

  // generate_pairs (rows.rs:767-943).
  void generate_pairs() {
    index_joints();
    std::unordered_map<uint64_t, int> index;
    bool use_index = !brute;
    if (use_index)
      for (size_t i = 0; i < pairs.size(); ++i) index[(uint64_t(pairs[i].a) << 32) | uint32_t(pairs[i].b)] = (int)i;
    std::vector<std::pair<int, int>> cand;
    candidate_pairs(cand);
    std::vector<Disc> disc;
    // Discovery runs on `threads` threads above PAR_DISCOVERY_BODIES bodies
    // (rows.rs:427, :769-782); chunks are concatenated in candidate order, so
    // the result does not depend on the thread count.
    const int nthr = (threads > 1 && bodies.size() > 256) ? threads : 1;
    if (nthr == 1) {
      Disc d;
      for (auto& c : cand)
        if (discover_pair(c.first, c.second, use_index ? &index : nullptr, d)) disc.push_back(d);
    } else {
      std::vector<std::vector<Disc>> part(nthr);
      std::vector<std::thread> pool;
      const size_t chunk = (cand.size() + nthr - 1) / nthr;
      for (int t = 0; t < nthr; ++t)
        pool.emplace_back([&, t]() {
          Disc d;
          const size_t lo = std::min(cand.size(), t * chunk), hi = std::min(cand.size(), lo + chunk);
          for (size_t i = lo; i < hi; ++i)
            if (discover_pair(cand[i].first, cand[i].second, use_index ? &index : nullptr, d)) part[t].push_back(d);
        });
      for (auto& th : pool) th.join();
      for (auto& p : part)
        for (auto& x : p) disc.push_back(std::move(x));
    }
    // candidates are generated in (ia, ib) order already; sort defensively.
    std::stable_sort(disc.begin(), disc.end(),
                     [](const Disc& x, const Disc& y) { return x.ia != y.ia ? x.ia < y.ia : x.ib < y.ib; });
    std::vector<uint8_t> seen(pairs.size(), 0);
    std::vector<std::pair<int, int>> wake;
    trigger_now.clear();
    for (auto& dd : disc) {
      int ia = dd.ia, ib = dd.ib;
      if (dd.trigger) {  // rows.rs:803-807: overlapping trigger pairs
        if (dd.signed_ <= T(0)) trigger_now.push_back({ia, ib});
        continue;
      }
      if (dd.exists) {
        int idx = -1;
        if (use_index) {
          auto it = index.find((uint64_t(ia) << 32) | uint32_t(ib));
          if (it != index.end()) idx = it->second;
        } else {
          for (size_t i = 0; i < pairs.size(); ++i)
            if (pairs[i].a == ia && pairs[i].b == ib) {
              idx = (int)i;
              break;
            }
        }
        if (idx < 0) continue;
        seen[idx] = 1;
        Pair<T>& pair = pairs[idx];
        pair.n = dd.normal;
        pair.mu[0] = dd.mu[0];
        pair.mu[1] = dd.mu[1];
        pair.gap = dd.signed_;
        if (dd.wake) wake.push_back({ia, ib});
        std::vector<Point<T>> next;
        bool touching = dd.signed_ <= T(0);
        bool smooth = bodies[ia].shape == SH_SPHERE || bodies[ib].shape == SH_SPHERE;
        for (auto& f : dd.fresh) {
          const Point<T>* old = nullptr;
          for (auto& p : pair.pts)
            if (length(p.ra - f.first) < K<T>::LAMBDA_MATCH()) {
              old = &p;
              break;
            }
          Point<T> np;
          if (old) {
            np.src_pair = idx;
            np.src_pt = (int)(old - pair.pts.data());
            bool keep = old->stuck && touching && !smooth;
            np.ra = keep ? old->ra : f.first;
            np.rb = keep ? old->rb : f.second;
            for (int k = 0; k < 3; ++k) {
              np.lam[k] = old->lam[k];
              np.pen[k] = old->pen[k];
              np.roll[k] = old->roll[k];
            }
            np.stuck = old->stuck;
          } else {
            np.ra = f.first;
            np.rb = f.second;
            for (int k = 0; k < 3; ++k) {
              np.lam[k] = T(0);
              np.pen[k] = K<T>::PEN_INIT();
              np.roll[k] = T(0);
            }
            np.stuck = true;
          }
          next.push_back(np);
        }
        pair.pts = std::move(next);
      } else {
        seen.push_back(1);
        if (dd.wake) wake.push_back({ia, ib});
        Pair<T> p;
        p.a = ia;
        p.b = ib;
        p.n = dd.normal;
        p.mu[0] = dd.mu[0];
        p.mu[1] = dd.mu[1];
        p.gap = dd.signed_;
        for (auto& f : dd.fresh) {
          Point<T> np;
          np.ra = f.first;
          np.rb = f.second;
          for (int k = 0; k < 3; ++k) {
            np.lam[k] = T(0);
            np.pen[k] = K<T>::PEN_INIT();
            np.roll[k] = T(0);
          }
          np.stuck = true;
          p.pts.push_back(np);
        }
        pairs.push_back(std::move(p));
      }
    }
    for (int idx = (int)seen.size() - 1; idx >= 0; --idx) {
      if (!seen[idx]) {
        if (idx != (int)pairs.size() - 1) pairs[idx] = std::move(pairs.back());
        pairs.pop_back();
      }
    }
    for (auto& w : wake) {
      wake_body(w.first);
      wake_body(w.second);
    }
    if (order == ORDER_COLOUR) {
      std::sort(pairs.begin(), pairs.end(),
                [](const Pair<T>& x, const Pair<T>& y) { return x.a != y.a ? x.a < y.a : x.b < y.b; });
    }
  }

  // ---- row helpers (rows.rs:947-1111; joints_ext.rs:21-32) ----
  T gap_c0(const Pair<T>& p, const Point<T>& pt) const {
    const Body<T>& a = bodies[p.a];
    const Body<T>& b = bodies[p.b];
    if (a.shape == SH_SPHERE || b.shape == SH_SPHERE) return p.gap + K<T>::MARGIN();
    V3<T> pa = pos0[p.a] + rotate(rot0[p.a], pt.ra);
    V3<T> pb = pos0[p.b] + rotate(rot0[p.b], pt.rb);
    return dot(p.n, pa - pb) + K<T>::MARGIN();
  }
  void contact_levers(const Pair<T>& p, const Point<T>& pt, V3<T>* ra_w, V3<T>* rb_w) const {
    const Body<T>& a = bodies[p.a];
    const Body<T>& b = bodies[p.b];
    *ra_w = a.shape == SH_SPHERE ? (-p.n) * a.sp.x : rotate(a.rot, pt.ra);
    *rb_w = b.shape == SH_SPHERE ? p.n * b.sp.x : rotate(b.rot, pt.rb);
  }
  void friction_levers(const Pair<T>& p, const Point<T>& pt, V3<T>* ra_f, V3<T>* rb_f) const {
    V3<T> ra_w, rb_w;
    contact_levers(p, pt, &ra_w, &rb_w);
    const Body<T>& a = bodies[p.a];
    const Body<T>& b = bodies[p.b];
    T sep = dot((a.pos + ra_w) - (b.pos + rb_w), p.n);
    *ra_f = ra_w - p.n * (sep * K<T>::HALF());
    *rb_f = rb_w + p.n * (sep * K<T>::HALF());
  }
  T row_c_levers(const Pair<T>& p, V3<T> u, T c0, V3<T> ra_w, V3<T> rb_w) const {
    const Body<T>& a = bodies[p.a];
    const Body<T>& b = bodies[p.b];
    V3<T> d_a = (a.pos - pos0[p.a]) + u * dot(cross(ra_w, u), quat_diff_vec(a.rot, rot0[p.a], math));
    V3<T> d_b = (b.pos - pos0[p.b]) + u * dot(cross(rb_w, u), quat_diff_vec(b.rot, rot0[p.b], math));
    return c0 * (T(1) - K<T>::ALPHA()) + dot(u, d_a - d_b);
  }
  static void contact_force(T cn, const T pens[3], const T lam[3], T ct0, T ct1, const T mu[2], T f[3]) {
    T fn_c = fmin_(pens[0] * cn + lam[0], T(0));
    T r0 = pens[1] * ct0 + lam[1], r1 = pens[2] * ct1 + lam[2];
    T b1 = std::fabs(fn_c) * mu[0];
    T b2 = std::fabs(fn_c) * mu[1];
    if (b1 <= T(0) || b2 <= T(0)) {
      f[0] = fn_c;
      f[1] = b1 <= T(0) ? T(0) : clampf(r0, -b1, b1);
      f[2] = b2 <= T(0) ? T(0) : clampf(r1, -b2, b2);
      return;
    }
    T s = (r0 / b1) * (r0 / b1) + (r1 / b2) * (r1 / b2);
    T k = s > T(1) ? T(1) / std::sqrt(s) : T(1);
    f[0] = fn_c;
    f[1] = r0 * k;
    f[2] = r1 * k;
  }
  static bool row_live(T c, T f) { return std::fabs(c) >= K<T>::C_EPS() || f != T(0); }
  static void stamp_row(T lhs[6][6], T rhs[6], V3<T> axis, T pen, T f, V3<T> r, T sign) {
    V3<T> nn = sign * axis;
    V3<T> t = cross(r, nn);
    for (int x = 0; x < 3; ++x)
      for (int y = 0; y < 3; ++y) {
        lhs[x][y] += pen * (nn[x] * nn[y]);
        lhs[3 + x][3 + y] += pen * (t[x] * t[y]);
        lhs[x][3 + y] += pen * (nn[x] * t[y]);
        lhs[3 + x][y] += pen * (nn[y] * t[x]);
      }
    rhs[0] += f * nn.x;
    rhs[1] += f * nn.y;
    rhs[2] += f * nn.z;
    rhs[3] += f * t.x;
    rhs[4] += f * t.y;
    rhs[5] += f * t.z;
  }

  // solve_6x6 (assembly.rs:25-69).
  static bool solve_6x6(const T lhs[6][6], const T rhs[6], T x[6]) {
    T l[6][6] = {};
    T d[6] = {};
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j <= i; ++j) {
        T s = lhs[i][j];
        for (int k = 0; k < j; ++k) s -= l[i][k] * d[k] * l[j][k];
        if (i == j) {
          if (s <= K<T>::DEGENERATE()) return false;
          d[i] = s;
          l[i][i] = T(1);
        } else {
          l[i][j] = s / d[j];
        }
      }
    T y[6];
    for (int i = 0; i < 6; ++i) {
      T s = rhs[i];
      for (int k = 0; k < i; ++k) s -= l[i][k] * y[k];
      y[i] = s;
    }
    T z[6];
    for (int i = 0; i < 6; ++i) z[i] = y[i] / d[i];
    for (int i = 5; i >= 0; --i) {
      T s = z[i];
      for (int k = i + 1; k < 6; ++k) s -= l[k][i] * x[k];
      x[i] = s;
    }
    return true;
  }

  // Row accumulation of one body solve.  ORDER_REFERENCE stamps every row
  // straight into the 6x6 system in item order (stamp_row; the reference).
  // ORDER_COLOUR is the GPU's 16-lane body solve: the rows of item k (the
  // body's k-th point-side in pair order, then its joint sides) are summed
  // into slot k mod 16, in item order within the slot, each slot starting at
  // +0; the slots are added by a pairwise tree (((s0 + s1) + (s2 + s3)) +
  // ((s4 + s5) + (s6 + s7))) + (the same over s8-s15), and that total is
  // added to the inertial system.  A stamp adds pen * (x_i * x_j) to
  // lower-triangle entry (i, j) and f * x_k to rhs k, with x = (sign * axis,
  // r x (sign * axis)), as stamp_row does per entry.
  static const int SLOTS = 16;
  struct RowSum {
    bool sliced = false;
    int item = 0;      // index of the current item in the body's list
    T part[SLOTS][27];  // per slot: lower triangle (tri(i, j)), then rhs
    T (*lhs)[6] = nullptr;
    T* rhs = nullptr;
    static int tri(int i, int j) { return i * (i + 1) / 2 + j; }
    void init(bool s, T (*l)[6], T* r) {
      sliced = s;
      lhs = l;
      rhs = r;
      item = 0;
      for (auto& p : part)
        for (auto& v : p) v = T(0);
    }
    T* slot() { return part[item % SLOTS]; }
    void stamp(V3<T> axis, T pen, T f, V3<T> r, T sign) {
      if (!sliced) return stamp_row(lhs, rhs, axis, pen, f, r, sign);
      V3<T> nn = sign * axis;
      V3<T> t = cross(r, nn);
      const T x[6] = {nn.x, nn.y, nn.z, t.x, t.y, t.z};
      T* p = slot();
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) p[tri(i, j)] += pen * (x[i] * x[j]);
      for (int k = 0; k < 6; ++k) p[21 + k] += f * x[k];
    }
    // angular couple (rolling / torsion row): axis g, stiffness pen, force f
    void couple(V3<T> g, T pen, T f) {
      if (!sliced) {
        for (int x = 0; x < 3; ++x)
          for (int y = 0; y < 3; ++y) lhs[3 + x][3 + y] += pen * (g[x] * g[y]);
        rhs[3] += f * g.x;
        rhs[4] += f * g.y;
        rhs[5] += f * g.z;
        return;
      }
      T* p = slot();
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j <= i; ++j) p[tri(3 + i, 3 + j)] += pen * (g[i] * g[j]);
      for (int k = 0; k < 3; ++k) p[24 + k] += f * g[k];
    }
    // diagonal add to the angular block (geometric stiffness)
    void diag(V3<T> hd) {
      if (!sliced) {
        lhs[3][3] += hd.x;
        lhs[4][4] += hd.y;
        lhs[5][5] += hd.z;
        return;
      }
      T* p = slot();
      p[tri(3, 3)] += hd.x;
      p[tri(4, 4)] += hd.y;
      p[tri(5, 5)] += hd.z;
    }
    // ORDER_COLOUR: the slot tree added to the inertial system (lower
    // triangle and rhs; the upper triangle is not read by solve_6x6).
    void finish() {
      if (!sliced) return;
      for (int e = 0; e < 27; ++e) {
        T v[SLOTS];
        for (int k = 0; k < SLOTS; ++k) v[k] = part[k][e];
        for (int w = SLOTS / 2; w >= 1; w /= 2)
          for (int k = 0; k < w; ++k) v[k] = v[2 * k] + v[2 * k + 1];
        T tot = v[0];
        if (e < 21) {
          int i = 0;
          while (tri(i + 1, 0) <= e) ++i;
          int j = e - tri(i, 0);
          lhs[i][j] = lhs[i][j] + tot;
        } else {
          rhs[e - 21] = rhs[e - 21] + tot;
        }
      }
    }
  };

  // solve_body (assembly.rs:73-903), contact + Ball arms.
  void solve_body(int h) {
    const T DT = K<T>::DT();
    const T DT2 = DT * DT;
    Body<T>& body = bodies[h];
    T m_dt2 = T(1) / eff_inv_mass(h) / DT2;
    T iw[3][3];
    world_inertia(body.inertia, body.rot, iw);
    T lhs[6][6] = {};
    T rhs[6] = {};
    for (int i = 0; i < 3; ++i) lhs[i][i] = m_dt2;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) lhs[3 + a][3 + b] = iw[a][b] / DT2;
    V3<T> rl = m_dt2 * (body.pos - inertial[h]);
    rhs[0] = rl.x;
    rhs[1] = rl.y;
    rhs[2] = rl.z;
    V3<T> ra = mat3_vec(iw, quat_diff_vec(body.rot, inertial_rot[h], math)) / DT2;
    rhs[3] = ra.x;
    rhs[4] = ra.y;
    rhs[5] = ra.z;
    RowSum acc;
    acc.init(order == ORDER_COLOUR, lhs, rhs);

    for (int pi : inc_pairs[h]) {
      const Pair<T>& p = pairs[pi];
      bool is_a = p.a == h;
      T sign = is_a ? T(1) : T(-1);
      bool touching = p.gap <= T(0);
      for (const Point<T>& pt : p.pts) {
        struct NextItem {
          RowSum& a;
          ~NextItem() { ++a.item; }
        } next_item{acc};
        T c0 = touching ? gap_c0(p, pt) : T(0);
        V3<T> r_up, r_lo;
        contact_levers(p, pt, &r_up, &r_lo);
        T cn = row_c_levers(p, p.n, c0, r_up, r_lo);
        V3<T> r_side = is_a ? r_up : r_lo;
        if (!touching) {
          V3<T> a_vel = bodies[p.a].vel, b_vel = bodies[p.b].vel;
          T approach = fmax_(dot(b_vel - a_vel, p.n), T(0));
          const bool imminent = freeze_mode == 2 ? frozen_damped[pi] != 0 : !(p.gap - approach * DT > T(0));
          if (!imminent) continue;
          if (br(std::fabs(cn) >= K<T>::C_EPS())) {
            T npts = T(std::max<size_t>(p.pts.size(), 1));
            T ka = eff_inv_mass(p.a), kb = eff_inv_mass(p.b);
            T m_red = ka + kb > K<T>::NEAR_ZERO() ? T(1) / (ka + kb) : T(0);
            if (m_red > T(0)) {
              T pen_d = fmin_(T(2) * m_red / DT2 / npts, K<T>::PEN_MAX());
              T fd0 = pen_d * cn;
              T f_d = br(fd0 < T(0)) ? fd0 : T(0);
              acc.stamp(p.n, pen_d, f_d, r_side, sign);
            }
          }
          continue;
        }
        // the friction frame from the live orientations (assembly.rs:176-181)
        V3<T> t1 = friction_t1(bodies[p.a], bodies[p.b], p.n);
        V3<T> t2 = cross(t1, p.n);
        V3<T> ra_f, rb_f;
        friction_levers(p, pt, &ra_f, &rb_f);
        V3<T> r_side_f = is_a ? ra_f : rb_f;
        T ct1 = row_c_levers(p, t1, T(0), ra_f, rb_f);
        T ct2 = row_c_levers(p, t2, T(0), ra_f, rb_f);
        T f[3];
        contact_force_b(cn, pt.pen, pt.lam, ct1, ct2, p.mu, f);
        V3<T> axes[3] = {p.n, t1, t2};
        T cvs[3] = {cn, ct1, ct2};
        V3<T> rs[3] = {r_side, r_side_f, r_side_f};
        for (int row = 0; row < 3; ++row) {
          if (!br(row_live(cvs[row], f[row]))) continue;
          acc.stamp(axes[row], pt.pen[row], f[row], rs[row], sign);
        }
        T mu_roll = fmax_(bodies[p.a].rolling, bodies[p.b].rolling);
        T mu_spin = fmax_(bodies[p.a].torsion, bodies[p.b].torsion);
        if (mu_roll > T(0) || mu_spin > T(0)) {
          V3<T> drot_a = quat_diff_vec(bodies[p.a].rot, rot0[p.a], math);
          V3<T> drot_b = quat_diff_vec(bodies[p.b].rot, rot0[p.b], math);
          V3<T> wrel = drot_a - drot_b;
          V3<T> rax[3] = {t1, t2, p.n};
          T mur[3] = {mu_roll, mu_roll, mu_spin};
          for (int li = 0; li < 3; ++li) {
            if (mur[li] <= T(0)) continue;
            T c = dot(wrel, rax[li]);
            T cap = mur[li] * std::fabs(f[0]);
            T fr = clampf(K<T>::ROLL_PEN() * c + pt.roll[li], -cap, cap);
            V3<T> g = is_a ? rax[li] : -rax[li];
            acc.couple(g, K<T>::ROLL_PEN(), fr);
          }
        }
      }
    }

    for (int ji : inc_joints[h]) {
      struct NextItem {
        RowSum& a;
        ~NextItem() { ++a.item; }
      } next_item{acc};
      const Joint<T>& j = joints[ji];
      if (j.kind == JK_GEAR) {
        // gear rows (assembly.rs:426-449, :650-679): C = ca + ratio cb -
        // const, the slide sides stamped, then the hinge sides coupled
        const GearSide sd[2] = {gear_side(j, 0), gear_side(j, 1)};
        const T c = sd[0].coord + j.gratio * sd[1].coord - j.ref_val;
        const T f = j.pen_l[0] * c + j.lam_l[0];
        if (!row_live(c, f)) continue;
        const T coef[2] = {T(1), j.gratio};
        for (int k = 0; k < 2; ++k)
          if (!sd[k].ang && (h == sd[k].a || h == sd[k].b))
            acc.stamp(sd[k].axis, j.pen_l[0] * coef[k] * coef[k], f * coef[k], h == sd[k].a ? sd[k].ra : sd[k].rb,
                      h == sd[k].a ? T(-1) : T(1));
        for (int k = 0; k < 2; ++k)
          if (sd[k].ang && (h == sd[k].a || h == sd[k].b))
            acc.couple(h == sd[k].a ? -sd[k].axis : sd[k].axis, j.pen_l[0] * coef[k] * coef[k], f * coef[k]);
        continue;
      }
      bool is_a = j.a == h;
      T sign = is_a ? T(1) : T(-1);
      const Body<T>& a = bodies[j.a];
      const Body<T>& b = bodies[j.b];
      V3<T> pa = a.pos + rotate(a.rot, j.la);
      V3<T> pb = b.pos + rotate(b.rot, j.lb);
      V3<T> pa0 = pos0[j.a] + rotate(rot0[j.a], j.la);
      V3<T> pb0 = pos0[j.b] + rotate(rot0[j.b], j.lb);
      V3<T> live = pa - pb;
      V3<T> c0v = pa0 - pb0;
      V3<T> fl{T(0), T(0), T(0)};
      const V3<T> Z{T(0), T(0), T(0)};
      const bool fixed = j.kind == JK_FIXED;
      // live slide / hinge axis, from A's side (assembly.rs:273-275)
      const V3<T> wa = normalize_or(rotate(a.rot, j.ax_a), V3<T>{T(0), T(0), T(1)});
      if (j.kind == JK_MOTOR) {
        // free drive: no position rows (its impulse ran at the step start)
      } else if (is_tether(j.kind)) {
        // Distance / Rope / Spring: one row along the live anchor delta
        // (assembly.rs:328-425)
        const T len = length(live);
        const V3<T> dir = len > K<T>::NEAR_ZERO() ? live / len : V3<T>{T(0), T(1), T(0)};
        const V3<T> r_side = is_a ? rotate(a.rot, j.la) : rotate(b.rot, j.lb);
        if (j.kind != JK_SPRING) {
          // a rope stamps only while stretched
          if (j.kind == JK_DISTANCE || len > j.ref_val) {
            T c = (len - j.ref_val) - K<T>::ALPHA() * (length(c0v) - j.ref_val);
            T f = j.pen_l[0] * c + j.lam_l[0];
            if (row_live(c, f)) {
              acc.stamp(dir, j.pen_l[0], f, r_side, sign);
              fl = dir * f;
            }
          }
        } else {
          // position-level spring about the rest length: a fixed penalty
          // from the stiffness (scaled by the reduced linear mass for the
          // acceleration model), damping from the step delta, no dual state
          const T sp = len - j.ref_val;
          const T s0 = length(c0v) - j.ref_val;
          const T ka = a.type == BT_DYNAMIC ? a.inv_mass : T(0);
          const T kb = b.type == BT_DYNAMIC ? b.inv_mass : T(0);
          const T m = ka + kb > K<T>::NEAR_ZERO() ? T(1) / (ka + kb) : T(0);
          if (m <= T(0)) continue;
          const bool force = j.m_model == MM_FORCE;
          const T stiff = force ? j.m_stiff : j.m_stiff * m;
          const T dampc = force ? j.m_damp : j.m_damp * m;
          if (stiff <= T(0) && dampc <= T(0)) continue;
          if (stiff > T(0)) {
            const T f = stiff * sp + dampc * (sp - s0) / K<T>::DT();
            acc.stamp(dir, stiff, f, r_side, sign);
          } else {  // a pure damper
            const T pen_v = dampc / K<T>::DT();
            const T f = dampc * (sp - s0) / K<T>::DT();
            acc.stamp(dir, pen_v, f, r_side, sign);
          }
        }
      } else if (j.kind == JK_SIXDOF) {
        // SixDof (assembly.rs:452-519): per axis of A's live frame, a
        // locked row on the drift from the assembly delta or a one-sided
        // row on the raw separation inside its window (warmed by sacc)
        const V3<T> r_side = is_a ? rotate(a.rot, j.la) : rotate(b.rot, j.lb);
        const T lsign = is_a ? T(-1) : T(1);  // B minus A
        const V3<T> evec = (pb - pa) - rotate(a.rot, j.dref);
        for (int i = 0; i < 3; ++i) {
          V3<T> e{T(0), T(0), T(0)};
          e[i] = T(1);
          const V3<T> dir = normalize_or(rotate(a.rot, e), e);
          if (j.six[i] == AX_LOCKED) {
            const T c = dot(evec, dir);
            const T f = j.pen_l[i] * c + j.lam_l[i];
            if (!row_live(c, f)) continue;
            acc.stamp(dir, j.pen_l[i], f, r_side, lsign);
          } else if (j.six[i] == AX_LIMITED) {
            const T sep = dot(evec, dir);
            const int st = warm_limit_state(sep, j.six_lo[i], j.six_hi[i], K<T>::LIMIT_SLOP_LIN(), j.sacc[i]);
            if (st < 0) continue;
            const bool lower = st == 1;
            const V3<T> axis0 = normalize_or(rotate(rot0[j.a], e), e);
            const T initial = dot((pb0 - pa0) - rotate(rot0[j.a], j.dref), axis0);
            const T c = regularized_limit(sep, initial, lower ? j.six_lo[i] : j.six_hi[i], lower);
            const T f_raw = j.pen_l[i] * c + j.sacc[i];
            const T f = lower ? fmin_(f_raw, T(0)) : fmax_(f_raw, T(0));
            if (!row_live(c, f)) continue;
            acc.stamp(dir, j.pen_l[i], f, r_side, lsign);
          }
        }
      } else if (j.kind == JK_WHEEL) {
        // Wheel (assembly.rs:521-580): the two rows across the live
        // suspension axis, then the suspension spring along it (a fixed
        // penalty m w^2 from the frequency, damping 2 m zeta w, no dual),
        // or a rigid row on slot 2 when the frequency is not positive
        V3<T> u, v;
        tangent_basis(wa, &u, &v);
        const V3<T> axes[2] = {u, v};
        const V3<T> r_side = is_a ? rotate(a.rot, j.la) : rotate(b.rot, j.lb);
        for (int li = 0; li < 2; ++li) {
          const V3<T> ax = axes[li];
          T c = dot(live - c0v * K<T>::ALPHA(), ax);
          T f = j.pen_l[li] * c + j.lam_l[li];
          if (!row_live(c, f)) continue;
          acc.stamp(ax, j.pen_l[li], f, r_side, sign);
          fl = fl + ax * f;
        }
        const T sv = dot(pb - pa, wa) - j.ref_val;
        const T lsign = is_a ? T(-1) : T(1);  // C = (B - A).wa
        const V3<T> wa0 = normalize_or(rotate(rot0[j.a], j.ax_a), V3<T>{T(0), T(0), T(1)});
        if (j.susp_hz > T(0)) {
          const T ka = a.type == BT_DYNAMIC ? a.inv_mass : T(0);
          const T kb = b.type == BT_DYNAMIC ? b.inv_mass : T(0);
          const T m = ka + kb > K<T>::NEAR_ZERO() ? T(1) / (ka + kb) : T(0);
          if (m > T(0)) {
            const T omega = K<T>::TAU() * j.susp_hz;
            const T stiff = m * omega * omega;
            const T dampc = T(2) * m * j.susp_zeta * omega;
            const T s0 = dot(pb0 - pa0, wa0) - j.ref_val;
            const T f = stiff * sv + dampc * (sv - s0) / K<T>::DT();
            acc.stamp(wa, stiff, f, r_side, lsign);
          }
        } else {
          const T c = sv - K<T>::ALPHA() * (dot(pb0 - pa0, wa0) - j.ref_val);
          const T f = j.pen_l[2] * c + j.lam_l[2];
          if (row_live(c, f)) acc.stamp(wa, j.pen_l[2], f, r_side, lsign);
        }
      } else if (j.kind != JK_PRISMATIC) {
        // Ball / Revolute / Fixed: three world-axis rows at the anchors
        // (assembly.rs:279-310), Fixed offset by the assembly delta
        const V3<T> lv = fixed ? live + rotate(a.rot, j.dref) : live;
        const V3<T> in0 = fixed ? c0v + rotate(rot0[j.a], j.dref) : c0v;
        for (int k = 0; k < 3; ++k) {
          V3<T> axis{T(0), T(0), T(0)};
          axis[k] = T(1);
          T c = lv[k] - K<T>::ALPHA() * in0[k];
          T f = j.pen_l[k] * c + j.lam_l[k];
          if (!row_live(c, f)) continue;
          V3<T> r_side = is_a ? rotate(a.rot, j.la + (fixed ? j.dref : Z)) : rotate(b.rot, j.lb);
          acc.stamp(axis, j.pen_l[k], f, r_side, sign);
          fl[k] = f;
        }
      } else {
        // Prismatic: the two rows across the live slide axis (:311-327)
        V3<T> u, v;
        tangent_basis(wa, &u, &v);
        const V3<T> axes[2] = {u, v};
        for (int li = 0; li < 2; ++li) {
          const V3<T> ax = axes[li];
          T c = dot(live - c0v * K<T>::ALPHA(), ax);
          T f = j.pen_l[li] * c + j.lam_l[li];
          if (!row_live(c, f)) continue;
          V3<T> r_side = is_a ? rotate(a.rot, j.la) : rotate(b.rot, j.lb);
          acc.stamp(ax, j.pen_l[li], f, r_side, sign);
          fl = fl + ax * f;
        }
      }
      // geometric stiffness (assembly.rs:589-618; rows.rs:69-88)
      V3<T> r = is_a ? rotate(a.rot, j.la + (fixed ? j.dref : Z)) : -rotate(b.rot, j.lb);
      T hm[3][3] = {};
      for (int k = 0; k < 3; ++k) {
        T g[3][3] = {};
        g[0][0] = -r[k];
        g[1][1] = -r[k];
        g[2][2] = -r[k];
        g[0][k] += r[0];
        g[1][k] += r[1];
        g[2][k] += r[2];
        for (int x = 0; x < 3; ++x)
          for (int y = 0; y < 3; ++y) hm[x][y] += g[x][y] * fl[k];
      }
      V3<T> hd{length(V3<T>{hm[0][0], hm[1][0], hm[2][0]}), length(V3<T>{hm[0][1], hm[1][1], hm[2][1]}),
               length(V3<T>{hm[0][2], hm[1][2], hm[2][2]})};
      acc.diag(hd);
      // angular equality rows (assembly.rs:620-648, :787-808)
      if (j.kind == JK_REVOLUTE || j.kind == JK_PRISMATIC || j.kind == JK_WHEEL) {
        // a wheel locks the two directions across its axle (assembly.rs:753-781)
        const V3<T> ea = j.kind == JK_WHEEL ? j.bx_a : j.ax_a, eb = j.kind == JK_WHEEL ? j.bx_b : j.ax_b;
        const V3<T> axa = rotate(a.rot, ea), axb = rotate(b.rot, eb);
        const V3<T> axa0 = rotate(rot0[j.a], ea), axb0 = rotate(rot0[j.b], eb);
        V3<T> t1, t2;
        tangent_basis(axa0, &t1, &t2);
        const V3<T> ts[2] = {t1, t2};
        for (int li = 0; li < 2; ++li) {
          const V3<T> t = ts[li];
          T c = dot(t, axa - axb) - K<T>::ALPHA() * dot(t, axa0 - axb0);
          T f = j.pen_a[li] * c + j.lam_a[li];
          if (!row_live(c, f)) continue;
          V3<T> g = is_a ? cross(axa, t) : -cross(axb, t);
          acc.couple(g, j.pen_a[li], f);
        }
      } else if (j.kind == JK_SIXDOF) {
        // angular rows about A's live axes (assembly.rs:682-740): locked on
        // the relative-rotation error, limited on the twist window
        const V3<T> diff = quat_diff_vec(quat_mul(conj(a.rot), b.rot), j.q_ref, math);
        for (int i = 0; i < 3; ++i) {
          V3<T> e{T(0), T(0), T(0)};
          e[i] = T(1);
          const V3<T> dir = normalize_or(rotate(a.rot, e), e);
          const int k = 3 + i;
          T c, f;
          if (j.six[k] == AX_LOCKED) {
            c = dot(diff, e);
            f = j.pen_a[i] * c + j.lam_a[i];
          } else if (j.six[k] == AX_LIMITED) {
            const T travel = hinge_twist(a.rot, b.rot, e) - quat_twist(j.q_ref, e);
            const int st = warm_limit_state(travel, j.six_lo[k], j.six_hi[k], K<T>::LIMIT_SLOP_ANG(), j.sacc[k]);
            if (st < 0) continue;
            const bool lower = st == 1;
            const T initial = hinge_twist(rot0[j.a], rot0[j.b], e) - quat_twist(j.q_ref, e);
            c = regularized_limit(travel, initial, lower ? j.six_lo[k] : j.six_hi[k], lower);
            const T f_raw = j.pen_a[i] * c + j.sacc[k];
            f = lower ? fmin_(f_raw, T(0)) : fmax_(f_raw, T(0));
          } else {
            continue;
          }
          if (!row_live(c, f)) continue;
          acc.couple(is_a ? -dir : dir, j.pen_a[i], f);
        }
      } else if (fixed) {
        const V3<T> diff = quat_diff_vec(quat_mul(conj(a.rot), b.rot), j.q_ref, math);
        const V3<T> diff0 = quat_diff_vec(quat_mul(conj(rot0[j.a]), rot0[j.b]), j.q_ref, math);
        for (int k = 0; k < 3; ++k) {
          T c = diff[k] - K<T>::ALPHA() * diff0[k];
          T f = j.pen_a[k] * c + j.lam_a[k];
          if (!row_live(c, f)) continue;
          V3<T> e{T(0), T(0), T(0)};
          e[k] = T(1);
          const V3<T> dir = rotate(a.rot, e);
          acc.couple(is_a ? -dir : dir, j.pen_a[k], f);
        }
      }
      // the travel limit (assembly.rs:809-874): a one-sided row warmed by
      // the dual's accumulator, never written here
      if (j.has_lim && (j.kind == JK_REVOLUTE || j.kind == JK_PRISMATIC)) {
        const bool rev = j.kind == JK_REVOLUTE;
        const T v = limit_value(j, a.rot, b.rot, pa, pb, wa);
        const T acc_f = rev ? j.acc_lim : j.lim_dual;
        const int st = warm_limit_state(v, j.lim_lo, j.lim_hi, rev ? K<T>::LIMIT_SLOP_ANG() : K<T>::LIMIT_SLOP_LIN(), acc_f);
        if (st >= 0) {
          const bool lower = st == 1;
          const V3<T> wa0 = normalize_or(rotate(rot0[j.a], j.ax_a), V3<T>{T(0), T(0), T(1)});
          const T initial = limit_value(j, rot0[j.a], rot0[j.b], pa0, pb0, wa0);
          const T c = regularized_limit(v, initial, lower ? j.lim_lo : j.lim_hi, lower);
          const T pen = rev ? j.pen_a[2] : j.pen_l[2];
          const T f_raw = pen * c + acc_f;
          const T f = lower ? fmin_(f_raw, T(0)) : fmax_(f_raw, T(0));
          if (row_live(c, f)) {
            if (rev) {
              acc.couple(is_a ? -wa : wa, pen, f);
            } else {
              const V3<T> r_side = is_a ? rotate(a.rot, j.la) : rotate(b.rot, j.lb);
              acc.stamp(wa, pen, f, r_side, is_a ? T(-1) : T(1));  // C = (B - A).wa
            }
          }
        }
      }
    }
    acc.finish();

    T neg[6] = {-rhs[0], -rhs[1], -rhs[2], -rhs[3], -rhs[4], -rhs[5]};
    T dx[6];
    if (solve_6x6(lhs, neg, dx)) {
      body.pos = body.pos + V3<T>{dx[0], dx[1], dx[2]};
      body.rot = quat_integrate(body.rot, V3<T>{dx[3], dx[4], dx[5]}, math);
    } else {
      stats.breakdowns++;
    }
  }

  // dual_update (joints_ext.rs:464-598), contact + Ball arms.
  void dual_update() {
    for (auto& p : pairs) {
      V3<T> t1 = friction_t1(bodies[p.a], bodies[p.b], p.n);  // joints_ext.rs:470-474
      V3<T> t2 = cross(t1, p.n);
      bool touching = p.gap <= T(0);
      for (auto& pt : p.pts) {
        T c0 = touching ? gap_c0(p, pt) : T(0);
        V3<T> r_up, r_lo;
        contact_levers(p, pt, &r_up, &r_lo);
        T cn = row_c_levers(p, p.n, c0, r_up, r_lo);
        V3<T> ra_f, rb_f;
        friction_levers(p, pt, &ra_f, &rb_f);
        T ct1 = row_c_levers(p, t1, T(0), ra_f, rb_f);
        T ct2 = row_c_levers(p, t2, T(0), ra_f, rb_f);
        T pen[3] = {pt.pen[0], pt.pen[1], pt.pen[2]};
        T lam[3] = {pt.lam[0], pt.lam[1], pt.lam[2]};
        T f[3];
        contact_force_b(cn, pen, lam, ct1, ct2, p.mu, f);
        {
          T bnd = std::fabs(f[0]) * fmax_(p.mu[0], p.mu[1]);
          T u0 = pen[1] * ct1 + lam[1], u1 = pen[2] * ct2 + lam[2];
          T fs = std::sqrt(u0 * u0 + u1 * u1);
          pt.stuck = fs <= bnd && std::sqrt(ct1 * ct1 + ct2 * ct2) < K<T>::STICK();
        }
        if (touching && br(std::fabs(cn) >= K<T>::C_EPS())) {
          pt.lam[0] = f[0];
          if (br(f[0] < T(0))) {
            T pn = pt.pen[0] + K<T>::BETA() * std::fabs(cn);
            pt.pen[0] = br(pn < K<T>::PEN_MAX()) ? pn : K<T>::PEN_MAX();
          }
        }
        T b1 = std::fabs(f[0]) * p.mu[0];
        T b2 = std::fabs(f[0]) * p.mu[1];
        T u0 = pen[1] * ct1 + lam[1], u1 = pen[2] * ct2 + lam[2];
        T s_ell = (b1 > T(0) && b2 > T(0)) ? (u0 / b1) * (u0 / b1) + (u1 / b2) * (u1 / b2)
                                           : std::numeric_limits<T>::infinity();
        if (touching) {
          pt.lam[1] = f[1];
          pt.lam[2] = f[2];
          if (br(s_ell <= T(1))) {
            if (br(std::fabs(ct1) >= K<T>::C_EPS())) {
              T pn = pt.pen[1] + K<T>::BETA() * std::fabs(ct1);
              pt.pen[1] = br(pn < K<T>::PEN_MAX()) ? pn : K<T>::PEN_MAX();
            }
            if (br(std::fabs(ct2) >= K<T>::C_EPS())) {
              T pn = pt.pen[2] + K<T>::BETA() * std::fabs(ct2);
              pt.pen[2] = br(pn < K<T>::PEN_MAX()) ? pn : K<T>::PEN_MAX();
            }
          }
        }
        T mu_roll = fmax_(bodies[p.a].rolling, bodies[p.b].rolling);
        T mu_spin = fmax_(bodies[p.a].torsion, bodies[p.b].torsion);
        if (touching && (mu_roll > T(0) || mu_spin > T(0))) {
          V3<T> drot_a = quat_diff_vec(bodies[p.a].rot, rot0[p.a], math);
          V3<T> drot_b = quat_diff_vec(bodies[p.b].rot, rot0[p.b], math);
          V3<T> wrel = drot_a - drot_b;
          T lam_n = std::fabs(pt.lam[0]);
          T rl[3] = {pt.roll[0], pt.roll[1], pt.roll[2]};
          V3<T> rax[3] = {t1, t2, p.n};
          T mur[3] = {mu_roll, mu_roll, mu_spin};
          for (int li = 0; li < 3; ++li) {
            if (mur[li] <= T(0)) continue;
            T c = dot(wrel, rax[li]);
            T cap = mur[li] * lam_n;
            pt.roll[li] = clampf(K<T>::ROLL_PEN() * c + rl[li], -cap, cap);
          }
        }
      }
    }
    for (auto& j : joints) {
      if (j.kind == JK_GEAR) {
        // the gear pass (joints_ext.rs:897-917) reads only poses and the
        // coupled joints' fixed data, so it may run in joint order here
        const T c = gear_error(j);
        if (std::fabs(c) >= K<T>::C_EPS()) {
          j.lam_l[0] += j.pen_l[0] * c;
          j.pen_l[0] = fmin_(j.pen_l[0] + K<T>::BETA() * std::fabs(c), K<T>::LIM_PEN_MAX());
        }
        continue;
      }
      const Body<T>& a = bodies[j.a];
      const Body<T>& b = bodies[j.b];
      V3<T> pa = a.pos + rotate(a.rot, j.la);
      V3<T> pb = b.pos + rotate(b.rot, j.lb);
      V3<T> pa0 = pos0[j.a] + rotate(rot0[j.a], j.la);
      V3<T> pb0 = pos0[j.b] + rotate(rot0[j.b], j.lb);
      V3<T> live = pa - pb, c0v = pa0 - pb0;
      const bool fixed = j.kind == JK_FIXED;
      if (j.kind == JK_DISTANCE) {
        const T len = length(live);
        T c = (len - j.ref_val) - K<T>::ALPHA() * (length(c0v) - j.ref_val);
        if (std::fabs(c) >= K<T>::C_EPS()) {
          j.lam_l[0] += j.pen_l[0] * c;
          j.pen_l[0] = fmin_(j.pen_l[0] + K<T>::BETA() * std::fabs(c), K<T>::PEN_MAX());
        }
      } else if (j.kind == JK_ROPE) {
        // one-sided: tension only, forgotten when slack; the penalty ramp
        // capped like a limit row's (joints_ext.rs:618-641)
        const T len = length(live);
        if (len > j.ref_val) {
          T c = (len - j.ref_val) - K<T>::ALPHA() * (length(c0v) - j.ref_val);
          if (std::fabs(c) >= K<T>::C_EPS()) {
            j.lam_l[0] = fmax_(j.lam_l[0] + j.pen_l[0] * c, T(0));
            j.pen_l[0] = fmin_(j.pen_l[0] + K<T>::BETA() * std::fabs(c), K<T>::LIM_PEN_MAX());
          }
        } else {
          j.lam_l[0] = T(0);
        }
      } else if (j.kind == JK_SPRING || j.kind == JK_MOTOR) {
        // no dual state
      } else if (j.kind == JK_SIXDOF) {
        // per axis: locked accumulates, limited commits into sacc (zeroed
        // when clear), penalties ramp capped (joints_ext.rs:646-686)
        const V3<T> evec = (pb - pa) - rotate(a.rot, j.dref);
        for (int i = 0; i < 3; ++i) {
          V3<T> e{T(0), T(0), T(0)};
          e[i] = T(1);
          const V3<T> dir = normalize_or(rotate(a.rot, e), e);
          if (j.six[i] == AX_LOCKED) {
            const T c = dot(evec, dir);
            if (std::fabs(c) >= K<T>::C_EPS()) {
              j.lam_l[i] += j.pen_l[i] * c;
              j.pen_l[i] = fmin_(j.pen_l[i] + K<T>::BETA() * std::fabs(c), K<T>::PEN_MAX());
            }
          } else if (j.six[i] == AX_LIMITED) {
            const T sep = dot(evec, dir);
            const int st = warm_limit_state(sep, j.six_lo[i], j.six_hi[i], K<T>::LIMIT_SLOP_LIN(), j.sacc[i]);
            if (st >= 0) {
              const bool lower = st == 1;
              const V3<T> axis0 = normalize_or(rotate(rot0[j.a], e), e);
              const T initial = dot((pb0 - pa0) - rotate(rot0[j.a], j.dref), axis0);
              const T c = regularized_limit(sep, initial, lower ? j.six_lo[i] : j.six_hi[i], lower);
              if (std::fabs(c) >= K<T>::C_EPS()) {
                const T f = j.pen_l[i] * c + j.sacc[i];
                j.sacc[i] = lower ? fmin_(f, T(0)) : fmax_(f, T(0));
                j.pen_l[i] = fmin_(j.pen_l[i] + K<T>::BETA() * std::fabs(c), K<T>::LIM_PEN_MAX());
              }
            } else {
              j.sacc[i] = T(0);
            }
          }
        }
      } else if (j.kind == JK_WHEEL) {
        // the perp rows, and the rigid row on slot 2 when the suspension
        // has no frequency (joints_ext.rs:738-766)
        const V3<T> wa = normalize_or(rotate(a.rot, j.ax_a), V3<T>{T(0), T(0), T(1)});
        V3<T> u, v;
        tangent_basis(wa, &u, &v);
        const V3<T> axes[2] = {u, v};
        for (int li = 0; li < 2; ++li) {
          T c = dot(live - c0v * K<T>::ALPHA(), axes[li]);
          if (std::fabs(c) >= K<T>::C_EPS()) {
            j.lam_l[li] += j.pen_l[li] * c;
            j.pen_l[li] = fmin_(j.pen_l[li] + K<T>::BETA() * std::fabs(c), K<T>::PEN_MAX());
          }
        }
        if (j.susp_hz <= T(0)) {
          const V3<T> wa0 = normalize_or(rotate(rot0[j.a], j.ax_a), V3<T>{T(0), T(0), T(1)});
          T c = (-dot(live, wa) - j.ref_val) - K<T>::ALPHA() * (-dot(pa0 - pb0, wa0) - j.ref_val);
          if (std::fabs(c) >= K<T>::C_EPS()) {
            j.lam_l[2] += j.pen_l[2] * c;
            j.pen_l[2] = fmin_(j.pen_l[2] + K<T>::BETA() * std::fabs(c), K<T>::LIM_PEN_MAX());
          }
        }
      } else if (j.kind != JK_PRISMATIC) {
        const V3<T> lv = fixed ? live + rotate(a.rot, j.dref) : live;
        const V3<T> in0 = fixed ? c0v + rotate(rot0[j.a], j.dref) : c0v;
        for (int k = 0; k < 3; ++k) {
          T c = lv[k] - K<T>::ALPHA() * in0[k];
          if (std::fabs(c) >= K<T>::C_EPS()) {
            j.lam_l[k] += j.pen_l[k] * c;
            j.pen_l[k] = fmin_(j.pen_l[k] + K<T>::BETA() * std::fabs(c), K<T>::PEN_MAX());
          }
        }
      } else {
        const V3<T> wa = normalize_or(rotate(a.rot, j.ax_a), V3<T>{T(0), T(0), T(1)});
        V3<T> u, v;
        tangent_basis(wa, &u, &v);
        const V3<T> axes[2] = {u, v};
        for (int li = 0; li < 2; ++li) {
          T c = dot(live - c0v * K<T>::ALPHA(), axes[li]);
          if (std::fabs(c) >= K<T>::C_EPS()) {
            j.lam_l[li] += j.pen_l[li] * c;
            j.pen_l[li] = fmin_(j.pen_l[li] + K<T>::BETA() * std::fabs(c), K<T>::PEN_MAX());
          }
        }
      }
      // angular equality rows (joints_ext.rs:767-822), a wheel's about its axle
      if (j.kind == JK_REVOLUTE || j.kind == JK_PRISMATIC || j.kind == JK_WHEEL) {
        const V3<T> ea = j.kind == JK_WHEEL ? j.bx_a : j.ax_a, eb = j.kind == JK_WHEEL ? j.bx_b : j.ax_b;
        const V3<T> axa = rotate(a.rot, ea), axb = rotate(b.rot, eb);
        const V3<T> axa0 = rotate(rot0[j.a], ea), axb0 = rotate(rot0[j.b], eb);
        V3<T> t1, t2;
        tangent_basis(axa0, &t1, &t2);
        const V3<T> ts[2] = {t1, t2};
        for (int li = 0; li < 2; ++li) {
          T c = dot(ts[li], axa - axb) - K<T>::ALPHA() * dot(ts[li], axa0 - axb0);
          if (std::fabs(c) < K<T>::C_EPS()) continue;
          j.lam_a[li] = j.pen_a[li] * c + j.lam_a[li];
          j.pen_a[li] = fmin_(j.pen_a[li] + K<T>::BETA_ANG() * std::fabs(c), K<T>::PEN_MAX());
        }
      } else if (j.kind == JK_SIXDOF) {
        const V3<T> diff = quat_diff_vec(quat_mul(conj(a.rot), b.rot), j.q_ref, math);
        for (int i = 0; i < 3; ++i) {
          V3<T> e{T(0), T(0), T(0)};
          e[i] = T(1);
          const int k = 3 + i;
          if (j.six[k] == AX_LOCKED) {
            const T c = dot(diff, e);
            if (std::fabs(c) >= K<T>::C_EPS()) {
              j.lam_a[i] += j.pen_a[i] * c;
              j.pen_a[i] = fmin_(j.pen_a[i] + K<T>::BETA_ANG() * std::fabs(c), K<T>::PEN_MAX());
            }
          } else if (j.six[k] == AX_LIMITED) {
            const T travel = hinge_twist(a.rot, b.rot, e) - quat_twist(j.q_ref, e);
            const int st = warm_limit_state(travel, j.six_lo[k], j.six_hi[k], K<T>::LIMIT_SLOP_ANG(), j.sacc[k]);
            if (st >= 0) {
              const bool lower = st == 1;
              const T initial = hinge_twist(rot0[j.a], rot0[j.b], e) - quat_twist(j.q_ref, e);
              const T c = regularized_limit(travel, initial, lower ? j.six_lo[k] : j.six_hi[k], lower);
              if (std::fabs(c) >= K<T>::C_EPS()) {
                const T f = j.pen_a[i] * c + j.sacc[k];
                j.sacc[k] = lower ? fmin_(f, T(0)) : fmax_(f, T(0));
                j.pen_a[i] = fmin_(j.pen_a[i] + K<T>::BETA_ANG() * std::fabs(c), K<T>::LIM_PEN_MAX());
              }
            } else {
              j.sacc[k] = T(0);
            }
          }
        }
      } else if (fixed) {
        const V3<T> diff = quat_diff_vec(quat_mul(conj(a.rot), b.rot), j.q_ref, math);
        const V3<T> diff0 = quat_diff_vec(quat_mul(conj(rot0[j.a]), rot0[j.b]), j.q_ref, math);
        for (int k = 0; k < 3; ++k) {
          T c = diff[k] - K<T>::ALPHA() * diff0[k];
          if (std::fabs(c) < K<T>::C_EPS()) continue;
          j.lam_a[k] += j.pen_a[k] * c;
          j.pen_a[k] = fmin_(j.pen_a[k] + K<T>::BETA_ANG() * std::fabs(c), K<T>::PEN_MAX());
        }
      }
      // the limit's one-sided accumulator (joints_ext.rs:820-890), zeroed
      // when the joint is clear of both bounds
      if (j.has_lim && (j.kind == JK_REVOLUTE || j.kind == JK_PRISMATIC)) {
        const bool rev = j.kind == JK_REVOLUTE;
        const V3<T> Zv{T(0), T(0), T(1)};
        const V3<T> wa = normalize_or(rotate(a.rot, j.ax_a), Zv);
        const T v = limit_value(j, a.rot, b.rot, pa, pb, wa);
        const int st = warm_limit_state(v, j.lim_lo, j.lim_hi, rev ? K<T>::LIMIT_SLOP_ANG() : K<T>::LIMIT_SLOP_LIN(),
                                        rev ? j.acc_lim : j.lim_dual);
        if (st >= 0) {
          const bool lower = st == 1;
          const V3<T> wa0 = normalize_or(rotate(rot0[j.a], j.ax_a), Zv);
          const T initial = limit_value(j, rot0[j.a], rot0[j.b], pa0, pb0, wa0);
          const T c = regularized_limit(v, initial, lower ? j.lim_lo : j.lim_hi, lower);
          if (std::fabs(c) >= K<T>::C_EPS()) {
            if (rev) {
              const T f = j.pen_a[2] * c + j.acc_lim;
              j.acc_lim = lower ? fmin_(f, T(0)) : fmax_(f, T(0));
              j.pen_a[2] = fmin_(j.pen_a[2] + K<T>::BETA_ANG() * std::fabs(c), K<T>::LIM_PEN_MAX());
            } else {
              const T f_raw = j.pen_l[2] * c + j.lim_dual;
              j.lim_dual = lower ? fmin_(f_raw, T(0)) : fmax_(f_raw, T(0));
              j.pen_l[2] = fmin_(j.pen_l[2] + K<T>::BETA() * std::fabs(c), K<T>::LIM_PEN_MAX());
            }
          }
        } else {
          j.acc_lim = T(0);
          if (!rev) j.lim_dual = T(0);
        }
      }
    }
  }

  // Deterministic greedy colouring of the awake dynamic bodies: vertices
  // in descending (hash(h), then ascending h) priority take the smallest
  // colour unused by already-coloured neighbours; edges are pairs holding
  // at least one point, and joints, between two swept bodies.
  static uint32_t prio_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
  }
  void colour_bodies() {
    int n = (int)bodies.size();
    std::vector<uint8_t> vert(n);
    for (int h = 0; h < n; ++h) vert[h] = solvable(h) && !asleep[h];
    std::vector<std::vector<int>> adj(n);
    for (auto& p : pairs)
      if (!p.pts.empty() && vert[p.a] && vert[p.b]) {
        adj[p.a].push_back(p.b);
        adj[p.b].push_back(p.a);
      }
    for (auto& j : joints) {
      int m[4];
      joint_members(j, m);  // a gear's rows read all four of its members
      for (int x = 0; x < 4; ++x)
        for (int y = x + 1; y < 4; ++y)
          if (m[x] != m[y] && vert[m[x]] && vert[m[y]] && (j.kind == JK_GEAR || (x == 0 && y == 1))) {
            adj[m[x]].push_back(m[y]);
            adj[m[y]].push_back(m[x]);
          }
    }
    std::vector<int> orderv;
    for (int h = 0; h < n; ++h)
      if (vert[h]) orderv.push_back(h);
    std::sort(orderv.begin(), orderv.end(), [](int x, int y) {
      uint32_t hx = prio_hash((uint32_t)x), hy = prio_hash((uint32_t)y);
      return hx != hy ? hx > hy : x < y;
    });
    std::fill(colour.begin(), colour.end(), -1);
    ncolours = 0;
    std::vector<int> used;
    for (int h : orderv) {
      used.clear();
      for (int o : adj[h])
        if (colour[o] >= 0) used.push_back(colour[o]);
      std::sort(used.begin(), used.end());
      int c = 0;
      for (int u : used) {
        if (u == c)
          ++c;
        else if (u > c)
          break;
      }
      colour[h] = c;
      ncolours = std::max(ncolours, c + 1);
    }
  }

  void build_incidence() {
    size_t n = bodies.size();
    inc_pairs.assign(n, {});
    inc_joints.assign(n, {});
    for (size_t i = 0; i < pairs.size(); ++i) {
      inc_pairs[pairs[i].a].push_back((int)i);
      inc_pairs[pairs[i].b].push_back((int)i);
    }
    for (size_t i = 0; i < joints.size(); ++i) {
      int m[4];
      joint_members(joints[i], m);  // a gear joins each distinct member's rows
      for (int x = 0; x < 4; ++x)
        if ((x == 0 || m[x] != m[0]) && (x < 2 || m[x] != m[1]) && (x < 3 || m[x] != m[2]) &&
            (joints[i].kind == JK_GEAR || x < 2))
          inc_joints[m[x]].push_back((int)i);
    }
  }

  // step_inner (mod.rs:846-1143), state-changing phases (events omitted).
  void record_frozen_step() {
    const int n = (int)bodies.size();
    FrozenStep fs;
    fs.toi_hits = stats.toi_hits;
    fs.colour = colour;
    fs.awake.assign(n, 0);
    for (int h = 0; h < n; ++h) fs.awake[h] = solvable(h) && !asleep[h] ? 1 : 0;
    for (const Pair<T>& p : pairs) {
      FrozenPair fp;
      fp.a = p.a;
      fp.b = p.b;
      fp.n = p.n;
      fp.gap = p.gap;
      T approach = fmax_(dot(bodies[p.b].vel - bodies[p.a].vel, p.n), T(0));
      fp.damped = !(p.gap - approach * K<T>::DT() > T(0));
      for (const Point<T>& q : p.pts) fp.pts.push_back(FrozenPoint{q.ra, q.rb, q.stuck, q.src_pair, q.src_pt});
      fs.pairs.push_back(fp);
    }
    frozen.push_back(std::move(fs));
  }

  // One step through the recorded active set `replay_at` (colour order).
  void step_frozen() {
    const T DT = K<T>::DT();
    const int n = (int)bodies.size();
    const FrozenStep& fs = frozen[replay_at++];
    stats = Stats{};
    for (int h = 0; h < n; ++h)
      if (fs.awake[h] && asleep[h]) wake_body(h);
    std::vector<Pair<T>> next;
    frozen_damped.assign(fs.pairs.size(), 0);
    for (size_t i = 0; i < fs.pairs.size(); ++i) {
      const FrozenPair& fp = fs.pairs[i];
      Pair<T> p;
      p.a = fp.a;
      p.b = fp.b;
      p.n = fp.n;
      p.gap = fp.gap;
      p.mu[0] = fmax_(bodies[fp.a].friction, bodies[fp.b].friction);
      p.mu[1] = p.mu[0];
      frozen_damped[i] = fp.damped ? 1 : 0;
      for (const FrozenPoint& fq : fp.pts) {
        Point<T> q;
        q.ra = fq.ra;
        q.rb = fq.rb;
        q.stuck = fq.stuck;
        q.src_pair = fq.src_pair;
        q.src_pt = fq.src_pt;
        if (fq.src_pair >= 0) {
          const Point<T>& o = pairs[fq.src_pair].pts[fq.src_pt];
          for (int k = 0; k < 3; ++k) {
            q.lam[k] = o.lam[k];
            q.pen[k] = o.pen[k];
            q.roll[k] = o.roll[k];
          }
        } else {
          for (int k = 0; k < 3; ++k) {
            q.lam[k] = T(0);
            q.pen[k] = K<T>::PEN_INIT();
            q.roll[k] = T(0);
          }
        }
        p.pts.push_back(q);
      }
      next.push_back(std::move(p));
    }
    pairs = std::move(next);
    for (int h = 0; h < n; ++h) {
      pos0[h] = bodies[h].pos;
      rot0[h] = bodies[h].rot;
      if (!solvable(h)) {
        inertial[h] = bodies[h].pos;
        inertial_rot[h] = bodies[h].rot;
        continue;
      }
      const Body<T>& b = bodies[h];
      inertial[h] = (b.pos + b.vel * DT) + gravity * (DT * DT);
      inertial_rot[h] = quat_integrate(b.rot, b.angvel * DT, math);
    }
    const T AG = K<T>::ALPHA() * K<T>::GAMMA();
    for (auto& p : pairs)
      for (auto& pt : p.pts) {
        for (int k = 0; k < 3; ++k) {
          pt.lam[k] *= AG;
          pt.roll[k] *= AG;
        }
        for (int k = 0; k < 3; ++k) pt.pen[k] = clamp_b(pt.pen[k] * K<T>::GAMMA(), K<T>::PEN_MIN(), K<T>::PEN_MAX());
      }
    for (int h = 0; h < n; ++h) {
      if (!solvable(h) || asleep[h]) continue;
      bodies[h].pos = inertial[h];
      bodies[h].rot = inertial_rot[h];
    }
    build_incidence();
    colour = fs.colour;
    ncolours = 0;
    for (int c : colour) ncolours = std::max(ncolours, c + 1);
    std::vector<std::vector<int>> cls(ncolours);
    for (int h = 0; h < n; ++h)
      if (colour[h] >= 0) cls[colour[h]].push_back(h);
    for (int it = 0; it < iters; ++it) {
      for (int c = 0; c < ncolours; ++c)
        for (int h : cls[c]) solve_body(h);
      dual_update();
    }
    for (int h = 0; h < n; ++h) {
      if (!solvable(h) || asleep[h]) continue;
      Body<T>& b = bodies[h];
      b.vel = (b.pos - pos0[h]) / DT;
      V3<T> spin = quat_diff_vec(b.rot, rot0[h], math);
      b.angvel = br(length(spin) < K<T>::SPIN_REST()) ? V3<T>{T(0), T(0), T(0)} : spin / DT;
      b.rot = qnormalize(b.rot);
    }
    for (int h = 0; h < n; ++h)
      if (fs.slept[h]) sleep_body(h);
    for (int h = 0; h < n; ++h) {
      prev_pos[h] = bodies[h].pos;
      prev_rot[h] = bodies[h].rot;
    }
  }

  // emit_events (mod.rs:1606-1669): trigger transitions, then contact
  // Begin / End (both in pair order), then hits of newly touching pairs in
  // pair-vector order.
  static void set_transitions(const std::vector<std::pair<int, int>>& now, const std::vector<std::pair<int, int>>& prev,
                              int k_in, int k_out, std::vector<Event<T>>& out) {
    size_t i = 0, j = 0;
    while (i < now.size() || j < prev.size()) {
      if (j == prev.size() || (i < now.size() && now[i] < prev[j])) {
        out.push_back(Event<T>{k_in, now[i].first, now[i].second, {}, {}, T(0)});
        ++i;
      } else if (i == now.size() || prev[j] < now[i]) {
        out.push_back(Event<T>{k_out, prev[j].first, prev[j].second, {}, {}, T(0)});
        ++j;
      } else {
        ++i;
        ++j;
      }
    }
  }
  void emit_events() {
    set_transitions(trigger_now, prev_trigger, EV_ENTERED, EV_EXITED, trigger_events);
    prev_trigger = trigger_now;
    std::vector<std::pair<int, int>> touch_now;
    std::vector<Event<T>> hits;
    for (const auto& p : pairs) {
      const Body<T>& a = bodies[p.a];
      const Body<T>& b = bodies[p.b];
      T deepest = T(0);
      V3<T> witness = (a.pos + b.pos) * T(0.5);
      for (const auto& pt : p.pts) {
        const V3<T> pa = a.pos + rotate(a.rot, pt.ra);
        const V3<T> pb = b.pos + rotate(b.rot, pt.rb);
        const T gap = dot(p.n, pa - pb);
        if (-gap > deepest) {
          deepest = -gap;
          witness = pa;
        }
      }
      if (deepest > T(0.005f)) {  // CONTACT_BEGIN_SLOP (trigger.rs:68)
        touch_now.push_back({p.a, p.b});
        if (!std::binary_search(prev_touch.begin(), prev_touch.end(), std::make_pair(p.a, p.b))) {
          const T approach = -dot(pre_vel[p.a] - pre_vel[p.b], p.n);
          if (approach > T(1)) hits.push_back(Event<T>{EV_HIT, p.a, p.b, witness, -p.n, approach});
        }
      }
    }
    std::sort(touch_now.begin(), touch_now.end());
    set_transitions(touch_now, prev_touch, EV_BEGIN, EV_END, contact_events);
    for (auto& e : hits) contact_events.push_back(e);
    prev_touch = std::move(touch_now);
  }

  // One substep; `emit`: the last substep of a step call, which reports
  // contact and trigger transitions (mod.rs:1123-1125); TOI impacts are
  // reported by every substep (mod.rs:1126-1136).
  void step_inner(bool emit = true) {
    const T DT = K<T>::DT();
    int n = (int)bodies.size();
    stats = Stats{};
    const size_t event_start = contact_events.size();
    ccd_hits.clear();
    wake_joint_motion();
    update_gear_mem();
    for (int h = 0; h < n; ++h) {
      if (!solvable(h)) continue;
      Body<T>& b = bodies[h];
      T iw[3][3], iwi[3][3];
      world_inertia(b.inertia, b.rot, iw);
      V3<T> dg{b.inertia.x > T(0) ? b.inertia.x : T(0), b.inertia.y > T(0) ? b.inertia.y : T(0),
               b.inertia.z > T(0) ? b.inertia.z : T(0)};
      inverse_symmetric(iw, dg, iwi);
      V3<T> torque = b.torque;
      b.torque = V3<T>{T(0), T(0), T(0)};
      b.angvel = b.angvel + mat3_vec(iwi, torque) * DT;
    }
    motor_impulse();
    rope_impulse();
    pre_vel.resize(n);
    for (int h = 0; h < n; ++h) pre_vel[h] = bodies[h].vel;  // mod.rs:874-876
    toi_clamp();
    generate_pairs();
    wake_joint_motion();
    for (int h = 0; h < n; ++h)
      if (asleep[h]) {
        const Body<T>& b = bodies[h];
        if (length(b.vel) > K<T>::SLEEP_LIN() || length(b.angvel) > K<T>::SLEEP_ANG()) wake_body(h);
      }
    for (int h = 0; h < n; ++h) {
      if (bodies[h].type == BT_KINEMATIC) {
        pos0[h] = prev_pos[h];
        rot0[h] = prev_rot[h];
      } else {
        pos0[h] = bodies[h].pos;
        rot0[h] = bodies[h].rot;
      }
      if (!solvable(h)) {
        inertial[h] = bodies[h].pos;
        inertial_rot[h] = bodies[h].rot;
        continue;
      }
      const Body<T>& b = bodies[h];
      inertial[h] = (b.pos + b.vel * DT) + gravity * (DT * DT);
      inertial_rot[h] = quat_integrate(b.rot, b.angvel * DT, math);
    }
    const T AG = K<T>::ALPHA() * K<T>::GAMMA();
    for (auto& p : pairs)
      for (auto& pt : p.pts) {
        for (int k = 0; k < 3; ++k) {
          pt.lam[k] *= AG;
          pt.roll[k] *= AG;
        }
        for (int k = 0; k < 3; ++k) pt.pen[k] = clamp_b(pt.pen[k] * K<T>::GAMMA(), K<T>::PEN_MIN(), K<T>::PEN_MAX());
      }
    for (auto& j : joints)
      for (int k = 0; k < 3; ++k) {
        j.lam_l[k] *= AG;
        j.lam_a[k] *= AG;
        j.pen_l[k] = clampf(j.pen_l[k] * K<T>::GAMMA(), K<T>::PEN_MIN(), K<T>::PEN_MAX());
        j.pen_a[k] = clampf(j.pen_a[k] * K<T>::GAMMA(), K<T>::PEN_MIN(), K<T>::PEN_MAX());
      }
    for (int h = 0; h < n; ++h) {
      if (!solvable(h) || asleep[h]) continue;
      bodies[h].pos = inertial[h];
      bodies[h].rot = inertial_rot[h];
    }
    build_incidence();
    if (order == ORDER_COLOUR) {
      colour_bodies();
      if (freeze_mode == 1) record_frozen_step();
      std::vector<std::vector<int>> cls(ncolours);
      for (int h = 0; h < n; ++h)
        if (colour[h] >= 0) cls[colour[h]].push_back(h);
      for (int it = 0; it < iters; ++it) {
        for (int c = 0; c < ncolours; ++c)
          for (int h : cls[c]) solve_body(h);
        dual_update();
      }
      stats.colours = ncolours;
      for (auto& c : cls) stats.max_colour = std::max<int>(stats.max_colour, (int)c.size());
    } else {
      std::fill(colour.begin(), colour.end(), -1);
      ncolours = 0;
      for (int it = 0; it < iters; ++it) {
        for (int h = n - 1; h >= 0; --h)
          if (solvable(h) && !asleep[h]) solve_body(h);
        dual_update();
      }
    }
    for (int h = 0; h < n; ++h) {
      if (!solvable(h) || asleep[h]) continue;
      Body<T>& b = bodies[h];
      b.vel = (b.pos - pos0[h]) / DT;
      V3<T> spin = quat_diff_vec(b.rot, rot0[h], math);
      b.angvel = br(length(spin) < K<T>::SPIN_REST()) ? V3<T>{T(0), T(0), T(0)} : spin / DT;
      b.rot = qnormalize(b.rot);
    }
    std::vector<uint8_t> supported(n, 0);
    for (auto& p : pairs)
      if (p.gap <= T(0) && !p.pts.empty()) {
        supported[p.a] = 1;
        supported[p.b] = 1;
      }
    for (auto& j : joints) {
      supported[j.a] = 1;
      supported[j.b] = 1;
    }
    std::vector<uint8_t> ready = joint_sleep_ready();
    std::vector<uint8_t> freeze = asleep;
    for (int h = 0; h < n; ++h) {
      if (!solvable(h) || asleep[h]) continue;
      if (!supported[h] || !ready[h]) {
        sleep_timer[h] = T(0);
        continue;
      }
      const Body<T>& b = bodies[h];
      if (length(b.vel) < K<T>::SLEEP_LIN() && length(b.angvel) < K<T>::SLEEP_ANG()) {
        sleep_timer[h] += DT;
        freeze[h] = sleep_timer[h] >= K<T>::SLEEP_TIME();
      } else {
        sleep_timer[h] = T(0);
      }
    }
    propagate_joint_flags(freeze, 0);
    if (freeze_mode == 1) {
      frozen.back().slept.assign(n, 0);
      for (int h = 0; h < n; ++h) frozen.back().slept[h] = freeze[h] && !asleep[h] ? 1 : 0;
    }
    for (int h = 0; h < n; ++h)
      if (freeze[h] && !asleep[h]) sleep_body(h);
    for (int h = 0; h < n; ++h) {
      prev_pos[h] = bodies[h].pos;
      prev_rot[h] = bodies[h].rot;
    }
    stats.pairs = (int)pairs.size();
    for (auto& p : pairs) stats.points += (int)p.pts.size();
    for (int h = 0; h < n; ++h) stats.sleeping += asleep[h] ? 1 : 0;
    if (emit) emit_events();
    for (const auto& hit : ccd_hits) {
      bool dup = false;
      for (size_t i = event_start; i < contact_events.size(); ++i) {
        const auto& e = contact_events[i];
        dup = dup || (e.kind == EV_HIT && e.a == hit.a && e.b == hit.b);
      }
      if (!dup) contact_events.push_back(hit);
    }
  }

  // PhysicsEngine::step (mod.rs:1212-1235).
  void step(T dt) {
    if (!std::isfinite(dt) || dt <= T(0)) return;
    const T DT = K<T>::DT();
    time_debt = fmin_(time_debt + dt, DT * T(float(MAX_SUBSTEPS)));
    int k = 0;
    while (time_debt >= DT && k < MAX_SUBSTEPS) {
      time_debt -= DT;
      ++k;
      step_inner(time_debt < DT || k == MAX_SUBSTEPS);
    }
  }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C API (ctypes) — float in/out for both precisions.
// ---------------------------------------------------------------------------
using namespace orc;

struct OracleCtx {
  int prec;
  Engine<float>* f = nullptr;
  Engine<double>* d = nullptr;
};

template <class T>
static Engine<T>* mk(const float* g, int order, int math, int iters, int brute) {
  Engine<T>* e = new Engine<T>();
  e->gravity = V3<T>{T(g[0]), T(g[1]), T(g[2])};
  e->order = order;
  e->math = math;
  e->iters = iters;
  e->brute = brute != 0;
  return e;
}

#define DISPATCH(ctx, ...)      \
  do {                          \
    if ((ctx)->prec == 64) {    \
      auto* E = (ctx)->d;       \
      __VA_ARGS__;              \
    } else {                    \
      auto* E = (ctx)->f;       \
      __VA_ARGS__;              \
    }                           \
  } while (0)

template <class T>
static void add_bodies(Engine<T>* E, int n, const int* shape, const float* sp, const float* pos, const float* quat,
                       const float* vel, const float* angvel, const float* mass, const float* friction,
                       const float* restitution, const uint32_t* layer, const uint32_t* mask, const uint8_t* trigger,
                       const int* type, const float* rolling, const float* torsion) {
  for (int i = 0; i < n; ++i) {
    Body<T> b;
    b.shape = shape[i];
    b.sp = V3<T>{T(sp[3 * i]), T(sp[3 * i + 1]), T(sp[3 * i + 2])};
    b.pos = V3<T>{T(pos[3 * i]), T(pos[3 * i + 1]), T(pos[3 * i + 2])};
    b.rot = Q<T>{T(quat[4 * i]), T(quat[4 * i + 1]), T(quat[4 * i + 2]), T(quat[4 * i + 3])};
    b.vel = V3<T>{T(vel[3 * i]), T(vel[3 * i + 1]), T(vel[3 * i + 2])};
    b.angvel = V3<T>{T(angvel[3 * i]), T(angvel[3 * i + 1]), T(angvel[3 * i + 2])};
    b.torque = V3<T>{T(0), T(0), T(0)};
    T m = T(mass[i]);
    b.type = type[i];
    if (b.type == BT_DYNAMIC && !(m > T(0))) b.type = BT_STATIC;
    if (b.type == BT_DYNAMIC) {
      b.mass = m;
      b.inv_mass = T(1) / m;
      b.inertia = shape_inertia(b.shape, b.sp, m);
    } else {
      b.mass = T(0);
      b.inv_mass = T(0);
      b.inertia = V3<T>{T(0), T(0), T(0)};
    }
    b.friction = T(friction[i]);
    b.restitution = T(restitution[i]);
    b.rolling = rolling ? T(rolling[i]) : T(0);
    b.torsion = torsion ? T(torsion[i]) : T(0);
    b.layer = layer[i];
    b.mask = mask[i];
    b.trigger = trigger[i] != 0;
    E->add_body(b);
  }
}

template <class A, class B>
static void convert_frozen(const Engine<A>& src, Engine<B>& dst) {
  dst.frozen.clear();
  for (const auto& fs : src.frozen) {
    typename Engine<B>::FrozenStep o;
    o.colour = fs.colour;
    o.awake = fs.awake;
    o.slept = fs.slept;
    o.toi_hits = fs.toi_hits;
    for (const auto& fp : fs.pairs) {
      typename Engine<B>::FrozenPair q;
      q.a = fp.a;
      q.b = fp.b;
      q.n = V3<B>{B(fp.n.x), B(fp.n.y), B(fp.n.z)};
      q.gap = B(fp.gap);
      q.damped = fp.damped;
      for (const auto& pt : fp.pts)
        q.pts.push_back(typename Engine<B>::FrozenPoint{V3<B>{B(pt.ra.x), B(pt.ra.y), B(pt.ra.z)},
                                                        V3<B>{B(pt.rb.x), B(pt.rb.y), B(pt.rb.z)}, pt.stuck,
                                                        pt.src_pair, pt.src_pt});
      o.pairs.push_back(std::move(q));
    }
    dst.frozen.push_back(std::move(o));
  }
  dst.btape = src.btape;
}

extern "C" {

void* orc_create(int prec, const float* gravity, int order, int math, int iters, int brute) {
  OracleCtx* c = new OracleCtx();
  c->prec = prec == 64 ? 64 : 32;
  if (c->prec == 64)
    c->d = mk<double>(gravity, order, math, iters, brute);
  else
    c->f = mk<float>(gravity, order, math, iters, brute);
  return c;
}

void orc_destroy(void* h) {
  OracleCtx* c = (OracleCtx*)h;
  delete c->f;
  delete c->d;
  delete c;
}

void orc_add_bodies(void* h, int n, const int* shape, const float* sp, const float* pos, const float* quat,
                    const float* vel, const float* angvel, const float* mass, const float* friction,
                    const float* restitution, const uint32_t* layer, const uint32_t* mask, const uint8_t* trigger,
                    const int* type, const float* rolling, const float* torsion) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, add_bodies(E, n, shape, sp, pos, quat, vel, angvel, mass, friction, restitution, layer, mask, trigger,
                         type, rolling, torsion));
}

int orc_add_joint(void* h, int kind, int a, int b, const float* la, const float* lb, const float* axa,
                  const float* axb) {
  OracleCtx* c = (OracleCtx*)h;
  int r = -1;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    r = E->add_joint(kind, a, b, V3<TT>{TT(la[0]), TT(la[1]), TT(la[2])}, V3<TT>{TT(lb[0]), TT(lb[1]), TT(lb[2])},
                     V3<TT>{TT(axa[0]), TT(axa[1]), TT(axa[2])}, V3<TT>{TT(axb[0]), TT(axb[1]), TT(axb[2])});
  });
  return r;
}

// add_joint with a parameter block (Rope: prm[0] = maximum; Spring:
// target position, stiffness, damping, max force, target velocity, motor
// kind, motor model); axa / axb unused by those kinds.
int orc_add_joint_prm(void* h, int kind, int a, int b, const float* la, const float* lb, const float* axa,
                      const float* axb, const float* prm) {
  OracleCtx* c = (OracleCtx*)h;
  int r = -1;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    r = E->add_joint(kind, a, b, V3<TT>{TT(la[0]), TT(la[1]), TT(la[2])}, V3<TT>{TT(lb[0]), TT(lb[1]), TT(lb[2])},
                     V3<TT>{TT(axa[0]), TT(axa[1]), TT(axa[2])}, V3<TT>{TT(axb[0]), TT(axb[1]), TT(axb[2])}, prm);
  });
  return r;
}

int orc_add_ball_joint(void* h, int a, int b, const float* la, const float* lb) {
  OracleCtx* c = (OracleCtx*)h;
  int r = -1;
  if (c->prec == 64)
    r = c->d->add_ball_joint(a, b, V3<double>{la[0], la[1], la[2]}, V3<double>{lb[0], lb[1], lb[2]});
  else
    r = c->f->add_ball_joint(a, b, V3<float>{la[0], la[1], la[2]}, V3<float>{lb[0], lb[1], lb[2]});
  return r;
}

void orc_step(void* h, float dt, int nsteps) {
  OracleCtx* c = (OracleCtx*)h;
  for (int i = 0; i < nsteps; ++i) DISPATCH(c, E->step(decltype(E->time_debt)(dt)));
}

void orc_step_inner(void* h, int nsteps) {
  OracleCtx* c = (OracleCtx*)h;
  // like a step call of nsteps substeps: the last one reports transitions
  for (int i = 0; i < nsteps; ++i) DISPATCH(c, {
      if (E->freeze_mode == 2)
        E->step_frozen();
      else
        E->step_inner(i + 1 == nsteps);
    });
}

int orc_body_count(void* h) {
  OracleCtx* c = (OracleCtx*)h;
  int n = 0;
  DISPATCH(c, n = (int)E->bodies.size());
  return n;
}

// out: n x 13 floats (pos3, quat4, vel3, angvel3), double precision cast.
void orc_get_state(void* h, float* out) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->bodies.size(); ++i) {
      auto& b = E->bodies[i];
      float* o = out + 13 * i;
      o[0] = (float)b.pos.x; o[1] = (float)b.pos.y; o[2] = (float)b.pos.z;
      o[3] = (float)b.rot.x; o[4] = (float)b.rot.y; o[5] = (float)b.rot.z; o[6] = (float)b.rot.w;
      o[7] = (float)b.vel.x; o[8] = (float)b.vel.y; o[9] = (float)b.vel.z;
      o[10] = (float)b.angvel.x; o[11] = (float)b.angvel.y; o[12] = (float)b.angvel.z;
    }
  });
}

void orc_get_state_f64(void* h, double* out) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->bodies.size(); ++i) {
      auto& b = E->bodies[i];
      double* o = out + 13 * i;
      o[0] = b.pos.x; o[1] = b.pos.y; o[2] = b.pos.z;
      o[3] = b.rot.x; o[4] = b.rot.y; o[5] = b.rot.z; o[6] = b.rot.w;
      o[7] = b.vel.x; o[8] = b.vel.y; o[9] = b.vel.z;
      o[10] = b.angvel.x; o[11] = b.angvel.y; o[12] = b.angvel.z;
    }
  });
}

int orc_pair_count(void* h) {
  OracleCtx* c = (OracleCtx*)h;
  int n = 0;
  DISPATCH(c, n = (int)E->pairs.size());
  return n;
}

// pairs in vector order: ab[2*i], ab[2*i+1]; npts[i]; gap[i]; normal[3*i].
void orc_get_pairs(void* h, uint32_t* ab, uint32_t* npts, float* gap, float* normal) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->pairs.size(); ++i) {
      auto& p = E->pairs[i];
      if (ab) { ab[2 * i] = (uint32_t)p.a; ab[2 * i + 1] = (uint32_t)p.b; }
      if (npts) npts[i] = (uint32_t)p.pts.size();
      if (gap) gap[i] = (float)p.gap;
      if (normal) { normal[3 * i] = (float)p.n.x; normal[3 * i + 1] = (float)p.n.y; normal[3 * i + 2] = (float)p.n.z; }
    }
  });
}

// points in pair-vector order, 16 floats each:
// ra3 rb3 lam3 pen3 roll3 stuck.
void orc_get_points(void* h, float* out) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    size_t k = 0;
    for (auto& p : E->pairs)
      for (auto& pt : p.pts) {
        float* o = out + 16 * k++;
        o[0] = (float)pt.ra.x; o[1] = (float)pt.ra.y; o[2] = (float)pt.ra.z;
        o[3] = (float)pt.rb.x; o[4] = (float)pt.rb.y; o[5] = (float)pt.rb.z;
        for (int q = 0; q < 3; ++q) { o[6 + q] = (float)pt.lam[q]; o[9 + q] = (float)pt.pen[q]; o[12 + q] = (float)pt.roll[q]; }
        o[15] = pt.stuck ? 1.0f : 0.0f;
      }
  });
}

void orc_get_colours(void* h, int32_t* out) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->colour.size(); ++i) out[i] = E->colour[i];
  });
}

void orc_get_sleep(void* h, uint8_t* asleep, float* timer) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->bodies.size(); ++i) {
      if (asleep) asleep[i] = E->asleep[i];
      if (timer) timer[i] = (float)E->sleep_timer[i];
    }
  });
}

// Anisotropic friction of the first n bodies: axis3 (body frame, normalised
// here; zero: isotropic) and the friction across it.
void orc_set_friction_axes(void* h, const float* axis3, const float* trans, int n) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    for (int i = 0; i < n && i < (int)E->bodies.size(); ++i) {
      // RigidBody::set_friction_axis (body.rs:452-460): normalize_checked in f32
      float v[3] = {axis3[3 * i], axis3[3 * i + 1], axis3[3 * i + 2]};
      if (v[0] != 0.0f || v[1] != 0.0f || v[2] != 0.0f) {
        const float r = 1.0f / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (float& x : v) x *= r;
      }
      E->bodies[i].fdir = V3<TT>{TT(v[0]), TT(v[1]), TT(v[2])};
      E->bodies[i].ftrans = TT(trans[i]);
    }
  });
}

// Events queued since the last drain: which = 0 contact, 1 trigger.  With
// out arrays, copies all of them (ints: kind, a, b; floats: point3,
// normal3, approach) and empties the queue; returns the count.
int orc_drain_events(void* h, int which, int32_t* out_i, float* out_f) {
  OracleCtx* c = (OracleCtx*)h;
  int cnt = 0;
  DISPATCH(c, {
    auto& q = which ? E->trigger_events : E->contact_events;
    cnt = (int)q.size();
    if (out_i) {
      for (int k = 0; k < cnt; ++k) {
        const auto& e = q[k];
        out_i[3 * k] = e.kind;
        out_i[3 * k + 1] = e.a;
        out_i[3 * k + 2] = e.b;
        const float f[7] = {float(e.point.x), float(e.point.y), float(e.point.z), float(e.normal.x),
                            float(e.normal.y), float(e.normal.z), float(e.approach)};
        for (int m = 0; m < 7; ++m) out_f[7 * k + m] = f[m];
      }
      q.clear();
    }
  });
  return cnt;
}

// The completed-step event baseline (event_state / restore_event_state,
// mod.rs:516-556): which = 0 touching pairs, 1 overlapping trigger pairs.
int orc_get_event_state(void* h, int which, uint32_t* ab) {
  OracleCtx* c = (OracleCtx*)h;
  int cnt = 0;
  DISPATCH(c, {
    const auto& v = which ? E->prev_trigger : E->prev_touch;
    cnt = (int)v.size();
    if (ab)
      for (int k = 0; k < cnt; ++k) {
        ab[2 * k] = (uint32_t)v[k].first;
        ab[2 * k + 1] = (uint32_t)v[k].second;
      }
  });
  return cnt;
}
void orc_set_event_state(void* h, int which, const uint32_t* ab, int cnt) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    auto& v = which ? E->prev_trigger : E->prev_touch;
    v.clear();
    for (int k = 0; k < cnt; ++k) v.push_back({(int)ab[2 * k], (int)ab[2 * k + 1]});
    std::sort(v.begin(), v.end());
  });
}

void orc_remove_body(void* h, int hi) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, E->remove_body(hi));
}

// Replace the pair table (the warm-start state carried between steps):
// pair i is (ab[2i], ab[2i+1]) with normal / gap n4[4i..], friction
// mu2[2i..] and points pts2[2i] .. pts2[2i] + pts2[2i+1] - 1 of `point20`,
// 20 floats each: ra3 _ rb3 _ lam3 _ pen3 stuck roll3 _.
void orc_set_contacts(void* h, int npairs, const uint32_t* ab, const float* n4, const float* mu2,
                      const uint32_t* pts2, const float* point20) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    E->pairs.clear();
    for (int i = 0; i < npairs; ++i) {
      Pair<TT> p;
      p.a = (int)ab[2 * i];
      p.b = (int)ab[2 * i + 1];
      p.n = {TT(n4[4 * i]), TT(n4[4 * i + 1]), TT(n4[4 * i + 2])};
      p.gap = TT(n4[4 * i + 3]);
      p.mu[0] = TT(mu2[2 * i]);
      p.mu[1] = TT(mu2[2 * i + 1]);
      for (uint32_t k = 0; k < pts2[2 * i + 1]; ++k) {
        const float* r = point20 + 20 * (size_t)(pts2[2 * i] + k);
        Point<TT> q;
        q.ra = {TT(r[0]), TT(r[1]), TT(r[2])};
        q.rb = {TT(r[4]), TT(r[5]), TT(r[6])};
        q.stuck = r[15] != 0.0f;
        for (int j = 0; j < 3; ++j) {
          q.lam[j] = TT(r[8 + j]);
          q.pen[j] = TT(r[12 + j]);
          q.roll[j] = TT(r[16 + j]);
        }
        p.pts.push_back(q);
      }
      E->pairs.push_back(std::move(p));
    }
  });
}

// Sleep flags and timers; a sleeping dynamic body also takes the state
// sleep_body leaves (zero velocities, inverse mass and inertia).
void orc_set_sleep(void* h, const uint8_t* asleep, const float* timer) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    for (size_t i = 0; i < E->bodies.size(); ++i) {
      E->sleep_timer[i] = TT(timer[i]);
      E->asleep[i] = 0;
      if (asleep[i]) E->sleep_body((int)i);
      E->asleep[i] = asleep[i] ? 1 : 0;
    }
  });
}

// joints: 12 floats each: lam_l3 pen_l3 lam_a3 pen_a3.
void orc_get_joints(void* h, float* out) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->joints.size(); ++i)
      for (int k = 0; k < 3; ++k) {
        out[12 * i + k] = (float)E->joints[i].lam_l[k];
        out[12 * i + 3 + k] = (float)E->joints[i].pen_l[k];
        out[12 * i + 6 + k] = (float)E->joints[i].lam_a[k];
        out[12 * i + 9 + k] = (float)E->joints[i].pen_a[k];
      }
  });
}

// stats: pairs, points, colours, max_colour, sleeping, breakdowns, toi_hits
void orc_get_stats(void* h, int32_t* out) {
  OracleCtx* c = (OracleCtx*)h;
  Stats s{};
  DISPATCH(c, s = E->stats);
  out[0] = s.pairs; out[1] = s.points; out[2] = s.colours; out[3] = s.max_colour;
  out[4] = s.sleeping; out[5] = s.breakdowns; out[6] = s.toi_hits;
}

// Last step's TOI clamps per body: t (-1 = none) and target index.
void orc_get_toi(void* h, float* t, int32_t* target) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->bodies.size(); ++i) {
      t[i] = i < E->toi_t.size() ? (float)E->toi_t[i] : -1.0f;
      target[i] = i < E->toi_tgt.size() ? E->toi_tgt[i] : -1;
    }
  });
}

// Unit hooks for analytic pins.
// avbd_inertial_hessian_diag (crates/physics/src/gpu/avbd.rs:88-107): the
// lumped diagonal [m/dt^2 x 3, I/dt^2 x 3] of a GpuAvbdMass {inertia[3],
// inv_mass} (:55-60); non-positive or non-finite inv_mass or dt, and a
// degenerate dt^2, give zeros.
static void stage_hessian_diag(const float* mass4, float dt, float diag[6]) {
  for (int k = 0; k < 6; ++k) diag[k] = 0.0f;
  const float inv_mass = mass4[3];
  if (!std::isfinite(inv_mass) || inv_mass <= 0.0f || !std::isfinite(dt) || dt <= 0.0f) return;
  const float dt2 = dt * dt;
  if (!std::isfinite(dt2) || dt2 <= K<float>::DEGENERATE()) return;
  const float lin = 1.0f / (inv_mass * dt2);
  diag[0] = lin;
  diag[1] = lin;
  diag[2] = lin;
  diag[3] = mass4[0] / dt2;
  diag[4] = mass4[1] / dt2;
  diag[5] = mass4[2] / dt2;
}

// Rung-1 staged solve, GpuAvbdStub::stage_diag_solve (crates/physics/src/
// gpu/avbd.rs:207-226): per body the lumped diagonal and avbd_diag_solve_cpu
// (:117-128), rhs / diag on axes above DEGENERATE_EPS and 0 elsewhere.  mass4
// holds GpuAvbdMass records, rhs8 and delta8 GpuBodyState records {lin[3],
// pad, ang[3], pad} (gpu/mod.rs:200-207).
void orc_stage_diag_solve_f32(const float* mass4, const float* rhs8, uint32_t n, float dt, float* delta8) {
  for (uint32_t i = 0; i < n; ++i) {
    float diag[6];
    stage_hessian_diag(mass4 + 4 * (size_t)i, dt, diag);
    const float* r = rhs8 + 8 * (size_t)i;
    const float rhs[6] = {r[0], r[1], r[2], r[4], r[5], r[6]};
    float x[6];
    for (int k = 0; k < 6; ++k) x[k] = diag[k] > K<float>::DEGENERATE() ? rhs[k] / diag[k] : 0.0f;
    float* d = delta8 + 8 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      d[k] = x[k];
      d[4 + k] = x[3 + k];
    }
    d[3] = 0.0f;
    d[7] = 0.0f;
  }
}

// Rung-2 staged solve, avbd_stage_contact_solve (crates/physics/src/gpu/
// avbd.rs:415-449) over packed GpuAvbdSystem records (:320-327; 24 floats:
// GpuAvbdMass inertia[3], inv_mass (:55-60); GpuBodyState residual
// {lin[3], pad, ang[3], pad} (gpu/mod.rs:200-207); GpuAvbdRow axis[3], pen,
// lever[3], force, sign, pad[3] (:264-281)).  Per system: the lumped
// inertial Hessian diagonal, one stamped row (avbd_stamp_row_cpu, :358-390,
// the engine's stamp_row) and solve_6x6 (avbd_ldl_6x6_cpu, :398-403: a
// breakdown gives a zero delta).  delta8 receives GpuBodyState records, ok
// 1.0 / 0.0.
void orc_stage_contact_solve_f32(const float* sys24, uint32_t n, float dt, float* delta8, float* ok) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* s = sys24 + 24 * (size_t)i;
    float diag[6];
    stage_hessian_diag(s, dt, diag);
    float lhs[6][6], rhs[6] = {s[4], s[5], s[6], s[8], s[9], s[10]};
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) lhs[r][c] = r == c ? diag[r] : 0.0f;
    Engine<float>::stamp_row(lhs, rhs, V3<float>{s[12], s[13], s[14]}, s[15], s[19], V3<float>{s[16], s[17], s[18]},
                             s[20]);
    float x[6];
    const bool solved = Engine<float>::solve_6x6(lhs, rhs, x);
    float* d = delta8 + 8 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      d[k] = solved ? x[k] : 0.0f;
      d[4 + k] = solved ? x[3 + k] : 0.0f;
    }
    d[3] = 0.0f;
    d[7] = 0.0f;
    ok[i] = solved ? 1.0f : 0.0f;
  }
}

int orc_solve_6x6_f32(const float* lhs36, const float* rhs6, float* x6) {
  float l[6][6], r[6];
  std::memcpy(l, lhs36, sizeof(l));
  std::memcpy(r, rhs6, sizeof(r));
  return Engine<float>::solve_6x6(l, r, x6) ? 1 : 0;
}

void orc_stamp_row_f32(float* lhs36, float* rhs6, const float* axis, float pen, float f, const float* r, float sign) {
  float l[6][6], rh[6];
  std::memcpy(l, lhs36, sizeof(l));
  std::memcpy(rh, rhs6, sizeof(rh));
  Engine<float>::stamp_row(l, rh, V3<float>{axis[0], axis[1], axis[2]}, pen, f, V3<float>{r[0], r[1], r[2]}, sign);
  std::memcpy(lhs36, l, sizeof(l));
  std::memcpy(rhs6, rh, sizeof(rh));
}

void orc_contact_force_f32(float cn, const float* pens, const float* lam, float ct0, float ct1, const float* mu,
                           float* f) {
  Engine<float>::contact_force(cn, pens, lam, ct0, ct1, mu, f);
}

void orc_quat_integrate_f32(const float* q, const float* v, int math, float* out) {
  Q<float> r = quat_integrate(Q<float>{q[0], q[1], q[2], q[3]}, V3<float>{v[0], v[1], v[2]}, math);
  out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void orc_quat_diff_vec_f32(const float* a, const float* b, int math, float* out) {
  V3<float> r = quat_diff_vec(Q<float>{a[0], a[1], a[2], a[3]}, Q<float>{b[0], b[1], b[2], b[3]}, math);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_tangent_basis_f32(const float* n, float* t1, float* t2) {
  V3<float> a, b;
  tangent_basis(V3<float>{n[0], n[1], n[2]}, &a, &b);
  t1[0] = a.x; t1[1] = a.y; t1[2] = a.z;
  t2[0] = b.x; t2[1] = b.y; t2[2] = b.z;
}

// shape_distance between two placed shapes: out = dist, pa3, pb3, sat.
void orc_shape_distance_f32(int sa, const float* pa_sp, const float* pa_pos, const float* pa_q, int sb,
                            const float* pb_sp, const float* pb_pos, const float* pb_q, float* out) {
  SRef<float> a{sa, {pa_sp[0], pa_sp[1], pa_sp[2]}, {pa_pos[0], pa_pos[1], pa_pos[2]}, {pa_q[0], pa_q[1], pa_q[2], pa_q[3]}};
  SRef<float> b{sb, {pb_sp[0], pb_sp[1], pb_sp[2]}, {pb_pos[0], pb_pos[1], pb_pos[2]}, {pb_q[0], pb_q[1], pb_q[2], pb_q[3]}};
  float sat;
  Dist<float> d = shape_distance(a, b, &sat);
  out[0] = d.dist;
  out[1] = d.pa.x; out[2] = d.pa.y; out[3] = d.pa.z;
  out[4] = d.pb.x; out[5] = d.pb.y; out[6] = d.pb.z;
  out[7] = sat;
}

// ---- internal queries the reference's own unit tests make -----------------
// (crates/physics/src/avbd/contact_tests.rs: pos0/rot0 + generate_pairs,
// gap_c0, row_c on hand-built pairs), float32 engines only.

// Step-start poses := current poses, then generate_pairs (no step).
void orc_prepare_pairs(void* h) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (size_t i = 0; i < E->bodies.size(); ++i) {
      E->pos0[i] = E->bodies[i].pos;
      E->rot0[i] = E->bodies[i].rot;
    }
    E->generate_pairs();
  });
}

// Overwrite body i's pose (a driver teleport: get_body_mut).
void orc_set_pose(void* h, int i, const float* pos, const float* quat) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    E->bodies[i].pos = {TT(pos[0]), TT(pos[1]), TT(pos[2])};
    E->bodies[i].rot = {TT(quat[0]), TT(quat[1]), TT(quat[2]), TT(quat[3])};
  });
}

// A Revolute / Prismatic joint's limit accumulators: out2 = acc_lim, lim_dual.
void orc_get_joint_limit_state(void* h, int j, float* out2) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    out2[0] = (float)E->joints[j].acc_lim;
    out2[1] = (float)E->joints[j].lim_dual;
  });
}
void orc_set_joint_limit_state(void* h, int j, float acc_lim, float lim_dual) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    E->joints[j].acc_lim = acc_lim;
    E->joints[j].lim_dual = lim_dual;
  });
}
// A SixDof joint's limited-axis accumulators (linear xyz, angular xyz).
void orc_get_joint_sacc(void* h, int j, float* out6) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (int k = 0; k < 6; ++k) out6[k] = (float)E->joints[j].sacc[k];
  });
}

// wake_body plus a zero sleep timer (a test's forced wake each step).
void orc_wake_body(void* h, int i) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    E->wake_body(i);
    E->sleep_timer[i] = 0;
  });
}
void orc_set_gravity(void* h, const float* g) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    E->gravity = V3<TT>{TT(g[0]), TT(g[1]), TT(g[2])};
  });
}

// Overwrite body i's velocities.
void orc_set_vel(void* h, int i, const float* vel, const float* angvel) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    E->bodies[i].vel = {TT(vel[0]), TT(vel[1]), TT(vel[2])};
    E->bodies[i].angvel = {TT(angvel[0]), TT(angvel[1]), TT(angvel[2])};
  });
}

// One primal body solve / one dual update outside a step (unit pins of
// the joint rows); the incidence lists are rebuilt first.
void orc_solve_body(void* h, int body) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    E->build_incidence();
    E->solve_body(body);
  });
}
void orc_dual_update(void* h) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, E->dual_update());
}
// Joint j's duals: 12 floats lam_l3 pen_l3 lam_a3 pen_a3.
void orc_set_joint_duals(void* h, int j, const float* d12) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    for (int k = 0; k < 3; ++k) {
      E->joints[j].lam_l[k] = TT(d12[k]);
      E->joints[j].pen_l[k] = TT(d12[3 + k]);
      E->joints[j].lam_a[k] = TT(d12[6 + k]);
      E->joints[j].pen_a[k] = TT(d12[9 + k]);
    }
  });
}
// Inertial target (position and rotation) of body i.
void orc_set_inertial(void* h, int i, const float* pos, const float* quat) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    E->inertial[i] = {TT(pos[0]), TT(pos[1]), TT(pos[2])};
    E->inertial_rot[i] = {TT(quat[0]), TT(quat[1]), TT(quat[2]), TT(quat[3])};
  });
}
// Body i's inertia diagonal (body frame).
void orc_get_inertia(void* h, int i, float* out3) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    for (int k = 0; k < 3; ++k) out3[k] = (float)E->bodies[i].inertia[k];
  });
}

// Step-start pose of body i.
void orc_set_step_start(void* h, int i, const float* pos, const float* quat) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    E->pos0[i] = {TT(pos[0]), TT(pos[1]), TT(pos[2])};
    E->rot0[i] = {TT(quat[0]), TT(quat[1]), TT(quat[2]), TT(quat[3])};
  });
}

// Replace the pair table by one pair (a, b, normal n, friction mu, gap) with
// one point (anchors ra, rb; unit penalties, zero duals).
void orc_set_single_pair(void* h, int a, int b, const float* n, const float* mu, float gap, const float* ra,
                         const float* rb) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    Pair<TT> p;
    p.a = a;
    p.b = b;
    p.n = {TT(n[0]), TT(n[1]), TT(n[2])};
    p.mu[0] = TT(mu[0]);
    p.mu[1] = TT(mu[1]);
    p.gap = TT(gap);
    Point<TT> q;
    q.ra = {TT(ra[0]), TT(ra[1]), TT(ra[2])};
    q.rb = {TT(rb[0]), TT(rb[1]), TT(rb[2])};
    for (int k = 0; k < 3; ++k) {
      q.lam[k] = TT(0);
      q.pen[k] = TT(1);
      q.roll[k] = TT(0);
    }
    q.stuck = true;
    p.pts.push_back(q);
    E->pairs.clear();
    E->pairs.push_back(p);
  });
}

// Overwrite the anchors of point q of pair p.
void orc_set_point(void* h, int p, int q, const float* ra, const float* rb) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    E->pairs[p].pts[q].ra = {TT(ra[0]), TT(ra[1]), TT(ra[2])};
    E->pairs[p].pts[q].rb = {TT(rb[0]), TT(rb[1]), TT(rb[2])};
  });
}

// gap_c0 of point q of pair p (joints_ext.rs:21-32).
double orc_gap_c0(void* h, int p, int q) {
  OracleCtx* c = (OracleCtx*)h;
  double r = 0.0;
  DISPATCH(c, r = (double)E->gap_c0(E->pairs[p], E->pairs[p].pts[q]));
  return r;
}

// row_c (rows.rs:947-958): out = C, ra_w (3), rb_w (3).
void orc_row_c(void* h, int p, int q, const float* axis, float c0, double* out7) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    const auto& pr = E->pairs[p];
    const auto& pt = pr.pts[q];
    V3<TT> ra_w, rb_w;
    E->contact_levers(pr, pt, &ra_w, &rb_w);
    TT cv = E->row_c_levers(pr, V3<TT>{TT(axis[0]), TT(axis[1]), TT(axis[2])}, TT(c0), ra_w, rb_w);
    out7[0] = (double)cv;
    out7[1] = (double)ra_w.x; out7[2] = (double)ra_w.y; out7[3] = (double)ra_w.z;
    out7[4] = (double)rb_w.x; out7[5] = (double)rb_w.y; out7[6] = (double)rb_w.z;
  });
}

// cast_shape against one target (distance.rs:841-891): out = t, normal3,
// valid (1/0).
void orc_cast_f32(int sa, const float* a_sp, const float* a_pos, const float* a_q, const float* delta, int sb,
                  const float* b_sp, const float* b_pos, const float* b_q, float* out5) {
  SRef<float> a{sa, {a_sp[0], a_sp[1], a_sp[2]}, {a_pos[0], a_pos[1], a_pos[2]}, {a_q[0], a_q[1], a_q[2], a_q[3]}};
  SRef<float> b{sb, {b_sp[0], b_sp[1], b_sp[2]}, {b_pos[0], b_pos[1], b_pos[2]}, {b_q[0], b_q[1], b_q[2], b_q[3]}};
  V3<float> d{delta[0], delta[1], delta[2]};
  float len = length(d);
  out5[0] = 0.0f; out5[1] = 0.0f; out5[2] = 0.0f; out5[3] = 0.0f; out5[4] = 0.0f;
  if (len < K<float>::NEAR_ZERO()) return;
  auto hit = Engine<float>::cast_one(a, d / len, len, 1, b);
  out5[0] = hit.t;
  out5[1] = hit.normal.x; out5[2] = hit.normal.y; out5[3] = hit.normal.z;
  out5[4] = hit.valid ? 1.0f : 0.0f;
}

// world_inertia (rows.rs:194-207): out9 row-major.
void orc_world_inertia_f32(const float* inertia, const float* q, float* out9) {
  float m[3][3];
  world_inertia(V3<float>{inertia[0], inertia[1], inertia[2]}, Q<float>{q[0], q[1], q[2], q[3]}, m);
  std::memcpy(out9, m, sizeof(m));
}

// Discovery threads (results are identical for every count).
void orc_set_threads(void* h, int threads) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, E->threads = threads < 1 ? 1 : threads);
}

// Frozen active set: mode 1 records every step's active set, mode 2
// replays the recorded steps from the first (orc_step_inner then steps
// through them); orc_copy_frozen gives one engine another's recording.
void orc_set_freeze(void* h, int mode) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    E->freeze_mode = mode;
    E->replay_at = 0;
    E->bpos = 0;
    if (mode == 1) {
      E->frozen.clear();
      E->btape.clear();
    }
  });
}
int orc_frozen_steps(void* h) {
  OracleCtx* c = (OracleCtx*)h;
  int r = 0;
  DISPATCH(c, r = (int)E->frozen.size());
  return r;
}
// TOI clamps recorded in step k (a frozen step must have none).
int orc_frozen_toi(void* h, int k) {
  OracleCtx* c = (OracleCtx*)h;
  int r = 0;
  DISPATCH(c, r = E->frozen[k].toi_hits);
  return r;
}
// Pair list of recorded step k: (a, b) per pair into ab (cap pairs), count returned.
int orc_frozen_pairs(void* h, int k, uint32_t* ab, int cap) {
  OracleCtx* c = (OracleCtx*)h;
  int r = 0;
  DISPATCH(c, {
    const auto& ps = E->frozen[k].pairs;
    r = (int)ps.size();
    for (int i = 0; i < r && i < cap; ++i) {
      ab[2 * i] = (uint32_t)ps[i].a;
      ab[2 * i + 1] = (uint32_t)ps[i].b;
    }
  });
  return r;
}
void orc_copy_frozen(void* dst, void* src) {
  OracleCtx* d = (OracleCtx*)dst;
  OracleCtx* s = (OracleCtx*)src;
  if (d->prec == 64 && s->prec == 32) return convert_frozen(*s->f, *d->d);
  if (d->prec == 64 && s->prec == 64) {
    d->d->frozen = s->d->frozen;
    d->d->btape = s->d->btape;
  }
  if (d->prec == 32 && s->prec == 32) {
    d->f->frozen = s->f->frozen;
    d->f->btape = s->f->btape;
  }
}

// Exact parameters of body i in the engine's precision (finite-difference
// checks perturb them by less than a float's resolution): initial velocity,
// mass (inverse mass and inertia follow, as RigidBody::build sets them) and
// friction.
void orc_set_params_f64(void* h, int i, const double* vel3, double mass, double friction) {
  OracleCtx* c = (OracleCtx*)h;
  DISPATCH(c, {
    using TT = std::remove_reference_t<decltype(E->bodies[0].pos.x)>;
    auto& b = E->bodies[i];
    if (vel3) b.vel = V3<TT>{TT(vel3[0]), TT(vel3[1]), TT(vel3[2])};
    if (b.type == BT_DYNAMIC && mass > 0.0) {
      b.mass = TT(mass);
      b.inv_mass = TT(1) / b.mass;
      b.inertia = shape_inertia(b.shape, b.sp, b.mass);
    }
    if (friction == friction) b.friction = TT(friction);
  });
}

void orc_fmath_sincos(float x, float* s, float* c) { oavbd_sincosf(x, s, c); }
float orc_fmath_asin(float x) { return oavbd_asinf(x); }
float orc_fmath_atan2(float y, float x) { return oavbd_atan2f(y, x); }

}  // extern "C"
