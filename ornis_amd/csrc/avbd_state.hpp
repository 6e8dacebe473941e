// avbd_state.hpp — device-resident state of one AVBD context (SoA in HBM).
//
// Bodies, the canonical pair table (sorted by (a, b)), contact points,
// Ball joints and the per-step scratch of the pipeline.  All arrays live on
// one GPU; a context's kernels read and write them through this struct,
// passed by value to every launch.
#pragma once

#include <stdint.h>

#include "avbd_math.hpp"

// Step constants (crates/physics/src/avbd/mod.rs:154-247, :1205).
#define K_DT (1.0f / 60.0f)
#define K_ALPHA 0.99f
#define K_GAMMA 0.999f
#define K_BETA 10000.0f
#define K_MARGIN 0.01f
#define K_PEN_MIN 1.0f
#define K_PEN_MAX 1.0e10f
#define K_PEN_INIT 1.0f
#define K_GEN_MARGIN 0.05f
#define K_EXPAND_SLOP 0.005f
#define K_POINT_MATCH 0.03f
#define K_LAMBDA_MATCH 0.05f
#define K_C_EPS 1e-7f
#define K_STICK 0.00001f
#define K_SLEEP_LIN 0.15f
#define K_SLEEP_ANG 0.15f
#define K_SLEEP_TIME 0.5f
#define K_WAKE_IMPACT 0.5f
#define K_WAKE_PEN 0.01f
#define K_ROLL_PEN 100.0f
#define K_MIN_TOI_LEN 1e-9f
#define K_SPIN_REST 1e-9f
#define K_CCD_BOUNCE 1.0f
#define K_NEAR_ZERO 1e-9f
#define K_AXIS_REST_LEN2 1e-16f
#define K_DEGENERATE 1e-12f
#define K_SHAPE_TOUCH 1e-3f
#define K_CCD_GATE 0.5f
#define K_JOINT_REST 0.005f
#define K_LIM_PEN_MAX 1e4f
#define K_MIN_EFF_MASS 1e-9f
#define K_MAX_POINTS 8
#define COLOUR_RING 64   // frontier sizes kept per round number modulo this (a batch is shorter)
#define CSTAT_FRONT 96   // words of the three frontier counters (own cache line)
#define CSTAT_MAX 128    // word of the colour count (own cache line)
#define CSTAT_ROUNDS 129 // word of the number of colouring rounds
#define COLOUR_TAIL_TPB 1024  // threads of the one-workgroup colouring tail
#define COLOUR_PASS_TPB 1024  // threads per workgroup of a colouring pass (one count atomic each)
#define COLOUR_TAIL_FRONTIER 512  // frontier size from which the tail takes over
#define CSTAT_HEAD 160   // words before the colour sizes in cstat
#define CSTAT_EAGER 64   // colour sizes fetched with every colouring check
#define CCLS_COLOURS 64  // colours whose lists are ordered by size class
#define CCLS (CCLS_COLOURS * 8)  // (colour, size class) counters
static_assert(CSTAT_FRONT >= COLOUR_RING && CSTAT_MAX >= CSTAT_FRONT + 32 && CSTAT_HEAD >= CSTAT_MAX + 32,
              "cstat layout");
#define TOI_LANES 64    // lanes cooperating on one TOI cast
#define INC_LANES 8    // lanes per body building its incidence list
#define PREP_LANES 4   // lanes per pair of the per-pair step constants
#define TOI_GATE_LANES 16  // lanes per cast of the TOI gate (15 separating axes)
#define NARROW_LANES 16    // lanes per box-box candidate of the narrowphase distance
#ifndef BP_LANES
#define BP_LANES 16       // lanes per body of the broadphase and TOI target searches
#endif
#define BP_CACHE 32     // broadphase candidates kept per body between count and fill
#define TOI_CACHE 32    // TOI targets kept per mover between count and fill

enum { BT_STATIC = 0, BT_DYNAMIC = 1, BT_KINEMATIC = 2 };

struct Bodies {
  float4* pos;       // xyz, w = effective inverse mass (0 unless dynamic & awake)
  float4* rot;       // quaternion xyzw
  float4* vel;       // xyz
  float4* angvel;    // xyz
  float4* torque;    // xyz
  float4* inertia;   // body-frame diagonal xyz (0 while asleep), w = mass
  float4* shape;     // param xyz, w = bound radius
  float4* mat;       // friction, restitution, rolling, torsion
  float4* fdir;      // anisotropic friction: unit axis in the body frame (zero: isotropic), w = friction across it
  int* kind;         // shape kind
  int* type;         // body type
  uint32_t* layer;
  uint32_t* mask;
  uint32_t* scene;
  uint32_t* prio;    // colouring priority hash of the scene-local index
  uint32_t* local;   // scene-local index (priority tie-break)
  uint8_t* trigger;
  float4* pos0;      // step-start position, w = sphere radius (or -1)
  float4* rot0;
  float4* inertial;
  float4* inertial_rot;
  float4* drot;      // quat_diff_vec(rot, rot0), kept current
  float4* prev_pos;
  float4* prev_rot;
  uint8_t* asleep;
  float* timer;
  int* colour;
  uint32_t* done;    // sweep stamp of the body's last solve in a fused tail launch
  uint8_t* flag_a;   // scratch flags (wake / awake / ready / freeze)
  uint8_t* flag_b;
  uint8_t* supported;
  uint32_t* comp;    // joint component label (static)
  uint8_t* comp_any;
  uint32_t* jp_off;  // joint-partner CSR (static): partners for no-collide
  uint32_t* jp;
  uint32_t* ji_off;  // joint incidence CSR (static): (joint << 1) | side
  uint32_t* ji;
};

// Pairs and points keep the fields one solve reads together in one
// 64-byte record (4 x float4), so a gather touches one cache line.
struct Pairs {
  uint2* ab;         // body indices, a < b
  float4* rec;       // 4 per pair: see n / t1 / t2 / mu
  uint2* pts;        // first point, point count
  // normal xyz (B -> A), w = signed gap
  __host__ __device__ float4& n(size_t p) const { return rec[4 * p]; }
  // friction tangent t1, w = damper stiffness (separated pairs)
  __host__ __device__ float4& t1(size_t p) const { return rec[4 * p + 1]; }
  // t2 = t1 x n
  __host__ __device__ float4& t2(size_t p) const { return rec[4 * p + 2]; }
  // mu1, mu2, rolling, torsion coefficients
  __host__ __device__ float4& mu(size_t p) const { return rec[4 * p + 3]; }
};
struct Points {
  float4* rec;       // 4 per point: see ra / rb / lam / pen
  float4* roll;      // rolling/torsion duals xyz
  uint32_t* pair;    // owning pair
  int* src;          // previous step's point this one inherits its duals from (-1: fresh)
  // anchor on A (body frame), w = step-start violation c0
  __host__ __device__ float4& ra(size_t q) const { return rec[4 * q]; }
  // anchor on B (body frame)
  __host__ __device__ float4& rb(size_t q) const { return rec[4 * q + 1]; }
  __host__ __device__ float4& lam(size_t q) const { return rec[4 * q + 2]; }  // duals xyz
  // penalties xyz, w = stuck (0/1): the dual writes lam and pen, one
  // 32-byte half of the record
  __host__ __device__ float4& pen(size_t q) const { return rec[4 * q + 3]; }
};

// Joint kinds (AvbdJointKind, mod.rs:286-300; the others are out of scope).
enum {
  JK_BALL = 0, JK_REVOLUTE = 1, JK_PRISMATIC = 2, JK_FIXED = 3, JK_DISTANCE = 4, JK_ROPE = 5, JK_SPRING = 6,
  JK_WHEEL = 7, JK_GEAR = 8, JK_SIXDOF = 9, JK_MOTOR = 10  // the reference's AvbdJointKind numbering
};
__host__ __device__ inline bool is_tether(int kind) { return kind == JK_DISTANCE || kind == JK_ROPE || kind == JK_SPRING; }
// SixDof axis configurations (AxisConfig), two bits per axis (linear x y z,
// angular x y z) in the w of the joint's first six record
enum { AX_FREE = 0, AX_LOCKED = 1, AX_LIMITED = 2 };
__host__ __device__ inline int six_cfg(float packed, int k) { return ((uint32_t)packed >> (2 * k)) & 3; }
// Flags of a Spring joint (the w of its parameter record): the motor's
// force-based model; the motor keeps its bodies awake (JointMotor::keeps_awake)
enum { SPRING_FORCE_MODEL = 1, SPRING_KEEPS_AWAKE = 2 };
// Flags of a Revolute / Prismatic joint (the w of its parameter record):
// a travel limit, a velocity motor
enum { HINGE_LIMIT = 1, HINGE_MOTOR = 2 };
#define K_LIMIT_SLOP_ANG 0.005f
#define K_LIMIT_SLOP_LIN 0.002f

struct Joints {
  uint2* ab;
  int* kind;         // JK_*
  float4* la;        // local anchor on A
  float4* lb;        // local anchor on B
  float4* la0;       // A's lever as the reference stamps it: la + 0 (Fixed: la + dref)
  float4* sus;       // Wheel: the unit suspension axis in A's frame
  float4* six;       // SixDof, 4 per joint: linear lo (w: packed AX_* per axis), linear hi, angular lo, hi
  float4* sacc;      // SixDof, 2 per joint: the limited axes' force accumulators (linear, angular)
  float4* ax;        // 2 per joint: unit hinge / slide axis (Wheel: axle) in A's and B's frames; Motor: the
                     // linear target and force budget, the angular target and torque budget
  float4* dref;      // Fixed / Motor / SixDof: assembly anchor offset in A's frame; Distance / Rope / Spring:
                     // (rest distance | maximum | rest length, stiffness, damping, SPRING_* flags);
                     // Revolute / Prismatic: (assembly angle | length, limit lo, hi, HINGE_* flags);
                     // Wheel: (assembly slide, suspension Hz, damping ratio, HINGE_* flags)
  float4* qref;      // Fixed / Motor / SixDof: assembly relative rotation conj(qa) qb; Revolute / Prismatic / Wheel:
                     // the motor (target speed, budget, 0, 0)
  float4* lam;       // lam_l xyz; w: a Prismatic limit's force accumulator
  float4* pen;       // pen_l xyz
  float4* lama;      // lam_a xyz; w: a Revolute limit's force accumulator
  float4* pena;      // pen_a xyz
  float4* c0v;       // per step: step-start anchor separation (Fixed: + rot0_a dref); w: the
                     // limit's step-start angle / slide
  float4* a0;        // per step, 2 per joint: hinge rows t1 | t.(axa0 - axb0), t2 | ...
                     // (tangent basis of the step-start axis); Fixed: step-start rotation error
};

// Broadphase grid + candidate scratch.
struct Grid {
  uint32_t* cnt;     // per bucket
  uint32_t* start;   // per bucket (exclusive scan)
  uint32_t* cursor;  // per bucket fill cursor
  uint32_t* items;   // body ids bucketed
  float4* ipos;      // per item: position xyz, bound radius (same order as items)
  unsigned long long* ikey;  // per item: its cell, packed (cell_key)
  unsigned long long* hkey;  // per body: its cell, packed
  uint32_t* bucket;  // per body bucket id (0xffffffff = large / not gridded)
  uint32_t nbuckets; // power of two
  float cell;
  uint32_t* large;   // large body ids, ascending
  uint32_t nlarge;
  uint32_t* lsc_off; // per dense scene: range of its large bodies in lsc
  uint32_t* lsc;     // large bodies grouped by scene (ascending within a scene)
  uint32_t* sitems;  // all bodies sorted by (scene, index)
  uint32_t* soff;    // per dense scene: its range in sitems
  uint4* lrange;     // per large body k: scene range in sitems (begin, end), flag base, body id
  uint32_t lflat;    // sum over large bodies of their scene's size (flag rows)
  uint8_t* is_large;
};

struct Cands {
  uint2* ab;
  uint32_t* acc;     // accepted (becomes a pair)
  uint32_t* npts;    // fresh point count of accepted candidates
  float4* disc0;     // normal xyz, signed gap
  float4* disc1;     // mu1, mu2, old pair index (as int bits), wake vote
  float4* fresh;     // 8 x (ra xyz, rb.x), 8 x (rb.y, rb.z, -, -) per candidate
};

struct Dev {
  uint32_t* nlist;     // per candidate capacity: the narrowphase gate's survivors
  Bodies b;
  uint32_t n;
  Pairs P;           // current pair table
  Pairs Pold;        // previous step's table (merge source)
  Points Q;
  Points Qold;
  uint32_t npairs, npts, npairs_old, npts_old;
  Joints J;
  uint32_t njoints;
  uint32_t* ropes;   // Rope joint indices, ascending (the velocity catch runs in joint order)
  uint32_t nropes;
  uint32_t* motors;  // Motor joints and Revolute / Prismatic joints with a motor, ascending
  uint32_t nmotors;
  Grid g;
  Cands C;
  float gravity[3];
  uint32_t* counters;  // device counters (see CTR_*)
  uint32_t* tail_err;  // raised by a fused tail launch whose wait gave up (never cleared)
  // per-step incidence
  uint32_t* inc_off;   // n+1
  uint4* inc;          // items per body in pair order: (point | side << 31, pair, a, b)
  uint32_t* bcnt;      // b-side pair counts
  uint32_t* boff;      // n+1
  uint32_t* bcur;
  uint32_t* blist;
  uint32_t* cand_off;  // n+1 candidate offsets per first body
  uint32_t* acc_off;   // ncand+1 exclusive scan of accepted
  uint32_t* poff;      // n+1 first pair of each body's a-run in the current table
  float4* sb;          // 4 per body during the primal/dual loop: pos, rot, pos0, drot
  float* inb;           // 28 per body: inertial block (lower triangle, rhs) of the next sweep
                       // (one 64-byte record, so a partner's state is one cache line)
  uint32_t* poff_old;  // n+1 the same for the previous step's table
  uint32_t* pt_off;    // ncand+1 exclusive scan of fresh points
  uint32_t* ptcnt;     // per body point-side counts
  uint8_t* sw;         // body is swept this step (dynamic, finite mass, awake)
  unsigned long long* ckey;  // colouring priority: hash << 32 | ~scene-local index
  uint32_t* nbr_cnt;   // n+1 higher-priority colouring neighbours (swept partners via points or joints)
  uint32_t* nbr_off;   // n+1
  uint32_t* nbr_cur;   // fill cursor (the number of higher-priority neighbours after the fill)
  uint32_t* nbr;
  uint32_t* remain;    // cstat: frontier size of round r at r % COLOUR_RING
  uint32_t* fcnt;      // cstat + CSTAT_FRONT: frontier counters (ring of three)
  uint32_t* front;     // 2 x n frontier lists (ring of two)
  uint32_t* lflag;     // per large body, one flag per body of its scene (lflat in all)
  uint32_t* lrank;     // exclusive scan of lflag (lflat + 1)
  uint32_t* cstat;     // colouring status: frontier sizes by round, frontier counters, colour count, sizes
  uint32_t* cmax;      // cstat + CSTAT_MAX: number of colours
  uint32_t* colour_cnt;  // cstat + CSTAT_HEAD: bodies per colour
  uint32_t* colour_off;
  uint32_t* colour_cur;
  uint32_t* ccls;        // 2 * CCLS: bodies per (colour, size class), then fill cursors
  uint4* colour_list;       // per coloured body: h | joints << 31, first item, items, first joint side
  float4* toi_hit;     // normal xyz, w = travel to the hit (< 0: none)
  uint32_t* toi_tgt;
  uint32_t* toi_cnt;   // n+1 cast candidates per mover
  uint32_t* toi_off;   // n+1
  uint2* toi_pairs;    // (mover, target) casts
  float4* toi_res;     // per cast: normal xyz, w = travel to the hit (< 0: none)
  uint32_t* pair_cand;  // candidate of each new pair (from the acceptance scan)
  uint32_t* toi_live;  // box-box casts left by the first swept bound (CTR_TOI_LIVE of them)
  uint32_t* bp_cache;  // n x BP_CACHE candidates found by the counting pass
  uint32_t* toi_cache; // n x TOI_CACHE targets found by the counting pass
  unsigned long long* toi_best;  // per mover min (t bits << 32 | target)
  uint8_t* fast;
  // Capacities of the buffers sized by step counts (TOI casts, candidates,
  // and this step's pair and point tables), and the counts the device
  // derives itself, so a step can run without the host reading them: a
  // count past its capacity reads as zero downstream and is recorded in
  // the overflow words (the host then repeats the steps with larger
  // buffers).
  uint32_t cap_toi, cap_cand, cap_pairs, cap_pts;
  uint32_t* dc;   // DC_*: this step's pair and point counts (k_pair_offsets)
  uint32_t* ovf;  // OVF_*: overflow bits and the largest counts that did not fit
  // contact / trigger events (emit_events, mod.rs:1606-1669)
  uint32_t ev_on;      // events enabled
  uint32_t ev_emit;    // this substep reports transitions (the last of a step call)
  float4* pre_vel;     // per body: velocity before the TOI clamp (mod.rs:874-876)
  float4* toi_pt;      // per cast: the target's witness at the hit
  float4* toi_hpt;     // per body: its clamp's witness
  uint32_t* ev_cflag;  // per body: its clamp reports an impact
  uint4* ev_crec;      // per body, 3 words: that impact event
  uint32_t* ev_gflag;  // per candidate: an overlapping trigger pair
};

// A device event record: (kind, a, b, 0), (point xyz, approach speed),
// (normal xyz, 0).
enum { EVK_BEGIN = 0, EVK_END = 1, EVK_HIT = 2, EVK_ENTERED = 3, EVK_EXITED = 4 };

enum { DC_NPAIRS = 0, DC_NPTS = 1, DC_WORDS = 4 };
enum { OVF_FLAGS = 0, OVF_TOI = 1, OVF_CAND = 2, OVF_PAIRS = 3, OVF_PTS = 4, OVF_COLOURS = 5, OVF_WORDS = 8 };

// CTR_NARROW and CTR_NARROW_PEN share one 64-bit word (even index) that the
// narrowphase gate bumps with one atomic.
enum {
  CTR_FAST = 0,
  CTR_REMAIN = 1,
  CTR_MAXCOL = 2,
  CTR_SLEEPING = 3,
  CTR_BREAKDOWN = 4,
  CTR_TOIHITS = 5,
  CTR_SOLVED = 6,
  CTR_PTSIDES = 7,
  CTR_PAIRSIDES = 8,
  CTR_JSIDES = 9,
  CTR_NARROW = 10,  // candidates past the narrowphase gate, listed from the front
  CTR_NARROW_PEN = 11,  // overlapping box pairs past the gate, listed from the back
  CTR_TOI_LIVE = 12,  // casts past the TOI gate
  CTR_COUNT = 13
};
static_assert(CTR_NARROW % 2 == 0 && CTR_NARROW_PEN == CTR_NARROW + 1, "narrowphase list counters share a word");
