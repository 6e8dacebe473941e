// oavbd.hip — liboavbd.so: C ABI, device memory and step orchestration of
// the MI355X AVBD rigid-body step (declarations: include/oavbd.h).
//
// One context = one HIP stream on one GPU.  All solver state stays in HBM
// between steps; a step is a fixed sequence of kernel launches
// (avbd_kernels.hpp) with a handful of small device->host count reads that
// size the next launches (candidate pairs, pairs/points, colour classes).

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/oavbd.h"
#include "avbd_diff.hpp"
#include "avbd_kernels.hpp"
#include "avbd_scan.hpp"

namespace {

struct HipError {
  std::string msg;
};
struct ArgError {  // an invalid argument found while working (EINVAL)
  std::string msg;
};

#define HIPC(expr)                                                                                   \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)};         \
  } while (0)

inline uint32_t blocks(uint64_t n, uint32_t tpb = TPB) { return (uint32_t)((n + tpb - 1) / tpb); }

template <class T>
void dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) throw HipError{std::string("hipMalloc: ") + hipGetErrorString(e)};
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// lowbias32 integer hash (C. Wellons): colouring priority of a local index.
inline uint32_t prio_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

struct HostJoint {
  uint32_t a, b;
  float la[3], lb[3];
  int kind = 0;                      // JK_*
  float axa[3] = {1.0f, 0.0f, 0.0f};  // unit hinge / slide axes (body frames)
  float axb[3] = {1.0f, 0.0f, 0.0f};
  float dref[3] = {0.0f, 0.0f, 0.0f};  // Fixed: assembly anchor offset in A's frame
  float qref[4] = {0.0f, 0.0f, 0.0f, 1.0f};  // Fixed: assembly relative rotation
  // Distance: assembly distance; Rope: maximum; Spring: rest length,
  // stiffness, damping, SPRING_* flags; Revolute / Prismatic: assembly
  // angle / length, limit lo, hi, HINGE_* flags
  float prm[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  float mot[2] = {0.0f, 0.0f};  // Revolute / Prismatic / Wheel motor: target speed, budget
  uint32_t gb[2] = {0u, 0u};    // Gear: the coupled joints (prm: ratio, constant)
  float sus[3] = {0.0f, 1.0f, 0.0f};  // Wheel: unit suspension axis in A's frame (axa / axb hold the axles)
  // SixDof: AX_* per axis (linear x y z, angular x y z) and the windows
  int six[6] = {0, 0, 0, 0, 0, 0};
  float six_lo[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, six_hi[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
};

}  // namespace

// Staging memory of oavbd_solve_systems_noalloc: a file-scope device array
// (one copy per GPU, held until the process ends) so that entry point never
// allocates.  It holds up to NOALLOC_MAX_SYSTEMS systems: the 96-byte records,
// then the 32-byte deltas, then the ok flags.  Host code never passes
// `scratch` itself to a kernel: it uses the device address resolved per GPU
// by noalloc_resolve (at context creation, with that GPU current).
constexpr uint32_t NOALLOC_MAX_SYSTEMS = 65536;
constexpr size_t NOALLOC_BYTES = (size_t)NOALLOC_MAX_SYSTEMS * (96 + 32 + 4);
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current GPU (`device`), resolved once.
static void* noalloc_resolve(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw HipError{"device index out of range for the staging table"};
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw HipError{std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

// Word offset of the colouring check in the pinned readback buffer (the
// words below it hold the counters and single-word reads).
constexpr uint32_t HEAD_H = 64;
constexpr uint32_t TAILERR_H = CTR_COUNT + 2;  // copy of the fused tail's error word
// Default bound on the bodies of the last colour classes that one fused
// launch (k_primal_tail) sweeps: off.  On the 100k-box drop, the chains and
// the batched scenes the fused launch is slower than a launch per class:
// each of its waits and releases is an agent-scope acquire / release, which
// on this GPU invalidates / writes back the XCD's L2.
constexpr uint32_t TAIL_DEFAULT = 0;
constexpr uint32_t POST_FLAG = 63;  // word of the readback sequence flag
constexpr uint32_t OVF_H = 40;      // copy of the overflow words (OVF_WORDS)
constexpr uint32_t DC_H = 48;       // copy of the device counts (DC_WORDS)
constexpr uint32_t TOTALS_H = 52;   // TOI cast and candidate totals of the last step
static_assert(OVF_H >= TAILERR_H + 1 && DC_H >= OVF_H + OVF_WORDS && TOTALS_H >= DC_H + DC_WORDS &&
                  TOTALS_H + 2 <= POST_FLAG,
              "readback words overlap");

// Device copy of everything a step reads from earlier steps (bodies, the
// pair table with its points, joint duals): restoring one makes the next
// steps repeat bit for bit.
struct Snap {
  bool valid = false;
  uint32_t alloc_n = 0, alloc_nj = 0;  // bodies and joints the per-body / per-joint copies hold
  size_t cap_pairs = 0, cap_pts = 0;
  uint32_t npairs = 0, npts = 0, sleeping = 0;
  float time_debt = 0.0f;
  float4 *pos = nullptr, *rot = nullptr, *vel = nullptr, *angvel = nullptr, *torque = nullptr, *inertia = nullptr,
         *prev_pos = nullptr, *prev_rot = nullptr;
  uint8_t *asleep = nullptr, *flag_a = nullptr;
  float* timer = nullptr;
  uint2* ab = nullptr;
  float4* prec = nullptr;
  uint2* ppts = nullptr;
  uint32_t* poff = nullptr;
  float4 *qrec = nullptr, *qroll = nullptr;
  uint32_t* qpair = nullptr;
  int* qsrc = nullptr;
  float4 *jlam = nullptr, *jpen = nullptr, *jlama = nullptr, *jpena = nullptr;
};

// The arrays oavbd_set_bodies was given, defaults filled in: bodies are
// appended or swap-removed here and the device table rebuilt from them.
struct BodyInputs {
  uint32_t n = 0;
  std::vector<int32_t> shape, type;
  std::vector<float> sp3, pos3, quat4, vel3, ang3, mass, friction, restitution, rolling, torsion;
  std::vector<float> fdir3, ftrans;  // anisotropic friction: body-frame unit axis (zero: none), friction across it
  std::vector<uint32_t> layer, mask, scene;
  std::vector<uint8_t> trigger;
};

// One recorded step of the backward pass: device buffers (grown on demand)
// and the host copies of its sizes and colour ranges.
struct TapeHost {
  Tape t{};
  size_t cap_n = 0, cap_pairs = 0, cap_pts = 0;
  uint32_t iters = 0;
  std::vector<uint32_t> col_off, col_cnt;
};

// Primal launch variants (see launch_primal) and colour-size buckets.
enum { PV_W1 = 0, PV_W2, PV_W4, PV_W8, PV_L1, PV_L2, PV_L4, PV_P1, PV_P2, PV_P4, PV_COUNT };
constexpr int PRIMAL_BUCKETS = 24;

struct oavbd_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  oavbd_params prm{};
  uint32_t max_bodies = 0, max_joints = 0;
  Dev d{};
  size_t cap_cand = 0, cap_pairs = 0, cap_pts = 0, cap_pairs_old = 0, cap_pts_old = 0, cap_blist = 0, cap_inc = 0, cap_toi = 0, cap_nbr = 0;
  uint32_t ncand = 0;
  uint32_t* cand_cnt = nullptr;  // n + 1
  TileScan ts;
  uint32_t* counters_h = nullptr;  // pinned, coherent: counters, read_u32 words, colouring check from HEAD_H
  uint32_t* counters_hd = nullptr;  // the same memory as the device addresses it
  uint32_t post_seq = 0, colour_seq = 0;
  std::vector<HostJoint> joints;
  bool joints_dirty = false;
  uint32_t ngears = 0;  // Gear joints in the table (their memory refreshes each step)
  BodyInputs inputs;  // the body table as last uploaded (set_bodies / add / remove)
  std::vector<int32_t> kind_h, type_h;
  float time_debt = 0.0f;
  uint32_t sleeping = 0;
  bool may_sleep = false;  // bodies may be asleep (wake passes needed)
  uint32_t solved = 0;  // bodies coloured (swept) this step
  int grid_rounds = 8;  // colouring rounds launched over the grid before the one-workgroup tail
  uint32_t colour_left1 = 0xffffffffu;  // bodies left after the first colouring pass, previous step
  oavbd_stats stats{};
  std::vector<uint32_t> colour_cnt_h, colour_off_h;
  int waves = 2;    // primal launch shape: wavefronts per workgroup
  int8_t ptable[PRIMAL_BUCKETS];  // primal variant per colour-size bucket (-1: default rule)
  int force_variant = -1;         // every primal launch takes this variant (autotuning)
  uint32_t tail_max = TAIL_DEFAULT;  // bodies of the last colour classes fused into one launch (0: none)
  uint32_t tail_waves = 0;           // wavefront limit of the fused tail launch (0: resident capacity)
  uint32_t tail_stamp = 0;           // stamp of the last fused tail launch
  uint32_t tail_cap[2] = {0u, 0u};   // resident wavefronts of k_primal_tail<4> / <6>
  std::vector<std::pair<uint32_t, int>> plog;  // (class size, variant) of each timed primal launch
  bool tuning = false;
  double tune_ms[PV_COUNT][PRIMAL_BUCKETS] = {};
  uint32_t tune_n[PV_COUNT][PRIMAL_BUCKETS] = {};
  bool rows6 = false;  // rolling/torsion rows present: 6 rows per item
  bool timing = false;
  std::vector<hipEvent_t> pev, dev_ev;  // per-launch event pools
  oavbd_kernel_times kt{};
  hipEvent_t ev[5] = {};
  float ms[5] = {};
  void* noalloc = nullptr;  // this GPU's address of `scratch`
  // differentiable stepping
  std::vector<Snap> snaps;        // checkpoints of oavbd_diff_forward (last: the final state)
  std::vector<TapeHost> tapes;    // recorded steps of one checkpoint interval
  TapeHost* rec = nullptr;        // step being recorded
  uint32_t diff_steps = 0, diff_every = 0, diff_max_pts = 0, diff_max_pairs = 0;
  Adj adj{};
  size_t adj_cap_n = 0, adj_cap_pts = 0, adj_cap_pairs = 0;
  // unwaited steps (step_inner with async): the counts the host last read,
  // which size the buffers and grids of the next batch, the primal
  // launches queued per sweep, and the batch's starting state
  bool async_enabled = false;  // off: an unwaited batch faulted the GPU (cause not yet found)
  bool sizes_known = false;  // a step's counts have been read back since the last upload
  uint32_t last_ncast = 0, last_live = 0, last_ncand = 0, last_npairs = 0, last_npts = 0;
  uint32_t slots = 0;        // colour classes launched per sweep by an unwaited step
  uint32_t slot_spare = 1;   // classes queued beyond the last step's colour count
  uint32_t replays = 0;      // batches repeated with exact sizes after an overflow
  Snap bsnap;
  // contact / trigger events: the queues the host drains, the sets of the
  // last report (pairs a << 32 | b, ascending), and the compaction scratch
  bool any_axis = false;  // some body has an anisotropic friction axis
  bool ev_on = false;
  std::vector<oavbd_event> ev_contact, ev_trigger;
  std::vector<uint64_t> prev_touch, prev_trigger;
  struct EvBuf {
    size_t cap = 0;
    uint32_t *fa = nullptr, *fb = nullptr, *oa = nullptr, *ob = nullptr;
    uint4 *rec = nullptr, *rec_out = nullptr;
    uint2* ab_out = nullptr;
  } evb;
  std::string err;
};

static int fail(oavbd_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

#define GUARD(c, ...)                                       \
  do {                                                      \
    try {                                                   \
      __VA_ARGS__;                                          \
    } catch (const HipError& e_) {                          \
      return fail((c), OAVBD_EDEVICE, e_.msg);              \
    } catch (const ArgError& e_) {                          \
      return fail((c), OAVBD_EINVAL, e_.msg);               \
    } catch (const std::bad_alloc&) {                       \
      return fail((c), OAVBD_ENOMEM, "host allocation failed"); \
    }                                                       \
  } while (0)

// ---------------------------------------------------------------------------
// memory
// ---------------------------------------------------------------------------
static void alloc_bodies(oavbd_ctx* c, uint32_t n) {
  Bodies& b = c->d.b;
  dalloc(&b.pos, n); dalloc(&b.rot, n); dalloc(&b.vel, n); dalloc(&b.angvel, n); dalloc(&b.torque, n);
  dalloc(&b.inertia, n); dalloc(&b.shape, n); dalloc(&b.mat, n); dalloc(&b.fdir, n); dalloc(&b.kind, n); dalloc(&b.type, n);
  dalloc(&b.layer, n); dalloc(&b.mask, n); dalloc(&b.scene, n); dalloc(&b.prio, n); dalloc(&b.local, n);
  dalloc(&b.trigger, n); dalloc(&b.pos0, n); dalloc(&b.rot0, n); dalloc(&b.inertial, n);
  dalloc(&b.inertial_rot, n); dalloc(&b.drot, n); dalloc(&b.prev_pos, n); dalloc(&b.prev_rot, n);
  dalloc(&b.asleep, n); dalloc(&b.timer, n); dalloc(&b.colour, n); dalloc(&b.done, n);
  HIPC(hipMemset(b.done, 0, (size_t)n * sizeof(uint32_t)));
  dalloc(&b.flag_a, n); dalloc(&b.flag_b, n); dalloc(&b.supported, n); dalloc(&b.comp, n);
  dalloc(&b.comp_any, n); dalloc(&b.jp_off, n + 1); dalloc(&b.ji_off, n + 1);
  b.jp = nullptr;
  b.ji = nullptr;
  Dev& d = c->d;
  dalloc(&d.counters, CTR_COUNT);
  dalloc(&d.dc, DC_WORDS);
  dalloc(&d.ovf, OVF_WORDS);
  HIPC(hipMemset(d.dc, 0, DC_WORDS * sizeof(uint32_t)));
  HIPC(hipMemset(d.ovf, 0, OVF_WORDS * sizeof(uint32_t)));
  d.cap_toi = d.cap_cand = d.cap_pairs = d.cap_pts = 0;
  dalloc(&d.tail_err, 1);
  HIPC(hipMemset(d.tail_err, 0, sizeof(uint32_t)));
  dalloc(&d.inc_off, n + 1); dalloc(&d.bcnt, n + 1); dalloc(&d.boff, n + 1); dalloc(&d.bcur, n + 1);
  dalloc(&d.cand_off, n + 1); dalloc(&d.ptcnt, n + 1);
  dalloc(&d.poff, n + 1); dalloc(&d.poff_old, n + 1); dalloc(&d.sb, (size_t)n * 4); dalloc(&d.inb, (size_t)n * 28);
  dalloc(&d.sw, n); dalloc(&d.ckey, n); dalloc(&d.nbr_cnt, n + 1); dalloc(&d.nbr_off, n + 1);
  dalloc(&d.nbr_cur, n + 1);
  d.nbr = nullptr;
  dalloc(&c->cand_cnt, n + 1);
  dalloc(&d.colour_list, n); dalloc(&d.cstat, CSTAT_HEAD + 65536); d.remain = d.cstat; d.cmax = d.cstat + CSTAT_MAX; d.fcnt = d.cstat + CSTAT_FRONT; dalloc(&d.front, 2 * (size_t)n); d.colour_cnt = d.cstat + CSTAT_HEAD; dalloc(&d.colour_off, 65537); dalloc(&d.colour_cur, 65536); dalloc(&d.ccls, 2 * CCLS);
  dalloc(&d.toi_hit, n); dalloc(&d.toi_tgt, n); dalloc(&d.fast, n);
  dalloc(&d.pre_vel, n); dalloc(&d.toi_hpt, n); dalloc(&d.ev_cflag, n); dalloc(&d.ev_crec, 3 * (size_t)n);
  d.toi_pt = nullptr;
  d.ev_gflag = nullptr;
  d.ev_on = d.ev_emit = 0u;
  dalloc(&d.toi_cnt, n + 1); dalloc(&d.toi_off, n + 1); dalloc(&d.toi_best, n);
  d.toi_pairs = nullptr;
  d.toi_res = nullptr;
  d.toi_live = nullptr;
  uint32_t nb = 1;
  while (nb < 2 * n) nb <<= 1;
  d.g.nbuckets = nb;
  dalloc(&d.g.cnt, nb); dalloc(&d.g.start, nb + 1); dalloc(&d.g.cursor, nb);
  dalloc(&d.g.items, n); dalloc(&d.g.ipos, n); dalloc(&d.g.ikey, n); dalloc(&d.g.hkey, n); dalloc(&d.g.bucket, n); dalloc(&d.g.is_large, n);
  dalloc(&d.bp_cache, (size_t)n * BP_CACHE); dalloc(&d.toi_cache, (size_t)n * TOI_CACHE);
  d.g.large = nullptr;
  d.g.lsc_off = nullptr;
  d.g.lsc = nullptr;
  d.g.sitems = nullptr;
  d.g.soff = nullptr;
  d.g.lrange = nullptr;
  d.g.lflat = 0;
  d.lflag = nullptr;
  d.lrank = nullptr;
  dalloc(&c->ts.tile, (size_t)SCAN_JOBS * (SCAN_MAX_TILES + 1));
  dalloc(&c->ts.done, SCAN_JOBS + 1);  // tickets of the jobs, then the post ticket
  HIPC(hipMemset(c->ts.done, 0, (SCAN_JOBS + 1) * sizeof(uint32_t)));
  d.blist = nullptr;
  d.inc = nullptr;
  d.acc_off = nullptr;
  d.pt_off = nullptr;
}

static void free_pairs(Pairs& P) {
  dfree(P.ab); dfree(P.rec); dfree(P.pts);
}
static void alloc_pairs(Pairs& P, size_t cap) {
  dalloc(&P.ab, cap); dalloc(&P.rec, 4 * cap); dalloc(&P.pts, cap);
}
static void free_points(Points& Q) {
  dfree(Q.rec); dfree(Q.roll); dfree(Q.pair); dfree(Q.src);
}
static void alloc_points(Points& Q, size_t cap) {
  dalloc(&Q.rec, 4 * cap); dalloc(&Q.roll, cap); dalloc(&Q.pair, cap); dalloc(&Q.src, cap);
}

static size_t grow(size_t need) {
  size_t c = 1024;
  while (c < need) c = c + c / 2;
  return c;
}

static void ensure_cand(oavbd_ctx* c, size_t need) {
  if (need <= c->cap_cand && c->d.C.ab) return;
  size_t cap = grow(need);
  Cands& C = c->d.C;
  dfree(C.ab); dfree(C.acc); dfree(C.npts); dfree(C.disc0); dfree(C.disc1); dfree(C.fresh);
  dfree(c->d.acc_off); dfree(c->d.pt_off); dfree(c->d.nlist); dfree(c->d.pair_cand);
  dalloc(&C.ab, cap); dalloc(&C.acc, cap); dalloc(&C.npts, cap); dalloc(&C.disc0, cap); dalloc(&C.disc1, cap);
  dalloc(&c->d.nlist, cap);
  dalloc(&c->d.pair_cand, cap);
  dfree(c->d.ev_gflag);
  dalloc(&c->d.ev_gflag, cap);
  dalloc(&C.fresh, cap * 16); dalloc(&c->d.acc_off, cap + 1); dalloc(&c->d.pt_off, cap + 1);
  if (cap + 1 > (size_t)SCAN_MAX_TILES * SCAN_TILE) throw HipError{"candidate count exceeds the scan capacity"};
  c->cap_cand = cap;
  c->d.cap_cand = (uint32_t)cap;
}

static void ensure_pairs(oavbd_ctx* c, size_t need) {
  if (need <= c->cap_pairs && c->d.P.ab) return;
  free_pairs(c->d.P);
  c->cap_pairs = grow(need);
  alloc_pairs(c->d.P, c->cap_pairs);
  c->d.cap_pairs = (uint32_t)c->cap_pairs;
}

static void ensure_points(oavbd_ctx* c, size_t need) {
  if (need <= c->cap_pts && c->d.Q.rec) return;
  free_points(c->d.Q);
  c->cap_pts = grow(need);
  alloc_points(c->d.Q, c->cap_pts);
  c->d.cap_pts = (uint32_t)c->cap_pts;
}

// TOI cast buffers for `need` casts.
static void ensure_toi(oavbd_ctx* c, size_t need) {
  if (need <= c->cap_toi && c->d.toi_pairs) return;
  Dev& d = c->d;
  dfree(d.toi_pairs);
  dfree(d.toi_res);
  dfree(d.toi_live);
  dfree(d.toi_pt);
  c->cap_toi = grow(need);
  dalloc(&d.toi_pairs, c->cap_toi);
  dalloc(&d.toi_res, c->cap_toi);
  dalloc(&d.toi_live, c->cap_toi);
  dalloc(&d.toi_pt, c->cap_toi);
  d.cap_toi = (uint32_t)c->cap_toi;
}

// Per-step incidence scratch, sized for the current pair / point counts.
static void ensure_incidence(oavbd_ctx* c, size_t npairs, size_t npts) {
  size_t nn = 2 * (npairs + (size_t)c->d.njoints);
  if (nn > c->cap_nbr || !c->d.nbr) {
    dfree(c->d.nbr);
    c->cap_nbr = grow(nn);
    dalloc(&c->d.nbr, c->cap_nbr);
  }
  if (npairs > c->cap_blist || !c->d.blist) {
    dfree(c->d.blist);
    c->cap_blist = grow(npairs);
    dalloc(&c->d.blist, c->cap_blist);
  }
  if (2 * npts > c->cap_inc || !c->d.inc) {
    dfree(c->d.inc);
    c->cap_inc = grow(2 * npts);
    dalloc(&c->d.inc, c->cap_inc);
  }
}

static void snap_free(Snap& s);
static void tape_free(TapeHost& h);
static void adj_free(oavbd_ctx* c) {
  for (auto& s : c->snaps) snap_free(s);
  c->snaps.clear();
  for (auto& t : c->tapes) tape_free(t);
  c->tapes.clear();
  Adj& a = c->adj;
  dfree(a.xb); dfree(a.vb); dfree(a.x0b); dfree(a.yb); dfree(a.lb); dfree(a.lb_prev); dfree(a.pb); dfree(a.mass);
  dfree(a.fric);
  c->adj_cap_n = c->adj_cap_pts = c->adj_cap_pairs = 0;
}

static void free_all(oavbd_ctx* c) {
  Bodies& b = c->d.b;
  dfree(b.pos); dfree(b.rot); dfree(b.vel); dfree(b.angvel); dfree(b.torque); dfree(b.inertia); dfree(b.shape);
  dfree(b.mat); dfree(b.fdir); dfree(b.kind); dfree(b.type); dfree(b.layer); dfree(b.mask); dfree(b.scene); dfree(b.prio);
  dfree(b.local); dfree(b.trigger); dfree(b.pos0); dfree(b.rot0); dfree(b.inertial); dfree(b.inertial_rot);
  dfree(b.drot); dfree(b.prev_pos); dfree(b.prev_rot); dfree(b.asleep); dfree(b.timer); dfree(b.colour); dfree(b.done);
  dfree(b.flag_a); dfree(b.flag_b); dfree(b.supported); dfree(b.comp); dfree(b.comp_any);
  dfree(b.jp_off); dfree(b.jp); dfree(b.ji_off); dfree(b.ji);
  Dev& d = c->d;
  free_pairs(d.P); free_pairs(d.Pold); free_points(d.Q); free_points(d.Qold);
  dfree(d.J.ab); dfree(d.J.kind); dfree(d.J.la); dfree(d.J.lb); dfree(d.J.la0); dfree(d.J.ax); dfree(d.J.dref);
  dfree(d.J.sus); dfree(d.J.six); dfree(d.J.sacc);
  dfree(d.J.qref); dfree(d.J.lam); dfree(d.J.pen); dfree(d.J.lama); dfree(d.J.pena); dfree(d.J.c0v); dfree(d.J.a0);
  dfree(d.ropes); dfree(d.motors);
  dfree(d.g.cnt); dfree(d.g.start); dfree(d.g.cursor); dfree(d.g.items); dfree(d.g.ipos); dfree(d.g.ikey); dfree(d.g.hkey); dfree(d.bp_cache); dfree(d.toi_cache); dfree(d.g.bucket); dfree(d.g.large); dfree(d.g.lsc_off); dfree(d.g.lsc); dfree(d.g.sitems); dfree(d.g.soff); dfree(d.g.lrange);
  dfree(d.g.is_large);
  Cands& C = d.C;
  dfree(C.ab); dfree(C.acc); dfree(C.npts); dfree(C.disc0); dfree(C.disc1); dfree(C.fresh);
  dfree(d.counters); dfree(d.dc); dfree(d.ovf); dfree(d.tail_err); dfree(d.inc_off); dfree(d.inc); dfree(d.bcnt); dfree(d.boff); dfree(d.bcur); dfree(d.blist);
  dfree(d.cand_off); dfree(d.poff); dfree(d.poff_old); dfree(d.sb); dfree(d.inb); dfree(d.acc_off); dfree(d.pt_off); dfree(d.nlist); dfree(d.ptcnt); dfree(d.lflag); dfree(d.lrank);
  dfree(d.cstat); dfree(d.front); dfree(d.colour_off); dfree(d.colour_cur); dfree(d.ccls); dfree(d.colour_list); dfree(d.toi_hit);
  dfree(d.toi_tgt); dfree(d.fast); dfree(d.pre_vel); dfree(d.toi_hpt); dfree(d.ev_cflag); dfree(d.ev_crec);
  dfree(d.toi_pt); dfree(d.ev_gflag);
  {
    auto& e = c->evb;
    dfree(e.fa); dfree(e.fb); dfree(e.oa); dfree(e.ob); dfree(e.rec); dfree(e.rec_out); dfree(e.ab_out);
    e.cap = 0;
  } dfree(d.toi_cnt); dfree(d.toi_off); dfree(d.toi_best); dfree(d.toi_pairs); dfree(d.toi_res); dfree(d.toi_live);
  dfree(d.sw); dfree(d.ckey); dfree(d.nbr_cnt); dfree(d.nbr_off); dfree(d.nbr_cur); dfree(d.nbr); dfree(c->cand_cnt); dfree(c->ts.tile); dfree(c->ts.done);
  adj_free(c);
  snap_free(c->bsnap);
}

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = gid();
  if (i < n) p[i] = 0u;
}
__global__ void k_zero_u8(uint8_t* p, uint32_t n) {
  uint32_t i = gid();
  if (i < n) p[i] = 0;
}

static void zero_u32(oavbd_ctx* c, uint32_t* p, uint32_t n) {
  if (n) k_zero_u32<<<blocks(n), TPB, 0, c->st>>>(p, n);
}

// Exclusive scan on the context stream (two launches, see avbd_scan.hpp).
static void scan(oavbd_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n) {
  if ((size_t)n > (size_t)SCAN_MAX_TILES * SCAN_TILE) throw HipError{"scan longer than the scan capacity"};
  ScanJobs js{};
  js.in[0] = in;
  js.out[0] = out;
  scan_tiles(c->ts, js, 1, n, c->st);
}
// Two or three arrays of the same length in one launch pair.
// With n_dev, the length is the device word n_dev[0] (zero when it exceeds
// n, for which the launch is sized).
static void scan2(oavbd_ctx* c, const uint32_t* in0, uint32_t* out0, const uint32_t* in1, uint32_t* out1,
                  uint32_t n, uint32_t* pick0 = nullptr, uint32_t post_seq = 0, const uint32_t* n_dev = nullptr) {
  if ((size_t)n > (size_t)SCAN_MAX_TILES * SCAN_TILE) throw HipError{"scan longer than the scan capacity"};
  ScanJobs js{{in0, in1, nullptr}, {out0, out1, nullptr}, pick0};
  js.n_dev = n_dev;
  if (post_seq && n) {  // both totals to the pinned buffer from CTR_COUNT, then the flag
    js.post = c->counters_hd;
    js.post_ticket = c->ts.done + SCAN_JOBS;
    js.post_word = CTR_COUNT;
    js.post_flag = POST_FLAG;
    js.post_seq = post_seq;
  }
  scan_tiles(c->ts, js, 2, n, c->st);
}
static void scan3(oavbd_ctx* c, const uint32_t* in0, uint32_t* out0, const uint32_t* in1, uint32_t* out1,
                  const uint32_t* in2, uint32_t* out2, uint32_t n) {
  if ((size_t)n > (size_t)SCAN_MAX_TILES * SCAN_TILE) throw HipError{"scan longer than the scan capacity"};
  ScanJobs js{{in0, in1, in2}, {out0, out1, out2}};
  scan_tiles(c->ts, js, 3, n, c->st);
}


// Readbacks the step waits for mid-stream: a one-workgroup kernel writes the
// words straight into the pinned buffer and then raises a sequence flag
// there, which the host polls (a stream synchronisation would add the copy
// launch and the runtime's wake-up to every wait).
struct PostJob {
  const uint32_t* src[2] = {nullptr, nullptr};  // single words ...
  uint32_t dst[2] = {0u, 0u};                   // ... to these buffer words
  const uint32_t* run = nullptr;                // and a run of words
  uint32_t drun = 0u, nrun = 0u;
};
__global__ void __launch_bounds__(256) k_post(PostJob j, uint32_t* buf, uint32_t seq) {
  const uint32_t t = threadIdx.x;
  if (t == 0 && j.src[0]) buf[j.dst[0]] = *j.src[0];
  if (t == 1 && j.src[1]) buf[j.dst[1]] = *j.src[1];
  for (uint32_t i = t; i < j.nrun; i += blockDim.x) buf[j.drun + i] = j.run[i];
  __threadfence_system();
  __syncthreads();
  if (t == 0) __hip_atomic_store(buf + POST_FLAG, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

static uint32_t post(oavbd_ctx* c, const PostJob& j) {
  const uint32_t seq = ++c->post_seq;
  k_post<<<1, 256, 0, c->st>>>(j, c->counters_hd, seq);
  return seq;
}

// Waits for post `seq`; an idle stream without the flag, or a failed one,
// is an error rather than a hang.
static void post_wait(oavbd_ctx* c, uint32_t seq) {
  uint32_t spins = 0;
  while (__atomic_load_n(c->counters_h + POST_FLAG, __ATOMIC_ACQUIRE) != seq) {
    if ((++spins & 1023u) != 0u) continue;
    const hipError_t e = hipStreamQuery(c->st);
    if (e == hipErrorNotReady) continue;
    HIPC(e);
    if (__atomic_load_n(c->counters_h + POST_FLAG, __ATOMIC_ACQUIRE) != seq)
      throw HipError{"readback flag not raised by an idle stream"};
  }
}

static uint32_t read_u32(oavbd_ctx* c, const uint32_t* dptr) {
  PostJob j;
  j.src[0] = dptr;
  j.dst[0] = CTR_COUNT;
  post_wait(c, post(c, j));
  return c->counters_h[CTR_COUNT];
}

// Scan whose total the host waits for: the scan's last workgroup posts it.
static uint32_t scan_total(oavbd_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n) {
  if (n == 0) {
    scan(c, in, out, n);
    return read_u32(c, out);
  }
  if ((size_t)n > (size_t)SCAN_MAX_TILES * SCAN_TILE) throw HipError{"scan longer than the scan capacity"};
  ScanJobs js{};
  js.in[0] = in;
  js.out[0] = out;
  js.post = c->counters_hd;
  js.post_ticket = c->ts.done + SCAN_JOBS;
  js.post_word = CTR_COUNT;
  js.post_flag = POST_FLAG;
  js.post_seq = ++c->post_seq;
  scan_tiles(c->ts, js, 1, n, c->st);
  post_wait(c, js.post_seq);
  return c->counters_h[CTR_COUNT];
}

// ---------------------------------------------------------------------------
// static joint structures: partner CSR (no-collide), incidence CSR (rows,
// colouring), dynamic-joint components (sleep/wake propagation)
// ---------------------------------------------------------------------------
// The bodies whose rows joint j stamps, each once: a gear's are the
// distinct bodies of the two joints it couples (joint_members,
// sleep.rs:9-16), any other joint's its own two.
static uint32_t joint_rows_members(const oavbd_ctx* c, uint32_t j, uint32_t m[4]) {
  const HostJoint& h = c->joints[j];
  if (h.kind != JK_GEAR) {
    m[0] = h.a;
    m[1] = h.b;
    return 2;
  }
  const HostJoint& x = c->joints[h.gb[0]];
  const HostJoint& y = c->joints[h.gb[1]];
  const uint32_t all[4] = {x.a, x.b, y.a, y.b};
  uint32_t nm = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    bool seen = false;
    for (uint32_t l = 0; l < nm; ++l) seen = seen || m[l] == all[k];
    if (!seen) m[nm++] = all[k];
  }
  return nm;
}

static void upload_joints(oavbd_ctx* c) {
  Dev& d = c->d;
  uint32_t n = d.n, nj = (uint32_t)c->joints.size();
  uint32_t ngears = 0;
  // joints already on the device keep their duals (a new joint does not
  // reset the others' warm starts: add_joint only pushes, mod.rs:1308-1510)
  const uint32_t kept = std::min(d.njoints, nj);
  std::vector<float4> kept_lam(kept), kept_pen(kept), kept_lama(kept), kept_pena(kept), kept_sacc(2 * (size_t)kept);
  if (kept) {
    HIPC(hipStreamSynchronize(c->st));
    HIPC(hipMemcpy(kept_lam.data(), d.J.lam, kept * sizeof(float4), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(kept_pen.data(), d.J.pen, kept * sizeof(float4), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(kept_lama.data(), d.J.lama, kept * sizeof(float4), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(kept_pena.data(), d.J.pena, kept * sizeof(float4), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(kept_sacc.data(), d.J.sacc, 2 * kept * sizeof(float4), hipMemcpyDeviceToHost));
  }
  Joints& J = d.J;
  dfree(J.ab); dfree(J.kind); dfree(J.la); dfree(J.lb); dfree(J.la0); dfree(J.ax); dfree(J.dref); dfree(J.qref);
  dfree(J.sus); dfree(J.six); dfree(J.sacc);
  dfree(J.lam); dfree(J.pen); dfree(J.lama); dfree(J.pena); dfree(J.c0v); dfree(J.a0);
  dfree(d.b.jp); dfree(d.b.ji); dfree(d.ropes); dfree(d.motors);
  dalloc(&J.ab, nj); dalloc(&J.kind, nj); dalloc(&J.la, nj); dalloc(&J.lb, nj); dalloc(&J.la0, nj);
  dalloc(&J.ax, 2 * (size_t)nj); dalloc(&J.sus, nj); dalloc(&J.six, 4 * (size_t)nj); dalloc(&J.sacc, 2 * (size_t)nj); dalloc(&J.dref, nj); dalloc(&J.qref, nj); dalloc(&J.lam, nj); dalloc(&J.pen, nj);
  dalloc(&J.lama, nj); dalloc(&J.pena, nj); dalloc(&J.c0v, nj); dalloc(&J.a0, 2 * (size_t)nj);
  std::vector<uint2> ab(nj);
  std::vector<int> kind(nj);
  std::vector<uint32_t> ropes, motors;
  std::vector<float4> la(nj), lb(nj), la0(nj), ax(2 * (size_t)nj), dref(nj), qref(nj), lam(nj), pen(nj), lama(nj),
      pena(nj), sus(nj), six(4 * (size_t)nj), sacc(2 * (size_t)nj);
  // no-collide partners: every kind but Fixed (rows.rs:437-450); the
  // incidence lists (rows, colouring) hold every joint
  std::vector<uint32_t> jp_cnt(n + 1, 0), ji_cnt(n + 1, 0);
  for (uint32_t j = 0; j < nj; ++j) {
    const HostJoint& h = c->joints[j];
    ab[j] = make_uint2(h.a, h.b);
    kind[j] = h.kind;
    la[j] = make_float4(h.la[0], h.la[1], h.la[2], 0.0f);
    lb[j] = make_float4(h.lb[0], h.lb[1], h.lb[2], 0.0f);
    // the reference stamps A's lever as la + Vec3::ZERO, or la + dref for a
    // Fixed joint (assembly.rs:297-304)
    const bool fixed = h.kind == JK_FIXED;
    la0[j] = make_float4(h.la[0] + (fixed ? h.dref[0] : 0.0f), h.la[1] + (fixed ? h.dref[1] : 0.0f),
                         h.la[2] + (fixed ? h.dref[2] : 0.0f), 0.0f);
    const bool motor = h.kind == JK_MOTOR;  // targets and budgets
    ax[2 * j] = make_float4(h.axa[0], h.axa[1], h.axa[2], motor ? h.prm[0] : 0.0f);
    ax[2 * j + 1] = make_float4(h.axb[0], h.axb[1], h.axb[2], motor ? h.prm[1] : 0.0f);
    const bool hinge = h.kind == JK_REVOLUTE || h.kind == JK_PRISMATIC || h.kind == JK_WHEEL;
    sus[j] = make_float4(h.sus[0], h.sus[1], h.sus[2], 0.0f);
    uint32_t packed = 0;
    for (int k = 0; k < 6; ++k) packed |= (uint32_t)h.six[k] << (2 * k);
    six[4 * j] = make_float4(h.six_lo[0], h.six_lo[1], h.six_lo[2], (float)packed);
    six[4 * j + 1] = make_float4(h.six_hi[0], h.six_hi[1], h.six_hi[2], 0.0f);
    six[4 * j + 2] = make_float4(h.six_lo[3], h.six_lo[4], h.six_lo[5], 0.0f);
    six[4 * j + 3] = make_float4(h.six_hi[3], h.six_hi[4], h.six_hi[5], 0.0f);
    sacc[2 * j] = j < kept ? kept_sacc[2 * j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    sacc[2 * j + 1] = j < kept ? kept_sacc[2 * j + 1] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    dref[j] = (is_tether(h.kind) || hinge) ? make_float4(h.prm[0], h.prm[1], h.prm[2], h.prm[3])
                                           : make_float4(h.dref[0], h.dref[1], h.dref[2], 0.0f);
    if (h.kind == JK_GEAR) {  // ratio, constant, the coupled joints' indices as bit patterns
      float g0, g1;
      std::memcpy(&g0, &h.gb[0], sizeof g0);
      std::memcpy(&g1, &h.gb[1], sizeof g1);
      dref[j] = make_float4(h.prm[0], h.prm[1], g0, g1);
      ++ngears;
    }
    if (h.kind == JK_ROPE) ropes.push_back(j);
    if ((hinge && ((uint32_t)h.prm[3] & HINGE_MOTOR)) || h.kind == JK_MOTOR) motors.push_back(j);
    qref[j] = hinge ? make_float4(h.mot[0], h.mot[1], 0.0f, 0.0f) : make_float4(h.qref[0], h.qref[1], h.qref[2], h.qref[3]);
    lam[j] = j < kept ? kept_lam[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    pen[j] = j < kept ? kept_pen[j] : make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    lama[j] = j < kept ? kept_lama[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    pena[j] = j < kept ? kept_pena[j] : make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    if (!fixed && h.kind != JK_SIXDOF && h.kind != JK_GEAR) {  // weld-like and geared: their contacts stay
      jp_cnt[h.a]++;
      jp_cnt[h.b]++;
    }
    uint32_t m[4];
    const uint32_t nm = joint_rows_members(c, j, m);
    for (uint32_t k = 0; k < nm; ++k) ji_cnt[m[k]]++;
  }
  std::vector<uint32_t> jp_off(n + 1, 0), ji_off(n + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    jp_off[i + 1] = jp_off[i] + jp_cnt[i];
    ji_off[i + 1] = ji_off[i] + ji_cnt[i];
  }
  std::vector<uint32_t> jp(jp_off[n]), ji(ji_off[n]), pcur(n, 0), icur(n, 0);
  for (uint32_t j = 0; j < nj; ++j) {  // joint order per body = ascending joint index
    const HostJoint& h = c->joints[j];
    if (h.kind != JK_FIXED && h.kind != JK_SIXDOF && h.kind != JK_GEAR) {
      jp[jp_off[h.a] + pcur[h.a]++] = h.b;
      jp[jp_off[h.b] + pcur[h.b]++] = h.a;
    }
    uint32_t m[4];
    const uint32_t nm = joint_rows_members(c, j, m);  // a gear's sides carry bit 0 clear
    for (uint32_t k = 0; k < nm; ++k) ji[ji_off[m[k]] + icur[m[k]]++] = (j << 1) | (h.kind == JK_GEAR ? 0u : k);
  }
  // components of the graph of joints between two dynamic bodies
  std::vector<uint32_t> parent(n);
  std::iota(parent.begin(), parent.end(), 0u);
  auto find = [&](uint32_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (uint32_t j = 0; j < nj; ++j) {
    const HostJoint& h = c->joints[j];
    uint32_t m[4];
    const uint32_t nm = joint_rows_members(c, j, m);  // a gear joins all its dynamic members
    for (uint32_t k = 1; k < nm; ++k)
      for (uint32_t l = 0; l < k; ++l)
        if (c->type_h[m[k]] == BT_DYNAMIC && c->type_h[m[l]] == BT_DYNAMIC) {
          uint32_t x = find(m[k]), y = find(m[l]);
          if (x != y) parent[std::max(x, y)] = std::min(x, y);
        }
  }
  std::vector<uint32_t> comp(n);
  for (uint32_t i = 0; i < n; ++i) comp[i] = find(i);
  dalloc(&d.b.jp, jp.size());
  dalloc(&d.b.ji, ji.size());
  auto up = [](auto* dst, const auto& v) {
    if (!v.empty()) HIPC(hipMemcpy(dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
  };
  up(J.ab, ab); up(J.kind, kind); up(J.la, la); up(J.lb, lb); up(J.la0, la0); up(J.ax, ax); up(J.dref, dref);
  up(J.sus, sus); up(J.six, six); up(J.sacc, sacc);
  up(J.qref, qref); up(J.lam, lam); up(J.pen, pen); up(J.lama, lama); up(J.pena, pena);
  up(d.b.jp_off, jp_off); up(d.b.ji_off, ji_off); up(d.b.jp, jp); up(d.b.ji, ji); up(d.b.comp, comp);
  if (!ropes.empty()) dalloc(&d.ropes, ropes.size());
  up(d.ropes, ropes);
  d.nropes = (uint32_t)ropes.size();
  if (!motors.empty()) dalloc(&d.motors, motors.size());
  up(d.motors, motors);
  d.nmotors = (uint32_t)motors.size();
  c->ngears = ngears;
  HIPC(hipDeviceSynchronize());  // null-stream uploads land before the context stream reads them
  d.njoints = nj;
  c->joints_dirty = false;
}

// ---------------------------------------------------------------------------
// checkpoints and recorded steps (differentiable stepping)
// ---------------------------------------------------------------------------
template <class T>
static void dcopy(oavbd_ctx* c, T* dst, const T* src, size_t n) {
  if (n) HIPC(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, c->st));
}

static void snap_free(Snap& s) {
  dfree(s.pos); dfree(s.rot); dfree(s.vel); dfree(s.angvel); dfree(s.torque); dfree(s.inertia); dfree(s.prev_pos);
  dfree(s.prev_rot); dfree(s.asleep); dfree(s.flag_a); dfree(s.timer); dfree(s.ab); dfree(s.prec); dfree(s.ppts);
  dfree(s.poff); dfree(s.qrec); dfree(s.qroll); dfree(s.qpair); dfree(s.qsrc); dfree(s.jlam); dfree(s.jpen); dfree(s.jlama); dfree(s.jpena);
  s.valid = false;
  s.cap_pairs = s.cap_pts = 0;
  s.alloc_n = s.alloc_nj = 0;
}

static void snap_save(oavbd_ctx* c, Snap& s) {
  Dev& d = c->d;
  const uint32_t n = d.n, nj = d.njoints;
  if (s.pos && (s.alloc_n < n || s.alloc_nj < nj)) snap_free(s);  // the body or joint table grew
  if (!s.pos) {
    s.alloc_n = n;
    s.alloc_nj = nj;
    dalloc(&s.pos, n); dalloc(&s.rot, n); dalloc(&s.vel, n); dalloc(&s.angvel, n); dalloc(&s.torque, n);
    dalloc(&s.inertia, n); dalloc(&s.prev_pos, n); dalloc(&s.prev_rot, n); dalloc(&s.asleep, n); dalloc(&s.flag_a, n);
    dalloc(&s.timer, n); dalloc(&s.poff, n + 1); dalloc(&s.jlam, nj); dalloc(&s.jpen, nj);
    dalloc(&s.jlama, nj); dalloc(&s.jpena, nj);
  }
  if (d.npairs > s.cap_pairs) {
    dfree(s.ab); dfree(s.prec); dfree(s.ppts);
    s.cap_pairs = grow(d.npairs);
    dalloc(&s.ab, s.cap_pairs); dalloc(&s.prec, 4 * s.cap_pairs); dalloc(&s.ppts, s.cap_pairs);
  }
  if (d.npts > s.cap_pts) {
    dfree(s.qrec); dfree(s.qroll); dfree(s.qpair); dfree(s.qsrc);
    s.cap_pts = grow(d.npts);
    dalloc(&s.qrec, 4 * s.cap_pts); dalloc(&s.qroll, s.cap_pts); dalloc(&s.qpair, s.cap_pts); dalloc(&s.qsrc, s.cap_pts);
  }
  Bodies& b = d.b;
  dcopy(c, s.pos, b.pos, n); dcopy(c, s.rot, b.rot, n); dcopy(c, s.vel, b.vel, n); dcopy(c, s.angvel, b.angvel, n);
  dcopy(c, s.torque, b.torque, n); dcopy(c, s.inertia, b.inertia, n); dcopy(c, s.prev_pos, b.prev_pos, n);
  dcopy(c, s.prev_rot, b.prev_rot, n); dcopy(c, s.asleep, b.asleep, n); dcopy(c, s.flag_a, b.flag_a, n);
  dcopy(c, s.timer, b.timer, n); dcopy(c, s.poff, d.poff, n + 1);
  dcopy(c, s.ab, d.P.ab, d.npairs); dcopy(c, s.prec, d.P.rec, 4 * (size_t)d.npairs); dcopy(c, s.ppts, d.P.pts, d.npairs);
  dcopy(c, s.qrec, d.Q.rec, 4 * (size_t)d.npts); dcopy(c, s.qroll, d.Q.roll, d.npts); dcopy(c, s.qpair, d.Q.pair, d.npts);
  dcopy(c, s.qsrc, d.Q.src, d.npts);
  dcopy(c, s.jlam, d.J.lam, nj); dcopy(c, s.jpen, d.J.pen, nj);
  dcopy(c, s.jlama, d.J.lama, nj); dcopy(c, s.jpena, d.J.pena, nj);
  s.npairs = d.npairs;
  s.npts = d.npts;
  s.time_debt = c->time_debt;
  s.sleeping = c->sleeping;
  s.valid = true;
}

static void snap_load(oavbd_ctx* c, const Snap& s) {
  Dev& d = c->d;
  const uint32_t n = d.n, nj = d.njoints;
  ensure_pairs(c, s.npairs);
  ensure_points(c, s.npts);
  ensure_incidence(c, s.npairs, s.npts);
  Bodies& b = d.b;
  dcopy(c, b.pos, s.pos, n); dcopy(c, b.rot, s.rot, n); dcopy(c, b.vel, s.vel, n); dcopy(c, b.angvel, s.angvel, n);
  dcopy(c, b.torque, s.torque, n); dcopy(c, b.inertia, s.inertia, n); dcopy(c, b.prev_pos, s.prev_pos, n);
  dcopy(c, b.prev_rot, s.prev_rot, n); dcopy(c, b.asleep, s.asleep, n); dcopy(c, b.flag_a, s.flag_a, n);
  dcopy(c, b.timer, s.timer, n); dcopy(c, d.poff, s.poff, n + 1);
  dcopy(c, d.P.ab, s.ab, s.npairs); dcopy(c, d.P.rec, s.prec, 4 * (size_t)s.npairs); dcopy(c, d.P.pts, s.ppts, s.npairs);
  dcopy(c, d.Q.rec, s.qrec, 4 * (size_t)s.npts); dcopy(c, d.Q.roll, s.qroll, s.npts); dcopy(c, d.Q.pair, s.qpair, s.npts);
  dcopy(c, d.Q.src, s.qsrc, s.npts);
  dcopy(c, d.J.lam, s.jlam, nj); dcopy(c, d.J.pen, s.jpen, nj);
  dcopy(c, d.J.lama, s.jlama, nj); dcopy(c, d.J.pena, s.jpena, nj);
  d.npairs = s.npairs;
  d.npts = s.npts;
  c->time_debt = s.time_debt;
  c->sleeping = s.sleeping;
  c->may_sleep = true;  // conservative: the wake passes run
}

static void tape_free(TapeHost& h) {
  Tape& t = h.t;
  dfree(t.sb); dfree(t.qrec); dfree(t.ab); dfree(t.prec); dfree(t.ppts); dfree(t.qpair); dfree(t.qsrc); dfree(t.inc);
  dfree(t.clist); dfree(t.colour); dfree(t.rot0); dfree(t.vel); dfree(t.angvel); dfree(t.inertial);
  dfree(t.inertial_rot); dfree(t.inertia); dfree(t.slept);
  h.cap_n = h.cap_pairs = h.cap_pts = 0;
}

// Start recording the current step (called once its set-up is queued and
// the colour counts are on the host): the active set, the body data the
// reverse pass reads and the solver records before the sweep.
__global__ void k_tape_slept(Tape t, Dev d) {
  const uint32_t h = gid();
  if (h >= t.n) return;
  t.slept[h] = (t.colour[h] >= 0 && d.b.asleep[h]) ? 1 : 0;
}

static void tape_begin(oavbd_ctx* c) {
  TapeHost& h = *c->rec;
  Dev& d = c->d;
  Tape& t = h.t;
  const uint32_t n = d.n, iters = c->prm.iters;
  if (h.cap_n < n || h.iters != iters) {
    dfree(t.sb); dfree(t.clist); dfree(t.colour); dfree(t.rot0); dfree(t.vel); dfree(t.angvel); dfree(t.inertial);
    dfree(t.inertial_rot); dfree(t.inertia); dfree(t.slept);
    dalloc(&t.sb, (size_t)(iters + 1) * n * 4); dalloc(&t.clist, n); dalloc(&t.colour, n); dalloc(&t.rot0, n);
    dalloc(&t.vel, n); dalloc(&t.angvel, n); dalloc(&t.inertial, n); dalloc(&t.inertial_rot, n);
    dalloc(&t.inertia, n); dalloc(&t.slept, n);
    h.cap_n = n;
    h.iters = iters;
    dfree(t.qrec);
    h.cap_pts = 0;
  }
  if (d.npairs > h.cap_pairs) {
    dfree(t.ab); dfree(t.prec); dfree(t.ppts);
    h.cap_pairs = grow(d.npairs);
    dalloc(&t.ab, h.cap_pairs); dalloc(&t.prec, 4 * h.cap_pairs); dalloc(&t.ppts, h.cap_pairs);
  }
  if (d.npts > h.cap_pts) {
    dfree(t.qrec); dfree(t.qpair); dfree(t.qsrc); dfree(t.inc);
    h.cap_pts = grow(d.npts);
    dalloc(&t.qrec, (size_t)iters * h.cap_pts * 4); dalloc(&t.qpair, h.cap_pts); dalloc(&t.qsrc, h.cap_pts);
    dalloc(&t.inc, 2 * h.cap_pts);
  }
  t.qcap = h.cap_pts;
  t.n = n;
  t.npairs = d.npairs;
  t.npts = d.npts;
  t.iters = iters;
  dcopy(c, t.ab, d.P.ab, d.npairs); dcopy(c, t.prec, d.P.rec, 4 * (size_t)d.npairs); dcopy(c, t.ppts, d.P.pts, d.npairs);
  dcopy(c, t.qpair, d.Q.pair, d.npts); dcopy(c, t.qsrc, d.Q.src, d.npts);
  dcopy(c, t.inc, d.inc, 2 * (size_t)d.npts);
  dcopy(c, t.clist, d.colour_list, n); dcopy(c, t.colour, d.b.colour, n);
  dcopy(c, t.rot0, d.b.rot0, n); dcopy(c, t.vel, d.b.vel, n); dcopy(c, t.angvel, d.b.angvel, n);
  dcopy(c, t.inertial, d.b.inertial, n); dcopy(c, t.inertial_rot, d.b.inertial_rot, n);
  dcopy(c, t.inertia, d.b.inertia, n);
  dcopy(c, t.sb, d.sb, 4 * (size_t)n);
  h.col_off = c->colour_off_h;
  h.col_cnt = c->colour_cnt_h;
}
// At the start of iteration it: the point records (duals) it reads.
static void tape_iter_start(oavbd_ctx* c, uint32_t it) {
  Tape& t = c->rec->t;
  dcopy(c, t.qrec + (size_t)it * t.qcap * 4, c->d.Q.rec, 4 * (size_t)c->d.npts);
}
// After iteration it's sweep: the solver records.
static void tape_iter_end(oavbd_ctx* c, uint32_t it) {
  Tape& t = c->rec->t;
  dcopy(c, t.sb + (size_t)(it + 1) * t.n * 4, c->d.sb, 4 * (size_t)t.n);
}
static void tape_end(oavbd_ctx* c) {
  Tape& t = c->rec->t;
  k_tape_slept<<<blocks(t.n), TPB, 0, c->st>>>(t, c->d);
}

// ---------------------------------------------------------------------------
// step phases
// ---------------------------------------------------------------------------
static void propagate(oavbd_ctx* c, uint8_t* flags, uint8_t value) {
  uint32_t n = c->d.n;
  k_prop_clear<<<blocks(n), TPB, 0, c->st>>>(c->d);
  k_prop_mark<<<blocks(n), TPB, 0, c->st>>>(c->d, flags, value);
  k_prop_apply<<<blocks(n), TPB, 0, c->st>>>(c->d, flags, value);
}

static void wake_joint_motion(oavbd_ctx* c) {
  uint32_t n = c->d.n;
  k_awake_init<<<blocks(n), TPB, 0, c->st>>>(c->d);
  if (c->d.njoints) {
    k_awake_joints<<<blocks(c->d.njoints), TPB, 0, c->st>>>(c->d);
    propagate(c, c->d.b.flag_b, 1);
  }
  k_awake_apply<<<blocks(n), TPB, 0, c->st>>>(c->d);
}

// `after_toi`: the rebuild after the TOI clamps, skipped on the device
// when there were no casts.
static void build_grid(oavbd_ctx* c, bool after_toi = false) {
  Dev& d = c->d;
  const uint32_t g = after_toi ? 1u : 0u;
  k_grid_clear<<<blocks(d.g.nbuckets), TPB, 0, c->st>>>(d, g);
  k_grid_count<<<blocks(d.n), TPB, 0, c->st>>>(d, g);
  ScanJobs js{};
  js.in[0] = d.g.cnt;
  js.out[0] = d.g.start;
  if (after_toi) {
    js.gate = d.toi_off + d.n;
    js.gate_max = d.cap_toi;
  }
  scan_tiles(c->ts, js, 1, d.g.nbuckets, c->st);
  k_grid_fill<<<blocks(d.n), TPB, 0, c->st>>>(d, g);
}

// The TOI clamp (mod.rs:882-978) for `m` casts: the exact count when the
// host waited for it, else the buffer capacity (the device's count rules;
// launches are sized for the capacity, the full casts from the last count
// seen).
static void toi_phase(oavbd_ctx* c, uint32_t m, uint32_t m_pair) {
  Dev& d = c->d;
  const uint32_t n = d.n;
  k_toi_fill<<<blocks(n), TPB, 0, c->st>>>(d);
  // the gate settles most casts; the survivors' count stays on the device,
  // so the full casts run a grid-stride loop over a quarter-size grid
  k_toi_gate<TOI_GATE_LANES><<<blocks((size_t)m * TOI_GATE_LANES), TPB, 0, c->st>>>(d);
  k_toi_pair<TOI_LANES><<<std::max<uint32_t>(blocks((size_t)(m_pair + 3) / 4 * TOI_LANES), 1u), TPB, 0, c->st>>>(d);
  k_toi_win<<<blocks(m), TPB, 0, c->st>>>(d);
  k_toi_apply<<<blocks(n), TPB, 0, c->st>>>(d);
  k_apply_wake<<<blocks(n), TPB, 0, c->st>>>(d, 1u);
  build_grid(c, true);
}

static void broadphase(oavbd_ctx* c, bool async) {
  Dev& d = c->d;
  uint32_t n = d.n;
  k_bp_count<<<blocks((size_t)n * BP_LANES), TPB, 0, c->st>>>(d, c->cand_cnt);
  if (d.g.nlarge) {
    k_bp_large_flag<<<blocks(d.g.lflat), TPB, 0, c->st>>>(d);
    scan(c, d.lflag, d.lrank, d.g.lflat);  // one scan over all large bodies' flag rows
    k_bp_large_count<<<blocks(d.g.nlarge), TPB, 0, c->st>>>(d, c->cand_cnt);
  }
  if (async) {
    scan(c, c->cand_cnt, d.cand_off, n);
  } else {
    c->ncand = scan_total(c, c->cand_cnt, d.cand_off, n);
    ensure_cand(c, c->ncand);
  }
  k_bp_fill<<<blocks(n), TPB, 0, c->st>>>(d);
  if (d.g.nlarge) k_bp_large_fill<<<blocks(d.g.lflat), TPB, 0, c->st>>>(d);
}

static void narrow_and_merge(oavbd_ctx* c, bool async) {
  Dev& d = c->d;
  std::swap(d.P, d.Pold);
  std::swap(d.poff, d.poff_old);
  std::swap(d.Q, d.Qold);
  std::swap(c->cap_pairs, c->cap_pairs_old);
  std::swap(c->cap_pts, c->cap_pts_old);
  d.cap_pairs = (uint32_t)c->cap_pairs;
  d.cap_pts = (uint32_t)c->cap_pts;
  d.npairs_old = d.npairs;
  d.npts_old = d.npts;
  const uint32_t nc = async ? (uint32_t)c->cap_cand : c->ncand;  // launch size
  if (nc) {
    k_narrow_gate<<<blocks(nc), TPB, 0, c->st>>>(d);
    k_narrow_dist<NARROW_LANES><<<blocks((size_t)nc * NARROW_LANES), TPB, 0, c->st>>>(d);
    k_narrow<<<blocks(nc), TPB, 0, c->st>>>(d);
  }
  // the accepted candidates' flags scan to pair indices; the pair -> candidate
  // map they also give lets the merge run over pairs only
  if (async) {
    scan2(c, d.C.acc, d.acc_off, d.C.npts, d.pt_off, nc, d.pair_cand, 0, d.cand_off + d.n);
    k_pair_offsets<<<blocks(d.n + 1), TPB, 0, c->st>>>(d);
    k_merge<<<blocks((size_t)c->cap_pairs * K_MAX_POINTS), TPB, 0, c->st>>>(d);
    return;
  }
  const uint32_t seq = nc ? ++c->post_seq : 0u;
  scan2(c, d.C.acc, d.acc_off, d.C.npts, d.pt_off, nc, d.pair_cand, seq);
  if (nc) {
    post_wait(c, seq);  // the scan posted both totals
  } else {
    PostJob j;
    j.src[0] = d.acc_off + nc;
    j.src[1] = d.pt_off + nc;
    j.dst[0] = CTR_COUNT;
    j.dst[1] = CTR_COUNT + 1;
    post_wait(c, post(c, j));
  }
  d.npairs = c->counters_h[CTR_COUNT];
  d.npts = c->counters_h[CTR_COUNT + 1];
  c->diff_max_pairs = std::max(c->diff_max_pairs, d.npairs);
  c->diff_max_pts = std::max(c->diff_max_pts, d.npts);
  ensure_pairs(c, d.npairs);
  ensure_points(c, d.npts);
  ensure_incidence(c, d.npairs, d.npts);
  k_pair_offsets<<<blocks(d.n + 1), TPB, 0, c->st>>>(d);  // sized tables: the totals fit
  if (d.npairs) k_merge<<<blocks((size_t)d.npairs * K_MAX_POINTS), TPB, 0, c->st>>>(d);
}

// `gp`: the launch size of the per-pair kernels (the pair count, or the
// table's capacity when the host has not read the count).
static void incidence_and_colour(oavbd_ctx* c, uint32_t gp, uint32_t slots) {
  Dev& d = c->d;
  uint32_t n = d.n;
  k_colour_init<<<blocks(std::max<uint32_t>(n + 1, CSTAT_HEAD + 65536)), TPB, 0, c->st>>>(d);
  if (gp) k_inc_count<<<blocks(gp), TPB, 0, c->st>>>(d);
  if (d.njoints) k_nbr_joint_count<<<blocks(d.njoints), TPB, 0, c->st>>>(d);
  scan3(c, d.ptcnt, d.inc_off, d.bcnt, d.boff, d.nbr_cnt, d.nbr_off, n);
  if (gp) k_inc_fill<<<blocks(gp), TPB, 0, c->st>>>(d);
  if (d.njoints) k_nbr_joint_fill<<<blocks(d.njoints), TPB, 0, c->st>>>(d);
  k_inc_build<<<blocks((size_t)n * INC_LANES), TPB, 0, c->st>>>(d);
  // Jones-Plassmann colouring in passes.  The first `grid` passes run over
  // all bodies, one launch each; the bodies still uncoloured after them are
  // listed and finished by one single-workgroup launch.  `grid` is the pass
  // after which at most COLOUR_TAIL_FRONTIER bodies were left in the previous
  // step (the counts change slowly).  One check then copies the per-pass
  // counts, the pass and colour counts and the first colour sizes; the
  // colour offsets and lists are made on the device meanwhile, and the host
  // waits for the check (colour_wait) only once the rest of the set-up is
  // queued behind it.
  // The first pass runs over all bodies; the bodies it leaves are listed,
  // and the next passes run over the list of the previous pass's leftovers
  // (grid sized from the previous step's leftovers after the first pass;
  // the passes loop over whatever the list holds).
  const int grid = c->grid_rounds;
  k_colour_pass<<<blocks((size_t)n * CL, COLOUR_PASS_TPB), COLOUR_PASS_TPB, 0, c->st>>>(d, d.remain);
  k_colour_collect<<<blocks(n), TPB, 0, c->st>>>(d, d.front + (size_t)1 * n, d.fcnt + 1);
  const size_t lgrid = std::max<size_t>(blocks((size_t)std::min(c->colour_left1, n) * CL, COLOUR_PASS_TPB), 1);
  for (int round = 1; round < grid; ++round)
    k_colour_pass_list<<<lgrid, COLOUR_PASS_TPB, 0, c->st>>>(d, (uint32_t)round, d.remain + round);
  k_colour_tail<<<1, COLOUR_TAIL_TPB, 0, c->st>>>(d, (uint32_t)grid);
  k_colour_classes<<<blocks(n), TPB, 0, c->st>>>(d);
  {
    PostJob j;
    j.run = d.cstat;
    j.drun = HEAD_H;
    j.nrun = CSTAT_HEAD + CSTAT_EAGER;
    c->colour_seq = post(c, j);
  }
  k_colour_offsets<<<1, 64, 0, c->st>>>(d, slots);
  k_colour_fill<<<blocks(n), TPB, 0, c->st>>>(d);
}

// The colouring check (the posted head of the colouring status): pass and
// colour counts, colour sizes and offsets on the host (the device made the
// same offsets).
static void colour_check(oavbd_ctx* c) {
  Dev& d = c->d;
  const uint32_t* head = c->counters_h + HEAD_H;
  const int grid = c->grid_rounds;
  int rounds = -1, next_grid = -1;
  for (int r = 0; r < grid; ++r) {
    if (rounds < 0 && head[r] == 0) rounds = r + 1;
    if (next_grid < 0 && head[r] <= COLOUR_TAIL_FRONTIER) next_grid = r + 1;
  }
  if (rounds < 0) rounds = (int)head[CSTAT_ROUNDS];
  // no grid pass got down to few uncoloured bodies: start the tail later
  c->grid_rounds = std::max(1, std::min(next_grid >= 0 ? next_grid : grid + 2, COLOUR_RING - 16));
  c->stats.colour_rounds = (uint32_t)rounds;
  c->colour_left1 = std::max<uint32_t>(head[0], 1024u);
  uint32_t ncol = head[CSTAT_MAX];
  c->colour_cnt_h.assign(ncol, 0);
  c->colour_off_h.assign(ncol + 1, 0);
  if (ncol) {
    if (ncol <= CSTAT_EAGER)
      std::copy(head + CSTAT_HEAD, head + CSTAT_HEAD + ncol, c->colour_cnt_h.begin());
    else
      HIPC(hipMemcpy(c->colour_cnt_h.data(), d.colour_cnt, ncol * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < ncol; ++k) c->colour_off_h[k + 1] = c->colour_off_h[k] + c->colour_cnt_h[k];
  }
  c->stats.colours = ncol;
  c->stats.max_colour = 0;
  c->solved = 0;
  for (uint32_t k = 0; k < ncol; ++k) {
    c->stats.max_colour = std::max(c->stats.max_colour, c->colour_cnt_h[k]);
    c->solved += c->colour_cnt_h[k];
  }
}
static void colour_wait(oavbd_ctx* c) {
  post_wait(c, c->colour_seq);
  colour_check(c);
}

// A kernel launch on the context stream; with events (timing), the start
// and stop events ride on the dispatch itself (hipExtLaunchKernelGGL), so
// they bracket the kernel's execution and not its dispatch.
template <class K, class... A>
static void launch_k(oavbd_ctx* c, K kernel, size_t grid, int threads, hipEvent_t e0, hipEvent_t e1, A... args) {
  if (e0 || e1) {
    hipExtLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(threads), 0, c->st, e0, e1, 0, args...);
    HIPC(hipGetLastError());
  } else {
    kernel<<<grid, threads, 0, c->st>>>(args...);
  }
}

// The primal variants compile the joint rows in only for a context that
// holds joints (JT): the rows' code sets the register budget of a launch.
template <int W, int PF>
static void launch_pf_any(oavbd_ctx* c, const uint4* list, uint32_t cnt, hipEvent_t e0, hipEvent_t e1) {
  constexpr int NB = (WAVE / BODY_LANES) * W;
  const uint32_t g = blocks(cnt, NB);
  const bool jt = c->d.njoints != 0;
  if (c->rows6) {
    if (jt)
      launch_k(c, k_primal8<W, 6, PF, true>, g, WAVE * W, e0, e1, c->d, list, cnt);
    else
      launch_k(c, k_primal8<W, 6, PF, false>, g, WAVE * W, e0, e1, c->d, list, cnt);
  } else {
    if (jt)
      launch_k(c, k_primal8<W, 4, PF, true>, g, WAVE * W, e0, e1, c->d, list, cnt);
    else
      launch_k(c, k_primal8<W, 4, PF, false>, g, WAVE * W, e0, e1, c->d, list, cnt);
  }
}
template <int W>
static void launch_w(oavbd_ctx* c, const uint4* list, uint32_t cnt, hipEvent_t e0, hipEvent_t e1) {
  launch_pf_any<W, 0>(c, list, cnt, e0, e1);
}
template <int W>
static void launch_pf(oavbd_ctx* c, const uint4* list, uint32_t cnt, hipEvent_t e0, hipEvent_t e1) {
  launch_pf_any<W, 1>(c, list, cnt, e0, e1);
}
template <int W>
static void launch_lds(oavbd_ctx* c, const uint4* list, uint32_t cnt, hipEvent_t e0, hipEvent_t e1) {
  if (c->rows6)
    launch_k(c, k_primal_lds<W, 6>, blocks(cnt, W), WAVE * W, e0, e1, c->d, list, cnt);
  else
    launch_k(c, k_primal_lds<W, 4>, blocks(cnt, W), WAVE * W, e0, e1, c->d, list, cnt);
}

// Primal launch of one colour class.  The launch shapes (variants): sixteen
// lanes per body with 1, 2, 4 or 8 wavefronts per workgroup (k_primal8);
// one body per wavefront with its records staged in LDS, 1, 2 or 4
// wavefronts per workgroup (k_primal_lds); sixteen lanes per body with the
// next item's records in flight behind the current one, 1, 2 or 4
// wavefronts per workgroup (k_primal8 with PF = 1).  Every variant gives the same
// bits.  The variant comes from the autotuned table for the class's size
// (log2 bucket), else the default: below PRIMAL_TAIL bodies sixteen lanes
// with the next item's records prefetched, one wavefront per workgroup (a
// small class is bound by its slowest body's chain of item loads and
// arithmetic, which the prefetch overlaps, and is spread over as many
// compute units as possible); the configured wavefront count above.
// Returns the variant launched.
constexpr uint32_t PRIMAL_TAIL = 8192;
static int size_bucket(uint32_t cnt) {
  int b = 0;
  while ((cnt >> (b + 1)) != 0u && b + 1 < PRIMAL_BUCKETS) ++b;
  return b;
}
static int launch_primal(oavbd_ctx* c, const uint4* list, uint32_t cnt, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr) {
  int v = c->force_variant;
  if (v < 0) v = c->ptable[size_bucket(cnt)];
  if (v < 0) v = cnt < PRIMAL_TAIL ? PV_P1 : (c->waves == 1 ? PV_W1 : c->waves == 2 ? PV_W2 : c->waves == 8 ? PV_W8 : PV_W4);
  switch (v) {
    case PV_W1: launch_w<1>(c, list, cnt, e0, e1); break;
    case PV_W2: launch_w<2>(c, list, cnt, e0, e1); break;
    case PV_W4: launch_w<4>(c, list, cnt, e0, e1); break;
    case PV_W8: launch_w<8>(c, list, cnt, e0, e1); break;
    case PV_L1: launch_lds<1>(c, list, cnt, e0, e1); break;
    case PV_L2: launch_lds<2>(c, list, cnt, e0, e1); break;
    case PV_P1: launch_pf<1>(c, list, cnt, e0, e1); break;
    case PV_P2: launch_pf<2>(c, list, cnt, e0, e1); break;
    case PV_P4: launch_pf<4>(c, list, cnt, e0, e1); break;
    default: launch_lds<4>(c, list, cnt, e0, e1); v = PV_L4; break;
  }
  return v;
}

// Primal launch of colour class k whose range only the device holds
// (k_primal8_class): the variant and the grid follow `est`, the class's
// size at the last readback (the grid loops over whatever the class holds).
// The LDS-staged variants map to the prefetching ones of the same
// wavefront count.
template <int W, int PF>
static void launch_class_w(oavbd_ctx* c, uint32_t k, uint32_t est) {
  constexpr int NB = (WAVE / BODY_LANES) * W;
  const uint32_t grid = std::max<uint32_t>(blocks(est + est / 8u + 1u, NB), 1u);
  const bool jt = c->d.njoints != 0;
  if (c->rows6) {
    if (jt)
      k_primal8_class<W, 6, PF, true><<<grid, WAVE * W, 0, c->st>>>(c->d, k);
    else
      k_primal8_class<W, 6, PF, false><<<grid, WAVE * W, 0, c->st>>>(c->d, k);
  } else {
    if (jt)
      k_primal8_class<W, 4, PF, true><<<grid, WAVE * W, 0, c->st>>>(c->d, k);
    else
      k_primal8_class<W, 4, PF, false><<<grid, WAVE * W, 0, c->st>>>(c->d, k);
  }
}
static void launch_primal_class(oavbd_ctx* c, uint32_t k, uint32_t est) {
  int v = c->force_variant;
  if (v < 0) v = c->ptable[size_bucket(est)];
  if (v < 0) v = est < PRIMAL_TAIL ? PV_P1 : (c->waves == 1 ? PV_W1 : c->waves == 2 ? PV_W2 : c->waves == 8 ? PV_W8 : PV_W4);
  switch (v) {
    case PV_W1: launch_class_w<1, 0>(c, k, est); break;
    case PV_W2: launch_class_w<2, 0>(c, k, est); break;
    case PV_W4: launch_class_w<4, 0>(c, k, est); break;
    case PV_W8: launch_class_w<8, 0>(c, k, est); break;
    case PV_L1: case PV_P1: launch_class_w<1, 1>(c, k, est); break;
    case PV_L2: case PV_P2: launch_class_w<2, 1>(c, k, est); break;
    default: launch_class_w<4, 1>(c, k, est); break;
  }
}

// Wavefronts of k_primal_tail that fit on the device at once (the fused
// tail launch is sized within this, so all its wavefronts are resident).
static uint32_t tail_capacity(oavbd_ctx* c) {
  uint32_t& cap = c->tail_cap[c->rows6 ? 1 : 0];
  if (!cap) {
    int nb = 0, ncu = 0;
    if (c->rows6)
      HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_primal_tail<6>, WAVE * TAIL_W, 0));
    else
      HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_primal_tail<4>, WAVE * TAIL_W, 0));
    HIPC(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    if (nb <= 0 || ncu <= 0) throw HipError{"k_primal_tail does not fit on the device"};
    cap = (uint32_t)nb * (uint32_t)ncu * TAIL_W;
  }
  return cap;
}

// First colour class of the fused tail: the last classes holding at most
// tail_max bodies together, when there are at least two of them; ncol
// (no tail) otherwise, and while a variant is forced.
static uint32_t tail_first(const oavbd_ctx* c, uint32_t ncol) {
  if (!c->tail_max || c->force_variant >= 0 || ncol < 2) return ncol;
  uint32_t k = ncol;
  while (k > 0 && c->colour_off_h[ncol] - c->colour_off_h[k - 1] <= c->tail_max) --k;
  return ncol - k >= 2 ? k : ncol;
}

// One fused launch of classes ktail..ncol-1 (k_primal_tail): batches of
// four bodies of one class, one batch per wavefront at a time.
static void launch_tail(oavbd_ctx* c, uint32_t ktail, uint32_t ncol, hipEvent_t e0, hipEvent_t e1) {
  uint32_t nbatch = 0;
  for (uint32_t k = ktail; k < ncol; ++k) nbatch += (c->colour_cnt_h[k] + 3u) / 4u;
  if (!nbatch) return;
  uint32_t cap = tail_capacity(c);
  if (c->tail_waves) cap = std::min(cap, std::max<uint32_t>(c->tail_waves / TAIL_W, 1u) * TAIL_W);
  const uint32_t nblk = std::min<uint32_t>((nbatch + TAIL_W - 1) / TAIL_W, cap / TAIL_W);
  if (++c->tail_stamp == 0u) {  // the stamps wrapped: no done word may hold a live stamp
    zero_u32(c, c->d.b.done, c->d.n);
    c->tail_stamp = 1u;
  }
  if (c->rows6)
    launch_k(c, k_primal_tail<6>, nblk, WAVE * TAIL_W, e0, e1, c->d, ktail, ncol, nbatch, nblk * TAIL_W,
             c->tail_stamp);
  else
    launch_k(c, k_primal_tail<4>, nblk, WAVE * TAIL_W, e0, e1, c->d, ktail, ncol, nbatch, nblk * TAIL_W,
             c->tail_stamp);
}

static void record(oavbd_ctx* c, int i) {
  if (c->timing) HIPC(hipEventRecord(c->ev[i], c->st));
}

// The words a finished step reads back, gathered into the pinned buffer by
// one workgroup: counters, the fused tail's error word, overflow words,
// device counts and the TOI / candidate totals.
struct Readback {
  const uint32_t* src[6];
  uint32_t nwords[6];
  uint32_t dst[6];
};
__global__ void __launch_bounds__(64) k_readback(Readback r, uint32_t* buf) {
  for (int j = 0; j < 6; ++j)
    for (uint32_t i = threadIdx.x; i < r.nwords[j]; i += 64) buf[r.dst[j] + i] = r.src[j][i];
}

static void read_counters(oavbd_ctx* c) {
  Dev& d = c->d;
  Readback r{{d.counters, d.tail_err, d.ovf, d.dc, d.toi_off + d.n, d.cand_off + d.n},
             {CTR_COUNT, 1u, OVF_WORDS, DC_WORDS, 1u, 1u},
             {0u, TAILERR_H, OVF_H, DC_H, TOTALS_H, TOTALS_H + 1}};
  k_readback<<<1, 64, 0, c->st>>>(r, c->counters_hd);
  HIPC(hipStreamSynchronize(c->st));
  if (c->counters_h[TAILERR_H]) throw HipError{"a fused primal tail launch timed out waiting for a partner"};
}

// ---------------------------------------------------------------------------
// events: the device flags and compacts, the host keeps the sets of the
// last report and builds the queues (emit_events, mod.rs:1606-1669)
// ---------------------------------------------------------------------------
static void ensure_evbuf(oavbd_ctx* c, size_t need) {
  auto& e = c->evb;
  if (need <= e.cap && e.fa) return;
  dfree(e.fa); dfree(e.fb); dfree(e.oa); dfree(e.ob); dfree(e.rec); dfree(e.rec_out); dfree(e.ab_out);
  e.cap = grow(need);
  dalloc(&e.fa, e.cap); dalloc(&e.fb, e.cap); dalloc(&e.oa, e.cap + 1); dalloc(&e.ob, e.cap + 1);
  dalloc(&e.rec, 3 * e.cap); dalloc(&e.rec_out, 3 * e.cap); dalloc(&e.ab_out, e.cap);
}
static uint64_t ab_key(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }
// The flagged body pairs of `ab[0..n)`, ascending.
static void ev_pairs(oavbd_ctx* c, const uint32_t* flag, const uint2* ab, uint32_t n, std::vector<uint64_t>& out) {
  out.clear();
  if (!n) return;
  const uint32_t total = scan_total(c, flag, c->evb.oa, n);
  if (!total) return;
  k_ev_compact_ab<<<blocks(n), TPB, 0, c->st>>>(flag, c->evb.oa, ab, c->evb.ab_out, n);
  std::vector<uint2> h(total);
  HIPC(hipMemcpyAsync(h.data(), c->evb.ab_out, total * sizeof(uint2), hipMemcpyDeviceToHost, c->st));
  HIPC(hipStreamSynchronize(c->st));
  for (const uint2& v : h) out.push_back(ab_key(v.x, v.y));
  std::sort(out.begin(), out.end());
}
// The flagged event records of `rec[0..n)` in index order.
static void ev_records(oavbd_ctx* c, const uint32_t* flag, const uint4* rec, uint32_t n,
                       std::vector<oavbd_event>& out) {
  out.clear();
  if (!n) return;
  const uint32_t total = scan_total(c, flag, c->evb.ob, n);
  if (!total) return;
  k_ev_compact_rec<<<blocks(n), TPB, 0, c->st>>>(flag, c->evb.ob, rec, c->evb.rec_out, n);
  std::vector<uint4> h(3 * (size_t)total);
  HIPC(hipMemcpyAsync(h.data(), c->evb.rec_out, h.size() * sizeof(uint4), hipMemcpyDeviceToHost, c->st));
  HIPC(hipStreamSynchronize(c->st));
  auto bits = [](uint32_t u) {
    float f;
    std::memcpy(&f, &u, sizeof(f));
    return f;
  };
  for (uint32_t i = 0; i < total; ++i) {
    const uint4 w0 = h[3 * i], w1 = h[3 * i + 1], w2 = h[3 * i + 2];
    oavbd_event e{};
    e.kind = w0.x;
    e.a = w0.y;
    e.b = w0.z;
    e.point[0] = bits(w1.x);
    e.point[1] = bits(w1.y);
    e.point[2] = bits(w1.z);
    e.approach = bits(w1.w);
    e.normal[0] = bits(w2.x);
    e.normal[1] = bits(w2.y);
    e.normal[2] = bits(w2.z);
    out.push_back(e);
  }
}
// Pairs in exactly one of two ascending sets, in ascending order: `in` for
// those of `now`, `out` for those of `prev` (BTreeSet::symmetric_difference).
static void ev_transitions(const std::vector<uint64_t>& now, const std::vector<uint64_t>& prev, uint32_t k_in,
                           uint32_t k_out, std::vector<oavbd_event>& q) {
  size_t i = 0, j = 0;
  auto push = [&](uint64_t key, uint32_t kind) {
    oavbd_event e{};
    e.kind = kind;
    e.a = (uint32_t)(key >> 32);
    e.b = (uint32_t)key;
    q.push_back(e);
  };
  while (i < now.size() || j < prev.size()) {
    if (j == prev.size() || (i < now.size() && now[i] < prev[j]))
      push(now[i++], k_in);
    else if (i == now.size() || prev[j] < now[i])
      push(prev[j++], k_out);
    else
      ++i, ++j;
  }
}
// After a waited substep: its clamp impacts, and on the substep that
// reports (the last of a step call) the trigger and contact transitions
// and the hits of newly touching pairs; clamp impacts already reported as
// hits of the same pair in this substep are dropped (mod.rs:1123-1136).
static void collect_events(oavbd_ctx* c, bool emit, uint32_t ncast) {
  Dev& d = c->d;
  std::vector<oavbd_event> ccd;
  ensure_evbuf(c, std::max<size_t>({(size_t)d.n, (size_t)d.npairs, (size_t)c->ncand, 1}));
  if (ncast) ev_records(c, d.ev_cflag, d.ev_crec, d.n, ccd);
  const size_t start = c->ev_contact.size();
  if (emit) {
    std::vector<uint64_t> trig_now, touch_now;
    if (c->ncand) ev_pairs(c, d.ev_gflag, d.C.ab, c->ncand, trig_now);
    ev_transitions(trig_now, c->prev_trigger, OAVBD_EVENT_ENTERED, OAVBD_EVENT_EXITED, c->ev_trigger);
    c->prev_trigger = std::move(trig_now);
    std::vector<oavbd_event> hits;
    if (d.npairs) {
      k_ev_touch<<<blocks(d.npairs), TPB, 0, c->st>>>(d, c->evb.fa, c->evb.fb, c->evb.rec);
      ev_pairs(c, c->evb.fa, d.P.ab, d.npairs, touch_now);
      ev_records(c, c->evb.fb, c->evb.rec, d.npairs, hits);
    }
    ev_transitions(touch_now, c->prev_touch, OAVBD_EVENT_BEGIN, OAVBD_EVENT_END, c->ev_contact);
    for (const auto& e : hits)
      if (!std::binary_search(c->prev_touch.begin(), c->prev_touch.end(), ab_key(e.a, e.b)))
        c->ev_contact.push_back(e);
    c->prev_touch = std::move(touch_now);
  }
  for (const auto& e : ccd) {
    bool dup = false;
    for (size_t i = start; i < c->ev_contact.size(); ++i) {
      const auto& f = c->ev_contact[i];
      dup = dup || (f.kind == OAVBD_EVENT_HIT && f.a == e.a && f.b == e.b);
    }
    if (!dup) c->ev_contact.push_back(e);
  }
}

// One fixed substep.  `finish` reads the step's counters back (stats,
// sleep count); other steps leave the stream running into the next one.
// `async`: the host waits for no count inside the step -- launches are
// sized from the buffer capacities and the last counts read back, every
// count comes from the device, and a count past its buffer is recorded in
// the overflow words (the caller repeats the steps); otherwise the host
// reads each count as soon as it exists and sizes buffers and launches
// exactly.
static void step_inner(oavbd_ctx* c, bool finish, bool async = false) {
  Dev& d = c->d;
  uint32_t n = d.n;
  if (c->joints_dirty) upload_joints(c);
  d.ev_on = c->ev_on ? 1u : 0u;
  d.ev_emit = (c->ev_on && finish) ? 1u : 0u;
  record(c, 0);
  zero_u32(c, d.counters, CTR_COUNT);
  // per-body passes are fused where no joint propagation sits between them
  const uint32_t may_sleep = c->may_sleep ? 1u : 0u;
  if (d.njoints == 0) {
    k_wake_begin<<<blocks(n), TPB, 0, c->st>>>(d, may_sleep);
  } else {
    if (may_sleep) wake_joint_motion(c);
    if (c->ngears) k_gear_mem<<<blocks(d.njoints), TPB, 0, c->st>>>(d);
    k_begin<<<blocks(n), TPB, 0, c->st>>>(d);
    if (d.nmotors || d.nropes) {  // motors, then the rope catch: after the torques, before the TOI gate
      k_joint_impulses<<<1, 64, 0, c->st>>>(d);
      k_refast<<<blocks(n), TPB, 0, c->st>>>(d);
    }
  }
  build_grid(c);
  // TOI: every body resets its clamp; fast movers count their targets
  k_toi_count<<<blocks((size_t)n * BP_LANES), TPB, 0, c->st>>>(d);
  if (async) {
    scan(c, d.toi_cnt, d.toi_off, n);
    toi_phase(c, (uint32_t)c->cap_toi, std::min<uint32_t>(c->last_ncast + c->last_ncast / 4u + 256u, (uint32_t)c->cap_toi));
  } else {
    const uint32_t ncast = scan_total(c, d.toi_cnt, d.toi_off, n);
    c->stats.toi_casts = ncast;
    c->last_ncast = ncast;
    if (ncast) {
      ensure_toi(c, ncast);
      toi_phase(c, ncast, ncast);
    }
  }
  broadphase(c, async);
  narrow_and_merge(c, async);
  // per-pair launches: the pair count, or the table's capacity
  const uint32_t gp = async ? (uint32_t)c->cap_pairs : d.npairs;
  if (d.njoints == 0) {
    k_wake_inertial<<<blocks(n), TPB, 0, c->st>>>(d, may_sleep);
  } else {
    k_apply_wake<<<blocks(n), TPB, 0, c->st>>>(d, 0u);
    if (may_sleep) {
      wake_joint_motion(c);
      k_external_wake<<<blocks(n), TPB, 0, c->st>>>(d);
    }
    k_inertial<<<blocks(n), TPB, 0, c->st>>>(d);
  }
  if (d.njoints) k_joint_prep<<<blocks(d.njoints), TPB, 0, c->st>>>(d);
  if (gp) k_pair_prep<<<blocks((size_t)gp * PREP_LANES), TPB, 0, c->st>>>(d);
  incidence_and_colour(c, gp, async ? c->slots : 0xffffffffu);
  k_sb_pack<<<blocks(n), TPB, 0, c->st>>>(d);
  k_inertial_blocks<<<blocks(n), TPB, 0, c->st>>>(d);
  record(c, 1);
  uint32_t ncol = 0;
  if (async) {
    ncol = c->slots;  // classes past the last one are empty launches
  } else {
    colour_wait(c);
    ncol = c->stats.colours;
  }
  // In timing mode every primal / dual launch is bracketed by pool events;
  // they are resolved after the step's final synchronisation.
  size_t np_ev = 0, nd_ev = 0;
  auto pool_event = [&](std::vector<hipEvent_t>& pool, size_t i) {
    while (pool.size() <= i) {
      hipEvent_t e;
      HIPC(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[i];
  };
  if (c->rec) tape_begin(c);
  const uint32_t ktail = async ? ncol : tail_first(c, ncol);
  if (!async) c->stats.tail_colours = ncol - ktail;
  // the dual's launch size: the point count, or the last one seen (its loop
  // covers the device's count)
  const uint32_t gq = async ? c->last_npts + c->last_npts / 8u : d.npts;
  for (uint32_t it = 0; it < c->prm.iters; ++it) {
    if (c->rec) tape_iter_start(c, it);
    for (uint32_t k = 0; k <= ktail && k < ncol; ++k) {
      if (async) {
        launch_primal_class(c, k, k < c->colour_cnt_h.size() ? c->colour_cnt_h[k] : 64u);
        continue;
      }
      const uint32_t cnt = k < ktail ? c->colour_cnt_h[k] : c->colour_off_h[ncol] - c->colour_off_h[k];
      if (!cnt) continue;
      hipEvent_t e0 = c->timing ? pool_event(c->pev, np_ev) : nullptr;
      hipEvent_t e1 = c->timing ? pool_event(c->pev, np_ev + 1) : nullptr;
      int v = -1;  // the fused tail is no tuned variant
      if (k < ktail)
        v = launch_primal(c, d.colour_list + c->colour_off_h[k], cnt, e0, e1);
      else
        launch_tail(c, ktail, ncol, e0, e1);
      if (c->timing) {
        if (c->plog.size() <= np_ev / 2) c->plog.resize(np_ev / 2 + 1);
        c->plog[np_ev / 2] = {cnt, v};
      }
      np_ev += 2;
    }
    if (c->rec) tape_iter_end(c, it);
    // the dual update, and the next sweep's inertial blocks (timed from the
    // first dual kernel's start to the last one's end)
    const uint32_t next = it + 1 < c->prm.iters ? 1u : 0u;
    const bool dual_pts = async || d.npts || next, dual_j = d.njoints != 0;
    hipEvent_t de0 = c->timing ? pool_event(c->dev_ev, nd_ev) : nullptr;
    hipEvent_t de1 = c->timing ? pool_event(c->dev_ev, nd_ev + 1) : nullptr;
    if (dual_pts)
      launch_k(c, k_dual, std::max<uint32_t>(blocks((size_t)gq + (next ? n : 0u)), 1u), TPB, de0,
               dual_j ? nullptr : de1, d, next);
    if (dual_j) launch_k(c, k_dual_joint, blocks(d.njoints), TPB, dual_pts ? nullptr : de0, de1, d);
    if (c->timing && !dual_pts && !dual_j) {
      HIPC(hipEventRecord(de0, c->st));
      HIPC(hipEventRecord(de1, c->st));
    }
    nd_ev += 2;
  }
  record(c, 3);
  if (d.njoints == 0) {
    k_finish_bodies<<<blocks(n), TPB, 0, c->st>>>(d);
    if (gp) k_supported<<<blocks(gp), TPB, 0, c->st>>>(d);
    k_sleep_bodies<<<blocks(n), TPB, 0, c->st>>>(d);
  } else {
    k_sb_unpack<<<blocks(n), TPB, 0, c->st>>>(d);
    k_bdf1<<<blocks(n), TPB, 0, c->st>>>(d);
    uint32_t ns = std::max(gp, d.njoints);
    if (ns) k_supported<<<blocks(ns), TPB, 0, c->st>>>(d);
    k_ready_init<<<blocks(n), TPB, 0, c->st>>>(d);
    k_ready_joints<<<blocks(d.njoints), TPB, 0, c->st>>>(d);
    propagate(c, d.b.flag_b, 0);
    k_sleep_timer<<<blocks(n), TPB, 0, c->st>>>(d);
    propagate(c, d.b.flag_a, 0);
    k_sleep_apply<<<blocks(n), TPB, 0, c->st>>>(d);
  }
  record(c, 4);
  if (c->rec) tape_end(c);
  if (c->ev_on && !async) collect_events(c, finish, c->stats.toi_casts);
  if (!async) {
    c->stats.pairs = d.npairs;
    c->stats.points = d.npts;
    c->stats.candidates = c->ncand;
    c->stats.solved = c->solved;
    c->last_ncand = c->ncand;
    c->last_npairs = d.npairs;
    c->last_npts = d.npts;
  }
  if (!finish && !c->timing) {
    c->may_sleep = true;  // the sleep count is not read back: wake passes run next step
    return;
  }
  read_counters(c);
  if (async) {
    // the last step's counts (the caller checks the overflow words first)
    const uint32_t* h = c->counters_h;
    d.npairs = h[DC_H + DC_NPAIRS];
    d.npts = h[DC_H + DC_NPTS];
    c->ncand = h[TOTALS_H + 1];
    c->last_ncast = h[TOTALS_H];
    c->last_ncand = c->ncand;
    c->last_npairs = d.npairs;
    c->last_npts = d.npts;
    c->stats.toi_casts = h[TOTALS_H];
    colour_check(c);  // the posted colouring status has landed
    c->stats.tail_colours = 0;
    c->stats.pairs = d.npairs;
    c->stats.points = d.npts;
    c->stats.candidates = c->ncand;
    c->stats.solved = c->solved;
  }
  c->sizes_known = true;
  c->sleeping = c->counters_h[CTR_SLEEPING];
  c->may_sleep = c->sleeping != 0;
  c->stats.sleeping = c->sleeping;
  c->stats.breakdowns = c->counters_h[CTR_BREAKDOWN];
  c->stats.toi_hits = c->counters_h[CTR_TOIHITS];
  c->stats.point_sides = c->counters_h[CTR_PTSIDES];
  c->stats.pair_sides = c->counters_h[CTR_PAIRSIDES];
  c->stats.joint_sides = c->counters_h[CTR_JSIDES];
  if (c->timing) {
    float setup = 0, iter = 0, fin = 0, all = 0;
    HIPC(hipEventElapsedTime(&setup, c->ev[0], c->ev[1]));
    HIPC(hipEventElapsedTime(&iter, c->ev[1], c->ev[3]));
    HIPC(hipEventElapsedTime(&fin, c->ev[3], c->ev[4]));
    HIPC(hipEventElapsedTime(&all, c->ev[0], c->ev[4]));
    double pm = 0.0, dm = 0.0;
    for (size_t i = 0; i < np_ev; i += 2) {
      float t = 0;
      HIPC(hipEventElapsedTime(&t, c->pev[i], c->pev[i + 1]));
      pm += t;
      if (c->tuning && c->plog[i / 2].second >= 0) {
        const auto& l = c->plog[i / 2];
        const int b = size_bucket(l.first);
        c->tune_ms[l.second][b] += t;
        c->tune_n[l.second][b] += 1;
      }
    }
    for (size_t i = 0; i < nd_ev; i += 2) {
      float t = 0;
      HIPC(hipEventElapsedTime(&t, c->dev_ev[i], c->dev_ev[i + 1]));
      dm += t;
    }
    const oavbd_stats& s = c->stats;
    double iters = c->prm.iters;
    c->kt.primal_ms += pm;
    c->kt.dual_ms += dm;
    c->kt.primal_launches += np_ev / 2;
    c->kt.dual_launches += nd_ev / 2;
    c->kt.primal_bytes += iters * (128.0 * s.solved + 96.0 * s.pair_sides + 48.0 * s.point_sides +
                                   168.0 * s.joint_sides);
    c->kt.dual_bytes += iters * (136.0 * s.pairs + 76.0 * s.points + 264.0 * d.njoints);
    c->kt.steps += 1;
    c->ms[0] = setup;
    c->ms[1] = (float)pm;
    c->ms[2] = (float)dm;
    c->ms[3] = fin;
    c->ms[4] = all;
    (void)iter;
  }
}

// ---------------------------------------------------------------------------
// backward pass
// ---------------------------------------------------------------------------
__global__ void k_zero_f4(float4* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__global__ void k_zero_f(float* p, uint32_t n) {
  const uint32_t i = gid();
  if (i < n) p[i] = 0.0f;
}
static void zero_f4(oavbd_ctx* c, float4* p, size_t n) {
  if (n) k_zero_f4<<<(uint32_t)((n + TPB - 1) / TPB), TPB, 0, c->st>>>(p, n);
}

static void adj_alloc(oavbd_ctx* c) {
  Adj& a = c->adj;
  const uint32_t n = c->d.n;
  if (c->adj_cap_n < n) {
    dfree(a.xb); dfree(a.vb); dfree(a.x0b); dfree(a.yb); dfree(a.mass); dfree(a.fric);
    dalloc(&a.xb, 2 * (size_t)n); dalloc(&a.vb, 2 * (size_t)n); dalloc(&a.x0b, 2 * (size_t)n);
    dalloc(&a.yb, 2 * (size_t)n); dalloc(&a.mass, n); dalloc(&a.fric, n);
    c->adj_cap_n = n;
  }
  const size_t npts = std::max<size_t>(c->diff_max_pts, 1), npairs = std::max<size_t>(c->diff_max_pairs, 1);
  if (c->adj_cap_pts < npts) {
    dfree(a.lb); dfree(a.lb_prev);
    dalloc(&a.lb, 2 * npts); dalloc(&a.lb_prev, 2 * npts);
    c->adj_cap_pts = npts;
  }
  if (c->adj_cap_pairs < npairs) {
    dfree(a.pb);
    dalloc(&a.pb, npairs);
    c->adj_cap_pairs = npairs;
  }
}

// Reverse of one recorded step: the adjoints of its outputs (body state,
// final duals) become adjoints of its inputs; parameter adjoints accumulate.
static void adjoint_step(oavbd_ctx* c, TapeHost& h) {
  Adj& a = c->adj;
  const Tape& t = h.t;
  // the duals' adjoints: this step's final duals <- what the next step's
  // warm start sent back; the previous step's start from zero
  std::swap(a.lb, a.lb_prev);
  zero_f4(c, a.lb_prev, 2 * c->adj_cap_pts);
  k_adj_finish<<<blocks(t.n), TPB, 0, c->st>>>(t, a);
  for (int it = (int)t.iters - 1; it >= 0; --it) {
    if (t.npts) k_adj_dual<<<blocks(t.npts), TPB, 0, c->st>>>(c->d, t, a, (uint32_t)it);
    for (int k = (int)h.col_cnt.size() - 1; k >= 0; --k) {
      const uint32_t cnt = h.col_cnt[k];
      if (!cnt) continue;
      constexpr int NB = (WAVE / BODY_LANES) * 2;
      k_adj_colour<2><<<blocks(cnt, NB), WAVE * 2, 0, c->st>>>(c->d, t, a, (uint32_t)it, k, t.clist + h.col_off[k], cnt);
    }
  }
  if (t.npairs) k_adj_pairs<<<blocks(t.npairs), TPB, 0, c->st>>>(c->d, t, a);
  if (t.npts) k_adj_warm<<<blocks(t.npts), TPB, 0, c->st>>>(t, a);
  k_adj_prelude<<<blocks(t.n), TPB, 0, c->st>>>(c->d, t, a);
  HIPC(hipGetLastError());
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

void oavbd_default_params(oavbd_params* p) {
  if (!p) return;
  p->gravity[0] = 0.0f;
  p->gravity[1] = -9.81f;
  p->gravity[2] = 0.0f;
  p->iters = 10;
}

int oavbd_create(const oavbd_params* params, uint32_t max_bodies, uint32_t max_joints, int device, oavbd_ctx** out) {
  if (!out || max_bodies == 0) return OAVBD_EINVAL;
  *out = nullptr;
  oavbd_ctx* c = new (std::nothrow) oavbd_ctx();
  if (!c) return OAVBD_ENOMEM;
  std::fill(c->ptable, c->ptable + PRIMAL_BUCKETS, (int8_t)-1);
  if (params)
    c->prm = *params;
  else
    oavbd_default_params(&c->prm);
  c->max_bodies = max_bodies;
  c->max_joints = max_joints;
  c->device = device;
  try {
    HIPC(hipSetDevice(device));
    HIPC(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIPC(hipHostMalloc((void**)&c->counters_h, (HEAD_H + CSTAT_HEAD + CSTAT_EAGER) * sizeof(uint32_t),
                       hipHostMallocCoherent));
    HIPC(hipHostGetDevicePointer((void**)&c->counters_hd, c->counters_h, 0));
    c->counters_h[POST_FLAG] = 0u;
    for (auto& e : c->ev) HIPC(hipEventCreate(&e));
    c->noalloc = noalloc_resolve(device);
    alloc_bodies(c, max_bodies);
    HIPC(hipDeviceSynchronize());  // the null-stream clears land before the context stream runs
  } catch (const HipError& e) {
    std::string m = e.msg;
    oavbd_destroy(c);
    std::fprintf(stderr, "oavbd_create: %s\n", m.c_str());
    return OAVBD_EDEVICE;
  }
  for (int k = 0; k < 3; ++k) c->d.gravity[k] = c->prm.gravity[k];
  *out = c;
  return OAVBD_OK;
}

void oavbd_destroy(oavbd_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  free_all(c);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->pev) (void)hipEventDestroy(e);
  for (auto& e : c->dev_ev) (void)hipEventDestroy(e);
  if (c->counters_h) (void)hipHostFree(c->counters_h);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

// Awake mass properties (RigidBody::build / wake_restore, body.rs:346-376,
// :412-418; shape.rs:751-767): inverse mass and body-frame inertia
// diagonal of a dynamic body, zero otherwise.
static void awake_mass(int shape, f3 p, int type, float m, float* inv, f3* I) {
  *inv = 0.0f;
  *I = zero3();
  if (type != BT_DYNAMIC) return;
  *inv = 1.0f / m;
  if (shape == SH_SPHERE) {
    float v = 0.4f * m * p.x * p.x;
    *I = f3{v, v, v};
  } else if (shape == SH_BOX) {
    f3 s2 = mul(p, 2.0f);
    float x = s2.x * s2.x, y = s2.y * s2.y, z = s2.z * s2.z;
    float k = m / 12.0f;
    *I = f3{k * (y + z), k * (z + x), k * (x + y)};
  } else if (shape == SH_CAPSULE) {
    *I = capsule_inertia(p.x, p.y, m);
  }
}

// Uploads `in` as the whole body table: derived per-body data, scene ranks,
// the broadphase's large-body lists, and a cold start (no pairs, joints or
// warm starts, everything awake).
static void set_bodies_impl(oavbd_ctx* c, const BodyInputs& in) {
  const uint32_t n = in.n;
  const int32_t* shape = in.shape.data();
  const float *sp3 = in.sp3.data(), *pos3 = in.pos3.data(), *quat4 = in.quat4.data(), *vel3 = in.vel3.data(),
              *angvel3 = in.ang3.data(), *mass = in.mass.data(), *friction = in.friction.data(),
              *restitution = in.restitution.data(), *rolling = in.rolling.data(), *torsion = in.torsion.data();
  const int32_t* body_type = in.type.data();
  const uint32_t *layer = in.layer.data(), *mask = in.mask.data(), *scene = in.scene.data();
  const uint8_t* trigger = in.trigger.data();
  {
    HIPC(hipSetDevice(c->device));
    Dev& d = c->d;
    d.n = n;
    std::vector<float4> pos(n), rot(n), vel(n), ang(n), zero(n), inert(n), shp4(n), mat(n);
    std::vector<uint32_t> lay(n), msk(n), scn(n), prio(n), loc(n);
    std::vector<uint8_t> trig(n), isl(n, 0);
    c->kind_h.assign(shape, shape + n);
    c->type_h.assign(n, 0);
    std::vector<float> brad(n);
    std::vector<uint32_t> per_scene_count;
    for (uint32_t i = 0; i < n; ++i) {
      int t = body_type[i];
      float m = mass[i];
      if (t == BT_DYNAMIC && !(m > 0.0f)) t = BT_STATIC;  // RigidBody::build
      c->type_h[i] = t;
      f3 p{sp3[3 * i], sp3[3 * i + 1], sp3[3 * i + 2]};
      float inv = 0.0f, mm = t == BT_DYNAMIC ? m : 0.0f;
      f3 I = zero3();
      awake_mass(shape[i], p, t, m, &inv, &I);
      float br = bound_radius(shape[i], p);
      brad[i] = br;
      pos[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], inv);
      rot[i] = make_float4(quat4[4 * i], quat4[4 * i + 1], quat4[4 * i + 2], quat4[4 * i + 3]);
      vel[i] = make_float4(vel3[3 * i], vel3[3 * i + 1], vel3[3 * i + 2], 0.0f);
      ang[i] = make_float4(angvel3[3 * i], angvel3[3 * i + 1], angvel3[3 * i + 2], 0.0f);
      zero[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      inert[i] = make_float4(I.x, I.y, I.z, mm);
      shp4[i] = make_float4(p.x, p.y, p.z, br);
      mat[i] = make_float4(friction[i], restitution[i], rolling ? rolling[i] : 0.0f, torsion ? torsion[i] : 0.0f);
      lay[i] = layer ? layer[i] : 1u;
      msk[i] = mask ? mask[i] : 0xffffffffu;
      scn[i] = scene ? scene[i] : 0u;
      trig[i] = trigger ? trigger[i] : 0;
    }
    // dense scene ranks (the device compares ranks only) and scene-local
    // indices (the colouring priority)
    std::vector<uint32_t> ids(scn.begin(), scn.end());
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    const uint32_t nsc = (uint32_t)ids.size();
    per_scene_count.assign(nsc, 0);
    for (uint32_t i = 0; i < n; ++i) {
      scn[i] = (uint32_t)(std::lower_bound(ids.begin(), ids.end(), scn[i]) - ids.begin());
      loc[i] = per_scene_count[scn[i]]++;
      prio[i] = prio_hash(loc[i]);
    }
    // grid geometry: bodies with unbounded or large spheres (over 4x the
    // median bound) are tested against everything instead of gridded.
    std::vector<float> fin;
    for (float r : brad)
      if (std::isfinite(r)) fin.push_back(r);
    float cut = 1.0f;
    if (!fin.empty()) {
      std::vector<float> s = fin;
      std::nth_element(s.begin(), s.begin() + s.size() / 2, s.end());
      cut = std::max(4.0f * s[s.size() / 2], 1e-3f);
    }
    float rmax = 0.0f;
    for (float r : fin)
      if (r <= cut) rmax = std::max(rmax, r);
    std::vector<uint32_t> large;
    for (uint32_t i = 0; i < n; ++i)
      if (!std::isfinite(brad[i]) || brad[i] > cut) {
        large.push_back(i);
        isl[i] = 1;
      }
    d.g.cell = (2.0f * rmax + 0.05f) * 1.01f + 1e-6f;
    c->rows6 = false;
    for (uint32_t i = 0; i < n; ++i)
      if (mat[i].z > 0.0f || mat[i].w > 0.0f) c->rows6 = true;
    d.g.nlarge = (uint32_t)large.size();
    // bodies by scene, large bodies by scene, and each large body's flag row
    // over the bodies of its scene
    std::vector<uint32_t> soff(nsc + 1, 0), sitems(n), lsc_off(nsc + 1, 0), lsc(large.size());
    for (uint32_t i = 0; i < n; ++i) soff[scn[i] + 1]++;
    for (uint32_t k = 0; k < nsc; ++k) soff[k + 1] += soff[k];
    {
      std::vector<uint32_t> cur(soff.begin(), soff.end() - 1);
      for (uint32_t i = 0; i < n; ++i) sitems[cur[scn[i]]++] = i;
    }
    for (uint32_t L : large) lsc_off[scn[L] + 1]++;
    for (uint32_t k = 0; k < nsc; ++k) lsc_off[k + 1] += lsc_off[k];
    {
      std::vector<uint32_t> cur(lsc_off.begin(), lsc_off.end() - 1);
      for (uint32_t L : large) lsc[cur[scn[L]]++] = L;
    }
    std::vector<uint4> lrange(large.size());
    size_t lflat = 0;
    for (size_t k = 0; k < large.size(); ++k) {
      uint32_t sc = scn[large[k]];
      lrange[k] = make_uint4(soff[sc], soff[sc + 1], (uint32_t)lflat, large[k]);
      lflat += soff[sc + 1] - soff[sc];
    }
    if (lflat >= ((size_t)1 << 31)) throw HipError{"too many large-body candidates"};
    d.g.lflat = (uint32_t)lflat;
    dfree(d.g.large); dfree(d.g.lsc_off); dfree(d.g.lsc); dfree(d.g.sitems); dfree(d.g.soff); dfree(d.g.lrange);
    dalloc(&d.g.large, large.size());
    dalloc(&d.g.soff, nsc + 1);
    dalloc(&d.g.lsc_off, nsc + 1);
    dalloc(&d.g.lsc, large.size());
    dalloc(&d.g.sitems, n);
    dalloc(&d.g.lrange, large.size());
    if (!large.empty()) {
      HIPC(hipMemcpy(d.g.large, large.data(), large.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      HIPC(hipMemcpy(d.g.lsc, lsc.data(), large.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      HIPC(hipMemcpy(d.g.lrange, lrange.data(), large.size() * sizeof(uint4), hipMemcpyHostToDevice));
    }
    HIPC(hipMemcpy(d.g.lsc_off, lsc_off.data(), (nsc + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.g.sitems, sitems.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.g.soff, soff.data(), (nsc + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    dfree(d.lflag);
    dfree(d.lrank);
    dalloc(&d.lflag, lflat);
    dalloc(&d.lrank, lflat + 1);
    auto up4 = [&](float4* dst, const std::vector<float4>& v) {
      HIPC(hipMemcpy(dst, v.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    };
    up4(d.b.pos, pos); up4(d.b.rot, rot); up4(d.b.vel, vel); up4(d.b.angvel, ang); up4(d.b.torque, zero);
    {
      std::vector<float4> fd(n);
      c->any_axis = false;
      for (uint32_t i = 0; i < n; ++i) {
        fd[i] = make_float4(in.fdir3[3 * i], in.fdir3[3 * i + 1], in.fdir3[3 * i + 2], in.ftrans[i]);
        c->any_axis = c->any_axis || fd[i].x != 0.0f || fd[i].y != 0.0f || fd[i].z != 0.0f;
      }
      up4(d.b.fdir, fd);
    }
    up4(d.b.inertia, inert); up4(d.b.shape, shp4); up4(d.b.mat, mat); up4(d.b.prev_pos, pos); up4(d.b.prev_rot, rot);
    up4(d.b.pos0, pos); up4(d.b.rot0, rot); up4(d.b.drot, zero);
    HIPC(hipMemcpy(d.b.kind, shape, n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.b.type, c->type_h.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.b.layer, lay.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.b.mask, msk.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.b.scene, scn.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.b.prio, prio.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.b.local, loc.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.b.trigger, trig.data(), n, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d.g.is_large, isl.data(), n, hipMemcpyHostToDevice));
    HIPC(hipMemset(d.b.asleep, 0, n));
    HIPC(hipMemset(d.b.timer, 0, n * sizeof(float)));
    HIPC(hipMemset(d.b.flag_a, 0, n));
    HIPC(hipMemset(d.b.flag_b, 0, n));
    HIPC(hipMemset(d.b.colour, 0xff, n * sizeof(int32_t)));
    {  // no TOI clamp before the first step (travel < 0)
      std::vector<float4> nohit(n, make_float4(0.0f, 0.0f, 0.0f, -1.0f));
      HIPC(hipMemcpy(d.toi_hit, nohit.data(), n * sizeof(float4), hipMemcpyHostToDevice));
      HIPC(hipMemset(d.toi_tgt, 0, n * sizeof(uint32_t)));
    }
    HIPC(hipMemset(d.poff, 0, (n + 1) * sizeof(uint32_t)));
    HIPC(hipMemset(d.poff_old, 0, (n + 1) * sizeof(uint32_t)));
    d.npairs = d.npts = d.npairs_old = d.npts_old = 0;
    c->joints.clear();
    upload_joints(c);
    c->sleeping = 0;
    c->may_sleep = false;
    c->time_debt = 0.0f;
    c->stats = oavbd_stats{};
    c->sizes_known = false;  // the next steps wait for their counts until one has been read back
    ensure_pairs(c, 1024);
    ensure_points(c, 4096);
    ensure_incidence(c, 1024, 4096);
    // the uploads above ran on the null stream, which the context's
    // non-blocking stream does not wait for
    HIPC(hipDeviceSynchronize());
  }
}

// Validates n bodies given as the oavbd_set_bodies arrays and appends them,
// defaults filled in, to `in`.
static int append_inputs(oavbd_ctx* c, BodyInputs& in, uint32_t n, const int32_t* shape, const float* sp3,
                         const float* pos3, const float* quat4, const float* vel3, const float* angvel3,
                         const float* mass, const float* friction, const float* restitution,
                         const int32_t* body_type, const uint32_t* layer, const uint32_t* mask,
                         const uint8_t* trigger, const float* rolling, const float* torsion, const uint32_t* scene) {
  if ((size_t)in.n + n > c->max_bodies) return fail(c, OAVBD_ECAPACITY, "more bodies than max_bodies");
  if (n && (!shape || !sp3 || !pos3 || !quat4 || !vel3 || !angvel3 || !mass || !friction || !restitution ||
            !body_type))
    return fail(c, OAVBD_EINVAL, "null body array");
  for (uint32_t i = 0; i < n; ++i) {
    if (shape[i] < 0 || shape[i] > 3) return fail(c, OAVBD_EINVAL, "bad shape kind");
    if (body_type[i] < 0 || body_type[i] > 2) return fail(c, OAVBD_EINVAL, "bad body type");
    if (shape[i] == SH_HALFSPACE && body_type[i] == BT_DYNAMIC && mass[i] > 0.0f)
      return fail(c, OAVBD_EINVAL, "dynamic half-space");
    if (shape[i] == SH_CAPSULE && !(sp3[3 * i] > 0.0f && sp3[3 * i] < INFINITY && sp3[3 * i + 1] >= 0.0f &&
                                    sp3[3 * i + 1] < INFINITY))
      return fail(c, OAVBD_EINVAL, "capsule needs a finite radius > 0 and half height >= 0");
  }
  auto add = [n](auto& v, const auto* src, size_t per, auto dflt) {
    for (uint32_t i = 0; i < n; ++i)
      for (size_t k = 0; k < per; ++k) v.push_back(src ? src[per * i + k] : dflt);
  };
  add(in.shape, shape, 1, 0);
  add(in.type, body_type, 1, 0);
  add(in.sp3, sp3, 3, 0.0f);
  add(in.pos3, pos3, 3, 0.0f);
  add(in.quat4, quat4, 4, 0.0f);
  add(in.vel3, vel3, 3, 0.0f);
  add(in.ang3, angvel3, 3, 0.0f);
  add(in.mass, mass, 1, 0.0f);
  add(in.friction, friction, 1, 0.0f);
  add(in.restitution, restitution, 1, 0.0f);
  add(in.rolling, rolling, 1, 0.0f);
  add(in.torsion, torsion, 1, 0.0f);
  add(in.layer, layer, 1, 1u);
  add(in.mask, mask, 1, 0xffffffffu);
  add(in.scene, scene, 1, 0u);
  add(in.trigger, trigger, 1, (uint8_t)0);
  for (uint32_t i = 0; i < n; ++i) {  // isotropic until oavbd_set_friction_axes
    for (int k = 0; k < 3; ++k) in.fdir3.push_back(0.0f);
    in.ftrans.push_back(friction[i]);
  }
  in.n += n;
  return OAVBD_OK;
}

int oavbd_set_bodies(oavbd_ctx* c, uint32_t n, const int32_t* shape, const float* sp3, const float* pos3,
                     const float* quat4, const float* vel3, const float* angvel3, const float* mass,
                     const float* friction, const float* restitution, const int32_t* body_type,
                     const uint32_t* layer, const uint32_t* mask, const uint8_t* trigger, const float* rolling,
                     const float* torsion, const uint32_t* scene) {
  if (!c) return OAVBD_EINVAL;
  BodyInputs in;
  const int r = append_inputs(c, in, n, shape, sp3, pos3, quat4, vel3, angvel3, mass, friction, restitution,
                              body_type, layer, mask, trigger, rolling, torsion, scene);
  if (r != OAVBD_OK) return r;
  GUARD(c, {
    set_bodies_impl(c, in);
    c->inputs = std::move(in);
    // a new body table: no event baseline, nothing queued
    c->ev_contact.clear();
    c->ev_trigger.clear();
    c->prev_touch.clear();
    c->prev_trigger.clear();
  });
  return OAVBD_OK;
}

// ---------------------------------------------------------------------------
// persistent state on the host: contact export / import, sleep state, and
// incremental body edits (append, swap-remove) through a rebuild
// ---------------------------------------------------------------------------
}  // extern "C"

template <class T>
static void d2h(std::vector<T>& v, const T* src, size_t n) {
  v.resize(n);
  if (n) HIPC(hipMemcpy(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
}
template <class T>
static void h2d(T* dst, const std::vector<T>& v, size_t n) {
  if (n) HIPC(hipMemcpy(dst, v.data(), n * sizeof(T), hipMemcpyHostToDevice));
}

// The pair table with its points, as the device holds it between steps.
struct HostContacts {
  std::vector<uint2> ab, pts;  // per pair: bodies, (first point, count)
  std::vector<float4> prec;    // 4 per pair: normal+gap, t1, t2, friction
  std::vector<float4> qrec;    // 4 per point: ra, rb, lam, pen
  std::vector<float4> qroll;   // per point
};

// Everything per body and per joint a step reads from earlier steps.
struct HostBodies {
  std::vector<float4> pos, rot, vel, ang, torque, inertia, prev_pos, prev_rot, toi_hit;
  std::vector<uint8_t> asleep;
  std::vector<float> timer;
  std::vector<uint32_t> toi_tgt;
  std::vector<int32_t> colour;
  std::vector<float4> jlam, jpen, jlama, jpena;
};

static HostContacts download_contacts(oavbd_ctx* c) {
  const Dev& d = c->d;
  HostContacts h;
  d2h(h.ab, d.P.ab, d.npairs);
  d2h(h.pts, d.P.pts, d.npairs);
  d2h(h.prec, d.P.rec, 4 * (size_t)d.npairs);
  d2h(h.qrec, d.Q.rec, 4 * (size_t)d.npts);
  d2h(h.qroll, d.Q.roll, d.npts);
  return h;
}

static HostBodies download_bodies(oavbd_ctx* c) {
  const Dev& d = c->d;
  const uint32_t n = d.n;
  HostBodies h;
  d2h(h.pos, d.b.pos, n); d2h(h.rot, d.b.rot, n); d2h(h.vel, d.b.vel, n); d2h(h.ang, d.b.angvel, n);
  d2h(h.torque, d.b.torque, n); d2h(h.inertia, d.b.inertia, n); d2h(h.prev_pos, d.b.prev_pos, n);
  d2h(h.prev_rot, d.b.prev_rot, n); d2h(h.toi_hit, d.toi_hit, n); d2h(h.asleep, d.b.asleep, n);
  d2h(h.timer, d.b.timer, n); d2h(h.toi_tgt, d.toi_tgt, n); d2h(h.colour, d.b.colour, n);
  d2h(h.jlam, d.J.lam, d.njoints);
  d2h(h.jpen, d.J.pen, d.njoints);
  d2h(h.jlama, d.J.lama, d.njoints);
  d2h(h.jpena, d.J.pena, d.njoints);
  return h;
}

// Writes the first hb.pos.size() bodies and the joint duals back.
static void upload_bodies(oavbd_ctx* c, const HostBodies& h) {
  Dev& d = c->d;
  const size_t m = h.pos.size();
  h2d(d.b.pos, h.pos, m); h2d(d.b.rot, h.rot, m); h2d(d.b.vel, h.vel, m); h2d(d.b.angvel, h.ang, m);
  h2d(d.b.torque, h.torque, m); h2d(d.b.inertia, h.inertia, m); h2d(d.b.prev_pos, h.prev_pos, m);
  h2d(d.b.prev_rot, h.prev_rot, m); h2d(d.toi_hit, h.toi_hit, m); h2d(d.b.asleep, h.asleep, m);
  h2d(d.b.timer, h.timer, m); h2d(d.toi_tgt, h.toi_tgt, m); h2d(d.b.colour, h.colour, m);
  h2d(d.J.lam, h.jlam, std::min<size_t>(h.jlam.size(), d.njoints));
  h2d(d.J.pen, h.jpen, std::min<size_t>(h.jpen.size(), d.njoints));
  h2d(d.J.lama, h.jlama, std::min<size_t>(h.jlama.size(), d.njoints));
  h2d(d.J.pena, h.jpena, std::min<size_t>(h.jpena.size(), d.njoints));
  HIPC(hipDeviceSynchronize());  // null-stream uploads land before the context stream reads them
  uint32_t sleeping = 0;
  for (uint8_t a : h.asleep) sleeping += a ? 1u : 0u;
  c->sleeping = sleeping;
  c->may_sleep = sleeping != 0;
}

// Installs a pair table as the current one: stored sorted by (a, b) with
// each pair's points contiguous, plus the per-body offsets of the a-runs the
// next merge searches.  Throws (EINVAL) on an invalid table.
static void upload_contacts(oavbd_ctx* c, const HostContacts& h) {
  Dev& d = c->d;
  const uint32_t n = d.n, np = (uint32_t)h.ab.size();
  const size_t nq = h.qroll.size();
  for (uint32_t i = 0; i < np; ++i) {
    const uint2 ab = h.ab[i], pt = h.pts[i];
    if (ab.x >= n || ab.y >= n || ab.x == ab.y) throw ArgError{"contact pair with a bad body index"};
    if (pt.y > K_MAX_POINTS || (size_t)pt.x + pt.y > nq)
      throw ArgError{"contact pair with a bad point range"};
  }
  std::vector<uint32_t> order(np);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    return h.ab[x].x != h.ab[y].x ? h.ab[x].x < h.ab[y].x : h.ab[x].y < h.ab[y].y;
  });
  for (uint32_t i = 1; i < np; ++i)
    if (h.ab[order[i]].x == h.ab[order[i - 1]].x && h.ab[order[i]].y == h.ab[order[i - 1]].y)
      throw ArgError{"duplicate contact pair"};
  std::vector<uint2> ab(np), pts(np);
  std::vector<float4> prec(4 * (size_t)np);
  std::vector<float4> qrec, qroll;
  std::vector<uint32_t> qpair;
  for (uint32_t i = 0; i < np; ++i) {
    const uint32_t o = order[i];
    ab[i] = h.ab[o];
    pts[i] = make_uint2((uint32_t)qroll.size(), h.pts[o].y);
    for (int k = 0; k < 4; ++k) prec[4 * (size_t)i + k] = h.prec[4 * (size_t)o + k];
    for (uint32_t k = 0; k < h.pts[o].y; ++k) {
      const size_t q = (size_t)h.pts[o].x + k;
      for (int j = 0; j < 4; ++j) qrec.push_back(h.qrec[4 * q + j]);
      qroll.push_back(h.qroll[q]);
      qpair.push_back(i);
    }
  }
  const uint32_t nqs = (uint32_t)qroll.size();
  ensure_pairs(c, std::max<uint32_t>(np, 1));
  ensure_points(c, std::max<uint32_t>(nqs, 1));
  ensure_incidence(c, np, nqs);
  h2d(d.P.ab, ab, np);
  h2d(d.P.pts, pts, np);
  h2d(d.P.rec, prec, 4 * (size_t)np);
  h2d(d.Q.rec, qrec, 4 * (size_t)nqs);
  h2d(d.Q.roll, qroll, nqs);
  h2d(d.Q.pair, qpair, nqs);
  std::vector<int> src(nqs, -1);
  h2d(d.Q.src, src, nqs);
  std::vector<uint32_t> poff(n + 1);
  uint32_t i = 0;
  for (uint32_t b = 0; b <= n; ++b) {
    while (i < np && ab[i].x < b) ++i;
    poff[b] = i;
  }
  h2d(d.poff, poff, n + 1);
  HIPC(hipDeviceSynchronize());  // null-stream uploads land before the context stream reads them
  d.npairs = np;
  d.npts = nqs;
  c->stats.pairs = np;
  c->stats.points = nqs;
}

// Rebuilds the device table from `in` (set_bodies_impl), then restores the
// first bodies' state, the joints with their duals, and the pair table.
static void rebuild(oavbd_ctx* c, BodyInputs in, const HostBodies& hb, std::vector<HostJoint> joints,
                    const HostContacts& hc) {
  const float debt = c->time_debt;
  set_bodies_impl(c, in);
  c->inputs = std::move(in);
  c->time_debt = debt;
  c->joints = std::move(joints);
  upload_joints(c);
  upload_bodies(c, hb);
  upload_contacts(c, hc);
}

extern "C" {

int oavbd_add_bodies(oavbd_ctx* c, uint32_t n, const int32_t* shape, const float* sp3, const float* pos3,
                     const float* quat4, const float* vel3, const float* angvel3, const float* mass,
                     const float* friction, const float* restitution, const int32_t* body_type,
                     const uint32_t* layer, const uint32_t* mask, const uint8_t* trigger, const float* rolling,
                     const float* torsion, const uint32_t* scene, uint32_t* first_index) {
  if (!c) return OAVBD_EINVAL;
  BodyInputs in = c->inputs;
  const uint32_t first = in.n;
  const int r = append_inputs(c, in, n, shape, sp3, pos3, quat4, vel3, angvel3, mass, friction, restitution,
                              body_type, layer, mask, trigger, rolling, torsion, scene);
  if (r != OAVBD_OK) return r;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    if (c->joints_dirty) upload_joints(c);
    HostBodies hb = download_bodies(c);
    HostContacts hc = download_contacts(c);
    rebuild(c, std::move(in), hb, c->joints, hc);
  });
  if (first_index) *first_index = first;
  return OAVBD_OK;
}

int oavbd_remove_body(oavbd_ctx* c, uint32_t h) {
  if (!c) return OAVBD_EINVAL;
  if (h >= c->inputs.n) return OAVBD_OK;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    if (c->joints_dirty) upload_joints(c);
    const uint32_t last = c->inputs.n - 1;
    {  // event state (mod.rs:1246-1286): trigger pairs with h exit, the rest
       // follow the renumbering; touches and undrained contact events go
      auto mp = [h, last](uint32_t x) { return x == last ? h : x; };
      std::vector<uint64_t> kept;
      for (uint64_t k : c->prev_trigger) {
        const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
        if (a == h || b == h) {
          oavbd_event e{};
          e.kind = OAVBD_EVENT_EXITED;
          e.a = a;
          e.b = b;
          c->ev_trigger.push_back(e);
        } else {
          kept.push_back(ab_key(std::min(mp(a), mp(b)), std::max(mp(a), mp(b))));
        }
      }
      std::sort(kept.begin(), kept.end());
      kept.erase(std::unique(kept.begin(), kept.end()), kept.end());
      c->prev_trigger = std::move(kept);
      c->prev_touch.clear();
      c->ev_contact.clear();
    }
    BodyInputs in = c->inputs;
    auto rm = [h, last](auto& v, size_t per) {
      for (size_t k = 0; k < per; ++k) v[per * h + k] = v[per * last + k];
      v.resize(per * last);
    };
    rm(in.shape, 1); rm(in.type, 1); rm(in.sp3, 3); rm(in.pos3, 3); rm(in.quat4, 4); rm(in.vel3, 3);
    rm(in.ang3, 3); rm(in.mass, 1); rm(in.friction, 1); rm(in.restitution, 1); rm(in.rolling, 1);
    rm(in.torsion, 1); rm(in.layer, 1); rm(in.mask, 1); rm(in.scene, 1); rm(in.trigger, 1);
    rm(in.fdir3, 3); rm(in.ftrans, 1);
    in.n = last;
    HostBodies hb = download_bodies(c);
    rm(hb.pos, 1); rm(hb.rot, 1); rm(hb.vel, 1); rm(hb.ang, 1); rm(hb.torque, 1); rm(hb.inertia, 1);
    rm(hb.prev_pos, 1); rm(hb.prev_rot, 1); rm(hb.toi_hit, 1); rm(hb.asleep, 1); rm(hb.timer, 1);
    rm(hb.toi_tgt, 1); rm(hb.colour, 1);
    auto map = [h, last](uint32_t x) { return x == last ? h : x; };
    // pairs and joints on h go; the rest keep their order and orientation
    HostContacts old = download_contacts(c), hc;
    for (size_t i = 0; i < old.ab.size(); ++i) {
      const uint2 ab = old.ab[i];
      if (ab.x == h || ab.y == h) continue;
      hc.ab.push_back(make_uint2(map(ab.x), map(ab.y)));
      hc.pts.push_back(make_uint2((uint32_t)hc.qroll.size(), old.pts[i].y));
      for (int k = 0; k < 4; ++k) hc.prec.push_back(old.prec[4 * i + k]);
      for (uint32_t k = 0; k < old.pts[i].y; ++k) {
        const size_t q = (size_t)old.pts[i].x + k;
        for (int j = 0; j < 4; ++j) hc.qrec.push_back(old.qrec[4 * q + j]);
        hc.qroll.push_back(old.qroll[q]);
      }
    }
    std::vector<HostJoint> joints;
    std::vector<float4> jlam, jpen, jlama, jpena;
    for (size_t j = 0; j < c->joints.size(); ++j) {
      HostJoint J = c->joints[j];
      if (J.a == h || J.b == h) continue;
      J.a = map(J.a);
      J.b = map(J.b);
      joints.push_back(J);
      jlam.push_back(hb.jlam[j]);
      jpen.push_back(hb.jpen[j]);
      jlama.push_back(hb.jlama[j]);
      jpena.push_back(hb.jpena[j]);
    }
    hb.jlam = std::move(jlam);
    hb.jpen = std::move(jpen);
    hb.jlama = std::move(jlama);
    hb.jpena = std::move(jpena);
    rebuild(c, std::move(in), hb, std::move(joints), hc);
    // every sleeping body wakes (wake_body through the device's own path)
    Dev& d = c->d;
    if (c->sleeping) {
      std::vector<uint8_t> flags(hb.asleep);
      h2d(d.b.flag_a, flags, flags.size());
      HIPC(hipDeviceSynchronize());
      k_apply_wake<<<blocks(d.n), TPB, 0, c->st>>>(d, 0u);
      HIPC(hipStreamSynchronize(c->st));
      c->sleeping = 0;
    }
    c->may_sleep = false;
  });
  return OAVBD_OK;
}

int oavbd_get_contacts(oavbd_ctx* c, oavbd_contact_pair* pairs, uint32_t cap_pairs, oavbd_contact_point* points,
                       uint32_t cap_points, uint32_t* npairs, uint32_t* npoints) {
  if (!c || !npairs || !npoints) return OAVBD_EINVAL;
  *npairs = c->d.npairs;
  *npoints = c->d.npts;
  if ((pairs && cap_pairs < c->d.npairs) || (points && cap_points < c->d.npts))
    return fail(c, OAVBD_ECAPACITY, "contact export arrays too small");
  if (!pairs && !points) return OAVBD_OK;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    const HostContacts h = download_contacts(c);
    if (pairs)
      for (size_t i = 0; i < h.ab.size(); ++i) {
        oavbd_contact_pair& o = pairs[i];
        o.a = h.ab[i].x;
        o.b = h.ab[i].y;
        o.first_point = h.pts[i].x;
        o.npoints = h.pts[i].y;
        std::memcpy(o.normal, &h.prec[4 * i], sizeof(float4));
        std::memcpy(o.tangent1, &h.prec[4 * i + 1], sizeof(float4));
        std::memcpy(o.tangent2, &h.prec[4 * i + 2], sizeof(float4));
        std::memcpy(o.friction, &h.prec[4 * i + 3], sizeof(float4));
      }
    if (points)
      for (size_t q = 0; q < h.qroll.size(); ++q) {
        oavbd_contact_point& o = points[q];
        std::memcpy(o.ra, &h.qrec[4 * q], sizeof(float4));
        std::memcpy(o.rb, &h.qrec[4 * q + 1], sizeof(float4));
        std::memcpy(o.lam, &h.qrec[4 * q + 2], sizeof(float4));
        std::memcpy(o.pen, &h.qrec[4 * q + 3], sizeof(float4));
        std::memcpy(o.roll, &h.qroll[q], sizeof(float4));
      }
  });
  return OAVBD_OK;
}

int oavbd_set_contacts(oavbd_ctx* c, const oavbd_contact_pair* pairs, uint32_t npairs,
                       const oavbd_contact_point* points, uint32_t npoints) {
  if (!c || (npairs && !pairs) || (npoints && !points)) return OAVBD_EINVAL;
  HostContacts h;
  h.ab.resize(npairs);
  h.pts.resize(npairs);
  h.prec.resize(4 * (size_t)npairs);
  h.qrec.resize(4 * (size_t)npoints);
  h.qroll.resize(npoints);
  for (uint32_t i = 0; i < npairs; ++i) {
    const oavbd_contact_pair& p = pairs[i];
    h.ab[i] = make_uint2(p.a, p.b);
    h.pts[i] = make_uint2(p.first_point, p.npoints);
    std::memcpy(&h.prec[4 * (size_t)i], p.normal, sizeof(float4));
    std::memcpy(&h.prec[4 * (size_t)i + 1], p.tangent1, sizeof(float4));
    std::memcpy(&h.prec[4 * (size_t)i + 2], p.tangent2, sizeof(float4));
    std::memcpy(&h.prec[4 * (size_t)i + 3], p.friction, sizeof(float4));
  }
  for (uint32_t q = 0; q < npoints; ++q) {
    const oavbd_contact_point& p = points[q];
    std::memcpy(&h.qrec[4 * (size_t)q], p.ra, sizeof(float4));
    std::memcpy(&h.qrec[4 * (size_t)q + 1], p.rb, sizeof(float4));
    std::memcpy(&h.qrec[4 * (size_t)q + 2], p.lam, sizeof(float4));
    std::memcpy(&h.qrec[4 * (size_t)q + 3], p.pen, sizeof(float4));
    std::memcpy(&h.qroll[q], p.roll, sizeof(float4));
  }
  try {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    upload_contacts(c, h);
  } catch (const ArgError& e) {
    return fail(c, OAVBD_EINVAL, e.msg);
  } catch (const HipError& e) {
    return fail(c, OAVBD_EDEVICE, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(c, OAVBD_ENOMEM, "host allocation failed");
  }
  return OAVBD_OK;
}

int oavbd_set_sleep(oavbd_ctx* c, const uint8_t* asleep, const float* timer, uint32_t n) {
  if (!c || !asleep || !timer || n != c->d.n) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    HostBodies hb = download_bodies(c);
    const BodyInputs& in = c->inputs;
    for (uint32_t i = 0; i < n; ++i) {
      const bool dyn = c->type_h[i] == BT_DYNAMIC;
      const bool to_sleep = asleep[i] != 0;
      if (dyn && to_sleep) {  // sleep_body (mod.rs:822-832)
        hb.vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        hb.ang[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        hb.pos[i].w = 0.0f;
        hb.inertia[i] = make_float4(0.0f, 0.0f, 0.0f, hb.inertia[i].w);
      } else if (dyn && hb.asleep[i]) {  // wake_restore (body.rs:412-418)
        float inv;
        f3 I;
        awake_mass(in.shape[i], f3{in.sp3[3 * i], in.sp3[3 * i + 1], in.sp3[3 * i + 2]}, BT_DYNAMIC, in.mass[i],
                   &inv, &I);
        hb.pos[i].w = inv;
        hb.inertia[i] = make_float4(I.x, I.y, I.z, hb.inertia[i].w);
      }
      hb.asleep[i] = (dyn && to_sleep) ? 1 : 0;
      hb.timer[i] = timer[i];
    }
    upload_bodies(c, hb);
  });
  return OAVBD_OK;
}

// The common body of the add_joint entry points: `prm` carries a Rope's
// maximum or a Spring's motor (already validated), as HostJoint::prm.
static int add_joint_impl(oavbd_ctx* c, uint32_t kind, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                          const float axis_a[3], const float axis_b[3], const float prm[4], uint32_t* out_index,
                          const float mot[2] = nullptr, const float sus[3] = nullptr) {
  const uint32_t n = c->d.n;
  // the reference wakes both members before it validates the spec
  // (mod.rs:1317-1318)
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    uint8_t one = 1;
    if (a < n) HIPC(hipMemcpy(c->d.b.flag_a + a, &one, 1, hipMemcpyHostToDevice));
    if (b < n) HIPC(hipMemcpy(c->d.b.flag_a + b, &one, 1, hipMemcpyHostToDevice));
    HIPC(hipDeviceSynchronize());
    k_apply_wake<<<blocks(n), TPB, 0, c->st>>>(c->d, 0u);
    HIPC(hipStreamSynchronize(c->st));
  });
  if (a == b || a >= n || b >= n) return fail(c, OAVBD_EINVAL, "bad joint bodies");
  if (c->joints.size() >= c->max_joints) return fail(c, OAVBD_ECAPACITY, "more joints than max_joints");
  HostJoint j{a, b, {la[0], la[1], la[2]}, {lb[0], lb[1], lb[2]}};
  j.kind = (int)kind;
  if (kind == JK_REVOLUTE || kind == JK_PRISMATIC) {
    // resolve_joint (joint.rs:1280-1312): a degenerate axis rejects
    if (!axis_a || !axis_b) return OAVBD_EINVAL;
    const f3 xa{axis_a[0], axis_a[1], axis_a[2]}, xb{axis_b[0], axis_b[1], axis_b[2]};
    if (dot3(xa, xa) < K_DEGENERATE || dot3(xb, xb) < K_DEGENERATE)
      return fail(c, OAVBD_EINVAL, "degenerate hinge / slide axis");
    const f3 na = normalize_or(xa, f3{0.0f, 0.0f, 1.0f}), nb = normalize_or(xb, f3{0.0f, 0.0f, 1.0f});
    j.axa[0] = na.x; j.axa[1] = na.y; j.axa[2] = na.z;
    j.axb[0] = nb.x; j.axb[1] = nb.y; j.axb[2] = nb.z;
  }
  if (prm)
    for (int k = 0; k < 4; ++k) j.prm[k] = prm[k];
  if (mot) {
    j.mot[0] = mot[0];
    j.mot[1] = mot[1];
  }
  if (kind == JK_WHEEL) {  // unit axles (already orthogonalised) and suspension
    for (int k = 0; k < 3; ++k) {
      j.axa[k] = axis_a[k];
      j.axb[k] = axis_b[k];
      j.sus[k] = sus[k];
    }
  }
  if (kind == JK_FIXED || kind == JK_DISTANCE || kind == JK_REVOLUTE || kind == JK_PRISMATIC || kind == JK_MOTOR ||
      kind == JK_WHEEL || kind == JK_SIXDOF) {
    // resolve_joint (joint.rs:1313-1332): the assembly rotation and anchor
    // offset (Fixed) or anchor distance (Distance) from the current poses
    float4 p[2], q[2];
    GUARD(c, {
      HIPC(hipMemcpy(&p[0], c->d.b.pos + a, sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(&p[1], c->d.b.pos + b, sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(&q[0], c->d.b.rot + a, sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(&q[1], c->d.b.rot + b, sizeof(float4), hipMemcpyDeviceToHost));
    });
    const qf qa = asq(q[0]), qb = asq(q[1]);
    const qf qr = qmul(qconj(qa), qb);
    const f3 ra = qrot(qa, f3{la[0], la[1], la[2]}), rb = qrot(qb, f3{lb[0], lb[1], lb[2]});
    const f3 dr = qrot(qconj(qa), sub(add(xyz(p[1]), rb), add(xyz(p[0]), ra)));
    if (kind == JK_FIXED || kind == JK_MOTOR || kind == JK_SIXDOF) {  // Motor: la = lb = 0, so dr = conj(qa) (pb - pa)
      j.qref[0] = qr.x; j.qref[1] = qr.y; j.qref[2] = qr.z; j.qref[3] = qr.w;
      j.dref[0] = dr.x; j.dref[1] = dr.y; j.dref[2] = dr.z;
    } else if (kind == JK_DISTANCE) {
      j.prm[0] = len3(sub(add(xyz(p[1]), rb), add(xyz(p[0]), ra)));
    } else if (kind == JK_REVOLUTE) {  // the assembly twist about A's hinge axis
      j.prm[0] = hinge_twist(qa, qb, f3{j.axa[0], j.axa[1], j.axa[2]});
    } else {  // the assembly slide along A's axis (a wheel's suspension)
      const f3 sa = kind == JK_WHEEL ? f3{j.sus[0], j.sus[1], j.sus[2]} : f3{j.axa[0], j.axa[1], j.axa[2]};
      const f3 wa = normalize_or(qrot(qa, sa), f3{0.0f, 0.0f, 1.0f});
      j.prm[0] = dot3(sub(add(xyz(p[1]), rb), add(xyz(p[0]), ra)), wa);
    }
  }
  c->joints.push_back(j);
  c->joints_dirty = true;
  if (out_index) *out_index = (uint32_t)c->joints.size() - 1;
  return OAVBD_OK;
}

int oavbd_add_joint(oavbd_ctx* c, uint32_t kind, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                    const float axis_a[3], const float axis_b[3], uint32_t* out_index) {
  if (!c || !la || !lb) return OAVBD_EINVAL;
  if (kind == JK_ROPE || kind == JK_SPRING)
    return fail(c, OAVBD_EINVAL, "rope and spring joints take oavbd_add_rope_joint / oavbd_add_spring_joint");
  if (kind > JK_SPRING) return fail(c, OAVBD_EINVAL, "unsupported joint kind (Motor: oavbd_add_motor_joint)");
  return add_joint_impl(c, kind, a, b, la, lb, axis_a, axis_b, nullptr, out_index);
}

int oavbd_add_ball_joint(oavbd_ctx* c, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                         uint32_t* out_index) {
  return oavbd_add_joint(c, JK_BALL, a, b, la, lb, nullptr, nullptr, out_index);
}

int oavbd_add_joint_limited(oavbd_ctx* c, uint32_t kind, uint32_t a, uint32_t b, const float la[3],
                            const float lb[3], const float axis_a[3], const float axis_b[3], const float limit[2],
                            const float motor[2], uint32_t* out_index) {
  if (!c || !la || !lb) return OAVBD_EINVAL;
  if (kind != JK_REVOLUTE && kind != JK_PRISMATIC)
    return fail(c, OAVBD_EINVAL, "limits and motors belong to Revolute and Prismatic joints");
  // validate_joint (migration.rs:160-178), before any wake
  if (limit && !(std::isfinite(limit[0]) && std::isfinite(limit[1]) && limit[0] <= limit[1]))
    return fail(c, OAVBD_EINVAL, "joint limit needs finite lo <= hi");
  if (motor && !(std::isfinite(motor[0]) && std::isfinite(motor[1]) && motor[1] >= 0.0f))
    return fail(c, OAVBD_EINVAL, "joint motor needs a finite speed and budget >= 0");
  const float flags = (float)((limit ? HINGE_LIMIT : 0) | (motor ? HINGE_MOTOR : 0));
  const float prm[4] = {0.0f, limit ? limit[0] : 0.0f, limit ? limit[1] : 0.0f, flags};
  return add_joint_impl(c, kind, a, b, la, lb, axis_a, axis_b, prm, out_index, motor);
}

int oavbd_add_gear_joint(oavbd_ctx* c, uint32_t a, uint32_t b, uint32_t joint_a, uint32_t joint_b, float ratio,
                         uint32_t* out_index) {
  if (!c) return OAVBD_EINVAL;
  // mod.rs:1329-1353, before any wake: a finite ratio over two existing
  // Revolute / Prismatic joints
  if (!std::isfinite(ratio)) return fail(c, OAVBD_EINVAL, "gear ratio must be finite");
  const uint32_t nj = (uint32_t)c->joints.size();
  if (joint_a >= nj || joint_b >= nj) return fail(c, OAVBD_EINVAL, "gear references an unknown joint");
  for (uint32_t r : {joint_a, joint_b}) {
    const int k = c->joints[r].kind;
    if (k != JK_REVOLUTE && k != JK_PRISMATIC)
      return fail(c, OAVBD_EINVAL, "gear must couple revolute / prismatic joints");
  }
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  uint32_t idx = 0;
  const int rc = add_joint_impl(c, JK_GEAR, a, b, zero, zero, nullptr, nullptr, nullptr, &idx);
  if (rc != OAVBD_OK) return rc;
  // the constant from the coupled joints' current coordinates
  // (joint_coordinate, joints_ext.rs:38-51; mod.rs:1354-1378)
  auto coordinate = [&](const HostJoint& r) -> float {
    float4 p[2], q[2];
    GUARD(c, {
      HIPC(hipMemcpy(&p[0], c->d.b.pos + r.a, sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(&p[1], c->d.b.pos + r.b, sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(&q[0], c->d.b.rot + r.a, sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(&q[1], c->d.b.rot + r.b, sizeof(float4), hipMemcpyDeviceToHost));
    });
    const qf qa = asq(q[0]), qb = asq(q[1]);
    const f3 ax{r.axa[0], r.axa[1], r.axa[2]};
    if (r.kind == JK_REVOLUTE) return hinge_twist(qa, qb, ax) - r.prm[0];
    const f3 wa = normalize_or(qrot(qa, ax), f3{0.0f, 0.0f, 1.0f});
    const f3 pa = add(xyz(p[0]), qrot(qa, f3{r.la[0], r.la[1], r.la[2]}));
    const f3 pb = add(xyz(p[1]), qrot(qb, f3{r.lb[0], r.lb[1], r.lb[2]}));
    return dot3(sub(pb, pa), wa) - r.prm[0];
  };
  const float ca = coordinate(c->joints[joint_a]);
  const float cb = coordinate(c->joints[joint_b]);
  HostJoint& j = c->joints[idx];
  j.gb[0] = joint_a;
  j.gb[1] = joint_b;
  j.prm[0] = ratio;
  j.prm[1] = ca + ratio * cb;
  if (out_index) *out_index = idx;
  return OAVBD_OK;
}

int oavbd_add_motor_joint(oavbd_ctx* c, uint32_t a, uint32_t b, const float linear_target[3],
                          const float angular_target[3], float max_force, float max_torque, uint32_t* out_index) {
  if (!c || !linear_target || !angular_target) return OAVBD_EINVAL;
  // validate_joint (migration.rs:335-350), before any wake
  bool ok = std::isfinite(max_force) && max_force >= 0.0f && std::isfinite(max_torque) && max_torque >= 0.0f;
  for (int k = 0; k < 3; ++k) ok = ok && std::isfinite(linear_target[k]) && std::isfinite(angular_target[k]);
  if (!ok) return fail(c, OAVBD_EINVAL, "motor joint needs finite targets and budgets >= 0");
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  const float ax[3] = {linear_target[0], linear_target[1], linear_target[2]};
  const float bx[3] = {angular_target[0], angular_target[1], angular_target[2]};
  uint32_t idx = 0;
  const int rc = add_joint_impl(c, JK_MOTOR, a, b, zero, zero, nullptr, nullptr, nullptr, &idx);
  if (rc != OAVBD_OK) return rc;
  HostJoint& j = c->joints[idx];  // the targets and budgets ride in the axis records
  for (int k = 0; k < 3; ++k) {
    j.axa[k] = ax[k];
    j.axb[k] = bx[k];
  }
  j.prm[0] = max_force;
  j.prm[1] = max_torque;
  if (out_index) *out_index = idx;
  return OAVBD_OK;
}

int oavbd_add_wheel_joint(oavbd_ctx* c, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                          const float suspension_a[3], const float suspension_b[3], const float axle_a[3],
                          const float axle_b[3], float frequency_hz, float damping_ratio, const float motor[2],
                          uint32_t* out_index) {
  if (!c || !la || !lb || !suspension_a || !suspension_b || !axle_a || !axle_b) return OAVBD_EINVAL;
  // validate_joint (migration.rs:288-315), before any wake
  bool ok = std::isfinite(frequency_hz) && frequency_hz >= 0.0f && std::isfinite(damping_ratio) && damping_ratio >= 0.0f;
  for (int k = 0; k < 3; ++k)
    ok = ok && std::isfinite(suspension_a[k]) && std::isfinite(suspension_b[k]) && std::isfinite(axle_a[k]) &&
         std::isfinite(axle_b[k]) && std::isfinite(la[k]) && std::isfinite(lb[k]);
  if (motor) ok = ok && std::isfinite(motor[0]) && std::isfinite(motor[1]) && motor[1] >= 0.0f;
  if (!ok) return fail(c, OAVBD_EINVAL, "wheel joint needs finite axes, a frequency and damping >= 0 and a valid motor");
  // resolve_joint (joint.rs:1352-1380): unit suspensions (fallback +y), the
  // axles orthogonalised against them
  const f3 Y{0.0f, 1.0f, 0.0f};
  const f3 sa = normalize_or(f3{suspension_a[0], suspension_a[1], suspension_a[2]}, Y);
  const f3 sb = normalize_or(f3{suspension_b[0], suspension_b[1], suspension_b[2]}, Y);
  const f3 xa = orthogonalize_axle(sa, f3{axle_a[0], axle_a[1], axle_a[2]});
  const f3 xb = orthogonalize_axle(sb, f3{axle_b[0], axle_b[1], axle_b[2]});
  // the wheel is stored as a hinge about its axle plus the suspension axis
  const float prm[4] = {0.0f, frequency_hz, damping_ratio, (float)(motor ? HINGE_MOTOR : 0)};
  const float axa[3] = {xa.x, xa.y, xa.z}, axb[3] = {xb.x, xb.y, xb.z}, sus[3] = {sa.x, sa.y, sa.z};
  return add_joint_impl(c, JK_WHEEL, a, b, la, lb, axa, axb, prm, out_index, motor, sus);
}

int oavbd_add_sixdof_joint(oavbd_ctx* c, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                           const uint32_t axes[6], const float lo[6], const float hi[6], uint32_t* out_index) {
  if (!c || !la || !lb || !axes) return OAVBD_EINVAL;
  // validate_joint (migration.rs:188-191), before any wake
  for (int k = 0; k < 6; ++k) {
    if (axes[k] > (uint32_t)AX_LIMITED) return fail(c, OAVBD_EINVAL, "bad SixDof axis configuration");
    if (axes[k] == (uint32_t)AX_LIMITED &&
        !(lo && hi && std::isfinite(lo[k]) && std::isfinite(hi[k]) && lo[k] <= hi[k]))
      return fail(c, OAVBD_EINVAL, "SixDof window needs finite lo <= hi");
  }
  uint32_t idx = 0;
  const int rc = add_joint_impl(c, JK_SIXDOF, a, b, la, lb, nullptr, nullptr, nullptr, &idx);
  if (rc != OAVBD_OK) return rc;
  HostJoint& j = c->joints[idx];
  for (int k = 0; k < 6; ++k) {
    j.six[k] = (int)axes[k];
    j.six_lo[k] = axes[k] == (uint32_t)AX_LIMITED ? lo[k] : 0.0f;
    j.six_hi[k] = axes[k] == (uint32_t)AX_LIMITED ? hi[k] : 0.0f;
  }
  if (out_index) *out_index = idx;
  return OAVBD_OK;
}

int oavbd_add_rope_joint(oavbd_ctx* c, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                         float max_distance, uint32_t* out_index) {
  if (!c || !la || !lb) return OAVBD_EINVAL;
  // JointKind::rope_checked (joint.rs:539-553), before any wake
  if (!(std::isfinite(max_distance) && max_distance > 0.0f))
    return fail(c, OAVBD_EINVAL, "rope maximum must be finite and positive");
  const float prm[4] = {max_distance, 0.0f, 0.0f, 0.0f};
  return add_joint_impl(c, JK_ROPE, a, b, la, lb, nullptr, nullptr, prm, out_index);
}

int oavbd_add_spring_joint(oavbd_ctx* c, uint32_t a, uint32_t b, const float la[3], const float lb[3],
                           const oavbd_motor* m, uint32_t* out_index) {
  if (!c || !la || !lb || !m) return OAVBD_EINVAL;
  // JointMotor::check (joint.rs:1101-1110), before any wake
  const bool ok = std::isfinite(m->target_position) && std::isfinite(m->target_velocity) &&
                  std::isfinite(m->stiffness) && m->stiffness >= 0.0f && std::isfinite(m->damping) &&
                  m->damping >= 0.0f && std::isfinite(m->max_force) && m->max_force >= 0.0f &&
                  m->kind <= OAVBD_MOTOR_SERVO && m->model <= OAVBD_MOTOR_FORCE_BASED;
  if (!ok) return fail(c, OAVBD_EINVAL, "bad spring motor");
  // JointMotor::keeps_awake (joint.rs:1116-1126)
  const bool awake = m->max_force > 0.0f && (m->kind == OAVBD_MOTOR_VELOCITY
                                                 ? m->target_velocity != 0.0f
                                                 : (m->stiffness > 0.0f || m->target_velocity != 0.0f));
  const float flags = (float)((m->model == OAVBD_MOTOR_FORCE_BASED ? SPRING_FORCE_MODEL : 0) |
                              (awake ? SPRING_KEEPS_AWAKE : 0));
  const float prm[4] = {m->target_position, m->stiffness, m->damping, flags};
  return add_joint_impl(c, JK_SPRING, a, b, la, lb, nullptr, nullptr, prm, out_index);
}

// Both tables of a double-buffered pair (or point) table hold at least
// `need` entries: the stale one is reallocated, the current one (the last
// step's, the next step's merge source) moved into a larger allocation.
static void ensure_pair_tables(oavbd_ctx* c, size_t need) {
  Dev& d = c->d;
  if (need > c->cap_pairs_old || !d.Pold.ab) {
    free_pairs(d.Pold);
    c->cap_pairs_old = grow(need);
    alloc_pairs(d.Pold, c->cap_pairs_old);
  }
  if (need > c->cap_pairs || !d.P.ab) {
    Pairs P{};
    const size_t cap = grow(need);
    alloc_pairs(P, cap);
    if (d.P.ab) {
      dcopy(c, P.ab, d.P.ab, d.npairs);
      dcopy(c, P.rec, d.P.rec, 4 * (size_t)d.npairs);
      dcopy(c, P.pts, d.P.pts, d.npairs);
      HIPC(hipStreamSynchronize(c->st));
    }
    free_pairs(d.P);
    d.P = P;
    c->cap_pairs = cap;
  }
  d.cap_pairs = (uint32_t)c->cap_pairs;
}
static void ensure_point_tables(oavbd_ctx* c, size_t need) {
  Dev& d = c->d;
  if (need > c->cap_pts_old || !d.Qold.rec) {
    free_points(d.Qold);
    c->cap_pts_old = grow(need);
    alloc_points(d.Qold, c->cap_pts_old);
  }
  if (need > c->cap_pts || !d.Q.rec) {
    Points Q{};
    const size_t cap = grow(need);
    alloc_points(Q, cap);
    if (d.Q.rec) {
      dcopy(c, Q.rec, d.Q.rec, 4 * (size_t)d.npts);
      dcopy(c, Q.roll, d.Q.roll, d.npts);
      dcopy(c, Q.pair, d.Q.pair, d.npts);
      dcopy(c, Q.src, d.Q.src, d.npts);
      HIPC(hipStreamSynchronize(c->st));
    }
    free_points(d.Q);
    d.Q = Q;
    c->cap_pts = cap;
  }
  d.cap_pts = (uint32_t)c->cap_pts;
}

// Buffers of an unwaited batch: every count-sized buffer a quarter above
// the last counts read back (steps change them slowly; a batch that still
// outgrows one is repeated), the colour classes queued per sweep one above
// the last colour count, the overflow words cleared.
static void async_prepare(oavbd_ctx* c) {
  Dev& d = c->d;
  auto room = [](uint32_t x) { return (size_t)x + x / 4u + 1024u; };
  ensure_toi(c, room(c->last_ncast));
  ensure_cand(c, room(c->last_ncand));
  ensure_pair_tables(c, room(c->last_npairs));
  ensure_point_tables(c, room(c->last_npts));
  ensure_incidence(c, std::max(c->cap_pairs, c->cap_pairs_old), std::max(c->cap_pts, c->cap_pts_old));
  c->slots = (uint32_t)c->colour_cnt_h.size() + c->slot_spare;
  zero_u32(c, d.ovf, OVF_WORDS);
}

// `nsteps` substeps.  Once a step's counts have been read back (the first
// step after an upload waits for them), the steps of a call run unwaited
// from a device checkpoint; a batch in which some count outgrew its buffer
// is restored and repeated with the host reading every count, which gives
// the same bits (the checkpoint holds everything a step reads).  Recorded
// (differentiable), timed and tuning steps always wait for their counts.
static void step_batch(oavbd_ctx* c, uint32_t nsteps) {
  if (!nsteps) return;
  if (c->joints_dirty) upload_joints(c);
  const bool async = c->async_enabled && c->sizes_known && !c->rec && !c->timing && !c->tuning &&
                     c->force_variant < 0 && !c->ev_on;
  c->stats.unwaited = async ? 1u : 0u;
  if (!async) {
    for (uint32_t s = 0; s < nsteps; ++s) step_inner(c, s + 1 == nsteps);
    return;
  }
  async_prepare(c);
  snap_save(c, c->bsnap);
  for (uint32_t s = 0; s < nsteps; ++s) step_inner(c, s + 1 == nsteps, true);
  if (c->counters_h[OVF_H + OVF_FLAGS] == 0u) return;
  ++c->replays;
  snap_load(c, c->bsnap);
  for (uint32_t s = 0; s < nsteps; ++s) step_inner(c, s + 1 == nsteps);
}

int oavbd_step_n(oavbd_ctx* c, uint32_t nsteps) {
  if (!c) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    step_batch(c, nsteps);
  });
  return OAVBD_OK;
}

int oavbd_step(oavbd_ctx* c, float dt) {
  if (!c) return OAVBD_EINVAL;
  if (!std::isfinite(dt) || dt <= 0.0f) return OAVBD_OK;
  const float DT = 1.0f / 60.0f;
  c->time_debt = fmin_(c->time_debt + dt, DT * 4.0f);
  uint32_t k = 0;
  while (c->time_debt >= DT && k < 4) {
    c->time_debt -= DT;
    ++k;
  }
  return k ? oavbd_step_n(c, k) : OAVBD_OK;
}

int oavbd_sync(oavbd_ctx* c) {
  if (!c) return OAVBD_EINVAL;
  GUARD(c, HIPC(hipStreamSynchronize(c->st)));
  return OAVBD_OK;
}

int oavbd_get_state(oavbd_ctx* c, float* out13, uint32_t n) {
  if (!c || !out13 || n > c->d.n) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    std::vector<float4> p(n), q(n), v(n), w(n);
    HIPC(hipMemcpyAsync(p.data(), c->d.b.pos, n * sizeof(float4), hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(q.data(), c->d.b.rot, n * sizeof(float4), hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(v.data(), c->d.b.vel, n * sizeof(float4), hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(w.data(), c->d.b.angvel, n * sizeof(float4), hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    for (uint32_t i = 0; i < n; ++i) {
      float* o = out13 + 13 * (size_t)i;
      o[0] = p[i].x; o[1] = p[i].y; o[2] = p[i].z;
      o[3] = q[i].x; o[4] = q[i].y; o[5] = q[i].z; o[6] = q[i].w;
      o[7] = v[i].x; o[8] = v[i].y; o[9] = v[i].z;
      o[10] = w[i].x; o[11] = w[i].y; o[12] = w[i].z;
    }
  });
  return OAVBD_OK;
}

int oavbd_set_state(oavbd_ctx* c, const float* in13, uint32_t n) {
  if (!c || !in13 || n > c->d.n) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    std::vector<float4> p(n), q(n), v(n), w(n);
    HIPC(hipMemcpy(p.data(), c->d.b.pos, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
      const float* s = in13 + 13 * (size_t)i;
      p[i] = make_float4(s[0], s[1], s[2], p[i].w);
      q[i] = make_float4(s[3], s[4], s[5], s[6]);
      v[i] = make_float4(s[7], s[8], s[9], 0.0f);
      w[i] = make_float4(s[10], s[11], s[12], 0.0f);
    }
    HIPC(hipMemcpy(c->d.b.pos, p.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d.b.rot, q.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d.b.vel, v.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d.b.angvel, w.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(hipDeviceSynchronize());  // null-stream uploads land before the context stream reads them
  });
  return OAVBD_OK;
}

int oavbd_get_pairs(oavbd_ctx* c, uint32_t* ab, uint32_t cap, uint32_t* n) {
  if (!c || !n) return OAVBD_EINVAL;
  *n = c->d.npairs;
  if (!ab) return OAVBD_OK;
  uint32_t m = std::min(cap, c->d.npairs);
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    if (m) HIPC(hipMemcpy(ab, c->d.P.ab, m * sizeof(uint2), hipMemcpyDeviceToHost));
  });
  return OAVBD_OK;
}

int oavbd_get_points(oavbd_ctx* c, float* out16, uint32_t cap, uint32_t* n) {
  if (!c || !n) return OAVBD_EINVAL;
  *n = c->d.npts;
  if (!out16) return OAVBD_OK;
  uint32_t m = std::min(cap, c->d.npts);
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    std::vector<float4> rec(4 * (size_t)m), roll(m);
    if (m) {
      HIPC(hipMemcpy(rec.data(), c->d.Q.rec, 4 * (size_t)m * sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(roll.data(), c->d.Q.roll, m * sizeof(float4), hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < m; ++i) {
      const float4& ra = rec[4 * (size_t)i];
      const float4& rb = rec[4 * (size_t)i + 1];
      const float4& lam = rec[4 * (size_t)i + 2];
      const float4& pen = rec[4 * (size_t)i + 3];
      float* o = out16 + 16 * (size_t)i;
      o[0] = ra.x; o[1] = ra.y; o[2] = ra.z;
      o[3] = rb.x; o[4] = rb.y; o[5] = rb.z;
      o[6] = lam.x; o[7] = lam.y; o[8] = lam.z;
      o[9] = pen.x; o[10] = pen.y; o[11] = pen.z;
      o[12] = roll[i].x; o[13] = roll[i].y; o[14] = roll[i].z;
      o[15] = pen.w;
    }
  });
  return OAVBD_OK;
}

int oavbd_get_colours(oavbd_ctx* c, int32_t* colour, uint32_t n) {
  if (!c || !colour || n > c->d.n) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipMemcpy(colour, c->d.b.colour, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  });
  return OAVBD_OK;
}

int oavbd_get_sleep(oavbd_ctx* c, uint8_t* asleep, float* timer, uint32_t n) {
  if (!c || n > c->d.n) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    if (asleep) HIPC(hipMemcpy(asleep, c->d.b.asleep, n, hipMemcpyDeviceToHost));
    if (timer) HIPC(hipMemcpy(timer, c->d.b.timer, n * sizeof(float), hipMemcpyDeviceToHost));
  });
  return OAVBD_OK;
}

int oavbd_get_toi(oavbd_ctx* c, float* t, int32_t* target, uint32_t n) {
  if (!c || !t || !target || n > c->d.n) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    std::vector<float4> hit(n);
    std::vector<uint32_t> tg(n);
    if (n) {
      HIPC(hipMemcpy(hit.data(), c->d.toi_hit, n * sizeof(float4), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(tg.data(), c->d.toi_tgt, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < n; ++i) {
      bool hitp = hit[i].w >= 0.0f;
      t[i] = hitp ? hit[i].w : -1.0f;
      target[i] = hitp ? (int32_t)tg[i] : -1;
    }
  });
  return OAVBD_OK;
}

int oavbd_get_joints(oavbd_ctx* c, float* out12, uint32_t n) {
  if (!c || !out12 || n > c->d.njoints) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    std::vector<float4> v[4];
    const float4* src[4] = {c->d.J.lam, c->d.J.pen, c->d.J.lama, c->d.J.pena};
    for (int k = 0; k < 4; ++k) {
      v[k].resize(n);
      if (n) HIPC(hipMemcpy(v[k].data(), src[k], n * sizeof(float4), hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < 4; ++k) {
        out12[12 * i + 3 * k] = v[k][i].x;
        out12[12 * i + 3 * k + 1] = v[k][i].y;
        out12[12 * i + 3 * k + 2] = v[k][i].z;
      }
  });
  return OAVBD_OK;
}

int oavbd_get_stats(oavbd_ctx* c, oavbd_stats* out) {
  if (!c || !out) return OAVBD_EINVAL;
  *out = c->stats;
  out->joints = (uint32_t)c->joints.size();
  out->replays = c->replays;
  return OAVBD_OK;
}

int oavbd_set_friction_axes(oavbd_ctx* c, const float* axis3, const float* transverse, uint32_t n) {
  if (!c || (n && (!axis3 || !transverse))) return OAVBD_EINVAL;
  if (n > c->inputs.n) return fail(c, OAVBD_EINVAL, "more friction axes than bodies");
  std::vector<float> ax(axis3, axis3 + 3 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i) {
    float* v = &ax[3 * i];
    if (!std::isfinite(transverse[i])) return fail(c, OAVBD_EINVAL, "transverse friction not finite");
    if (v[0] == 0.0f && v[1] == 0.0f && v[2] == 0.0f) continue;  // isotropic
    // RigidBody::set_friction_axis (body.rs:452-460): normalize_checked
    const float l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (!(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2])) || l2 < 1e-12f)
      return fail(c, OAVBD_EINVAL, "friction axis not normalizable");
    const float r = 1.0f / std::sqrt(l2);
    v[0] *= r;
    v[1] *= r;
    v[2] *= r;
  }
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    BodyInputs& in = c->inputs;
    std::copy(ax.begin(), ax.end(), in.fdir3.begin());
    std::copy(transverse, transverse + n, in.ftrans.begin());
    std::vector<float4> fd(in.n);
    c->any_axis = false;
    for (uint32_t i = 0; i < in.n; ++i) {
      fd[i] = make_float4(in.fdir3[3 * i], in.fdir3[3 * i + 1], in.fdir3[3 * i + 2], in.ftrans[i]);
      c->any_axis = c->any_axis || fd[i].x != 0.0f || fd[i].y != 0.0f || fd[i].z != 0.0f;
    }
    HIPC(hipMemcpy(c->d.b.fdir, fd.data(), in.n * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(hipDeviceSynchronize());  // the null-stream upload lands before the context stream reads it
  });
  return OAVBD_OK;
}

int oavbd_set_events(oavbd_ctx* c, int enable) {
  if (!c) return OAVBD_EINVAL;
  c->ev_on = enable != 0;
  return OAVBD_OK;
}

int oavbd_drain_events(oavbd_ctx* c, int which, oavbd_event* out, uint32_t cap, uint32_t* n) {
  if (!c || !n || (which != 0 && which != 1)) return OAVBD_EINVAL;
  auto& q = which ? c->ev_trigger : c->ev_contact;
  if (!out) {
    *n = (uint32_t)q.size();
    return OAVBD_OK;
  }
  const uint32_t k = std::min<uint32_t>(cap, (uint32_t)q.size());
  std::copy(q.begin(), q.begin() + k, out);
  q.erase(q.begin(), q.begin() + k);
  *n = k;
  return OAVBD_OK;
}

int oavbd_get_event_state(oavbd_ctx* c, int which, uint32_t* ab, uint32_t cap, uint32_t* n) {
  if (!c || !n || (which != 0 && which != 1)) return OAVBD_EINVAL;
  const auto& v = which ? c->prev_trigger : c->prev_touch;
  *n = (uint32_t)v.size();
  if (!ab) return OAVBD_OK;
  if (cap < v.size()) return fail(c, OAVBD_ECAPACITY, "event state larger than the buffer");
  for (size_t i = 0; i < v.size(); ++i) {
    ab[2 * i] = (uint32_t)(v[i] >> 32);
    ab[2 * i + 1] = (uint32_t)v[i];
  }
  return OAVBD_OK;
}

int oavbd_set_event_state(oavbd_ctx* c, int which, const uint32_t* ab, uint32_t n) {
  if (!c || (which != 0 && which != 1) || (n && !ab)) return OAVBD_EINVAL;
  std::vector<uint64_t> v;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t a = ab[2 * i], b = ab[2 * i + 1];
    if (a >= b || b >= c->inputs.n) return fail(c, OAVBD_EINVAL, "event state pair not (a < b) of live bodies");
    v.push_back(ab_key(a, b));
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  (which ? c->prev_trigger : c->prev_touch) = std::move(v);
  return OAVBD_OK;
}

int oavbd_set_unwaited(oavbd_ctx* c, int enable, uint32_t spare_colours) {
  if (!c) return OAVBD_EINVAL;
  if (spare_colours > 64u) return fail(c, OAVBD_EINVAL, "spare_colours above 64");
  c->async_enabled = enable != 0;
  c->slot_spare = spare_colours;
  return OAVBD_OK;
}

int oavbd_solve_diag(oavbd_ctx* c, const void* masses16, const void* rhs32, uint32_t n, float dt, void* delta32) {
  if (!c || (n && (!masses16 || !rhs32 || !delta32))) return OAVBD_EINVAL;
  if (!n) return OAVBD_OK;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    // Buffers live for this call only and are released on every exit path.
    struct Staging {
      float4 *mass = nullptr, *rhs = nullptr, *dl = nullptr;
      ~Staging() { dfree(mass); dfree(rhs); dfree(dl); }
    } b;
    dalloc(&b.mass, n);
    dalloc(&b.rhs, 2 * (size_t)n);
    dalloc(&b.dl, 2 * (size_t)n);
    HIPC(hipMemcpyAsync(b.mass, masses16, 16 * (size_t)n, hipMemcpyHostToDevice, c->st));
    HIPC(hipMemcpyAsync(b.rhs, rhs32, 32 * (size_t)n, hipMemcpyHostToDevice, c->st));
    k_solve_diag<<<blocks(n), TPB, 0, c->st>>>(b.mass, b.rhs, n, dt, b.dl);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(delta32, b.dl, 32 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
  });
  return OAVBD_OK;
}

int oavbd_solve_systems(oavbd_ctx* c, const void* systems96, uint32_t n, float dt, void* delta32, float* ok) {
  if (!c || (n && (!systems96 || !delta32 || !ok))) return OAVBD_EINVAL;
  if (!n) return OAVBD_OK;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    // Buffers live for this call only and are released on every exit path.
    struct Staging {
      float4 *sys = nullptr, *dl = nullptr;
      float* ok = nullptr;
      ~Staging() { dfree(sys); dfree(dl); dfree(ok); }
    } b;
    dalloc(&b.sys, 6 * (size_t)n);
    dalloc(&b.dl, 2 * (size_t)n);
    dalloc(&b.ok, n);
    HIPC(hipMemcpyAsync(b.sys, systems96, 96 * (size_t)n, hipMemcpyHostToDevice, c->st));
    k_solve_systems<<<blocks(n), TPB, 0, c->st>>>(b.sys, n, dt, b.dl, b.ok);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(delta32, b.dl, 32 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(ok, b.ok, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
  });
  return OAVBD_OK;
}

int oavbd_solve_systems_noalloc(oavbd_ctx* c, const void* systems96, uint32_t n, float dt, void* delta32,
                                float* ok) {
  if (!c || (n && (!systems96 || !delta32 || !ok))) return OAVBD_EINVAL;
  if (n > NOALLOC_MAX_SYSTEMS)
    return fail(c, OAVBD_ECAPACITY, "oavbd_solve_systems_noalloc: more than 65536 systems per call");
  if (!n) return OAVBD_OK;
  if (!c->noalloc) return fail(c, OAVBD_EDEVICE, "staging memory not resolved for this device");
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    // the three regions of this GPU's staging array; every word a launch
    // reads is written by this call first, so nothing needs clearing
    unsigned char* base = static_cast<unsigned char*>(c->noalloc);
    float4* sys = reinterpret_cast<float4*>(base);
    float4* dl = reinterpret_cast<float4*>(base + (size_t)NOALLOC_MAX_SYSTEMS * 96);
    float* okd = reinterpret_cast<float*>(base + (size_t)NOALLOC_MAX_SYSTEMS * (96 + 32));
    HIPC(hipMemcpyAsync(sys, systems96, 96 * (size_t)n, hipMemcpyHostToDevice, c->st));
    k_solve_systems<<<blocks(n), TPB, 0, c->st>>>(sys, n, dt, dl, okd);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(delta32, dl, 32 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(ok, okd, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
  });
  return OAVBD_OK;
}

int oavbd_checkpoint_save(oavbd_ctx* c, uint32_t slot) {
  if (!c || slot > 4096) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    if (c->snaps.size() <= slot) c->snaps.resize(slot + 1);
    snap_save(c, c->snaps[slot]);
    HIPC(hipStreamSynchronize(c->st));
  });
  return OAVBD_OK;
}

int oavbd_checkpoint_load(oavbd_ctx* c, uint32_t slot) {
  if (!c) return OAVBD_EINVAL;
  if (slot >= c->snaps.size() || !c->snaps[slot].valid) return fail(c, OAVBD_EINVAL, "no checkpoint in that slot");
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    snap_load(c, c->snaps[slot]);
    HIPC(hipStreamSynchronize(c->st));
  });
  return OAVBD_OK;
}

int oavbd_diff_forward(oavbd_ctx* c, uint32_t nsteps, uint32_t every) {
  if (!c || every == 0) return OAVBD_EINVAL;
  if (c->d.njoints) return fail(c, OAVBD_EINVAL, "differentiable stepping covers contact scenes (no joints)");
  if (c->rows6) return fail(c, OAVBD_EINVAL, "differentiable stepping covers sliding friction (no rolling/torsion)");
  if (c->any_axis) return fail(c, OAVBD_EINVAL, "differentiable stepping covers isotropic friction (no axes)");
  for (int t : c->type_h)
    if (t == BT_KINEMATIC) return fail(c, OAVBD_EINVAL, "differentiable stepping has no kinematic bodies");
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    const uint32_t nseg = (nsteps + every - 1) / every;
    for (size_t k = nseg + 1; k < c->snaps.size(); ++k) snap_free(c->snaps[k]);
    c->snaps.resize(nseg + 1);
    c->diff_max_pairs = c->d.npairs;
    c->diff_max_pts = c->d.npts;
    for (uint32_t s = 0; s < nsteps; ++s) {
      if (s % every == 0) snap_save(c, c->snaps[s / every]);
      step_inner(c, s + 1 == nsteps);
    }
    snap_save(c, c->snaps[nseg]);
    HIPC(hipStreamSynchronize(c->st));
    c->diff_steps = nsteps;
    c->diff_every = every;
  });
  return OAVBD_OK;
}

int oavbd_diff_backward(oavbd_ctx* c, const float* dfinal13, float* dinit13, float* dmass, float* dfriction) {
  if (!c || !dfinal13) return OAVBD_EINVAL;
  if (!c->diff_steps) return fail(c, OAVBD_EINVAL, "oavbd_diff_forward has not run");
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    const uint32_t n = c->d.n, every = c->diff_every, nsteps = c->diff_steps;
    const uint32_t nseg = (nsteps + every - 1) / every;
    adj_alloc(c);
    Adj& a = c->adj;
    {
      std::vector<float4> xb(2 * (size_t)n), vb(2 * (size_t)n);
      for (uint32_t i = 0; i < n; ++i) {
        const float* g = dfinal13 + 13 * (size_t)i;
        xb[2 * i] = make_float4(g[0], g[1], g[2], 0.0f);
        xb[2 * i + 1] = make_float4(g[3], g[4], g[5], g[6]);
        vb[2 * i] = make_float4(g[7], g[8], g[9], 0.0f);
        vb[2 * i + 1] = make_float4(g[10], g[11], g[12], 0.0f);
      }
      HIPC(hipMemcpyAsync(a.xb, xb.data(), xb.size() * sizeof(float4), hipMemcpyHostToDevice, c->st));
      HIPC(hipMemcpyAsync(a.vb, vb.data(), vb.size() * sizeof(float4), hipMemcpyHostToDevice, c->st));
    }
    zero_f4(c, a.x0b, 2 * (size_t)n);
    zero_f4(c, a.yb, 2 * (size_t)n);
    zero_f4(c, a.lb, 2 * c->adj_cap_pts);
    zero_f4(c, a.lb_prev, 2 * c->adj_cap_pts);
    zero_f4(c, a.pb, c->adj_cap_pairs);
    k_zero_f<<<blocks(n), TPB, 0, c->st>>>(a.mass, n);
    k_zero_f<<<blocks(n), TPB, 0, c->st>>>(a.fric, n);
    if (c->tapes.size() < every) c->tapes.resize(every);
    for (int seg = (int)nseg - 1; seg >= 0; --seg) {
      const uint32_t len = std::min(every, nsteps - (uint32_t)seg * every);
      snap_load(c, c->snaps[seg]);
      for (uint32_t k = 0; k < len; ++k) {
        c->rec = &c->tapes[k];
        step_inner(c, true);
        c->rec = nullptr;
        if (c->stats.toi_hits) throw HipError{"a TOI clamp inside the differentiated steps (not differentiable)"};
      }
      // the recorded steps may have grown the pair / point maxima
      if (c->diff_max_pts > c->adj_cap_pts || c->diff_max_pairs > c->adj_cap_pairs)
        throw HipError{"recorded steps exceed the forward pass's pair / point counts"};
      for (int k = (int)len - 1; k >= 0; --k) adjoint_step(c, c->tapes[k]);
    }
    snap_load(c, c->snaps[nseg]);
    std::vector<float4> xb(2 * (size_t)n), vb(2 * (size_t)n);
    std::vector<float> m(n), f(n);
    HIPC(hipMemcpyAsync(xb.data(), a.xb, xb.size() * sizeof(float4), hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(vb.data(), a.vb, vb.size() * sizeof(float4), hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(m.data(), a.mass, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(f.data(), a.fric, n * sizeof(float), hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    for (uint32_t i = 0; i < n; ++i) {
      if (dinit13) {
        float* o = dinit13 + 13 * (size_t)i;
        o[0] = xb[2 * i].x; o[1] = xb[2 * i].y; o[2] = xb[2 * i].z;
        o[3] = xb[2 * i + 1].x; o[4] = xb[2 * i + 1].y; o[5] = xb[2 * i + 1].z; o[6] = xb[2 * i + 1].w;
        o[7] = vb[2 * i].x; o[8] = vb[2 * i].y; o[9] = vb[2 * i].z;
        o[10] = vb[2 * i + 1].x; o[11] = vb[2 * i + 1].y; o[12] = vb[2 * i + 1].z;
      }
      if (dmass) dmass[i] = m[i];
      if (dfriction) dfriction[i] = f[i];
    }
  });
  return OAVBD_OK;
}

int oavbd_get_kernel_times(oavbd_ctx* c, oavbd_kernel_times* out, int reset) {
  if (!c || !out) return OAVBD_EINVAL;
  *out = c->kt;
  if (reset) c->kt = oavbd_kernel_times{};
  return OAVBD_OK;
}

int oavbd_set_primal_shape(oavbd_ctx* c, uint32_t lanes, uint32_t waves) {
  if (!c) return OAVBD_EINVAL;
  if (lanes != BODY_LANES) return fail(c, OAVBD_EINVAL, "the primal solve runs 16 lanes per body");
  if (waves != 1 && waves != 2 && waves != 4 && waves != 8) return OAVBD_EINVAL;
  c->waves = (int)waves;
  return OAVBD_OK;
}

int oavbd_set_primal_table(oavbd_ctx* c, const int32_t* table, uint32_t n) {
  if (!c || (n && !table) || n > PRIMAL_BUCKETS) return OAVBD_EINVAL;
  for (uint32_t b = 0; b < n; ++b)
    if (table[b] < -1 || table[b] >= PV_COUNT) return fail(c, OAVBD_EINVAL, "bad primal variant");
  for (uint32_t b = 0; b < PRIMAL_BUCKETS; ++b) c->ptable[b] = b < n ? (int8_t)table[b] : (int8_t)-1;
  return OAVBD_OK;
}

int oavbd_set_primal_tail(oavbd_ctx* c, uint32_t max_bodies, uint32_t max_waves) {
  if (!c) return OAVBD_EINVAL;
  c->tail_max = max_bodies;
  c->tail_waves = max_waves;
  return OAVBD_OK;
}

int oavbd_get_primal_table(oavbd_ctx* c, int32_t* table, uint32_t n) {
  if (!c || !table || n > PRIMAL_BUCKETS) return OAVBD_EINVAL;
  for (uint32_t b = 0; b < n; ++b) table[b] = c->ptable[b];
  return OAVBD_OK;
}

// Runs the next `nsteps` steps once per primal variant from the same saved
// state, every primal launch timed with HIP events and credited to its
// class-size bucket; each bucket then takes the variant with the least mean
// launch time, and buckets no launch fell in take the nearest measured
// bucket's choice.  The state (bodies, pairs, warm starts, sleep), the
// statistics and the kernel-time accumulators are restored afterwards, so
// the call changes nothing but the table (all variants give the same bits).
int oavbd_autotune(oavbd_ctx* c, uint32_t nsteps, int32_t* table_out, float* mean_us_out) {
  if (!c || nsteps == 0 || nsteps > 64) return OAVBD_EINVAL;
  GUARD(c, {
    HIPC(hipSetDevice(c->device));
    HIPC(hipStreamSynchronize(c->st));
    if (c->joints_dirty) upload_joints(c);
    Snap snap;
    snap_save(c, snap);
    const oavbd_stats st0 = c->stats;
    const oavbd_kernel_times kt0 = c->kt;
    const bool timing0 = c->timing, may_sleep0 = c->may_sleep;
    const uint32_t sleeping0 = c->sleeping;
    float ms0[5];
    std::copy(c->ms, c->ms + 5, ms0);
    for (int v = 0; v < PV_COUNT; ++v)
      for (int b = 0; b < PRIMAL_BUCKETS; ++b) {
        c->tune_ms[v][b] = 0.0;
        c->tune_n[v][b] = 0;
      }
    c->tuning = true;
    c->timing = true;
    try {
      for (int v = 0; v < PV_COUNT; ++v) {
        snap_load(c, snap);
        c->may_sleep = may_sleep0;
        c->sleeping = sleeping0;
        c->force_variant = v;
        for (uint32_t k = 0; k < nsteps; ++k) step_inner(c, true);
      }
    } catch (...) {
      c->tuning = false;
      c->force_variant = -1;
      snap_free(snap);
      throw;
    }
    c->tuning = false;
    c->force_variant = -1;
    snap_load(c, snap);
    HIPC(hipStreamSynchronize(c->st));
    snap_free(snap);
    c->may_sleep = may_sleep0;
    c->sleeping = sleeping0;
    c->stats = st0;
    c->kt = kt0;
    c->timing = timing0;
    std::copy(ms0, ms0 + 5, c->ms);
    int8_t pick[PRIMAL_BUCKETS];
    for (int b = 0; b < PRIMAL_BUCKETS; ++b) {
      pick[b] = -1;
      double best = 0.0;
      for (int v = 0; v < PV_COUNT; ++v) {
        if (!c->tune_n[v][b]) continue;
        const double mean = c->tune_ms[v][b] / c->tune_n[v][b];
        if (pick[b] < 0 || mean < best) {
          best = mean;
          pick[b] = (int8_t)v;
        }
      }
    }
    for (int b = 0; b < PRIMAL_BUCKETS; ++b) {
      int8_t v = pick[b];
      for (int r = 1; v < 0 && r < PRIMAL_BUCKETS; ++r) {
        if (b - r >= 0 && pick[b - r] >= 0)
          v = pick[b - r];
        else if (b + r < PRIMAL_BUCKETS && pick[b + r] >= 0)
          v = pick[b + r];
      }
      c->ptable[b] = v;
    }
    if (table_out)
      for (int b = 0; b < PRIMAL_BUCKETS; ++b) table_out[b] = c->ptable[b];
    if (mean_us_out)
      for (int v = 0; v < PV_COUNT; ++v)
        for (int b = 0; b < PRIMAL_BUCKETS; ++b)
          mean_us_out[v * PRIMAL_BUCKETS + b] =
              c->tune_n[v][b] ? (float)(1e3 * c->tune_ms[v][b] / c->tune_n[v][b]) : -1.0f;
  });
  return OAVBD_OK;
}

int oavbd_set_timing(oavbd_ctx* c, int enable) {
  if (!c) return OAVBD_EINVAL;
  c->timing = enable != 0;
  return OAVBD_OK;
}

int oavbd_get_timing(oavbd_ctx* c, float* ms5) {
  if (!c || !ms5) return OAVBD_EINVAL;
  for (int k = 0; k < 5; ++k) ms5[k] = c->ms[k];
  return OAVBD_OK;
}

const char* oavbd_last_error(oavbd_ctx* c) {
  if (!c) return "null context";
  return c->err.c_str();
}

}  // extern "C"
