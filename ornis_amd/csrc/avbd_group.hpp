// avbd_group.hpp — box-box separating axes and closest features evaluated
// by a group of G lanes of one wavefront (G = 16, 32 or 64).
//
// The serial separating_axis / feature_distance (avbd_geom.hpp) scan a
// fixed list of candidates (15 axes; 8 + 8 corner-box and 144 edge-edge
// tests) and keep the first best one.  Here each lane evaluates some of the
// candidates with the same arithmetic and keeps its own first best, and the
// group picks (value, candidate index) lexicographically, so the winner —
// and every bit of the result — is the serial one.  Every lane of a group
// ends with the same result; all lanes of a group must execute these
// functions together.
#pragma once

#include "avbd_geom.hpp"

namespace grp {

// v of lane `src` (0..G-1) of the calling lane's group.  For a whole
// wavefront src is uniform (every caller takes it from a group pick), so
// a lane read replaces the permute.
template <int G>
__device__ __forceinline__ f3 from_lane(f3 v, int src) {
  if constexpr (G == 64) {
    return f3{__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), src)),
              __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), src)),
              __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.z), src))};
  } else {
    return f3{__shfl(v.x, src, G), __shfl(v.y, src, G), __shfl(v.z, src, G)};
  }
}

// Group reductions on the DPP paths (no LDS round trip), for groups of 16,
// 32 or 64 lanes aligned in the wavefront.  Within each row of 16 lanes,
// shifts by 1, 2, 4 and 8 leave the row's result in its lane 15 (a lane
// whose source lies outside the row keeps its own value); the row
// broadcasts then carry it to lane 31 / 63 for groups of 32 / 64, and every
// lane reads its group's result from that lane.  Only selections (min, max,
// or, lexicographic picks) are reduced, so the grouping cannot change a bit
// of the result.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xf, false);
}
// Every lane's copy of its group's result lane (15, 31 or 63 of the group).
template <int G>
__device__ __forceinline__ int group_last(int v) {
  if constexpr (G == 64) {
    return __builtin_amdgcn_readlane(v, 63);
  } else if constexpr (G == 32) {
    const int a = __builtin_amdgcn_readlane(v, 31), b = __builtin_amdgcn_readlane(v, 63);
    return (__lane_id() & 32u) ? b : a;
  } else {
    static_assert(G == 16, "groups of 16, 32 or 64 lanes");
    const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31);
    const int r2 = __builtin_amdgcn_readlane(v, 47), r3 = __builtin_amdgcn_readlane(v, 63);
    const uint32_t row = __lane_id() >> 4;
    return row == 0u ? r0 : row == 1u ? r1 : row == 2u ? r2 : r3;
  }
}
// Pick over the group of (v, i, l) triples: better(ov, oi, v, i) tells
// whether the other triple wins.  Every lane ends with the group's pick.
template <int G, class Better>
__device__ __forceinline__ void group_pick(int& v, int& i, int& l, Better better) {
#define OAVBD_PICK(CTRL, ROWS)                                                                    \
  {                                                                                              \
    const int ov = dpp_mov<CTRL, ROWS>(v), oi = dpp_mov<CTRL, ROWS>(i), ol = dpp_mov<CTRL, ROWS>(l); \
    if (better(ov, oi, v, i)) {                                                                  \
      v = ov;                                                                                    \
      i = oi;                                                                                    \
      l = ol;                                                                                    \
    }                                                                                            \
  }
  OAVBD_PICK(0x111, 0xf)
  OAVBD_PICK(0x112, 0xf)
  OAVBD_PICK(0x114, 0xf)
  OAVBD_PICK(0x118, 0xf)
  if constexpr (G >= 32) OAVBD_PICK(0x142, 0xa)  // row_bcast:15 into rows 1 and 3
  if constexpr (G == 64) OAVBD_PICK(0x143, 0xc)  // row_bcast:31 into rows 2 and 3
#undef OAVBD_PICK
  v = group_last<G>(v);
  i = group_last<G>(i);
  l = group_last<G>(l);
}
// Minimum (a NaN loses to any number) and "any" over the group.
template <int G>
__device__ __forceinline__ float group_fmin(float x) {
  int v = __float_as_int(x), i = 0, l = 0;
  group_pick<G>(v, i, l, [](int ov, int, int cv, int) {
    const float o = __int_as_float(ov), c = __int_as_float(cv);
    return o < c || (c != c && o == o);
  });
  return __int_as_float(v);
}
template <int G>
__device__ __forceinline__ bool group_any(bool x) {
  int v = x ? 1 : 0, i = 0, l = 0;
  group_pick<G>(v, i, l, [](int ov, int, int cv, int) { return ov > cv; });
  return v != 0;
}

// separating_axis over the group.  With sweep != nullptr it also tells
// whether box a, translated along `dir` by up to `rem`, stays farther than
// `reach` from b along one of the axes: along a fixed axis the projected
// separation is linear in the translation, so its minimum over the sweep is
// at an end or zero where it changes sign; `slack` bounds the rounding.
template <int G>
__device__ float sat(const shp& a, f3 ha, const shp& b, f3 hb, f3* normal_out, int s, bool* sweep = nullptr,
                     f3 dir = f3{0.0f, 0.0f, 0.0f}, float rem = 0.0f, float reach = 0.0f, float slack = 0.0f) {
  m3c ra = mat3_from_quat(a.rot), rb = mat3_from_quat(b.rot);
  f3 delta = sub(b.pos, a.pos);
  float best = -__builtin_inff();
  int bk = 99;
  f3 bn{0.0f, 1.0f, 0.0f};
  bool clear = false;
  for (int k = s; k < 15; k += G) {
    f3 n;
    if (k < 3) {
      n = mcol(ra, k);
    } else if (k < 6) {
      n = mcol(rb, k - 3);
    } else {
      int i = (k - 6) / 3, j = (k - 6) % 3;
      f3 c = cross3(mcol(ra, i), mcol(rb, j));
      if (!(len2(c) > 1e-12f)) continue;
      n = normalize3(c);
    }
    if (n.x == 0.0f && n.y == 0.0f && n.z == 0.0f) continue;
    float ea = ha.x * fabsf(dot3(n, ra.c0)) + ha.y * fabsf(dot3(n, ra.c1)) + ha.z * fabsf(dot3(n, ra.c2));
    float eb = hb.x * fabsf(dot3(n, rb.c0)) + hb.y * fabsf(dot3(n, rb.c1)) + hb.z * fabsf(dot3(n, rb.c2));
    float proj = dot3(delta, n);
    float gap = fabsf(proj) - ea - eb;
    if (gap > best) {
      best = gap;
      bk = k;
      bn = proj >= 0.0f ? n : neg(n);
    }
    if (sweep) {
      float q2 = proj - rem * dot3(dir, n);
      bool same = (proj > 0.0f && q2 > 0.0f) || (proj < 0.0f && q2 < 0.0f);
      float lo = same ? fminf(fabsf(proj), fabsf(q2)) : 0.0f;
      if (lo - ea - eb - slack > reach) clear = true;
    }
  }
  if (sweep) *sweep = group_any<G>(clear);
  // larger gap wins; equal gaps go to the earlier axis
  int gv = __float_as_int(best), k = bk, src = s;
  group_pick<G>(gv, k, src, [](int og, int ok, int cg, int ck) {
    const float o = __int_as_float(og), c = __int_as_float(cg);
    return o > c || (o == c && ok < ck);
  });
  if (k != 99) {
    bn = from_lane<G>(bn, k % G);
  } else {
    bn = f3{0.0f, 1.0f, 0.0f};
  }
  *normal_out = bn;
  return __int_as_float(gv);
}

// The calling lane's group's bits of a wave-wide ballot (bit k: lane k of
// the group).
template <int G>
__device__ __forceinline__ uint64_t ballot(bool p) {
  const uint64_t b = __ballot(p);
  if constexpr (G == 64) return b;
  return (b >> (__lane_id() & ~(unsigned)(G - 1))) & ((1ull << G) - 1ull);
}

// Index of the n-th set bit (n < popcount(m)).
__device__ __forceinline__ int nth_bit(uint32_t m, uint32_t n) {
  for (uint32_t q = 0; q < n; ++q) m &= m - 1u;
  return __builtin_ctz(m);
}

// feature_distance over the group with the edge pairs pruned (G >= 16).
// Lane s < 16 evaluates corner candidate s; their minimum U bounds the
// result from above.  An edge whose lower-bound distance to the other box
// (edge_box_gap) exceeds sqrt(U) by more than `eps` cannot hold the minimum,
// nor can any edge pair it belongs to: eps covers the rounding of the
// bound, of the box corners and of a candidate's closest points (a few ulps
// of the coordinate scale, against eps = 2e-5 of it), so a pruned pair's
// computed |y - x|^2 is strictly above U.  The bound takes both rotations
// as rigid, so nothing is pruned unless both quaternions are unit to 1e-6
// (a quaternion off unit length by d distorts the corners by up to 2d of
// the extents).  The surviving pairs are dealt out to the lanes in index
// order and the (value, index) reduction picks the candidate — and the
// bits — of the full scan.
template <int G>
__device__ dist_t feature_pruned(const shp& a, f3 ha, const shp& b, f3 hb, int s) {
  static_assert(G >= 16 && G <= 64, "one corner candidate per lane");
  const box_axes_t ea = box_axes(ha, a.rot), eb = box_axes(hb, b.rot);
  float best = __builtin_inff();
  int bc = 0x7fffffff;
  f3 bx = a.pos, by = b.pos;
  if (s < 16) {
    f3 x, y;
    const float sq = feature_candidate(a, ha, ea, b, hb, eb, s, &x, &y);
    if (sq < best) {
      best = sq;
      bc = s;
      bx = x;
      by = y;
    }
  }
  const float U = group_fmin<G>(best);
  const float pa = fmaxf(fmaxf(fabsf(a.pos.x), fabsf(a.pos.y)), fabsf(a.pos.z));
  const float pb = fmaxf(fmaxf(fabsf(b.pos.x), fabsf(b.pos.y)), fabsf(b.pos.z));
  const float eps = 2e-5f * (1.0f + fmaxf(pa, pb) + 2.0f * (ha.x + ha.y + ha.z + hb.x + hb.y + hb.z));
  const float na = (a.rot.x * a.rot.x + a.rot.y * a.rot.y) + (a.rot.z * a.rot.z + a.rot.w * a.rot.w);
  const float nq = (b.rot.x * b.rot.x + b.rot.y * b.rot.y) + (b.rot.z * b.rot.z + b.rot.w * b.rot.w);
  const bool rigid = fabsf(na - 1.0f) < 1e-6f && fabsf(nq - 1.0f) < 1e-6f;
  auto keep = [&](float gap) {
    const float r = gap - eps;
    return !(rigid && r > 0.0f && r * r * 0.99999f > U);
  };
  bool ka = false, kb = false;
  if (s < 12) {
    ka = keep(edge_box_gap(a.pos, ea, s, b, hb));
    kb = keep(edge_box_gap(b.pos, eb, s, a, ha));
  }
  const uint32_t ma = (uint32_t)ballot<G>(ka), mb = (uint32_t)ballot<G>(kb);
  const uint32_t nb = __popc(mb), total = __popc(ma) * nb;
  // r / nb for r < 144, nb <= 12: (r * ceil(2^16 / nb)) >> 16 is exact
  const uint32_t inv = nb ? (65536u + nb - 1u) / nb : 0u;
  for (uint32_t r = (uint32_t)s; r < total; r += G) {
    const uint32_t qa = (r * inv) >> 16;
    const int i = nth_bit(ma, qa), j = nth_bit(mb, r - qa * nb);
    int a0, a1, b0, b1;
    box_edge(i, &a0, &a1);
    box_edge(j, &b0, &b1);
    f3 x, y;
    seg_seg_closest(box_corner(a.pos, ea, a0), box_corner(a.pos, ea, a1), box_corner(b.pos, eb, b0),
                    box_corner(b.pos, eb, b1), &x, &y);
    const float sq = len2(sub(y, x));
    if (sq < best) {  // candidates rise along the lane's sequence
      best = sq;
      bc = 16 + i * 12 + j;
      bx = x;
      by = y;
    }
  }
  // smaller squared distance wins; ties go to the earlier candidate
  int qv = __float_as_int(best), c = bc, src = s;
  group_pick<G>(qv, c, src, [](int oq, int oc, int cq, int cc) {
    const float o = __int_as_float(oq), q = __int_as_float(cq);
    return o < q || (o == q && oc < cc);
  });
  if (c == 0x7fffffff) return dist_t{__builtin_inff(), a.pos, b.pos};
  return dist_t{sqrtf(__int_as_float(qv)), from_lane<G>(bx, src), from_lane<G>(by, src)};
}

// box_box_from_sat over the group (G >= 16: pruned feature scan).
template <int G>
__device__ dist_t box_box(const shp& a, f3 ha, const shp& b, f3 hb, float gap, f3 normal, int s) {
  if (gap > 0.0f) return feature_pruned<G>(a, ha, b, hb, s);
  shp fa, fb;
  f3 ea, eb;
  support_face(a, ha, normal, &fa, &ea);
  support_face(b, hb, neg(normal), &fb, &eb);
  dist_t r = feature_pruned<G>(fa, ea, fb, eb, s);
  r.dist = gap;
  return r;
}

}  // namespace grp
