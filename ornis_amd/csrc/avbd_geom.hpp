// avbd_geom.hpp — exact distance queries between placed boxes, spheres and
// half-spaces (the narrowphase of the AVBD step), device + host.
//
// Each function follows the operation order of its reference counterpart
// so that witnesses and gaps are bit-identical:
//   seg_seg_closest      crates/physics/src/collision/distance.rs:75-113
//   point_obb_closest    distance.rs:117-138
//   box_corner           obb_corners, distance.rs:142-156
//   sphere_sphere        distance.rs:180-189
//   sphere_obb           distance.rs:192-216
//   halfspace_distance   distance.rs:607-641 (+ gjk::support, gjk.rs:68-79)
//   separating_axis      crates/physics/src/collision/distance_box.rs:21-52
//   support_face         distance_box.rs:81-98
//   feature_distance     distance_box.rs:100-133
//   box_box              distance_box.rs:135-147
//   shape_distance       distance.rs:654-823 (box/sphere/half-space arms)
#pragma once

#include "avbd_math.hpp"

enum { SH_SPHERE = 0, SH_BOX = 1, SH_HALFSPACE = 2 };

struct shp {
  int kind;
  f3 p;  // sphere (r,0,0) / box half extents / half-space unit normal
  f3 pos;
  qf rot;
};

struct dist_t {
  float dist;
  f3 pa, pb;
};

HD float bound_radius(int kind, f3 p) {
  if (kind == SH_SPHERE) return p.x;
  if (kind == SH_BOX) return len3(p);
  return __builtin_inff();
}

HD float shape_min_dim(int kind, f3 p) {
  if (kind == SH_SPHERE) return p.x;
  if (kind == SH_BOX) return fmin_(p.x, fmin_(p.y, p.z));
  return __builtin_inff();
}

HD void seg_seg_closest(f3 a0, f3 a1, f3 b0, f3 b1, f3* x, f3* y) {
  const float EPS = 1e-12f;
  f3 d1 = sub(a1, a0), d2 = sub(b1, b0), r = sub(a0, b0);
  float a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
  if (a < EPS && e < EPS) {
    *x = a0;
    *y = b0;
    return;
  }
  if (a < EPS) {
    float t = clampf_(f / e, 0.0f, 1.0f);
    *x = a0;
    *y = add(b0, mul(d2, t));
    return;
  }
  float c = dot3(d1, r);
  if (e < EPS) {
    float t = clampf_(-c / a, 0.0f, 1.0f);
    *x = add(a0, mul(d1, t));
    *y = b0;
    return;
  }
  float b = dot3(d1, d2);
  float denom = a * e - b * b;
  float s = fabsf(denom) > EPS ? clampf_((b * f - c * e) / denom, 0.0f, 1.0f) : 0.0f;
  float t = e > EPS ? (b * s + f) / e : 0.0f;
  if (!(t >= 0.0f && t <= 1.0f)) {
    t = clampf_(t, 0.0f, 1.0f);
    s = clampf_((b * t - c) / a, 0.0f, 1.0f);
  }
  *x = add(a0, mul(d1, s));
  *y = add(b0, mul(d2, t));
}

HD f3 point_obb_closest(f3 p, f3 pos, f3 half, qf rot) {
  f3 l = qrot(qconj(rot), sub(p, pos));
  f3 cl = min3(max3(l, neg(half)), half);
  f3 q;
  if (eq3(cl, l)) {
    float dx = half.x - fabsf(l.x), dy = half.y - fabsf(l.y), dz = half.z - fabsf(l.z);
    q = l;
    if (dx <= dy && dx <= dz)
      q.x = copysignf(half.x, l.x);
    else if (dy <= dz)
      q.y = copysignf(half.y, l.y);
    else
      q.z = copysignf(half.z, l.z);
  } else {
    q = cl;
  }
  return add(pos, qrot(rot, q));
}

HD dist_t sphere_sphere(const shp& a, float ra, const shp& b, float rb) {
  f3 d = sub(b.pos, a.pos);
  float cd = len3(d);
  f3 n = normalize_or(d, f3{1.0f, 0.0f, 0.0f});
  return dist_t{cd - ra - rb, add(a.pos, mul(n, ra)), sub(b.pos, mul(n, rb))};
}

HD dist_t sphere_obb(const shp& a, float r, const shp& b, f3 half) {
  f3 pb = point_obb_closest(a.pos, b.pos, half, b.rot);
  f3 dv = sub(a.pos, pb);
  float core = len3(dv);
  f3 local = qrot(qconj(b.rot), sub(a.pos, b.pos));
  f3 al = abs3(local);
  bool inside = al.x <= half.x && al.y <= half.y && al.z <= half.z;
  f3 n;
  if (inside) {
    f3 cl = sub(half, al);
    f3 axis;
    if (cl.x <= cl.y && cl.x <= cl.z) {
      float s = signumf(local.x);
      axis = f3{1.0f * s, 0.0f * s, 0.0f * s};
    } else if (cl.y <= cl.z) {
      float s = signumf(local.y);
      axis = f3{0.0f * s, 1.0f * s, 0.0f * s};
    } else {
      float s = signumf(local.z);
      axis = f3{0.0f * s, 0.0f * s, 1.0f * s};
    }
    n = normalize_or(sub(pb, a.pos), qrot(b.rot, axis));
  } else {
    n = normalize_or(dv, f3{1.0f, 0.0f, 0.0f});
  }
  return dist_t{inside ? -core - r : core - r, sub(a.pos, mul(n, r)), pb};
}

HD f3 support_pt(const shp& s, f3 dir) {
  if (s.kind == SH_SPHERE) return add(s.pos, mul(normalize_or(dir, f3{1.0f, 0.0f, 0.0f}), s.p.x));
  f3 local = qrot(qconj(s.rot), dir);
  f3 sg{local.x >= 0.0f ? 1.0f : -1.0f, local.y >= 0.0f ? 1.0f : -1.0f, local.z >= 0.0f ? 1.0f : -1.0f};
  return add(s.pos, qrot(s.rot, cmul(s.p, sg)));
}

HD dist_t halfspace_distance(const shp& hs, const shp& other, bool first) {
  const float INF = __builtin_inff();
  dist_t none{INF, first ? hs.pos : other.pos, first ? other.pos : hs.pos};
  if (other.kind == SH_HALFSPACE) return none;
  f3 n = qrot(hs.rot, hs.p);
  if (!finite3(n)) return none;
  f3 s = support_pt(other, neg(n));
  float dist = dot3(sub(s, hs.pos), n);
  f3 on = sub(s, mul(n, dist));
  if (first) return dist_t{dist, on, s};
  return dist_t{dist, s, on};
}

HD float separating_axis(const shp& a, f3 ha, const shp& b, f3 hb, f3* normal_out) {
  m3c ra = mat3_from_quat(a.rot), rb = mat3_from_quat(b.rot);
  f3 delta = sub(b.pos, a.pos);
  float best = -__builtin_inff();
  f3 bn{0.0f, 1.0f, 0.0f};
  for (int k = 0; k < 15; ++k) {
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
      bn = proj >= 0.0f ? n : neg(n);
    }
  }
  if (normal_out) *normal_out = bn;
  return best;
}

HD void support_face(const shp& body, f3 half, f3 dir, shp* out, f3* ext) {
  f3 local = qrot(qconj(body.rot), dir);
  f3 off = zero3();
  f3 e = half;
  for (int i = 0; i < 3; ++i) {
    float li = comp(local, i);
    if (fabsf(li) > 1e-6f) {
      setc(off, i, signumf(li) * comp(half, i));
      setc(e, i, 0.0f);
    }
  }
  *out = body;
  out->pos = add(body.pos, qrot(body.rot, off));
  *ext = e;
}

// Box axes scaled by the half extents, as obb_corners (distance.rs:142-156)
// forms them.
struct box_axes_t {
  f3 x, y, z;
};
HD box_axes_t box_axes(f3 half, qf rot) {
  return box_axes_t{qrot(rot, f3{1.0f * half.x, 0.0f * half.x, 0.0f * half.x}),
                    qrot(rot, f3{0.0f * half.y, 1.0f * half.y, 0.0f * half.y}),
                    qrot(rot, f3{0.0f * half.z, 0.0f * half.z, 1.0f * half.z})};
}
// Corner k of obb_corners: bit 2 selects -x, bit 1 -y, bit 0 -z, summed in
// the same order, so every corner is bit-identical to the reference's.
HD f3 box_corner(f3 pos, const box_axes_t& e, int k) {
  f3 c = (k & 4) ? sub(pos, e.x) : add(pos, e.x);
  c = (k & 2) ? sub(c, e.y) : add(c, e.y);
  return (k & 1) ? sub(c, e.z) : add(c, e.z);
}
// Box edge i as the corner pair (ea_[i], eb_[i]) of feature_distance:
// ea_ = {0,1,2,3, 0,1,4,5, 0,2,4,6}, eb_ = {4,5,6,7, 2,3,6,7, 1,3,5,7}.
HD void box_edge(int i, int* c0, int* c1) {
  if (i < 4) {
    *c0 = i;
    *c1 = i + 4;
  } else if (i < 8) {
    int m = i - 4;
    *c0 = (m & 1) | ((m & 2) << 1);
    *c1 = *c0 + 2;
  } else {
    *c0 = 2 * (i - 8);
    *c1 = *c0 + 1;
  }
}

// Lower bound on the distance from edge i of a box (centre pos, scaled axes
// e) to the box o with half extents ho: the distance from the edge's
// midpoint to that box less the edge's half length.
HD float edge_box_gap(f3 pos, const box_axes_t& e, int i, const shp& o, f3 ho) {
  int c0, c1;
  box_edge(i, &c0, &c1);
  const f3 p0 = box_corner(pos, e, c0), p1 = box_corner(pos, e, c1);
  const f3 l = abs3(qrot(qconj(o.rot), sub(mul(add(p0, p1), 0.5f), o.pos)));
  const f3 out{fmaxf(l.x - ho.x, 0.0f), fmaxf(l.y - ho.y, 0.0f), fmaxf(l.z - ho.z, 0.0f)};
  return len3(out) - 0.5f * len3(sub(p1, p0));
}

// Candidate c of feature_distance's scan (0-7 corners of a against b,
// 8-15 corners of b against a, 16-159 edge pairs, a's edge major): the
// closest points x (on a) and y (on b) and |y - x|^2.
HD float feature_candidate(const shp& a, f3 ha, const box_axes_t& ea, const shp& b, f3 hb, const box_axes_t& eb,
                           int c, f3* x, f3* y) {
  if (c < 8) {
    *x = box_corner(a.pos, ea, c);
    *y = point_obb_closest(*x, b.pos, hb, b.rot);
  } else if (c < 16) {
    *y = box_corner(b.pos, eb, c - 8);
    *x = point_obb_closest(*y, a.pos, ha, a.rot);
  } else {
    int i = (c - 16) / 12, j = (c - 16) % 12;
    int a0, a1, b0, b1;
    box_edge(i, &a0, &a1);
    box_edge(j, &b0, &b1);
    seg_seg_closest(box_corner(a.pos, ea, a0), box_corner(a.pos, ea, a1), box_corner(b.pos, eb, b0),
                    box_corner(b.pos, eb, b1), x, y);
  }
  return len2(sub(*y, *x));
}

// Closest features of two boxes: the first minimum over the fixed scan
// (corners precomputed once; with the constant trip counts the compiler
// keeps them in registers).
HD void obb_corners(f3 pos, f3 half, qf rot, f3 out[8]) {
  f3 x = qrot(rot, f3{1.0f * half.x, 0.0f * half.x, 0.0f * half.x});
  f3 y = qrot(rot, f3{0.0f * half.y, 1.0f * half.y, 0.0f * half.y});
  f3 z = qrot(rot, f3{0.0f * half.z, 0.0f * half.z, 1.0f * half.z});
  f3 px = add(pos, x), mx = sub(pos, x);
  out[0] = add(add(px, y), z);
  out[1] = sub(add(px, y), z);
  out[2] = add(sub(px, y), z);
  out[3] = sub(sub(px, y), z);
  out[4] = add(add(mx, y), z);
  out[5] = sub(add(mx, y), z);
  out[6] = add(sub(mx, y), z);
  out[7] = sub(sub(mx, y), z);
}

HD dist_t feature_distance(const shp& a, f3 ha, const shp& b, f3 hb) {
  f3 ca[8], cb[8];
  obb_corners(a.pos, ha, a.rot, ca);
  obb_corners(b.pos, hb, b.rot, cb);
  dist_t res{__builtin_inff(), a.pos, b.pos};
  float best = __builtin_inff();
  for (int i = 0; i < 8; ++i) {
    f3 y = point_obb_closest(ca[i], b.pos, hb, b.rot);
    float sq = len2(sub(y, ca[i]));
    if (sq < best) {
      best = sq;
      res = dist_t{sqrtf(sq), ca[i], y};
    }
  }
  for (int i = 0; i < 8; ++i) {
    f3 x = point_obb_closest(cb[i], a.pos, ha, a.rot);
    float sq = len2(sub(cb[i], x));
    if (sq < best) {
      best = sq;
      res = dist_t{sqrtf(sq), x, cb[i]};
    }
  }
  const unsigned char ea_[12] = {0, 1, 2, 3, 0, 1, 4, 5, 0, 2, 4, 6};
  const unsigned char eb_[12] = {4, 5, 6, 7, 2, 3, 6, 7, 1, 3, 5, 7};
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      f3 x, y;
      seg_seg_closest(ca[ea_[i]], ca[eb_[i]], cb[ea_[j]], cb[eb_[j]], &x, &y);
      float sq = len2(sub(y, x));
      if (sq < best) {
        best = sq;
        res = dist_t{sqrtf(sq), x, y};
      }
    }
  return res;
}

// box_box with the SAT gap already computed (sat, normal).
HD dist_t box_box_from_sat(const shp& a, f3 ha, const shp& b, f3 hb, float gap, f3 normal) {
  if (gap > 0.0f) return feature_distance(a, ha, b, hb);
  shp fa, fb;
  f3 ea, eb;
  support_face(a, ha, normal, &fa, &ea);
  support_face(b, hb, neg(normal), &fb, &eb);
  dist_t r = feature_distance(fa, ea, fb, eb);
  r.dist = gap;
  return r;
}

// shape_distance; for box-box pairs *sat receives the SAT gap
// (box_box_signed_gap) the discovery reads as the signed gap.  BOXES =
// false drops the box-box arm for callers that never pass two boxes.
template <bool BOXES = true>
HD dist_t shape_distance(const shp& a, const shp& b, float* sat) {
  *sat = 0.0f;
  if (a.kind == SH_HALFSPACE) return halfspace_distance(a, b, true);
  if (b.kind == SH_HALFSPACE) return halfspace_distance(b, a, false);
  if (a.kind == SH_SPHERE && b.kind == SH_SPHERE) return sphere_sphere(a, a.p.x, b, b.p.x);
  if (a.kind == SH_SPHERE) return sphere_obb(a, a.p.x, b, b.p);
  if (b.kind == SH_SPHERE) {
    dist_t d = sphere_obb(b, b.p.x, a, a.p);
    return dist_t{d.dist, d.pb, d.pa};
  }
  if constexpr (!BOXES) {
    return dist_t{__builtin_inff(), a.pos, b.pos};
  } else {
    f3 normal;
    float gap = separating_axis(a, a.p, b, b.p, &normal);
    *sat = gap;
    return box_box_from_sat(a, a.p, b, b.p, gap, normal);
  }
}
