// avbd_geom.hpp — exact distance queries between placed boxes, spheres,
// capsules and half-spaces (the narrowphase of the AVBD step), device + host.
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
//   capsule_segment      distance.rs:174-178
//   point_segment_closest distance.rs:63-71
//   sphere_capsule       distance.rs:218-230
//   capsule_capsule      distance.rs:232-245
//   segment_box          distance_box.rs:150-197
//   box_capsule          distance_box.rs:199-240
//   spine_penetration    distance_box.rs:244-265
//   shape_distance       distance.rs:654-823 (box/sphere/capsule/half-space arms)
#pragma once

#include "avbd_math.hpp"

enum { SH_SPHERE = 0, SH_BOX = 1, SH_HALFSPACE = 2, SH_CAPSULE = 3 };

struct shp {
  int kind;
  f3 p;  // sphere (r,0,0) / box half extents / half-space unit normal / capsule (r, half height, 0)
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
  if (kind == SH_CAPSULE) return p.y + p.x;
  return __builtin_inff();
}

HD float shape_min_dim(int kind, f3 p) {
  if (kind == SH_SPHERE) return p.x;
  if (kind == SH_BOX) return fmin_(p.x, fmin_(p.y, p.z));
  if (kind == SH_CAPSULE) return p.x;
  return __builtin_inff();
}

// Shape::inertia's capsule arm (crates/physics/src/collision/shape.rs:769-781):
// about the core axis (body y) the disk term, across it the disk, stem
// coupling and stem terms.
HD f3 capsule_inertia(float r, float h, float mass) {
  const float iy = 0.5f * mass * r * r;
  const float ixz = 0.25f * mass * (r * r) + (mass / 3.0f) * h * r + 0.25f * mass * h * h;
  return f3{ixz, iy, ixz};
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
  if (s.kind == SH_CAPSULE) {
    const f3 axis = qrot(s.rot, f3{0.0f, 1.0f, 0.0f});
    const f3 tip = mul(dot3(dir, axis) >= 0.0f ? axis : neg(axis), s.p.y);
    return add(add(s.pos, tip), mul(normalize_or(dir, f3{1.0f, 0.0f, 0.0f}), s.p.x));
  }
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

// The capsule's core segment: its y axis, half height hh each way.
HD void capsule_segment(f3 pos, float hh, qf rot, f3* s0, f3* s1) {
  const f3 axis = qrot(rot, f3{0.0f, 1.0f, 0.0f});
  *s0 = sub(pos, mul(axis, hh));
  *s1 = add(pos, mul(axis, hh));
}

HD f3 point_segment_closest(f3 p, f3 a, f3 b) {
  const f3 ab = sub(b, a);
  const float len_sq = len2(ab);
  if (len_sq < 1e-12f) return a;
  const float t = clampf_(dot3(sub(p, a), ab) / len_sq, 0.0f, 1.0f);
  return add(a, mul(ab, t));
}

HD dist_t sphere_capsule(const shp& a, float r, const shp& b, float cr, float hh) {
  f3 s0, s1;
  capsule_segment(b.pos, hh, b.rot, &s0, &s1);
  const f3 core_b = point_segment_closest(a.pos, s0, s1);
  const f3 dv = sub(a.pos, core_b);
  const float core = len3(dv);
  const f3 n = normalize_or(dv, f3{1.0f, 0.0f, 0.0f});
  return dist_t{core - r - cr, sub(a.pos, mul(n, r)), add(core_b, mul(n, cr))};
}

HD dist_t capsule_capsule(const shp& a, float ra, float ha, const shp& b, float rb, float hb) {
  f3 a0, a1, b0, b1, ca, cb;
  capsule_segment(a.pos, ha, a.rot, &a0, &a1);
  capsule_segment(b.pos, hb, b.rot, &b0, &b1);
  seg_seg_closest(a0, a1, b0, b1, &ca, &cb);
  const f3 dv = sub(cb, ca);
  const float core = len3(dv);
  const f3 n = normalize_or(dv, f3{1.0f, 0.0f, 0.0f});
  return dist_t{core - ra - rb, add(ca, mul(n, ra)), sub(cb, mul(n, rb))};
}

// Closest points between the segment [start, end] and the solid box
// [lo, hi] (box frame) and their squared distance.  The segment is cut
// where it crosses the box's slab planes; on each piece the squared
// distance is one quadratic, whose clamped stationary point and the piece's
// ends are tried in turn (first strict minimum kept).  At most 6 interior
// cuts: the cut list is kept sorted by insertion.
HD void segment_box(f3 start, f3 end, f3 lo, f3 hi, f3* core, f3* surface, float* squared) {
  const f3 delta = sub(end, start);
  float cuts[8];
  int nc = 2;
  cuts[0] = 0.0f;
  cuts[1] = 1.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float di = comp(delta, i);
    if (fabsf(di) > 1e-12f) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const float t = ((k ? comp(hi, i) : comp(lo, i)) - comp(start, i)) / di;
        if (t > 0.0f && t < 1.0f) {
          int j = nc++;
          while (j > 0 && cuts[j - 1] > t) {
            cuts[j] = cuts[j - 1];
            --j;
          }
          cuts[j] = t;
        }
      }
    }
  }
  *core = start;
  *surface = min3(max3(start, lo), hi);
  *squared = __builtin_inff();
  for (int w = 0; w + 1 < nc; ++w) {
    const float r0 = cuts[w], r1 = cuts[w + 1];
    const float middle = (r0 + r1) * 0.5f;
    const f3 p = add(start, mul(delta, middle));
    float num = 0.0f, den = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float pi = comp(p, i), li = comp(lo, i), hi_ = comp(hi, i);
      if (!(pi < li) && !(pi > hi_)) continue;
      const float boundary = pi < li ? li : hi_;
      const float di = comp(delta, i);
      num += di * (comp(start, i) - boundary);
      den += di * di;
    }
    const float stationary = den > 0.0f ? clampf_(-num / den, r0, r1) : middle;
    const float ts[3] = {stationary, r0, r1};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const f3 point = add(start, mul(delta, ts[k]));
      const f3 closest = min3(max3(point, lo), hi);
      const float sq = len2(sub(point, closest));
      if (sq < *squared) {
        *core = point;
        *surface = closest;
        *squared = sq;
      }
    }
  }
}

// The deepest separating axis between a box (half extents `half`, at the
// origin) and a capsule whose spine meets it: the box's face normals and
// the spine's cross products with them, each face offset by the radius.
HD float spine_penetration(f3 center, f3 axis, f3 half, float radius, float height, f3* normal) {
  float best = -__builtin_inff();
  *normal = f3{1.0f, 0.0f, 0.0f};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const f3 ev{e == 0 ? 1.0f : 0.0f, e == 1 ? 1.0f : 0.0f, e == 2 ? 1.0f : 0.0f};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const f3 n = k ? normalize_or(cross3(axis, ev), zero3()) : ev;
      if (eq3(n, zero3())) continue;
      const float projection = dot3(center, n);
      const float gap = fabsf(projection) - dot3(half, abs3(n)) - height * fabsf(dot3(axis, n)) - radius;
      if (gap > best) {
        best = gap;
        *normal = projection >= 0.0f ? n : neg(n);
      }
    }
  }
  return best;
}

// Box a (half extents `half`) against capsule b.
HD dist_t box_capsule(const shp& a, f3 half, const shp& b, float radius, float height) {
  const qf inv = qconj(a.rot);
  const f3 center = qrot(inv, sub(b.pos, a.pos));
  const f3 axis = qrot(inv, qrot(b.rot, f3{0.0f, 1.0f, 0.0f}));
  const f3 start = sub(center, mul(axis, height)), end = add(center, mul(axis, height));
  f3 core, surface, normal;
  float squared, distance;
  segment_box(start, end, neg(half), half, &core, &surface, &squared);
  if (squared > 1e-12f) {
    const float len = sqrtf(squared);
    distance = len - radius;
    normal = divs(sub(core, surface), len);
  } else {
    distance = spine_penetration(center, axis, half, radius, height, &normal);
    f3 lo = neg(half), hi = half;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (fabsf(comp(normal, i)) > 1e-6f) {
        const float v = signumf(comp(normal, i)) * comp(half, i);
        setc(lo, i, v);
        setc(hi, i, v);
      }
    const float projection = dot3(axis, normal);
    f3 e0 = start, e1 = end;
    if (projection > 1e-6f)
      e1 = start;
    else if (projection < -1e-6f)
      e0 = end;
    float sq;
    segment_box(e0, e1, lo, hi, &core, &surface, &sq);
  }
  return dist_t{distance, add(a.pos, qrot(a.rot, surface)), add(a.pos, qrot(a.rot, sub(core, mul(normal, radius))))};
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
  if (a.kind == SH_CAPSULE || b.kind == SH_CAPSULE) {
    if (a.kind == SH_SPHERE) return sphere_capsule(a, a.p.x, b, b.p.x, b.p.y);
    if (b.kind == SH_SPHERE) {
      dist_t d = sphere_capsule(b, b.p.x, a, a.p.x, a.p.y);
      return dist_t{d.dist, d.pb, d.pa};
    }
    if (a.kind == SH_BOX) return box_capsule(a, a.p, b, b.p.x, b.p.y);
    if (b.kind == SH_BOX) {
      dist_t d = box_capsule(b, b.p, a, a.p.x, a.p.y);
      return dist_t{d.dist, d.pb, d.pa};
    }
    return capsule_capsule(a, a.p.x, a.p.y, b, b.p.x, b.p.y);
  }
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
