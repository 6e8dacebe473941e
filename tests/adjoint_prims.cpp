// Host build of the step's adjoint functions (ornis_amd/csrc/avbd_adjoint.hpp)
// in double precision, with the forward scalar functions they differentiate,
// for tests/test_adjoint_prims.py (central finite differences).
#include "../ornis_amd/csrc/avbd_adjoint.hpp"

using namespace adj;
using D = double;

static V3<D> v3(const D* p) { return {p[0], p[1], p[2]}; }
static Q4<D> q4(const D* p) { return {p[0], p[1], p[2], p[3]}; }
static void put3(D* o, V3<D> v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
static void put4(D* o, Q4<D> q) { o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w; }

// Contact layout (55 doubles): n3 t1_3 t2_3 mu1 mu2 damp touching ra3 rb3
// rad_a rad_b xa3 xb3 xa0_3 xb0_3 qa4 qb4 qa0_4 qb0_4 lam3 pen3.
static Contact<D> contact(const D* p) {
  Contact<D> c;
  c.n = v3(p + 0);
  c.t1 = v3(p + 3);
  c.t2 = v3(p + 6);
  c.mu1 = p[9];
  c.mu2 = p[10];
  c.damp = p[11];
  c.touching = p[12] != 0.0;
  c.ra = v3(p + 13);
  c.rb = v3(p + 16);
  c.rad_a = p[19];
  c.rad_b = p[20];
  c.xa = v3(p + 21);
  c.xb = v3(p + 24);
  c.xa0 = v3(p + 27);
  c.xb0 = v3(p + 30);
  c.qa = q4(p + 33);
  c.qb = q4(p + 37);
  c.qa0 = q4(p + 41);
  c.qb0 = q4(p + 45);
  for (int k = 0; k < 3; ++k) {
    c.lam[k] = p[49 + k];
    c.pen[k] = p[52 + k];
  }
  return c;
}
// ContactBar layout (37 doubles), the adjoints of contact entries 21..54
// (xa xb xa0 xb0 qa qb qa0 qb0 lam pen) then mu1 mu2 damp.
static void put_bar(D* o, const ContactBar<D>& b) {
  put3(o + 0, b.xa);
  put3(o + 3, b.xb);
  put3(o + 6, b.xa0);
  put3(o + 9, b.xb0);
  put4(o + 12, b.qa);
  put4(o + 16, b.qb);
  put4(o + 20, b.qa0);
  put4(o + 24, b.qb0);
  for (int k = 0; k < 3; ++k) {
    o[28 + k] = b.lam[k];
    o[31 + k] = b.pen[k];
  }
  o[34] = b.mu1;
  o[35] = b.mu2;
  o[36] = b.damp;
}

// The row terms of S = w^T g + w^T H dx for one point-side, restated from the
// primal's row construction (stamp_row rows.rs:1064-1092 with the levers and
// forces of rows.rs:947-1060).
static D point_side_S(const Contact<D>& c, bool is_a, const D w[6], const D dx[6]) {
  const D sign = is_a ? 1.0 : -1.0;
  RowVals<D> r = row_vals(c);
  auto term = [&](V3<D> u, V3<D> lever, D pen, D f) {
    V3<D> nn = u * sign;
    V3<D> t = cross(lever, nn);
    D a = w[0] * nn.x + w[1] * nn.y + w[2] * nn.z + w[3] * t.x + w[4] * t.y + w[5] * t.z;
    D b = dx[0] * nn.x + dx[1] * nn.y + dx[2] * nn.z + dx[3] * t.x + dx[4] * t.y + dx[5] * t.z;
    return f * a + pen * a * b;
  };
  V3<D> r_side = is_a ? r.ra_w : r.rb_w;
  if (!c.touching) {
    if (!(c.damp > 0.0 && std::fabs(r.cn) >= 1e-7)) return 0.0;
    D fd = c.damp * r.cn;
    return term(c.n, r_side, c.damp, fd < 0.0 ? fd : 0.0);
  }
  V3<D> r_side_f = is_a ? r.ra_f : r.rb_f;
  const V3<D> ax[3] = {c.n, c.t1, c.t2};
  const D cv[3] = {r.cn, r.ct1, r.ct2};
  const D fv[3] = {r.f.fn, r.f.f1, r.f.f2};
  D s = 0.0;
  for (int k = 0; k < 3; ++k) {
    if (!(std::fabs(cv[k]) >= 1e-7 || fv[k] != 0.0)) continue;
    s += term(ax[k], k == 0 ? r_side : r_side_f, c.pen[k], fv[k]);
  }
  return s;
}

extern "C" {

void ap_qrot(const D* q, const D* v, D* out) { put3(out, qrot(q4(q), v3(v))); }
void ap_qrot_vjp(const D* q, const D* v, const D* o, D* qb, D* vb) {
  Q4<D> a{0, 0, 0, 0};
  V3<D> b{0, 0, 0};
  qrot_vjp(q4(q), v3(v), v3(o), &a, &b);
  put4(qb, a);
  put3(vb, b);
}
void ap_qint(const D* q, const D* v, D* out) { put4(out, quat_integrate(q4(q), v3(v))); }
void ap_qint_vjp(const D* q, const D* v, const D* o, D* qb, D* vb) {
  Q4<D> a{0, 0, 0, 0};
  V3<D> b{0, 0, 0};
  quat_integrate_vjp(q4(q), v3(v), q4(o), &a, &b);
  put4(qb, a);
  put3(vb, b);
}
void ap_qdiff(const D* a, const D* b, D* out) { put3(out, quat_diff_vec(q4(a), q4(b))); }
void ap_qdiff_vjp(const D* a, const D* b, const D* o, D* ab, D* bb) {
  Q4<D> x{0, 0, 0, 0}, y{0, 0, 0, 0};
  quat_diff_vec_vjp(q4(a), q4(b), v3(o), &x, &y);
  put4(ab, x);
  put4(bb, y);
}
void ap_qnorm(const D* q, D* out) { put4(out, qnormalize(q4(q))); }
void ap_qnorm_vjp(const D* q, const D* o, D* qb) {
  Q4<D> a{0, 0, 0, 0};
  qnormalize_vjp(q4(q), q4(o), &a);
  put4(qb, a);
}
D ap_iform(const D* q, const D* Dg, const D* w, const D* z) { return inertia_form(q4(q), v3(Dg), v3(w), v3(z)); }
void ap_iform_vjp(const D* q, const D* Dg, const D* w, const D* z, D* qb, D* Db, D* wb, D* zb) {
  Q4<D> a{0, 0, 0, 0};
  V3<D> b{0, 0, 0}, c{0, 0, 0}, d{0, 0, 0};
  inertia_form_vjp(q4(q), v3(Dg), v3(w), v3(z), 1.0, &a, &b, &c, &d);
  put4(qb, a);
  put3(Db, b);
  put3(wb, c);
  put3(zb, d);
}
// contact_force: in = cn ct1 ct2 lam3 pen3 mu1 mu2 (11); out = fn f1 f2.
void ap_force(const D* in, D* out) {
  Force<D> f = contact_force(in[0], in[1], in[2], in + 3, in + 6, in[9], in[10]);
  out[0] = f.fn;
  out[1] = f.f1;
  out[2] = f.f2;
}
void ap_force_vjp(const D* in, const D* fb, D* inb) {
  for (int k = 0; k < 11; ++k) inb[k] = 0.0;
  contact_force_vjp(in[0], in[1], in[2], in + 3, in + 6, in[9], in[10], fb[0], fb[1], fb[2], inb + 0, inb + 1, inb + 2,
                    inb + 3, inb + 6, inb + 9, inb + 10);
}
D ap_side_S(const D* c, int is_a, const D* w, const D* dx) { return point_side_S(contact(c), is_a != 0, w, dx); }
// adds -dS/d(inputs): the test compares with -FD
void ap_side_vjp(const D* c, int is_a, const D* w, const D* dx, D* bar) {
  ContactBar<D> b;
  clear(b);
  point_side_vjp(contact(c), is_a != 0, w, dx, b);
  put_bar(bar, b);
}
void ap_dual(const D* c, D* lam_o, D* pen_o) { dual_forward(contact(c), lam_o, pen_o); }
void ap_dual_vjp(const D* c, const D* lb, const D* pb, D* bar) {
  ContactBar<D> b;
  clear(b);
  dual_vjp(contact(c), lb, pb, b);
  put_bar(bar, b);
}
// S_in = w^T g_in + w^T H_in dx (assembly.rs:74-96); in = x3 q4 y3 qy4 m
// dunit3 (18).
D ap_inertial_S(const D* in, const D* w, const D* dx, D dt) {
  V3<D> x = v3(in), y = v3(in + 7), du = v3(in + 15);
  Q4<D> q = q4(in + 3), qy = q4(in + 10);
  D m = in[14];
  D s = 0.0;
  for (int k = 0; k < 3; ++k) s += w[k] * (m / (dt * dt)) * ((k == 0 ? x.x - y.x : k == 1 ? x.y - y.y : x.z - y.z) + dx[k]);
  V3<D> dq = quat_diff_vec(q, qy);
  V3<D> z{(dq.x + dx[3]) / (dt * dt), (dq.y + dx[4]) / (dt * dt), (dq.z + dx[5]) / (dt * dt)};
  s += inertia_form(q, du * m, V3<D>{w[3], w[4], w[5]}, z);
  return s;
}
void ap_inertial_vjp(const D* in, const D* w, const D* dx, D dt, D* bar) {
  V3<D> xb{0, 0, 0}, yb{0, 0, 0};
  Q4<D> qb{0, 0, 0, 0}, qyb{0, 0, 0, 0};
  D mb = 0.0;
  inertial_vjp(v3(in), q4(in + 3), v3(in + 7), q4(in + 10), in[14], v3(in + 15), w, dx, dt, &xb, &qb, &yb, &qyb, &mb);
  put3(bar, xb);
  put4(bar + 3, qb);
  put3(bar + 7, yb);
  put4(bar + 10, qyb);
  bar[14] = mb;
}

}  // extern "C"
