// Entries of the co-kriging matrix in FP64, shared by the assembly, the
// local-block extraction of the learned preconditioner and the tests.
//   GG[(u,g),(v,g')] = -[d_uv C'/r + h_u h_v / r^2 (C'' - C'/r)] + d nugget_grad
//   GI[(u,g), i]     = gi_res [(q-p_i)_u C'/r(|q-p_i|) - (q-p0)_u C'/r(|q-p0|)]
//   II[i, j]         = i_res [C(pi,pj) - C(pi,p0j) - C(p0i,pj) + C(p0i,p0j)] + d (nu_i + nu_ref)
//   U_G[(u,g), l]    = d f_l / d u (q_g);  U_I[i, l] = gi_res (f_l(p_i) - f_l(p0_i))
//   fault column     = gi_res (F(p_i) - F(p0_i)) on interface rows, 0 on gradient rows
#pragma once
#include "gk_internal.h"

namespace gk_detail {

struct Geo {
  gk_system s;
  int64_t NG, Nc, N, L;
  double inv_a, c_o_a2, c_o_a;
};

__device__ __forceinline__ void load_pt(const double* p, const double* c, double ia, double* o) {
  o[0] = (p[0] - c[0]) * ia; o[1] = (p[1] - c[1]) * ia; o[2] = (p[2] - c[2]) * ia;
}

__device__ __forceinline__ double cov_n(const double* x, const double* y) {   // C / c_o
  const double dx = x[0] - y[0], dy = x[1] - y[1], dz = x[2] - y[2];
  const double t = dx * dx + dy * dy + dz * dz;
  if (t >= 1.0) return 0.0;
  const double rho = t * gk_rsqrt(t);
  return gk_cov_poly(t, rho);
}

// (x - y)_u C'/r * a / c_o  ->  h_u * dpoly (normalised coordinates)
__device__ __forceinline__ void dcov_vec(const double* x, const double* y, double* o) {
  const double dx = x[0] - y[0], dy = x[1] - y[1], dz = x[2] - y[2];
  const double t = dx * dx + dy * dy + dz * dz;
  double p = 0.0;
  if (t < 1.0) p = gk_dcov_poly(t, t * gk_rsqrt(t));
  o[0] = dx * p; o[1] = dy * p; o[2] = dz * p;
}

__device__ __forceinline__ double drift_f(int l, const double* p) {
  if (l < 3) return p[l];
  if (l < 6) return p[l - 3] * p[l - 3];
  if (l == 6) return p[0] * p[1];
  if (l == 7) return p[0] * p[2];
  return p[1] * p[2];
}

__device__ __forceinline__ double drift_df(int l, int u, const double* q) {
  if (l < 3) return l == u ? 1.0 : 0.0;
  if (l < 6) return (l - 3 == u) ? 2.0 * q[u] : 0.0;
  if (l == 6) return u == 0 ? q[1] : (u == 1 ? q[0] : 0.0);
  if (l == 7) return u == 0 ? q[2] : (u == 2 ? q[0] : 0.0);
  return u == 1 ? q[2] : (u == 2 ? q[1] : 0.0);
}

// One matrix entry in FP64.
__device__ inline double entry(const Geo& G, int64_t i, int64_t j) {
  const gk_system& s = G.s;
  const int64_t no = s.n_ori, NG = G.NG, Nc = G.Nc, N = G.N, L = G.L;
  if (i >= N || j >= N) return i == j ? 1.0 : 0.0;
  if (i >= Nc && j >= Nc) return 0.0;
  // order so that i is a covariance row when possible
  if (i >= Nc) { int64_t t = i; i = j; j = t; }
  const double* c = s.center;
  if (j >= Nc) {   // drift / fault column, i covariance row
    const int64_t l = j - Nc;
    if (i < NG) {
      const int64_t u = i / no, g = i % no;
      if (l >= L) return 0.0;
      return drift_df((int)l, (int)u, s.ori_xyz + 3 * g);
    }
    const int64_t r = i - NG;
    if (l >= L) return s.gi_res * s.fault_rest[r * s.n_fault + (l - L)];
    const double* p = s.rest_xyz + 3 * r;
    const double* p0 = s.ref_xyz + 3 * s.rest_surf[r];
    return s.gi_res * (drift_f((int)l, p) - drift_f((int)l, p0));
  }
  if (i >= NG && j < NG) { int64_t t = i; i = j; j = t; }
  if (i < NG && j < NG) {
    const int64_t u = i / no, g = i % no, v = j / no, g2 = j % no;
    double q1[3], q2[3];
    load_pt(s.ori_xyz + 3 * g, c, G.inv_a, q1);
    load_pt(s.ori_xyz + 3 * g2, c, G.inv_a, q2);
    const double h[3] = {q1[0] - q2[0], q1[1] - q2[1], q1[2] - q2[2]};
    const double t = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
    double val = 0.0;
    if (t < 1.0) {
      const double y = gk_rsqrt(t);
      const double rho = t * y;
      const double p = gk_dcov_poly(t, rho);
      const double om = 1.0 - t;
      const double dd_over_t = 26.25 * om * om * y;       // (C''-C'/r)/r^2 normalised, times t/rho
      val = -(h[u] * h[v] * dd_over_t + (u == v ? p : 0.0));
      val *= G.c_o_a2;
      if (t == 0.0 && u == v) val = 14.0 * G.c_o_a2;
    }
    if (i == j) val += s.ori_nugget[g];
    return val;
  }
  if (i < NG) {   // GI
    const int64_t u = i / no, g = i % no, r = j - NG;
    double q[3], p[3], p0[3], d1[3], d2[3];
    load_pt(s.ori_xyz + 3 * g, c, G.inv_a, q);
    load_pt(s.rest_xyz + 3 * r, c, G.inv_a, p);
    load_pt(s.ref_xyz + 3 * s.rest_surf[r], c, G.inv_a, p0);
    dcov_vec(q, p, d1);
    dcov_vec(q, p0, d2);
    return s.gi_res * G.c_o_a * (d1[u] - d2[u]);
  }
  // II
  const int64_t r1 = i - NG, r2 = j - NG;
  double p1[3], p2[3], q1[3], q2[3];
  load_pt(s.rest_xyz + 3 * r1, c, G.inv_a, p1);
  load_pt(s.rest_xyz + 3 * r2, c, G.inv_a, p2);
  load_pt(s.ref_xyz + 3 * s.rest_surf[r1], c, G.inv_a, q1);
  load_pt(s.ref_xyz + 3 * s.rest_surf[r2], c, G.inv_a, q2);
  double val = cov_n(p1, p2) - cov_n(p1, q2) - cov_n(q1, p2) + cov_n(q1, q2);
  val *= s.i_res * s.c_o;
  if (i == j) val += s.rest_nugget[r1];
  return val;
}


inline Geo make_geo(const gk_system* s) {
  Geo G;
  G.s = *s;
  G.L = s->uni_degree == 0 ? 0 : (s->uni_degree == 1 ? 3 : 9);
  G.NG = 3 * s->n_ori;
  G.Nc = G.NG + s->n_rest;
  G.N = G.Nc + G.L + s->n_fault;
  G.inv_a = 1.0 / s->range;
  G.c_o_a2 = s->c_o / (s->range * s->range);
  G.c_o_a = s->c_o / s->range;
  return G;
}


}  // namespace gk_detail
