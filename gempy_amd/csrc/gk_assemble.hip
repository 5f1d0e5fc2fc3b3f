// Co-kriging matrix assembly (the engine's kernel_constructor).
//
// Blocks (h = x - y, r = |h|, all in FP64, written as FP64 or FP32):
//   GG[(u,g),(v,g')] = -[d_uv C'/r + h_u h_v / r^2 (C'' - C'/r)] + d nugget_grad
//   GI[(u,g), i]     = gi_res [(q-p_i)_u C'/r(|q-p_i|) - (q-p0)_u C'/r(|q-p0|)]
//   II[i, j]         = i_res [C(pi,pj) - C(pi,p0j) - C(p0i,pj) + C(p0i,p0j)] + d (nu_i + nu_ref)
//   U_G[(u,g), l]    = d f_l / d u (q_g);  U_I[i, l] = gi_res (f_l(p_i) - f_l(p0_i))
//   fault column     = gi_res (F(p_i) - F(p0_i)) on interface rows, 0 on gradient rows
// The four-term rest-minus-reference differencing of II is evaluated in FP64
// whatever the output precision (it cancels heavily for nearby points).
#include <cmath>
#include "gk_internal.h"

namespace {

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
__device__ double entry(const Geo& G, int64_t i, int64_t j) {
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

// Tile kernel: block (64, 4) computes a 64 x 64 tile, 16 rows per thread row.
template <typename OUT, bool EQUIL, bool LOWER>
__global__ __launch_bounds__(256) void assemble_kernel(Geo G, OUT* A, int64_t lda, int64_t nrows,
                                                       const double* dscale) {
  const int64_t tj = blockIdx.x, ti = blockIdx.y;
  if (LOWER && tj > ti) return;
  const int64_t j = tj * 64 + threadIdx.x;
  for (int k = 0; k < 16; ++k) {
    const int64_t i = ti * 64 + threadIdx.y * 16 + k;
    if (i >= nrows || j >= lda) continue;
    double v;
    if (EQUIL) {
      if (i >= G.Nc || j >= G.Nc) v = (i == j) ? 1.0 : 0.0;
      else v = entry(G, i, j) * dscale[i] * dscale[j];
    } else {
      v = entry(G, i, j);
    }
    A[i * lda + j] = (OUT)v;
  }
}

__global__ void diag_scale_kernel(Geo G, double* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.Nc) return;
  d[i] = 1.0 / sqrt(entry(G, i, i));
}

__global__ void rhs_kernel(Geo G, double* b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.N) return;
  double v = 0.0;
  if (i < G.NG) {
    const int64_t no = G.s.n_ori, u = i / no, g = i % no;
    v = G.s.ori_grad[3 * g + u];
  }
  b[i] = v;
}

__global__ void drift_block_kernel(Geo G, double* U) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t Lt = G.N - G.Nc;
  if (idx >= G.Nc * Lt) return;
  const int64_t i = idx / Lt, l = idx % Lt;
  U[idx] = entry(G, i, G.Nc + l);
}

Geo make_geo(const gk_system* s) {
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

}  // namespace

extern "C" int64_t gk_system_n(const gk_system* s) {
  if (!s) return -1;
  const int L = s->uni_degree == 0 ? 0 : (s->uni_degree == 1 ? 3 : 9);
  return 3 * s->n_ori + s->n_rest + L + s->n_fault;
}

extern "C" int64_t gk_system_padded(const gk_system* s) {
  const int64_t n = gk_system_n(s);
  return n < 0 ? n : gk_round_up(n > 0 ? n : 1, GK_BLOCK);
}

extern "C" int gk_assemble_impl(gk_handle h, const gk_system* sys, int32_t dtype, void* A, int64_t lda,
                                double* rhs, double* dscale, hipStream_t st, int lower_only) {
  (void)h;
  Geo G = make_geo(sys);
  GK_REQUIRE(dtype == GK_F32 || dtype == GK_F64, GK_ERR_ARG, "gk_assemble: dtype must be GK_F32 or GK_F64");
  GK_REQUIRE(A != nullptr, GK_ERR_ARG, "gk_assemble: A is NULL");
  const bool equil = dscale != nullptr;
  GK_REQUIRE(equil || dtype == GK_F64, GK_ERR_ARG, "gk_assemble: the FP32 matrix is the equilibrated block");
  const int64_t need = equil ? G.Nc : G.N;
  GK_REQUIRE(lda >= need && lda % GK_BLOCK == 0, GK_ERR_ARG,
             "gk_assemble: lda (%lld) must be >= %lld and a multiple of %d", (long long)lda, (long long)need,
             GK_BLOCK);
  if (equil && G.Nc > 0)
    hipLaunchKernelGGL(diag_scale_kernel, dim3((unsigned)gk_ceil_div(G.Nc, 256)), dim3(256), 0, st, G, dscale);
  dim3 grid((unsigned)gk_ceil_div(lda, 64), (unsigned)gk_ceil_div(lda, 64));
  if (dtype == GK_F64) {
    double* M = static_cast<double*>(A);
    if (equil) {
      if (lower_only) hipLaunchKernelGGL((assemble_kernel<double, true, true>), grid, dim3(64, 4), 0, st, G, M, lda, lda, dscale);
      else hipLaunchKernelGGL((assemble_kernel<double, true, false>), grid, dim3(64, 4), 0, st, G, M, lda, lda, dscale);
    } else {
      if (lower_only) hipLaunchKernelGGL((assemble_kernel<double, false, true>), grid, dim3(64, 4), 0, st, G, M, lda, lda, nullptr);
      else hipLaunchKernelGGL((assemble_kernel<double, false, false>), grid, dim3(64, 4), 0, st, G, M, lda, lda, nullptr);
    }
  } else {
    float* M = static_cast<float*>(A);
    if (lower_only) hipLaunchKernelGGL((assemble_kernel<float, true, true>), grid, dim3(64, 4), 0, st, G, M, lda, lda, dscale);
    else hipLaunchKernelGGL((assemble_kernel<float, true, false>), grid, dim3(64, 4), 0, st, G, M, lda, lda, dscale);
  }
  if (rhs && G.N > 0)
    hipLaunchKernelGGL(rhs_kernel, dim3((unsigned)gk_ceil_div(G.N, 256)), dim3(256), 0, st, G, rhs);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_rhs_impl(const gk_system* sys, double* b, hipStream_t st) {
  Geo G = make_geo(sys);
  if (G.N > 0)
    hipLaunchKernelGGL(rhs_kernel, dim3((unsigned)gk_ceil_div(G.N, 256)), dim3(256), 0, st, G, b);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_assemble_drift_impl(const gk_system* sys, double* U, hipStream_t st) {
  Geo G = make_geo(sys);
  const int64_t Lt = G.N - G.Nc;
  if (Lt == 0 || G.Nc == 0) return GK_OK;
  hipLaunchKernelGGL(drift_block_kernel, dim3((unsigned)gk_ceil_div(G.Nc * Lt, 256)), dim3(256), 0, st,
                     G, U);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_diag_scale_impl(const gk_system* sys, double* d, hipStream_t st) {
  Geo G = make_geo(sys);
  if (G.Nc == 0) return GK_OK;
  hipLaunchKernelGGL(diag_scale_kernel, dim3((unsigned)gk_ceil_div(G.Nc, 256)), dim3(256), 0, st, G, d);
  GK_HIP(hipGetLastError());
  return GK_OK;
}
