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
#include "gk_entry.h"

using namespace gk_detail;

namespace {

// Terms that involve a reference point depend on the surface, not on the rest
// point: tables, per stack, of
//   tabI[r][s] = C(p_r, ref_s) / c_o          (rest point r, surface s)
//   tabG[g][s][u] = (q_g - ref_s)_u C'/r a/c_o (orientation g)
//   tabB[s][s'] = C(ref_s, ref_s') / c_o
// so an II entry evaluates one covariance instead of four and a GI entry one
// derivative instead of two (the same FP64 values, looked up).
__global__ void assemble_tables_kernel(Geo G, double* tabI, double* tabG, double* tabB) {
  const gk_system& s = G.s;
  const int64_t S = s.n_surf, R = s.n_rest, no = s.n_ori;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double* c = s.center;
  double x[3], y[3];
  if (t < R * S) {
    const int64_t r = t / S, q = t % S;
    load_pt(s.rest_xyz + 3 * r, c, G.inv_a, x);
    load_pt(s.ref_xyz + 3 * q, c, G.inv_a, y);
    tabI[t] = cov_n(x, y);
  } else if (t < R * S + no * S) {
    const int64_t e = t - R * S, g = e / S, q = e % S;
    double d[3];
    load_pt(s.ori_xyz + 3 * g, c, G.inv_a, x);
    load_pt(s.ref_xyz + 3 * q, c, G.inv_a, y);
    dcov_vec(x, y, d);
    tabG[3 * e] = d[0]; tabG[3 * e + 1] = d[1]; tabG[3 * e + 2] = d[2];
  } else if (t < R * S + no * S + S * S) {
    const int64_t e = t - R * S - no * S, q1 = e / S, q2 = e % S;
    load_pt(s.ref_xyz + 3 * q1, c, G.inv_a, x);
    load_pt(s.ref_xyz + 3 * q2, c, G.inv_a, y);
    tabB[e] = cov_n(x, y);
  }
}

// Index of the covariance block: an orientation component (u, g) or a rest
// point r with its surface, and the point in the normalised frame.
struct Unk {
  int type;          // 0 orientation component, 1 rest point
  int u, g, r, srf;
  double x[3];
};
__device__ __forceinline__ Unk unknown(const Geo& G, int64_t i) {
  const gk_system& s = G.s;
  Unk k;
  if (i < G.NG) {
    const int no = (int)s.n_ori, ii = (int)i;
    k.type = 0; k.u = ii / no; k.g = ii - k.u * no; k.r = 0; k.srf = 0;
    load_pt(s.ori_xyz + 3 * k.g, s.center, G.inv_a, k.x);
  } else {
    k.type = 1; k.u = 0; k.g = 0; k.r = (int)(i - G.NG); k.srf = s.rest_surf[k.r];
    load_pt(s.rest_xyz + 3 * k.r, s.center, G.inv_a, k.x);
  }
  return k;
}

// Covariance-block entry (i, j < Nc) from the unknowns and the tables; the
// formulas of entry() (gk_entry.h).
__device__ __forceinline__ double cov_entry(const Geo& G, const Unk& a, const Unk& b, bool diag,
                                            const double* tabI, const double* tabG, const double* tabB) {
  const gk_system& s = G.s;
  const int64_t S = s.n_surf;
  if (a.type == 0 && b.type == 0) {           // GG
    const double h[3] = {a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]};
    const double t = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
    double val = 0.0;
    if (t < 1.0) {
      const double y = gk_rsqrt(t);
      const double rho = t * y;
      const double p = gk_dcov_poly(t, rho);
      const double om = 1.0 - t;
      const double dd_over_t = 26.25 * om * om * y;
      val = -(h[a.u] * h[b.u] * dd_over_t + (a.u == b.u ? p : 0.0));
      val *= G.c_o_a2;
      if (t == 0.0 && a.u == b.u) val = 14.0 * G.c_o_a2;
    }
    if (diag) val += s.ori_nugget[a.g];
    return val;
  }
  if (a.type == 0 || b.type == 0) {           // GI (orientation q, rest point p)
    const Unk& q = a.type == 0 ? a : b;
    const Unk& p = a.type == 0 ? b : a;
    double d1[3];
    dcov_vec(q.x, p.x, d1);
    const double d2 = tabG[(q.g * S + p.srf) * 3 + q.u];
    return s.gi_res * G.c_o_a * (d1[q.u] - d2);
  }
  // II
  double val = cov_n(a.x, b.x) - tabI[(int64_t)a.r * S + b.srf] - tabI[(int64_t)b.r * S + a.srf] +
               tabB[a.srf * S + b.srf];
  val *= s.i_res * s.c_o;
  if (diag) val += s.rest_nugget[a.r];
  return val;
}

// Tile kernel: block (64, 4) computes a 64 x 64 tile, 16 rows per thread row;
// the column's unknown is set up once per thread, the rows' once per tile (LDS).
template <typename OUT, bool EQUIL, bool LOWER>
__global__ __launch_bounds__(256) void assemble_kernel(Geo G, OUT* A, int64_t lda, int64_t nrows,
                                                       const double* dscale, const double* tabI,
                                                       const double* tabG, const double* tabB) {
  __shared__ Unk rows[64];
  __shared__ double rscale[64];
  const int64_t tj = blockIdx.x, ti = blockIdx.y;
  if (LOWER && tj > ti) return;
  const int tid = threadIdx.y * 64 + threadIdx.x;
  if (tid < 64) {
    const int64_t i = ti * 64 + tid;
    if (i < G.Nc) {
      rows[tid] = unknown(G, i);
      rscale[tid] = EQUIL ? dscale[i] : 1.0;
    }
  }
  __syncthreads();
  const int64_t j = tj * 64 + threadIdx.x;
  if (j >= lda) return;
  const bool jcov = j < G.Nc;
  Unk cj{};
  double dj = 1.0;
  if (jcov) {
    cj = unknown(G, j);
    if (EQUIL) dj = dscale[j];
  }
  for (int k = 0; k < 16; ++k) {
    const int il = threadIdx.y * 16 + k;
    const int64_t i = ti * 64 + il;
    if (i >= nrows) continue;
    double v;
    if (jcov && i < G.Nc) {
      v = cov_entry(G, rows[il], cj, i == j, tabI, tabG, tabB);
      if (EQUIL) v = v * rscale[il] * dj;
    } else if (EQUIL) {
      v = (i == j) ? 1.0 : 0.0;
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
  GK_REQUIRE(h != nullptr && sys != nullptr, GK_ERR_ARG, "gk_assemble: NULL handle or system");
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
  const int64_t S = sys->n_surf, ntab = sys->n_rest * S + sys->n_ori * S + S * S;
  void* wt = nullptr;
  GK_TRY(gk_workspace(h, WS_ASSEMBLE, (size_t)(sys->n_rest * S + 3 * sys->n_ori * S + S * S + 3) * sizeof(double),
                      &wt));
  double* tabI = static_cast<double*>(wt);
  double* tabG = tabI + sys->n_rest * S + 1;
  double* tabB = tabG + 3 * sys->n_ori * S + 1;
  if (ntab > 0)
    hipLaunchKernelGGL(assemble_tables_kernel, dim3((unsigned)gk_ceil_div(ntab, 256)), dim3(256), 0, st, G, tabI,
                       tabG, tabB);
  dim3 grid((unsigned)gk_ceil_div(lda, 64), (unsigned)gk_ceil_div(lda, 64));
  const dim3 blk(64, 4);
  if (dtype == GK_F64) {
    double* M = static_cast<double*>(A);
    if (equil) {
      if (lower_only) hipLaunchKernelGGL((assemble_kernel<double, true, true>), grid, blk, 0, st, G, M, lda, lda, dscale, tabI, tabG, tabB);
      else hipLaunchKernelGGL((assemble_kernel<double, true, false>), grid, blk, 0, st, G, M, lda, lda, dscale, tabI, tabG, tabB);
    } else {
      if (lower_only) hipLaunchKernelGGL((assemble_kernel<double, false, true>), grid, blk, 0, st, G, M, lda, lda, nullptr, tabI, tabG, tabB);
      else hipLaunchKernelGGL((assemble_kernel<double, false, false>), grid, blk, 0, st, G, M, lda, lda, nullptr, tabI, tabG, tabB);
    }
  } else {
    float* M = static_cast<float*>(A);
    if (lower_only) hipLaunchKernelGGL((assemble_kernel<float, true, true>), grid, blk, 0, st, G, M, lda, lda, dscale, tabI, tabG, tabB);
    else hipLaunchKernelGGL((assemble_kernel<float, true, false>), grid, blk, 0, st, G, M, lda, lda, dscale, tabI, tabG, tabB);
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
