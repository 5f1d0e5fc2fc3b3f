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
