// Solve of one stack's co-kriging system (the engine's kernel_solver).
//
// GK_SOLVE_FP64_LU    : FP64 assembly + blocked LU (FP64 MFMA) + triangular solves.
// GK_SOLVE_FP32_REFINE: FP32 equilibrated Cholesky (FP32 MFMA) of the SPD
//   covariance block, exact Schur complement on the drift block, used as the
//   right preconditioner of FP64 restarted flexible GMRES whose matvec is the
//   matrix-free FP64 operator (kernel tiles recomputed, no FP64 matrix).  The
//   iteration stops on the true FP64 residual (tolerance reached, or no more
//   progress at the FP64 rounding floor); if it is still above `accept` the
//   solve falls back to the FP64 LU.
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>
#include <algorithm>
#include "gk_internal.h"

extern "C" int gk_assemble_impl(gk_handle h, const gk_system* sys, int32_t dtype, void* A, int64_t lda,
                                double* rhs, double* dscale, hipStream_t st, int lower_only);
extern "C" int gk_assemble_drift_impl(const gk_system* sys, double* U, hipStream_t st);
extern "C" int gk_rhs_impl(const gk_system* sys, double* b, hipStream_t st);
extern "C" int gk_diag_scale_impl(const gk_system* sys, double* d, hipStream_t st);
extern "C" int gk_spai_apply_impl(const gk_system* sys, const int32_t* idx, const float* Y, const double* u, double* y,
                                  hipStream_t st);
extern "C" int gk_matvec_boxed(gk_handle h, const gk_system* sys, int32_t precision, const double* x, double* y,
                               gk_stream stream, const double* src_box);
extern "C" int gk_source_box(gk_handle h, const gk_system* sys, double* box, gk_stream stream);
extern "C" int gk_knn(gk_handle h, const gk_system* sys, int32_t* idx, float* hk, gk_stream stream);
extern "C" int gk_spai_local_inverse(gk_handle h, const gk_system* sys, const int32_t* idx, const double* dscale,
                                     float* Y, gk_stream stream);
extern "C" int gk_cholesky_f64(gk_handle h, double* A, int64_t n, int64_t lda, double* dinv, int32_t* info,
                               gk_stream stream);
extern "C" int gk_cholesky_solve_f64(gk_handle h, const double* L, const double* dinv, int64_t lda, double* b,
                                     gk_stream stream);

namespace {

constexpr int RED_ROWS = 4096;   // rows per block of the column-dot kernels

// partial[b][c] = sum over rows [b*RED_ROWS, (b+1)*RED_ROWS) of V[:, c] * w; one
// workgroup per (row block, column), so a Gram-Schmidt step's dots fill the chip.
__global__ __launch_bounds__(256) void dots_partial_kernel(const double* V, int64_t ldv, int ncols, const double* w,
                                                           int64_t n, double* partial) {
  __shared__ double red[4];
  const int c = blockIdx.y;
  const int64_t r0 = (int64_t)blockIdx.x * RED_ROWS;
  const int64_t r1 = std::min<int64_t>(n, r0 + RED_ROWS);
  const double* v = V + (int64_t)c * ldv;
  double acc = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) acc = fma(v[i], w[i], acc);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * ncols + c] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void reduce_cols_kernel(const double* partial, int nblk, int ncols, double* out) {
  const int c = threadIdx.x;
  if (c >= ncols) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * ncols + c];
  out[c] = s;
}

// Second Gram-Schmidt pass: tmp = column sums (the pass's coefficients) and
// acc += tmp (the Hessenberg column).
__global__ void reduce_cols_acc_kernel(const double* partial, int nblk, int ncols, double* tmp, double* acc) {
  const int c = threadIdx.x;
  if (c >= ncols) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * ncols + c];
  tmp[c] = s;
  acc[c] += s;
}

// v = w / |w| with |w|^2 read from device memory (0 for a zero vector).
__global__ void normalize_kernel(const double* w, const double* nrm2, double* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hn = sqrt(fmax(*nrm2, 0.0));
  v[i] = hn > 0.0 ? w[i] * (1.0 / hn) : 0.0;
}

// w -= V h   (V columns of length n, h on device)
__global__ void gemv_sub_kernel(const double* V, int64_t ldv, int ncols, const double* h, double* w,
                                int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int c = 0; c < ncols; ++c) s = fma(-V[(int64_t)c * ldv + i], h[c], s);
  w[i] = s;
}

__global__ void scale_kernel(const double* w, double s, double* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = w[i] * s;
}

__global__ void sub_kernel(const double* b, const double* ax, double* r, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - ax[i];
}

template <typename T>
__global__ void to_scaled_kernel(const double* r, const double* d, T* u, int64_t nc, int64_t npad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npad) u[i] = i < nc ? (T)(d[i] * r[i]) : T(0);
}

template <typename T>
__global__ void from_scaled_kernel(const T* u, const double* d, double* y, int64_t nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) y[i] = d[i] * (double)u[i];
}

__global__ void axpy_kernel(const double* a, double* x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += a[i];
}

template <typename T>
__global__ void add_diag_kernel(T* H, int64_t lda, int64_t n, T s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) H[i * lda + i] += s;
}

// U^T y partial sums (U row-major Nc x Lt)
__global__ void drift_dots_kernel(const double* U, int Lt, const double* y, int64_t nc, double* partial) {
  __shared__ double red[256];
  const int64_t r0 = (int64_t)blockIdx.x * RED_ROWS;
  const int64_t r1 = std::min<int64_t>(nc, r0 + RED_ROWS);
  for (int l = 0; l < Lt; ++l) {
    double acc = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) acc = fma(U[i * Lt + l], y[i], acc);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * Lt + l] = red[0];
    __syncthreads();
  }
}

// lambda = Sinv (sum_b partial - r2)
__global__ void drift_lambda_kernel(const double* partial, int nblk, int Lt, const double* r2,
                                    const double* Sinv, double* lam) {
  __shared__ double t[32];
  const int l = threadIdx.x;
  if (l < Lt) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[(int64_t)b * Lt + l];
    t[l] = s - (r2 ? r2[l] : 0.0);
  }
  __syncthreads();
  if (l < Lt) {
    double v = 0.0;
    for (int k = 0; k < Lt; ++k) v = fma(Sinv[l * Lt + k], t[k], v);
    lam[l] = v;
  }
}

// x[:Nc] = y - Y lam ; x[Nc:] = lam
__global__ void saddle_final_kernel(const double* y, const double* Y, int Lt, const double* lam, double* x,
                                    int64_t nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) {
    double s = y[i];
    for (int l = 0; l < Lt; ++l) s = fma(-Y[i * Lt + l], lam[l], s);
    x[i] = s;
  } else if (i < nc + Lt) {
    x[i] = lam[i - nc];
  }
}

__global__ void set_col_kernel(const double* src, double* Y, int Lt, int l, int64_t nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) Y[i * Lt + l] = src[i];
}
__global__ void get_col_kernel(const double* U, int Lt, int l, double* dst, int64_t nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) dst[i] = U[i * Lt + l];
}

// GK_VERBOSE=1 in the environment prints the refinement's residual history
// GK_FACTOR_UPDATE=fp32 runs the FP32 form's rank-256 updates on the FP32
// MFMA GEMM instead of the split-fp16 product.
inline bool factor_update_fp32() {
  static const bool v = [] {
    const char* e = std::getenv("GK_FACTOR_UPDATE");
    return e && std::string(e) == "fp32";
  }();
  return v;
}

inline bool gk_verbose() {
  static const bool v = [] { const char* e = getenv("GK_VERBOSE"); return e && *e && *e != '0'; }();
  return v;
}

// float copy of a scaled vector (padding rows zero)
__global__ void to_float_kernel(const double* u, float* f, int64_t nc, int64_t npad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npad) f[i] = i < nc ? (float)u[i] : 0.0f;
}

// y (+)= (double) f
__global__ void add_float_kernel(const float* f, double* y, int64_t nc, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) y[i] = accumulate ? y[i] + (double)f[i] : (double)f[i];
}

// x = [D y ; 0] (covariance unknowns scaled back, drift unknowns zero)
__global__ void unscale_pad_kernel(const double* y, const double* d, double* x, int64_t nc, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = i < nc ? d[i] * y[i] : 0.0;
}

// v = u - D w  (scaled residual of the covariance block)
__global__ void scaled_residual_kernel(const double* u, const double* d, const double* w, double* v, int64_t nc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) v[i] = u[i] - d[i] * w[i];
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)std::max<int64_t>(1, gk_ceil_div(n, b)); }

struct Solver {
  gk_handle h;
  gk_system sys_copy;
  const gk_system* sys;
  hipStream_t st;
  int64_t N, Nc, Lt, Npad, Ncpad;
  // preconditioner (FP32 or FP64 factor of the equilibrated covariance block,
  // or the learned sparse approximate inverse)
  bool f64 = false;
  bool spai = false;
  bool pcp = false;                  // local-inverse correction with the FP32 factor (needs spai blocks and a factor)
  bool pcp_sandwich = false;         // factor, correction, factor (for an approximate correction)
  hipEvent_t pc_ready = nullptr;     // recorded when spai_idx / spai_Y are built on another stream
  double* pc_u = nullptr;            // scaled right-hand side of the covariance block
  double* pc_y = nullptr;            // scaled iterate
  double* pc_v = nullptr;            // scaled residual / correction
  double* pc_x = nullptr;            // full-length operator input / output (N)
  double* pc_w = nullptr;
  float* pc_f = nullptr;             // FP32 right-hand side of the triangular solves
  int32_t* own_idx = nullptr;        // neighbour lists and exact local-inverse blocks of the FP32 form
  float* own_Y = nullptr;
  float* own_hk = nullptr;
  const int32_t* spai_idx = nullptr;
  int32_t spai_k = GK_SPAI_K;        // neighbours per point of spai_idx / spai_Y
  const float* spai_Y = nullptr;
  int32_t mv_prec = GK_F64;          // precision of the Krylov matvec
  double box[6] = {0, 0, 0, 0, 0, 0}; // bounding box of the system's points (drops the operator's range mask)
  const double* mv_box = nullptr;
  double shift = 0.0;                // diagonal shift of the factored block
  void* H = nullptr;
  void* dinv = nullptr;
  double* dscale = nullptr;
  double* U = nullptr;
  double* Y = nullptr;
  double* Sinv = nullptr;
  void* uvec = nullptr;
  float* ucol3 = nullptr;            // FP32 drift columns solved at once: uvec holds two, ucol3 the third
  double* ytmp = nullptr;
  double* ytmp2 = nullptr;
  double* lam = nullptr;
  double* partial = nullptr;
  // Krylov
  int m = 50;
  double* V = nullptr;
  double* Z = nullptr;
  double* w = nullptr;
  double* r = nullptr;
  double* hdev = nullptr;
  double* hcol = nullptr;            // Hessenberg columns on the device: column j at j (m + 2)
  double* b = nullptr;
  double* xt = nullptr;

  int host_dots(const double* Vc, int64_t ldv, int ncols, const double* vec, int64_t n, double* out_dev,
                double* out_host) {
    const unsigned nb = nblk(n, RED_ROWS);
    hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, (unsigned)ncols), dim3(256), 0, st, Vc, ldv, ncols, vec, n,
                       partial);
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(1), dim3(128), 0, st, partial, (int)nb, ncols, out_dev);
    if (out_host) {
      GK_HIP(hipMemcpyAsync(out_host, out_dev, ncols * sizeof(double), hipMemcpyDeviceToHost, st));
      GK_HIP(hipStreamSynchronize(st));
    }
    return GK_OK;
  }

  // out_dev[c] = <V[:, c], vec> (c < ncols) on the stream, no host copy.
  void dev_dots(const double* Vc, int64_t ldv, int ncols, const double* vec, int64_t n, double* out_dev,
                double* acc_dev = nullptr) {
    const unsigned nb = nblk(n, RED_ROWS);
    hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, (unsigned)ncols), dim3(256), 0, st, Vc, ldv, ncols, vec, n,
                       partial);
    if (acc_dev)
      hipLaunchKernelGGL(reduce_cols_acc_kernel, dim3(1), dim3(128), 0, st, partial, (int)nb, ncols, out_dev,
                         acc_dev);
    else
      hipLaunchKernelGGL(reduce_cols_kernel, dim3(1), dim3(128), 0, st, partial, (int)nb, ncols, out_dev);
  }

  int norm(const double* v, int64_t n, double* out) {
    GK_TRY(host_dots(v, n, 1, v, n, hdev + 120, out));
    *out = std::sqrt(std::max(*out, 0.0));
    return GK_OK;
  }

  // FP32 shifted factor applied to a scaled vector: f = (L L^T)^-1 u
  int factor_apply_f32(const double* u, double* y, bool accumulate) {
    hipLaunchKernelGGL(to_float_kernel, dim3(nblk(Ncpad)), dim3(256), 0, st, u, pc_f, Nc, Ncpad);
    GK_TRY(gk_trsv_chol_impl<float>(h, static_cast<float*>(H), static_cast<float*>(dinv), Ncpad, pc_f, st));
    hipLaunchKernelGGL(add_float_kernel, dim3(nblk(Nc)), dim3(256), 0, st, pc_f, y, Nc, accumulate ? 1 : 0);
    return GK_OK;
  }

  // v = u - H y on the equilibrated covariance block (H = D A_cc D), matrix-free
  int scaled_residual(const double* u, const double* y, double* v) {
    hipLaunchKernelGGL(unscale_pad_kernel, dim3(nblk(N)), dim3(256), 0, st, y, dscale, pc_x, Nc, N);
    GK_TRY(gk_matvec_boxed(h, sys, GK_F64, pc_x, pc_w, st, mv_box));
    hipLaunchKernelGGL(scaled_residual_kernel, dim3(nblk(Nc)), dim3(256), 0, st, u, dscale, pc_w, v, Nc);
    return GK_OK;
  }

  // K^-1 r (covariance block).  Forms: the Cholesky factor of the equilibrated
  // block (FP32 shifted or FP64); the sparse approximate inverse M alone; or
  // the multiplicative combination of M and the FP32 shifted factor P:
  //   y = M u;  y += P (u - H y)
  // M removes the local modes that the shift leaves (eigenvalues below the
  // shift), P the smooth (large-eigenvalue) error.  One factor application
  // and one operator product per call.  The sandwich P, M, P costs two of
  // each; it needs a few iterations fewer and tolerates an inexact M (the
  // last factor step removes the smooth error M introduces).
  int apply_K(const double* rin, double* yout) {
    if (pcp && !f64) {
      hipLaunchKernelGGL(to_scaled_kernel<double>, dim3(nblk(Nc)), dim3(256), 0, st, rin, dscale, pc_u, Nc, Nc);
      if (pcp_sandwich) {         // y = P u;  y += M (u - H y);  y += P (u - H y)
        GK_TRY(factor_apply_f32(pc_u, pc_y, false));
        GK_TRY(scaled_residual(pc_u, pc_y, pc_v));
        GK_TRY(gk_spai_apply_k(sys, spai_k, spai_idx, spai_Y, pc_v, ytmp2, st));
        hipLaunchKernelGGL(axpy_kernel, dim3(nblk(Nc)), dim3(256), 0, st, ytmp2, pc_y, Nc);
      } else {
        GK_TRY(gk_spai_apply_k(sys, spai_k, spai_idx, spai_Y, pc_u, pc_y, st));
      }
      GK_TRY(scaled_residual(pc_u, pc_y, pc_v));
      GK_TRY(factor_apply_f32(pc_v, pc_y, true));
      hipLaunchKernelGGL(from_scaled_kernel<double>, dim3(nblk(Nc)), dim3(256), 0, st, pc_y, dscale, yout, Nc);
      return GK_OK;
    }
    if (spai) {
      double* u = static_cast<double*>(uvec);
      hipLaunchKernelGGL(to_scaled_kernel<double>, dim3(nblk(Nc)), dim3(256), 0, st, rin, dscale, u, Nc, Nc);
      GK_TRY(gk_spai_apply_k(sys, spai_k, spai_idx, spai_Y, u, ytmp2, st));
      hipLaunchKernelGGL(from_scaled_kernel<double>, dim3(nblk(Nc)), dim3(256), 0, st, ytmp2, dscale, yout, Nc);
      return GK_OK;
    }
    if (f64) {
      double* u = static_cast<double*>(uvec);
      hipLaunchKernelGGL(to_scaled_kernel<double>, dim3(nblk(Ncpad)), dim3(256), 0, st, rin, dscale, u, Nc, Ncpad);
      GK_TRY(gk_trsv_chol_impl<double>(h, static_cast<double*>(H), static_cast<double*>(dinv), Ncpad, u, st));
      hipLaunchKernelGGL(from_scaled_kernel<double>, dim3(nblk(Nc)), dim3(256), 0, st, u, dscale, yout, Nc);
    } else {
      float* u = static_cast<float*>(uvec);
      hipLaunchKernelGGL(to_scaled_kernel<float>, dim3(nblk(Ncpad)), dim3(256), 0, st, rin, dscale, u, Nc, Ncpad);
      GK_TRY(gk_trsv_chol_impl<float>(h, static_cast<float*>(H), static_cast<float*>(dinv), Ncpad, u, st));
      hipLaunchKernelGGL(from_scaled_kernel<float>, dim3(nblk(Nc)), dim3(256), 0, st, u, dscale, yout, Nc);
    }
    return GK_OK;
  }

  // M^-1 r for the saddle system
  int apply_M(const double* rin, double* xout) {
    GK_TRY(apply_K(rin, ytmp));
    if (Lt > 0) {
      const unsigned nb = nblk(Nc, RED_ROWS);
      hipLaunchKernelGGL(drift_dots_kernel, dim3(nb), dim3(256), 0, st, U, (int)Lt, ytmp, Nc, partial);
      hipLaunchKernelGGL(drift_lambda_kernel, dim3(1), dim3(32), 0, st, partial, (int)nb, (int)Lt, rin + Nc,
                         Sinv, lam);
    }
    hipLaunchKernelGGL(saddle_final_kernel, dim3(nblk(Nc + Lt)), dim3(256), 0, st, ytmp, Y, (int)Lt, lam, xout,
                       Nc);
    GK_HIP(hipGetLastError());
    return GK_OK;
  }

  // Assemble and factor the equilibrated covariance block (unit diagonal).
  // FP32: the block is factored with a diagonal shift, H + sigma I (shifted
  // low-precision Cholesky, Higham & Pranesh, SIAM J. Sci. Comput. 43(1),
  // 2021): rounding H to FP32 alone perturbs it by ~u32 sqrt(n), more than its
  // smallest eigenvalues, so the unshifted FP32 factorization breaks down.
  // sigma starts at 4 u32 sqrt(n) and grows 4x per breakdown (4 attempts);
  // the FP64 GMRES outer iteration removes the shift's effect.  FP64: no
  // shift unless the factorization breaks down (then 1e-13, 1e-11).
  int setup_preconditioner(bool use_f64, int* info_out) {
    f64 = use_f64;
    const size_t esz = f64 ? sizeof(double) : sizeof(float);
    void* p = nullptr;
    GK_TRY(gk_workspace(h, WS_MATRIX, (size_t)Ncpad * Ncpad * esz, &p));
    H = p;
    int32_t info = 0;
    shift = f64 ? 0.0 : 4.0 * 5.9604644775390625e-08 * std::sqrt((double)Nc);
    for (int attempt = 0; attempt < (f64 ? 3 : 4); ++attempt) {
      GK_TRY(gk_assemble_impl(h, sys, f64 ? GK_F64 : GK_F32, H, Ncpad, nullptr, dscale, st, 1));
      int rc;
      if (f64) {
        if (shift > 0.0)
          hipLaunchKernelGGL(add_diag_kernel<double>, dim3(nblk(Nc)), dim3(256), 0, st, static_cast<double*>(H),
                             Ncpad, Nc, shift);
        rc = gk_cholesky_f64(h, static_cast<double*>(H), Nc, Ncpad, static_cast<double*>(dinv), &info, st);
      } else {
        hipLaunchKernelGGL(add_diag_kernel<float>, dim3(nblk(Nc)), dim3(256), 0, st, static_cast<float*>(H),
                           Ncpad, Nc, (float)shift);
        // the equilibrated block has a unit diagonal: its factor's entries
        // lie within [-1.01, 1.01], inside the split-fp16 update's range
        rc = factor_update_fp32()
                 ? gk_cholesky_f32(h, static_cast<float*>(H), Nc, Ncpad, static_cast<float*>(dinv), &info, st)
                 : gk_cholesky_f32_split16(h, static_cast<float*>(H), Nc, Ncpad, static_cast<float*>(dinv), &info,
                                           st);
      }
      if (gk_verbose())
        fprintf(stderr, "[gk] %s Cholesky of the covariance block, shift %.3e: %s (info %d)\n", f64 ? "FP64" : "FP32",
                shift, rc == GK_OK ? "ok" : "breakdown", info);
      if (rc == GK_OK) break;
      if (rc != GK_ERR_NOT_SPD) return rc;
      shift = f64 ? (shift == 0.0 ? 1e-13 : shift * 100.0) : shift * 4.0;
    }
    *info_out = info;
    if (info != 0) return GK_ERR_NOT_SPD;
    if (pc_ready) GK_HIP(hipStreamWaitEvent(st, pc_ready, 0));    // local preconditioner built beside the factor
    return setup_drift();
  }

  // Y = K^-1 U with apply_K (the factor without the local correction) and the
  // inverse of S = U^T Y (Lt x Lt, host)
  int setup_drift() {
    if (Lt > 0) {
      GK_TRY(gk_assemble_drift_impl(sys, U, st));
      // The drift columns take the factor alone (without the local-inverse
      // correction and its operator product): the flexible iteration does not
      // need Y to come from the same approximation of K, and at N = 35k the
      // iteration count is the same.
      if (!f64 && !spai && Lt > 1) {
        // FP32 factor: three columns at a time, their solves on the caller's,
        // the side and the auxiliary stream with solve sets 0, 1, 2 (each
        // solve is bound by its dependency chain, not by HBM, so they overlap)
        hipStream_t sx[3] = {st, nullptr, nullptr};
        hipEvent_t* e2 = nullptr;
        hipEvent_t* e3 = nullptr;
        GK_TRY(gk_side_stream(h, &sx[1], &e2));
        GK_TRY(gk_aux_stream(h, &sx[2], &e3));
        float* ub[3] = {static_cast<float*>(uvec), static_cast<float*>(uvec) + Ncpad, ucol3};
        for (int l0 = 0; l0 < Lt; l0 += 3) {
          const int nl = (int)std::min<int64_t>(3, Lt - l0);
          for (int c = 0; c < nl; ++c) {
            hipLaunchKernelGGL(get_col_kernel, dim3(nblk(Nc)), dim3(256), 0, st, U, (int)Lt, l0 + c, r, Nc);
            hipLaunchKernelGGL(to_scaled_kernel<float>, dim3(nblk(Ncpad)), dim3(256), 0, st, r, dscale, ub[c], Nc,
                               Ncpad);
          }
          GK_HIP(hipEventRecord(e2[0], st));
          for (int c = 1; c < nl; ++c) GK_HIP(hipStreamWaitEvent(sx[c], e2[0], 0));
          for (int c = 0; c < nl; ++c)
            GK_TRY(gk_trsv_chol_impl<float>(h, static_cast<float*>(H), static_cast<float*>(dinv), Ncpad, ub[c],
                                            sx[c], c));
          if (nl > 1) {
            GK_HIP(hipEventRecord(e2[1], sx[1]));
            GK_HIP(hipStreamWaitEvent(st, e2[1], 0));
          }
          if (nl > 2) {
            GK_HIP(hipEventRecord(e3[0], sx[2]));
            GK_HIP(hipStreamWaitEvent(st, e3[0], 0));
          }
          for (int c = 0; c < nl; ++c) {
            hipLaunchKernelGGL(from_scaled_kernel<float>, dim3(nblk(Nc)), dim3(256), 0, st, ub[c], dscale, w, Nc);
            hipLaunchKernelGGL(set_col_kernel, dim3(nblk(Nc)), dim3(256), 0, st, w, Y, (int)Lt, l0 + c, Nc);
          }
        }
        GK_HIP(hipGetLastError());
      } else {
        for (int l = 0; l < Lt; ++l) {
          hipLaunchKernelGGL(get_col_kernel, dim3(nblk(Nc)), dim3(256), 0, st, U, (int)Lt, l, r, Nc);
          const bool keep = pcp;
          pcp = false;
          const int rc_k = apply_K(r, w);
          pcp = keep;
          GK_TRY(rc_k);
          hipLaunchKernelGGL(set_col_kernel, dim3(nblk(Nc)), dim3(256), 0, st, w, Y, (int)Lt, l, Nc);
        }
      }
      // S = U^T Y (Lt x Lt, column l at hdev + 512 + l Lt), one read-back,
      // inverted on the host
      GK_REQUIRE(Lt <= 22, GK_ERR_ARG, "gk_solve: at most 22 drift and fault columns (got %d)", (int)Lt);
      std::vector<double> S(Lt * Lt), cols(Lt * Lt);
      for (int l = 0; l < Lt; ++l) {
        hipLaunchKernelGGL(get_col_kernel, dim3(nblk(Nc)), dim3(256), 0, st, Y, (int)Lt, l, w, Nc);
        const unsigned nb = nblk(Nc, RED_ROWS);
        hipLaunchKernelGGL(drift_dots_kernel, dim3(nb), dim3(256), 0, st, U, (int)Lt, w, Nc, partial);
        hipLaunchKernelGGL(reduce_cols_kernel, dim3(1), dim3(128), 0, st, partial, (int)nb, (int)Lt,
                           hdev + 512 + l * Lt);
      }
      GK_HIP(hipMemcpyAsync(cols.data(), hdev + 512, Lt * Lt * sizeof(double), hipMemcpyDeviceToHost, st));
      GK_HIP(hipStreamSynchronize(st));
      for (int l = 0; l < Lt; ++l)
        for (int k = 0; k < Lt; ++k) S[k * Lt + l] = cols[l * Lt + k];
      std::vector<double> inv(Lt * Lt, 0.0);
      for (int i = 0; i < Lt; ++i) inv[i * Lt + i] = 1.0;
      for (int c = 0; c < Lt; ++c) {           // Gauss-Jordan with partial pivoting
        int piv = c;
        for (int i = c + 1; i < Lt; ++i)
          if (std::fabs(S[i * Lt + c]) > std::fabs(S[piv * Lt + c])) piv = i;
        GK_REQUIRE(S[piv * Lt + c] != 0.0, GK_ERR_SINGULAR, "gk_solve: singular drift Schur complement");
        for (int k = 0; k < Lt; ++k) { std::swap(S[c * Lt + k], S[piv * Lt + k]); std::swap(inv[c * Lt + k], inv[piv * Lt + k]); }
        const double d = 1.0 / S[c * Lt + c];
        for (int k = 0; k < Lt; ++k) { S[c * Lt + k] *= d; inv[c * Lt + k] *= d; }
        for (int i = 0; i < Lt; ++i) {
          if (i == c) continue;
          const double f = S[i * Lt + c];
          if (f == 0.0) continue;
          for (int k = 0; k < Lt; ++k) { S[i * Lt + k] -= f * S[c * Lt + k]; inv[i * Lt + k] -= f * inv[c * Lt + k]; }
        }
      }
      GK_HIP(hipMemcpyAsync(Sinv, inv.data(), Lt * Lt * sizeof(double), hipMemcpyHostToDevice, st));
      GK_HIP(hipStreamSynchronize(st));
    }
    return GK_OK;
  }

  // Right-preconditioned restarted flexible GMRES on the FP64 matrix-free operator.
  // The true FP64 residual is checked every `check` inner iterations: the
  // iteration stops when it is below tol, or when it no longer halves between
  // checks (the FP64 rounding floor of the operator has been reached).
  // x0 (nullable): initial guess (the previous solution of a related system).
  int fgmres(double* x, double tol, double accept, int max_iter, int* iters, double* relres, bool* ok,
             const double* x0 = nullptr) {
    const int check = 8;
    double nb = 0.0;
    GK_TRY(norm(b, N, &nb));
    if (nb == 0.0) nb = 1.0;
    if (x0)
      GK_HIP(hipMemcpyAsync(x, x0, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    else
      GK_HIP(hipMemsetAsync(x, 0, N * sizeof(double), st));
    int total = 0;
    double last_true = 1.0;
    std::vector<double> hist;                 // true residual at each check
    std::vector<double> Hh((m + 1) * m), cs(m), sn(m), g(m + 1), yv(m), hv_cols((size_t)m * (m + 2));
    *ok = false;
    bool done = false;
    // out = x + Z y(j), y from the first j columns of the rotated Hessenberg
    auto combine = [&](int j, double* out) -> int {
      for (int i = j - 1; i >= 0; --i) {
        double sacc = g[i];
        for (int k2 = i + 1; k2 < j; ++k2) sacc -= Hh[i * m + k2] * yv[k2];
        yv[i] = sacc / Hh[i * m + i];
      }
      std::vector<double> ny(j);
      for (int i = 0; i < j; ++i) ny[i] = -yv[i];
      GK_HIP(hipMemcpyAsync(hdev + 64, ny.data(), j * sizeof(double), hipMemcpyHostToDevice, st));
      if (out != x) GK_HIP(hipMemcpyAsync(out, x, N * sizeof(double), hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(gemv_sub_kernel, dim3(nblk(N)), dim3(256), 0, st, Z, Npad, j, hdev + 64, out, N);
      GK_HIP(hipStreamSynchronize(st));
      return GK_OK;
    };
    auto true_rel = [&](const double* xv, double* out) -> int {
      GK_TRY(gk_matvec_boxed(h, sys, GK_F64, xv, w, st, mv_box));
      hipLaunchKernelGGL(sub_kernel, dim3(nblk(N)), dim3(256), 0, st, b, w, r, N);
      double nr = 0.0;
      GK_TRY(norm(r, N, &nr));
      GK_TRY(gk_trsv_check(h, st));
      *out = nr / nb;
      return GK_OK;
    };
    bool x_zero = x0 == nullptr;           // the first cycle from x = 0 has r = b
    while (!done) {
      double rel0 = 1.0;
      if (x_zero) GK_HIP(hipMemcpyAsync(r, b, N * sizeof(double), hipMemcpyDeviceToDevice, st));
      else GK_TRY(true_rel(x, &rel0));     // r = b - A x
      x_zero = false;
      last_true = rel0;
      if (rel0 <= tol || total >= max_iter) { *ok = rel0 <= accept; break; }
      const double beta = rel0 * nb;
      hipLaunchKernelGGL(scale_kernel, dim3(nblk(N)), dim3(256), 0, st, r, 1.0 / beta, V, N);
      std::fill(g.begin(), g.end(), 0.0);
      g[0] = beta;
      int j = 0;
      bool restart = false;
      int jdone = 0;                        // Hessenberg columns already rotated on the host
      while (!done && !restart) {
        double* vj = V + (int64_t)j * Npad;
        double* zj = Z + (int64_t)j * Npad;
        GK_TRY(apply_M(vj, zj));
        GK_TRY(gk_matvec_boxed(h, sys, mv_prec, zj, w, st, mv_box));
        // classical Gram-Schmidt, two passes, and the normalisation, all on the
        // device: the Hessenberg column goes to the host at the next check
        double* hc = hcol + (int64_t)j * (m + 2);
        dev_dots(V, Npad, j + 1, w, N, hc);
        hipLaunchKernelGGL(gemv_sub_kernel, dim3(nblk(N)), dim3(256), 0, st, V, Npad, j + 1, hc, w, N);
        dev_dots(V, Npad, j + 1, w, N, hdev, hc);
        hipLaunchKernelGGL(gemv_sub_kernel, dim3(nblk(N)), dim3(256), 0, st, V, Npad, j + 1, hdev, w, N);
        dev_dots(w, N, 1, w, N, hc + j + 1);
        hipLaunchKernelGGL(normalize_kernel, dim3(nblk(N)), dim3(256), 0, st, w, hc + j + 1,
                           V + (int64_t)(j + 1) * Npad, N);
        ++total;
        ++j;
        // Between checks, past the first two, the Hessenberg column is also
        // rotated on the host after every iteration ("peek"): the true residual
        // is then checked as soon as the Arnoldi estimate reaches tol.
        const bool peek = total >= 2 * check;
        if (!(j == m || total >= max_iter || j % check == 0 || peek)) continue;
        // rotate the new columns (j - jdone of them) on the host
        GK_HIP(hipMemcpyAsync(hv_cols.data(), hcol + (int64_t)jdone * (m + 2),
                              (size_t)(j - jdone) * (m + 2) * sizeof(double), hipMemcpyDeviceToHost, st));
        GK_HIP(hipStreamSynchronize(st));
        double hn = 1.0;
        for (int jj = jdone; jj < j; ++jj) {
          const double* col = hv_cols.data() + (int64_t)(jj - jdone) * (m + 2);
          hn = std::sqrt(std::max(col[jj + 1], 0.0));
          for (int i = 0; i <= jj; ++i) Hh[i * m + jj] = col[i];
          Hh[(jj + 1) * m + jj] = hn;
          for (int i = 0; i < jj; ++i) {
            const double a = Hh[i * m + jj], c = Hh[(i + 1) * m + jj];
            Hh[i * m + jj] = cs[i] * a + sn[i] * c;
            Hh[(i + 1) * m + jj] = -sn[i] * a + cs[i] * c;
          }
          const double a = Hh[jj * m + jj], c = Hh[(jj + 1) * m + jj];
          const double den = std::hypot(a, c);
          cs[jj] = den > 0 ? a / den : 1.0;
          sn[jj] = den > 0 ? c / den : 0.0;
          Hh[jj * m + jj] = den;
          Hh[(jj + 1) * m + jj] = 0.0;
          g[jj + 1] = -sn[jj] * g[jj];
          g[jj] = cs[jj] * g[jj];
          if (hn == 0.0) {                  // breakdown: the Krylov space ends at column jj
            j = jj + 1;
            break;
          }
        }
        jdone = j;
        const bool full = (j == m) || hn == 0.0 || total >= max_iter;
        if (!full && j % check != 0 && std::fabs(g[j]) / nb > tol) continue;   // a peek above tol
        {
          GK_TRY(combine(j, xt));
          double rel = 0.0;
          GK_TRY(true_rel(xt, &rel));
          const bool converged = rel <= tol;
          // The Arnoldi estimate |g_j| tracks the true residual until the
          // true one reaches its floor (FP64 rounding of the operator, or an
          // FP32 Arnoldi matvec); once it runs 10x below, the current cycle
          // cannot lower the true residual any more.
          const double est = std::fabs(g[j]) / nb;
          const bool floor_hit = est < 0.1 * rel;
          // Below `accept` a check that does not halve the residual, or a
          // floor hit, ends the iteration; above it, allow 3 checks for a
          // halving, and restart the cycle from the true residual at a floor.
          const bool stalled = rel <= accept ? (rel > 0.5 * last_true || floor_hit)
                                             : hist.size() >= 3 && rel > 0.5 * hist[hist.size() - 3];
          last_true = rel;
          hist.push_back(rel);
          if (gk_verbose())
            fprintf(stderr, "[gk] gmres it %d: true relres %.3e (estimate %.3e)%s%s\n", total, rel, est,
                    converged ? " (converged)" : "", stalled ? " (stalled)" : "");
          if (converged || stalled || total >= max_iter || hn == 0.0) {
            GK_HIP(hipMemcpyAsync(x, xt, N * sizeof(double), hipMemcpyDeviceToDevice, st));
            *ok = rel <= accept;
            done = true;
          } else if (full || floor_hit) {
            GK_HIP(hipMemcpyAsync(x, xt, N * sizeof(double), hipMemcpyDeviceToDevice, st));
            restart = true;
          }
        }
      }
    }
    GK_HIP(hipStreamSynchronize(st));
    *iters = total;
    *relres = last_true;
    return GK_OK;
  }

  // x = M^-1 b with the FP64 factor, then refinement steps x += M^-1 (b - A x)
  // while the FP64 residual keeps decreasing.
  int direct_refine(double* x, int max_steps, int* steps, double* relres) {
    double nb = 0.0;
    GK_TRY(norm(b, N, &nb));
    if (nb == 0.0) nb = 1.0;
    GK_TRY(apply_M(b, x));
    double prev = 1e300, rel = 0.0;
    int k = 0;
    for (; k < max_steps; ++k) {
      GK_TRY(gk_matvec_boxed(h, sys, GK_F64, x, w, st, mv_box));
      hipLaunchKernelGGL(sub_kernel, dim3(nblk(N)), dim3(256), 0, st, b, w, r, N);
      double nr = 0.0;
      GK_TRY(norm(r, N, &nr));
      rel = nr / nb;
      if (rel > 0.5 * prev) break;
      prev = rel;
      GK_TRY(apply_M(r, xt));
      hipLaunchKernelGGL(axpy_kernel, dim3(nblk(N)), dim3(256), 0, st, xt, x, N);
    }
    *steps = k;
    *relres = rel;
    // a triangular solve whose wait reached its spin limit leaves x unusable
    GK_TRY(gk_trsv_check(h, st));
    return GK_OK;
  }

  // ||b - A x|| / ||b|| with the FP64 matrix-free operator
  int true_relres(const double* x, double* out) {
    double nb = 0.0, nr = 0.0;
    GK_TRY(norm(b, N, &nb));
    GK_TRY(gk_matvec_boxed(h, sys, GK_F64, x, w, st, mv_box));
    hipLaunchKernelGGL(sub_kernel, dim3(nblk(N)), dim3(256), 0, st, b, w, r, N);
    GK_TRY(norm(r, N, &nr));
    *out = nb > 0 ? nr / nb : nr;
    return GK_OK;
  }
};

double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

}  // namespace

static int solve_lu(gk_handle h, const gk_system* sys, hipStream_t st, int64_t N, int64_t Npad, double* x,
                    gk_solve_info* info, hipEvent_t* ev) {
  void* p = nullptr;
  GK_TRY(gk_workspace(h, WS_MATRIX2, (size_t)Npad * Npad * sizeof(double) + (size_t)3 * Npad * GK_BLOCK * sizeof(double), &p));
  double* A = static_cast<double*>(p);
  double* dinv = A + (size_t)Npad * Npad;
  double* bpad = dinv + (size_t)2 * Npad * GK_BLOCK;
  GK_HIP(hipEventRecord(ev[0], st));
  GK_HIP(hipMemsetAsync(bpad, 0, Npad * sizeof(double), st));
  GK_TRY(gk_assemble_impl(h, sys, GK_F64, A, Npad, bpad, nullptr, st, 0));
  GK_HIP(hipEventRecord(ev[1], st));
  int32_t finfo = 0;
  int rc = gk_lu_f64(h, A, N, Npad, dinv, &finfo, st);
  info->info = finfo;
  if (rc != GK_OK) return rc;
  GK_HIP(hipEventRecord(ev[2], st));
  GK_TRY(gk_lu_solve_f64(h, A, dinv, Npad, bpad, st));
  GK_HIP(hipMemcpyAsync(x, bpad, N * sizeof(double), hipMemcpyDeviceToDevice, st));
  GK_HIP(hipEventRecord(ev[3], st));
  GK_HIP(hipEventSynchronize(ev[3]));
  info->t_assemble_ms += elapsed_ms(ev[0], ev[1]);
  info->t_factor_ms += elapsed_ms(ev[1], ev[2]);
  info->t_refine_ms += elapsed_ms(ev[2], ev[3]);
  info->form = GK_SOLVE_FP64_LU;
  return GK_OK;
}

// Carve the solver's vectors out of the WS_SOLVER workspace and set b.
static int init_solver(gk_handle h, const gk_system* sys, hipStream_t st, int form, Solver& S) {
  const int64_t N = gk_system_n(sys);
  const int64_t Npad = gk_round_up(N, GK_BLOCK);
  const int L = sys->uni_degree == 0 ? 0 : (sys->uni_degree == 1 ? 3 : 9);
  const int64_t Lt = L + sys->n_fault;
  const int64_t Nc = N - Lt;
  S.h = h; S.sys_copy = *sys; S.sys = &S.sys_copy; S.st = st;
  if (sys->n_ori + sys->n_rest + sys->n_surf > 0) {
    GK_TRY(gk_source_box(h, sys, S.box, st));
    S.mv_box = S.box;
  }
  S.N = N; S.Nc = Nc; S.Lt = Lt; S.Npad = Npad; S.Ncpad = gk_round_up(std::max<int64_t>(Nc, 1), GK_BLOCK);
  const bool fast = form != GK_SOLVE_FP64_LU && form != GK_SOLVE_FP64_CHOL;   // needs the Krylov basis
  const int m = fast ? S.m : 1;
  size_t need = 0;
  auto take = [&](size_t bytes) { size_t o = need; need += gk_round_up(bytes, 256); return o; };
  const size_t o_dinv = take(form != GK_SOLVE_FP64_LU ? (size_t)GK_CHOL_PLANES * S.Ncpad * GK_BLOCK * sizeof(double) : 0);
  const size_t o_dsc = take(S.Ncpad * sizeof(double));
  const size_t o_U = take((size_t)S.Ncpad * std::max<int64_t>(Lt, 1) * sizeof(double));
  const size_t o_Y = take((size_t)S.Ncpad * std::max<int64_t>(Lt, 1) * sizeof(double));
  const size_t o_Si = take(32 * 32 * sizeof(double));
  const size_t o_u32 = take(S.Ncpad * sizeof(double));
  const size_t o_u3 = take(S.Ncpad * sizeof(float));   // third FP32 drift column (two fit in o_u32)
  const size_t o_yt = take(Npad * sizeof(double));
  const size_t o_yt2 = take(Npad * sizeof(double));
  const size_t o_lam = take(64 * sizeof(double));
  const size_t o_part = take((size_t)(gk_ceil_div(Npad, RED_ROWS) + 1) * 64 * sizeof(double));
  const size_t o_V = take((size_t)(m + 1) * Npad * sizeof(double));
  const size_t o_Z = take((size_t)m * Npad * sizeof(double));
  const size_t o_w = take(Npad * sizeof(double));
  const size_t o_r = take(Npad * sizeof(double));
  const size_t o_h = take(1024 * sizeof(double));     // [0, 256): dots and coefficients; [512, 801): S
  const size_t o_hc = take((size_t)m * (m + 2) * sizeof(double));
  const size_t o_b = take(Npad * sizeof(double));
  const size_t o_xt = take(Npad * sizeof(double));
  const bool pc = form == GK_SOLVE_FP32_REFINE || form == 3;
  const size_t o_pu = take(pc ? S.Ncpad * sizeof(double) : 0);
  const size_t o_py = take(pc ? S.Ncpad * sizeof(double) : 0);
  const size_t o_pv = take(pc ? S.Ncpad * sizeof(double) : 0);
  const size_t o_px = take(pc ? Npad * sizeof(double) : 0);
  const size_t o_pw = take(pc ? Npad * sizeof(double) : 0);
  const size_t o_pf = take(pc ? S.Ncpad * sizeof(float) : 0);
  const size_t o_si = take(pc ? (size_t)(sys->n_ori + sys->n_rest) * GK_LOCAL_K * sizeof(int32_t) : 0);
  const size_t o_sy = take(pc ? (size_t)(sys->n_ori + sys->n_rest) * GK_LOCAL_K * 9 * sizeof(float) : 0);
  const size_t o_hk = take(pc ? (size_t)(sys->n_ori + sys->n_rest) * sizeof(float) : 0);
  void* p = nullptr;
  GK_TRY(gk_workspace(h, WS_SOLVER, need, &p));
  char* base = static_cast<char*>(p);
  S.dinv = base + o_dinv;
  S.dscale = reinterpret_cast<double*>(base + o_dsc);
  S.U = reinterpret_cast<double*>(base + o_U);
  S.Y = reinterpret_cast<double*>(base + o_Y);
  S.Sinv = reinterpret_cast<double*>(base + o_Si);
  S.uvec = base + o_u32;
  S.ucol3 = reinterpret_cast<float*>(base + o_u3);
  S.ytmp = reinterpret_cast<double*>(base + o_yt);
  S.ytmp2 = reinterpret_cast<double*>(base + o_yt2);
  S.lam = reinterpret_cast<double*>(base + o_lam);
  S.partial = reinterpret_cast<double*>(base + o_part);
  S.V = reinterpret_cast<double*>(base + o_V);
  S.Z = reinterpret_cast<double*>(base + o_Z);
  S.w = reinterpret_cast<double*>(base + o_w);
  S.r = reinterpret_cast<double*>(base + o_r);
  S.hdev = reinterpret_cast<double*>(base + o_h);
  S.hcol = reinterpret_cast<double*>(base + o_hc);
  S.b = reinterpret_cast<double*>(base + o_b);
  S.xt = reinterpret_cast<double*>(base + o_xt);
  if (pc) {
    S.pc_u = reinterpret_cast<double*>(base + o_pu);
    S.pc_y = reinterpret_cast<double*>(base + o_py);
    S.pc_v = reinterpret_cast<double*>(base + o_pv);
    S.pc_x = reinterpret_cast<double*>(base + o_px);
    S.pc_w = reinterpret_cast<double*>(base + o_pw);
    S.pc_f = reinterpret_cast<float*>(base + o_pf);
    S.own_idx = reinterpret_cast<int32_t*>(base + o_si);
    S.own_Y = reinterpret_cast<float*>(base + o_sy);
    S.own_hk = reinterpret_cast<float*>(base + o_hk);
  }
  GK_TRY(gk_rhs_impl(sys, S.b, st));
  return GK_OK;
}

static int solve_impl(gk_handle h, const gk_system* sys, int32_t form, double tol, int32_t max_iter,
                      const double* x0, double* x, gk_solve_info* info, gk_stream stream) {
  GK_REQUIRE(h && sys && x && info, GK_ERR_ARG, "gk_solve: NULL argument");
  GK_REQUIRE(x0 != x || x0 == nullptr, GK_ERR_ARG, "gk_solve_x0: x0 and x must not alias");
  GK_REQUIRE(form == GK_SOLVE_FP64_LU || form == GK_SOLVE_FP32_REFINE || form == GK_SOLVE_FP64_CHOL, GK_ERR_ARG,
             "gk_solve: bad form");
  hipStream_t st = (hipStream_t)stream;
  *info = gk_solve_info{};
  const int64_t N = gk_system_n(sys);
  GK_REQUIRE(N > 0, GK_ERR_ARG, "gk_solve: empty system");
  const int64_t Npad = gk_round_up(N, GK_BLOCK);
  const int L = sys->uni_degree == 0 ? 0 : (sys->uni_degree == 1 ? 3 : 9);
  const int64_t Lt = L + sys->n_fault;
  const int64_t Nc = N - Lt;
  hipEvent_t ev[4];
  for (auto& e : ev) GK_HIP(hipEventCreate(&e));
  struct EvGuard { hipEvent_t* e; ~EvGuard() { for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]); } } guard{ev};
  if (tol <= 0) tol = 1e-10;
  if (max_iter <= 0) max_iter = 300;

  if (h->solver_state && h->solver_free) h->solver_free(h->solver_state);
  h->solver_state = nullptr;
  Solver S;
  GK_TRY(init_solver(h, sys, st, form, S));

  if (form == GK_SOLVE_FP64_LU) {
    GK_TRY(solve_lu(h, sys, st, N, Npad, x, info, ev));
    GK_TRY(S.true_relres(x, &info->relres));
    return GK_OK;
  }
  auto run_chol64 = [&]() -> int {
    GK_HIP(hipEventRecord(ev[0], st));
    int32_t finfo = 0;
    GK_TRY(S.setup_preconditioner(true, &finfo));
    info->info = finfo;
    GK_HIP(hipEventRecord(ev[1], st));
    int steps = 0;
    GK_TRY(S.direct_refine(x, 3, &steps, &info->relres));
    GK_HIP(hipEventRecord(ev[2], st));
    GK_HIP(hipEventSynchronize(ev[2]));
    info->t_factor_ms += elapsed_ms(ev[0], ev[1]);
    info->t_refine_ms += elapsed_ms(ev[1], ev[2]);
    info->iterations += steps;
    info->form = GK_SOLVE_FP64_CHOL;
    return GK_OK;
  };
  auto keep_state = [&]() {
    Solver* k = new Solver(S);
    k->sys = &k->sys_copy;
    h->solver_state = k;
    h->solver_free = [](void* p) { delete static_cast<Solver*>(p); };
  };
  if (form == GK_SOLVE_FP64_CHOL) {
    GK_TRY(run_chol64());
    keep_state();
    return GK_OK;
  }
  // FP32 shifted factor + exact local-inverse correction, FP64 GMRES refinement
  GK_HIP(hipEventRecord(ev[0], st));
  GK_TRY(gk_diag_scale_impl(sys, S.dscale, st));
  // The local-inverse correction needs GK_LOCAL_K data points for its
  // neighbourhoods; smaller systems use the shifted factor alone (their shift,
  // 4 u32 sqrt(n), is tiny, and FGMRES removes it).
  const bool local = sys->n_ori + sys->n_rest >= GK_LOCAL_K;
  if (local) {               // neighbourhoods and local inverses overlap the factorization
    hipStream_t s3;
    hipEvent_t* aev;
    GK_TRY(gk_aux_stream(h, &s3, &aev));
    GK_HIP(hipEventRecord(aev[0], st));
    GK_HIP(hipStreamWaitEvent(s3, aev[0], 0));
    GK_TRY(gk_knn_local(h, sys, S.box, S.own_idx, S.own_hk, (gk_stream)s3));
    GK_TRY(gk_local_inverse_local(h, sys, S.own_idx, S.dscale, S.own_Y, (gk_stream)s3));
    GK_HIP(hipEventRecord(aev[1], s3));
    S.pc_ready = aev[1];
    S.spai_idx = S.own_idx;
    S.spai_k = GK_LOCAL_K;
    S.spai_Y = S.own_Y;
  }
  S.pcp = local;
  int32_t finfo = 0;
  int rc = S.setup_preconditioner(false, &finfo);
  info->info = finfo;
  GK_HIP(hipEventRecord(ev[1], st));
  bool ok = false;
  if (rc == GK_OK) {
    GK_TRY(S.fgmres(x, tol, 1e-6, max_iter, &info->iterations, &info->relres, &ok, x0));
  } else if (rc != GK_ERR_NOT_SPD) {
    return rc;
  }
  GK_HIP(hipEventRecord(ev[2], st));
  GK_HIP(hipEventSynchronize(ev[2]));
  info->t_factor_ms = elapsed_ms(ev[0], ev[1]);
  info->t_refine_ms = elapsed_ms(ev[1], ev[2]);
  info->form = GK_SOLVE_FP32_REFINE;
  info->shift = S.shift;
  if (!ok) {
    info->fallback = 1;
    GK_TRY(run_chol64());
  }
  keep_state();
  return GK_OK;
}

extern "C" int gk_solve(gk_handle h, const gk_system* sys, int32_t form, double tol, int32_t max_iter,
                        double* x, gk_solve_info* info, gk_stream stream) {
  return solve_impl(h, sys, form, tol, max_iter, nullptr, x, info, stream);
}

extern "C" int gk_solve_x0(gk_handle h, const gk_system* sys, int32_t form, double tol, int32_t max_iter,
                           const double* x0, double* x, gk_solve_info* info, gk_stream stream) {
  return solve_impl(h, sys, form, tol, max_iter, x0, x, info, stream);
}

// ---------------------------------------------------------------- adjoint
namespace {

__global__ void gather_all_points_kernel(gk_system sys, double* pts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t no = sys.n_ori, R = sys.n_rest, S = sys.n_surf;
  if (i >= no + R + S) return;
  const double* p = i < no ? sys.ori_xyz + 3 * i
                    : (i < no + R ? sys.rest_xyz + 3 * (i - no) : sys.ref_xyz + 3 * (i - no - R));
  pts[3 * i] = p[0]; pts[3 * i + 1] = p[1]; pts[3 * i + 2] = p[2];
}

// c = Phi^T e on the covariance rows: gi grad E(q) and i_res (E(p_i) - E(p_ref))
__global__ void adjoint_rhs_kernel(gk_system sys, const double* E, const double* gE, double* c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t no = sys.n_ori, R = sys.n_rest, NG = 3 * no;
  if (i < NG) {
    const int64_t u = i / no, g = i % no;
    c[i] = sys.gi_res * gE[3 * g + u];
  } else if (i < NG + R) {
    const int64_t r = i - NG;
    c[i] = sys.i_res * (E[no + r] - E[no + R + sys.rest_surf[r]]);
  }
}

// drift rows of Phi^T e: gi_res sum_o e_o f_l(x_o), one block per drift function
__global__ void adjoint_drift_rhs_kernel(const double* obs, const double* e, int64_t n_obs, double gi, double* c) {
  __shared__ double red[256];
  const int l = blockIdx.x;
  double acc = 0.0;
  for (int64_t o = threadIdx.x; o < n_obs; o += 256) {
    const double* p = obs + 3 * o;
    double f;
    if (l < 3) f = p[l];
    else if (l < 6) f = p[l - 3] * p[l - 3];
    else if (l == 6) f = p[0] * p[1];
    else if (l == 7) f = p[0] * p[2];
    else f = p[1] * p[2];
    acc = fma(e[o], f, acc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) c[l] = gi * red[0];
}

__global__ void adjoint_fault_rhs_kernel(const double* F, int nf, int f, const double* e, int64_t n_obs, double gi,
                                         double* c) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t o = threadIdx.x; o < n_obs; o += 256) acc = fma(e[o], F[o * nf + f], acc);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *c = gi * red[0];
}

__global__ void surface_sum2_kernel(const double* a, const double* b, const int32_t* surf, int64_t R, double* Wa,
                                    double* Wb) {
  __shared__ double ra[256], rb[256];
  const int s = blockIdx.x;
  double x = 0.0, y = 0.0;
  for (int64_t i = threadIdx.x; i < R; i += 256)
    if (surf[i] == s) { x += a[i]; y += b[i]; }
  ra[threadIdx.x] = x; rb[threadIdx.x] = y;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) { ra[threadIdx.x] += ra[threadIdx.x + w]; rb[threadIdx.x] += rb[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { Wa[s] = ra[0]; Wb[s] = rb[0]; }
}

// dL/dp_a = i_res Omega_a grad E(p_a) - Lambda_a grad S_w(p_a) - Omega_a grad S_lam(p_a)
// with Omega (Lambda) the signed per-point weights of w (lam): w_i on a rest
// point, -sum over the surface on its reference point.
__global__ void adjoint_points_kernel(gk_system sys, const double* w, const double* lam, const double* Ww,
                                      const double* Wl, const double* gE, const double* gSw, const double* gSl,
                                      double* out) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t R = sys.n_rest, S = sys.n_surf, NG = 3 * sys.n_ori;
  if (a >= R + S) return;
  double Om, La;
  if (a < R) { Om = w[NG + a]; La = lam[NG + a]; }
  else { Om = -Ww[a - R]; La = -Wl[a - R]; }
  for (int u = 0; u < 3; ++u)
    out[3 * a + u] = sys.i_res * Om * gE[3 * a + u] - La * gSw[3 * a + u] - Om * gSl[3 * a + u];
}

// dL/dq_g = gi_res Hess E(q_g) w_g - (Hess S_w(q_g) lam_g + Hess S_lam(q_g) w_g) / gi_res
// (Hessians as xx, yy, zz, xy, xz, yz; w_g, lam_g the gradient weights of q_g)
__global__ void adjoint_ori_pos_kernel(int64_t no, double gi, const double* w, const double* lam,
                                       const double* HE, const double* HSw, const double* HSl, double* out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= no) return;
  const double wg[3] = {w[g], w[no + g], w[2 * no + g]};
  const double lg[3] = {lam[g], lam[no + g], lam[2 * no + g]};
  const int ix[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
  for (int v = 0; v < 3; ++v) {
    double e = 0.0, m = 0.0;
    for (int u = 0; u < 3; ++u) {
      const int c = ix[v][u];
      e += HE[6 * g + c] * wg[u];
      m += HSw[6 * g + c] * lg[u] + HSl[6 * g + c] * wg[u];
    }
    out[3 * g + v] = gi * e - m / gi;
  }
}

__global__ void adjoint_grad_kernel(int64_t no, const double* lam, double* d_grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * no) return;
  const int64_t u = i / no, g = i % no;
  d_grad[3 * g + u] = lam[i];
}

}  // namespace

extern "C" int gk_eval_sources(gk_handle h, double range, double c_o, const double* center, int64_t n_src,
                               const double* src_xyz, const double* src_w, int64_t n_points, const double* xyz,
                               double* scalar, double* gradient, gk_stream stream);
extern "C" int gk_eval_impl(gk_handle h, const gk_system* sys, const double* weights, gk_eval_desc* d,
                            hipStream_t st, bool force_mask, const double* src_box, bool reuse_sources);

// The kept factorization belongs to the system of the last gk_solve: every
// field of the descriptor (sizes, options, frame centre and the device
// pointers of the data) must match.  Data changed in place at the same
// addresses cannot be detected; callers solve again after such a change.
static Solver* kept_solver(gk_handle h, const gk_system* sys) {
  Solver* S = static_cast<Solver*>(h->solver_state);
  if (!S) return nullptr;
  const gk_system& k = S->sys_copy;
  const bool same = gk_system_n(sys) == S->N && sys->n_ori == k.n_ori && sys->n_rest == k.n_rest &&
                    sys->n_surf == k.n_surf && sys->uni_degree == k.uni_degree && sys->n_fault == k.n_fault &&
                    sys->range == k.range && sys->c_o == k.c_o && sys->i_res == k.i_res && sys->gi_res == k.gi_res &&
                    sys->ori_xyz == k.ori_xyz && sys->ori_grad == k.ori_grad && sys->ori_nugget == k.ori_nugget &&
                    sys->rest_xyz == k.rest_xyz && sys->rest_surf == k.rest_surf &&
                    sys->rest_nugget == k.rest_nugget && sys->ref_xyz == k.ref_xyz && sys->fault_rest == k.fault_rest &&
                    sys->center[0] == k.center[0] && sys->center[1] == k.center[1] && sys->center[2] == k.center[2];
  return same ? S : nullptr;
}

extern "C" int gk_solve_rhs(gk_handle h, const gk_system* sys, const double* rhs, double* x, gk_solve_info* info,
                            gk_stream stream) {
  GK_REQUIRE(h && sys && rhs && x && info, GK_ERR_ARG, "gk_solve_rhs: NULL argument");
  Solver* S = kept_solver(h, sys);
  GK_REQUIRE(S != nullptr, GK_ERR_ARG,
             "gk_solve_rhs: no factorization of this system is kept (call gk_solve with the FP64_CHOL or "
             "FP32_REFINE form first)");
  hipStream_t st = (hipStream_t)stream;
  S->st = st;
  *info = gk_solve_info{};
  GK_HIP(hipMemcpyAsync(S->b, rhs, S->N * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (S->f64) {
    int steps = 0;
    GK_TRY(S->direct_refine(x, 3, &steps, &info->relres));
    info->iterations = steps;
    info->form = GK_SOLVE_FP64_CHOL;
  } else {
    bool ok = false;
    GK_TRY(S->fgmres(x, 1e-10, 1e-6, 300, &info->iterations, &info->relres, &ok));
    info->form = GK_SOLVE_FP32_REFINE;
    GK_REQUIRE(ok, GK_ERR_NOT_CONVERGED, "gk_solve_rhs: refinement stagnated at relres %.3e", info->relres);
  }
  return GK_OK;
}

namespace {

// Number of eigenvalues of the symmetric tridiagonal T (diagonal a, off-diagonal
// b) below x (Sturm sequence of the LDL^T pivots).
int sturm_count(const std::vector<double>& a, const std::vector<double>& b, int k, double x) {
  int c = 0;
  double d = 1.0;
  for (int i = 0; i < k; ++i) {
    const double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
    d = a[i] - x - (i > 0 ? bb / d : 0.0);
    if (d == 0.0) d = -1e-300;
    if (d < 0.0) ++c;
  }
  return c;
}

// Smallest and largest eigenvalue of T by bisection inside Gershgorin bounds.
void tridiag_extremes(const std::vector<double>& a, const std::vector<double>& b, int k, double* lo, double* hi) {
  double glo = 1e300, ghi = -1e300;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    glo = std::min(glo, a[i] - r);
    ghi = std::max(ghi, a[i] + r);
  }
  auto bisect = [&](int target) {          // smallest x with count(x) > target
    double l = glo, u = ghi;
    for (int it = 0; it < 200 && u - l > 1e-16 * std::max(std::fabs(l), std::fabs(u)); ++it) {
      const double mid = 0.5 * (l + u);
      if (sturm_count(a, b, k, mid) > target) u = mid; else l = mid;
    }
    return 0.5 * (l + u);
  };
  *lo = bisect(0);
  *hi = bisect(k - 1);
}

__global__ void lanczos_start_kernel(double* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // deterministic start vector with no special alignment to the system's structure
  uint64_t z = (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  v[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

// Largest |eigenvalue| of the symmetric operator `op` by Lanczos with full
// reorthogonalisation (classical Gram-Schmidt, twice); Vb holds kmax + 1
// columns of leading dimension ld.  Stops when the estimate changes by less
// than tol (relative) over 3 steps, at an invariant subspace, or at kmax.
// part: (ceil(N / RED_ROWS) + 1) * 64 partial sums; coef: 2 (kmax + 1).
template <class Op>
int lanczos_sigma_max(Solver& S, Op op, int kmax, double tol, double* Vb, int64_t ld, double* wv, double* part,
                      double* coef, double* sig, int* steps) {
  const int64_t N = S.N;
  hipStream_t st = S.st;
  std::vector<double> a, bt, c(kmax + 1);
  const unsigned nb = nblk(N, RED_ROWS);
  // c[0..n) = V[:, 0..n)^T w, 64 columns per launch
  auto dots = [&](int n) -> int {
    for (int c0 = 0; c0 < n; c0 += 64) {
      const int nc = std::min(64, n - c0);
      hipLaunchKernelGGL(dots_partial_kernel, dim3(nb, (unsigned)nc), dim3(256), 0, st, Vb + (int64_t)c0 * ld, ld, nc,
                         wv, N, part);
      hipLaunchKernelGGL(reduce_cols_kernel, dim3(1), dim3(128), 0, st, part, (int)nb, nc, coef + c0);
    }
    GK_HIP(hipMemcpyAsync(c.data(), coef, n * sizeof(double), hipMemcpyDeviceToHost, st));
    GK_HIP(hipStreamSynchronize(st));
    return GK_OK;
  };
  hipLaunchKernelGGL(lanczos_start_kernel, dim3(nblk(N)), dim3(256), 0, st, Vb, N);
  double nv = 0.0;
  GK_TRY(S.norm(Vb, N, &nv));
  hipLaunchKernelGGL(scale_kernel, dim3(nblk(N)), dim3(256), 0, st, Vb, 1.0 / nv, Vb, N);
  double est = 0.0;
  int stable = 0, j = 0;
  for (; j < kmax; ++j) {
    double* vj = Vb + (int64_t)j * ld;
    GK_TRY(op(vj, wv));
    // alpha_j = <v_j, w>; then w -= V c over all previous columns, twice
    double alpha = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      GK_TRY(dots(j + 1));
      if (pass == 0) alpha = c[j]; else alpha += c[j];
      GK_HIP(hipMemcpyAsync(coef + kmax + 1, c.data(), (j + 1) * sizeof(double), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(gemv_sub_kernel, dim3(nblk(N)), dim3(256), 0, st, Vb, ld, j + 1, coef + kmax + 1, wv, N);
    }
    a.push_back(alpha);
    double beta = 0.0;
    GK_TRY(S.norm(wv, N, &beta));
    double lo = 0.0, hi = 0.0;
    tridiag_extremes(a, bt, j + 1, &lo, &hi);
    const double cur = std::max(std::fabs(lo), std::fabs(hi));
    stable = (j > 0 && std::fabs(cur - est) <= tol * cur) ? stable + 1 : 0;
    est = cur;
    if (stable >= 3 || beta <= 1e-14 * cur || j + 1 >= N) { ++j; break; }
    bt.push_back(beta);
    hipLaunchKernelGGL(scale_kernel, dim3(nblk(N)), dim3(256), 0, st, wv, 1.0 / beta, Vb + (int64_t)(j + 1) * ld, N);
  }
  GK_HIP(hipStreamSynchronize(st));
  *sig = est;
  *steps = j;
  return GK_OK;
}

}  // namespace

extern "C" int gk_condition_number(gk_handle h, const gk_system* sys, int32_t max_steps, double tol, double* cond,
                                   double* sigma, int32_t* steps, gk_stream stream) {
  GK_REQUIRE(h && sys && cond, GK_ERR_ARG, "gk_condition_number: NULL argument");
  GK_REQUIRE(max_steps >= 2 && max_steps <= 300 && tol > 0.0, GK_ERR_ARG,
             "gk_condition_number: max_steps must be in [2, 300] and tol > 0");
  Solver* S = kept_solver(h, sys);
  GK_REQUIRE(S != nullptr, GK_ERR_ARG,
             "gk_condition_number: no factorization of this system is kept (call gk_solve with the FP64_CHOL or "
             "FP32_REFINE form first)");
  hipStream_t st = (hipStream_t)stream;
  S->st = st;
  const int64_t N = S->N, ld = S->Npad;
  const int k = (int)std::min<int64_t>(max_steps, N);
  void* p = nullptr;
  const size_t n_part = (size_t)(gk_ceil_div(N, RED_ROWS) + 1) * 64;
  GK_TRY(gk_workspace(h, WS_LANCZOS, ((size_t)(k + 1) * ld + 2 * ld + n_part + 2 * (k + 1)) * sizeof(double), &p));
  double* Vb = static_cast<double*>(p);
  double* wv = Vb + (size_t)(k + 1) * ld;
  double* xo = wv + ld;
  double* part = xo + ld;
  double* coef = part + n_part;
  auto op_A = [&](const double* v, double* out) -> int {
    return gk_matvec_boxed(h, S->sys, GK_F64, v, out, st, S->mv_box);
  };
  auto op_inv = [&](const double* v, double* out) -> int {
    GK_HIP(hipMemcpyAsync(S->b, v, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (S->f64) {
      int it = 0;
      double rel = 0.0;
      GK_TRY(S->direct_refine(xo, 3, &it, &rel));
    } else {
      int it = 0;
      double rel = 0.0;
      bool ok = false;
      // A^-1 v is dominated by the smallest singular direction: the FP64
      // operator's residual floor is then near eps * cond(A), not eps
      GK_TRY(S->fgmres(xo, 1e-12, 1e-5, 300, &it, &rel, &ok));
      GK_REQUIRE(ok, GK_ERR_NOT_CONVERGED, "gk_condition_number: inner solve stagnated at relres %.3e", rel);
    }
    GK_HIP(hipMemcpyAsync(out, xo, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    return GK_OK;
  };
  double smax = 0.0, sinv = 0.0;
  int s1 = 0, s2 = 0;
  GK_TRY(lanczos_sigma_max(*S, op_A, k, tol, Vb, ld, wv, part, coef, &smax, &s1));
  GK_TRY(lanczos_sigma_max(*S, op_inv, k, tol, Vb, ld, wv, part, coef, &sinv, &s2));
  *cond = smax * sinv;
  if (sigma) { sigma[0] = smax; sigma[1] = sinv > 0.0 ? 1.0 / sinv : 0.0; }
  if (steps) { steps[0] = s1; steps[1] = s2; }
  return GK_OK;
}

static int adjoint_impl(gk_handle h, const gk_system* sys, const double* w, int64_t n_obs, const double* obs_xyz,
                        const double* obs_fault, const double* e, double* d_points, double* d_grad, double* d_ori_xyz,
                        double* lam_out, gk_stream stream);

extern "C" int gk_adjoint(gk_handle h, const gk_system* sys, const double* w, int64_t n_obs, const double* obs_xyz,
                          const double* obs_fault, const double* e, double* d_points, double* d_grad, double* lam_out,
                          gk_stream stream) {
  return adjoint_impl(h, sys, w, n_obs, obs_xyz, obs_fault, e, d_points, d_grad, nullptr, lam_out, stream);
}

extern "C" int gk_adjoint_full(gk_handle h, const gk_system* sys, const double* w, int64_t n_obs,
                               const double* obs_xyz, const double* obs_fault, const double* e, double* d_points,
                               double* d_grad, double* d_ori_xyz, double* lam_out, gk_stream stream) {
  return adjoint_impl(h, sys, w, n_obs, obs_xyz, obs_fault, e, d_points, d_grad, d_ori_xyz, lam_out, stream);
}

static int adjoint_impl(gk_handle h, const gk_system* sys, const double* w, int64_t n_obs, const double* obs_xyz,
                        const double* obs_fault, const double* e, double* d_points, double* d_grad, double* d_ori_xyz,
                        double* lam_out, gk_stream stream) {
  GK_REQUIRE(h && sys && w && d_points && d_grad && n_obs >= 0 && (n_obs == 0 || (obs_xyz && e)), GK_ERR_ARG,
             "gk_adjoint: NULL argument");
  GK_REQUIRE(sys->n_fault == 0 || n_obs == 0 || obs_fault, GK_ERR_ARG,
             "gk_adjoint: obs_fault (fault-block values at the observations) required with fault columns");
  Solver* S = kept_solver(h, sys);
  GK_REQUIRE(S != nullptr, GK_ERR_ARG, "gk_adjoint: no factorization of this system is kept");
  hipStream_t st = (hipStream_t)stream;
  const int64_t no = sys->n_ori, R = sys->n_rest, Sf = sys->n_surf, N = S->N, NP = no + R + Sf;
  const int L = sys->uni_degree == 0 ? 0 : (sys->uni_degree == 1 ? 3 : 9);
  // buffers: pts 3NP | E NP | gE 3NP | c N | lam N | gSw 3(R+S) | gSl 3(R+S) | Ww S | Wl S
  size_t need = 0;
  auto take = [&](size_t n) { size_t o = need; need += gk_round_up(n * sizeof(double), 256); return o; };
  const size_t o_pts = take(3 * NP), o_E = take(NP), o_gE = take(3 * NP), o_c = take(N), o_lam = take(N),
               o_gSw = take(3 * (R + Sf)), o_gSl = take(3 * (R + Sf)), o_Ww = take(Sf + 1), o_Wl = take(Sf + 1),
               o_fz = take((R + Sf) * std::max(sys->n_fault, 1));
  // orientation positions: three Hessian sets and the source-split partials
  const int64_t nh = d_ori_xyz ? no : 0;
  const size_t part_doubles = (size_t)64 * 6 * std::max<int64_t>(nh, 1);
  const size_t o_H = take(18 * std::max<int64_t>(nh, 1)), o_hp = take(d_ori_xyz ? part_doubles : 1);
  void* p = nullptr;
  GK_TRY(gk_workspace(h, WS_ADJOINT, need, &p));
  char* base = static_cast<char*>(p);
  double *pts = (double*)(base + o_pts), *E = (double*)(base + o_E), *gE = (double*)(base + o_gE),
         *c = (double*)(base + o_c), *lam = (double*)(base + o_lam), *gSw = (double*)(base + o_gSw),
         *gSl = (double*)(base + o_gSl), *Ww = (double*)(base + o_Ww), *Wl = (double*)(base + o_Wl),
         *fz = (double*)(base + o_fz);
  if (NP > 0)
    hipLaunchKernelGGL(gather_all_points_kernel, dim3(nblk(NP)), dim3(256), 0, st, *sys, pts);
  // 1. c = Phi^T e: the observation field E and its gradient at the data points
  GK_TRY(gk_eval_sources(h, sys->range, sys->c_o, sys->center, n_obs, obs_xyz, e, NP, pts, E, gE, st));
  GK_HIP(hipMemsetAsync(c, 0, N * sizeof(double), st));
  hipLaunchKernelGGL(adjoint_rhs_kernel, dim3(nblk(3 * no + R)), dim3(256), 0, st, *sys, E, gE, c);
  if (L > 0 && n_obs > 0)
    hipLaunchKernelGGL(adjoint_drift_rhs_kernel, dim3(L), dim3(256), 0, st, obs_xyz, e, n_obs, sys->gi_res,
                       c + 3 * no + R);
  // fault rows: gi_res sum_o e_o F_f(x_o) (the fault block is constant in the
  // interface coordinates almost everywhere: these columns have no other term)
  for (int f = 0; f < sys->n_fault && n_obs > 0; ++f)
    hipLaunchKernelGGL(adjoint_fault_rhs_kernel, dim3(1), dim3(256), 0, st, obs_fault, sys->n_fault, f, e, n_obs,
                       sys->gi_res, c + 3 * no + R + L + f);
  if (sys->n_fault > 0) GK_HIP(hipMemsetAsync(fz, 0, (R + Sf) * sys->n_fault * sizeof(double), st));
  // 2. lam = A^-1 c with the kept factorization
  gk_solve_info info{};
  GK_TRY(gk_solve_rhs(h, sys, c, lam, &info, st));
  // 3. grad S_w and grad S_lam at the rest + reference points
  double* prs = pts + 3 * no;
  for (int pass = 0; pass < 2; ++pass) {
    gk_eval_desc d{};
    d.kind = GK_POINTS_EXPLICIT;
    d.precision = GK_F64;
    d.n_points = R + Sf;
    d.xyz = prs;
    d.scalar = E;                 // scratch (E is no longer needed)
    d.gradient = pass == 0 ? gSw : gSl;
    d.fault_values = sys->n_fault > 0 ? fz : nullptr;   // no gradient from the fault block
    GK_TRY(gk_eval_impl(h, sys, pass == 0 ? w : lam, &d, st, true, nullptr, false));
  }
  // 4. combine
  if (Sf > 0)
    hipLaunchKernelGGL(surface_sum2_kernel, dim3((unsigned)Sf), dim3(256), 0, st, w + 3 * no, lam + 3 * no,
                       sys->rest_surf, R, Ww, Wl);
  if (R + Sf > 0)
    hipLaunchKernelGGL(adjoint_points_kernel, dim3(nblk(R + Sf)), dim3(256), 0, st, *sys, w, lam, Ww, Wl,
                       gE + 3 * no, gSw, gSl, d_points);
  if (no > 0)
    hipLaunchKernelGGL(adjoint_grad_kernel, dim3(nblk(3 * no)), dim3(256), 0, st, no, lam, d_grad);
  // 5. orientation positions: Hessians of E, S_w and S_lam at the orientations
  if (d_ori_xyz && no > 0) {
    double* H = (double*)(base + o_H);
    double* hp = (double*)(base + o_hp);
    const size_t cap = part_doubles * sizeof(double);
    GK_TRY(gk_hessian_sources(h, sys->range, sys->c_o, sys->center, n_obs, obs_xyz, e, no, sys->ori_xyz, H, hp, cap,
                              st));
    GK_TRY(gk_hessian_system(h, sys, w, no, sys->ori_xyz, H + 6 * no, hp, cap, st));
    GK_TRY(gk_hessian_system(h, sys, lam, no, sys->ori_xyz, H + 12 * no, hp, cap, st));
    hipLaunchKernelGGL(adjoint_ori_pos_kernel, dim3(nblk(no)), dim3(256), 0, st, no, sys->gi_res, w, lam, H,
                       H + 6 * no, H + 12 * no, d_ori_xyz);
  }
  if (lam_out) GK_HIP(hipMemcpyAsync(lam_out, lam, N * sizeof(double), hipMemcpyDeviceToDevice, st));
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_solve_spai(gk_handle h, const gk_system* sys, const int32_t* idx, const float* Y, double tol,
                             int32_t max_iter, int32_t krylov_precision, int32_t flags, double* x,
                             gk_solve_info* info, gk_stream stream) {
  GK_REQUIRE(h && sys && x && info, GK_ERR_ARG, "gk_solve_spai: NULL argument");
  const bool with_factor = (flags & GK_SPAI_WITH_FACTOR) != 0;
  GK_REQUIRE(with_factor || (idx && Y), GK_ERR_ARG, "gk_solve_spai: blocks required without the factor");
  GK_REQUIRE(krylov_precision == GK_F32 || krylov_precision == GK_F64, GK_ERR_ARG, "gk_solve_spai: bad precision");
  hipStream_t st = (hipStream_t)stream;
  *info = gk_solve_info{};
  if (tol <= 0) tol = 1e-10;
  if (max_iter <= 0) max_iter = 500;
  if (h->solver_state && h->solver_free) h->solver_free(h->solver_state);
  h->solver_state = nullptr;
  Solver S;
  GK_TRY(init_solver(h, sys, st, 3, S));
  S.spai_idx = idx;
  S.spai_Y = Y;
  S.mv_prec = krylov_precision;
  hipEvent_t ev[3];
  for (auto& e : ev) GK_HIP(hipEventCreate(&e));
  struct EvGuard { hipEvent_t* e; ~EvGuard() { for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]); } } guard{ev};
  GK_HIP(hipEventRecord(ev[0], st));
  GK_TRY(gk_diag_scale_impl(sys, S.dscale, st));
  bool have_factor = false;
  if (with_factor) {              // shifted FP32 factor, with M as its correction when given
    S.pcp = idx && Y;
    S.pcp_sandwich = (flags & GK_SPAI_SANDWICH) != 0;
    int32_t finfo = 0;
    const int rc = S.setup_preconditioner(false, &finfo);
    if (rc != GK_OK && rc != GK_ERR_NOT_SPD) return rc;
    have_factor = rc == GK_OK;
    info->shift = S.shift;
    info->info = finfo;
  } else {
    S.spai = true;
    GK_TRY(S.setup_drift());
  }
  GK_HIP(hipEventRecord(ev[1], st));
  bool ok = false;
  if (!with_factor || have_factor)
    GK_TRY(S.fgmres(x, tol, 1e-6, max_iter, &info->iterations, &info->relres, &ok));
  GK_HIP(hipEventRecord(ev[2], st));
  GK_HIP(hipEventSynchronize(ev[2]));
  info->t_factor_ms = elapsed_ms(ev[0], ev[1]);
  info->t_refine_ms = elapsed_ms(ev[1], ev[2]);
  info->form = with_factor ? GK_SOLVE_FP32_REFINE : 3;
  if (!ok && !(flags & GK_SPAI_NO_FALLBACK)) {    // stagnation: the FP64 direct solve takes over
    info->fallback = 1;
    S.spai = false;
    S.pcp = false;
    S.mv_prec = GK_F64;
    int32_t finfo = 0;
    GK_TRY(S.setup_preconditioner(true, &finfo));
    int steps = 0;
    GK_TRY(S.direct_refine(x, 3, &steps, &info->relres));
    info->info = finfo;
    info->form = GK_SOLVE_FP64_CHOL;
  }
  return GK_OK;
}

extern "C" int gk_diag_scale(gk_handle h, const gk_system* sys, double* dscale, gk_stream stream) {
  GK_REQUIRE(h && sys && dscale, GK_ERR_ARG, "gk_diag_scale: NULL argument");
  return gk_diag_scale_impl(sys, dscale, (hipStream_t)stream);
}
