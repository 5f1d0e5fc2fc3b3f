// Dense blocked factorizations for the co-kriging solve on CDNA4 MFMA.
//
//  * GEMM C = beta C + alpha A op(B) on v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64:
//    128x128 block tile, 4 waves of 64x64, LDS double buffer.  The K index is
//    permuted so that each lane reads its 4 (f64) / 8 (f32) consecutive k
//    values with 128-bit LDS loads; MFMA only needs A and B to agree on it.
//  * Right-looking blocked Cholesky (FP32) and LU without pivoting (FP64) with
//    256-wide panels: the diagonal block is factored and inverted by one
//    workgroup, the panel is a GEMM with the inverted diagonal block, and the
//    trailing update is the MFMA GEMM (lower tiles only for Cholesky).
//  * Blocked triangular solves with the inverted diagonal blocks.
#include <cmath>
#include "gk_internal.h"

namespace {

constexpr int NB = GK_BLOCK;   // 256

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
  using acc_t = f4v;
  static constexpr int BK = 32;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<double> {
  using acc_t = d4v;
  static constexpr int BK = 16;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// ------------------------------------------------------------------ GEMM
// Requirements (checked on the host): m, n multiples of 128, k multiple of BK,
// leading dimensions multiples of 4 (16-byte rows), 16-byte aligned bases.
template <typename T, bool TRANS_B, bool LOWER>
__global__ __launch_bounds__(256) void gemm_kernel(int64_t m, int64_t n, int64_t k, const T* __restrict__ A,
                                                   int64_t lda, const T* __restrict__ B, int64_t ldb,
                                                   T alpha, T beta, T* C, int64_t ldc, int64_t tiles_n,
                                                   int64_t n_tiles) {
  using M = Mfma<T>;
  constexpr int BK = M::BK;
  constexpr int KS = BK / 4;              // k values per lane per stage
  constexpr int PAD = 16 / sizeof(T);     // keep rows 16-byte aligned
  constexpr int LDS_LD = BK + PAD;
  constexpr int VEC = 16 / sizeof(T);     // elements per 16-byte vector
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  __shared__ __attribute__((aligned(16))) T As[2][128 * LDS_LD];
  __shared__ __attribute__((aligned(16))) T Bs[2][128 * LDS_LD];

  // XCD-aware tile order: consecutive hardware block ids go to different
  // XCDs, so give each XCD a contiguous range of logical tiles.
  int64_t bid = blockIdx.x;
  const int64_t nblk = gridDim.x;
  if (nblk % 8 == 0) bid = (bid % 8) * (nblk / 8) + bid / 8;
  int64_t tm, tn;
  if (LOWER) {
    tm = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while ((tm + 1) * (tm + 2) / 2 <= bid) ++tm;
    while (tm * (tm + 1) / 2 > bid) --tm;
    tn = bid - tm * (tm + 1) / 2;
  } else {
    tm = bid / tiles_n;
    tn = bid % tiles_n;
  }
  (void)n_tiles;
  const int64_t m0 = tm * 128, n0 = tn * 128;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lrow = lane & 15, lk = lane >> 4;

  typename M::acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = typename M::acc_t{0, 0, 0, 0};

  // global -> register staging: 128 x BK tile = 128*BK/VEC vectors = 4 per thread
  constexpr int VPR = BK / VEC;           // vectors per row
  vec_t ra[4], rb[4];
  auto load_global = [&](int64_t k0) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + 256 * it;
      const int row = idx / VPR, cv = idx % VPR;
      ra[it] = *reinterpret_cast<const vec_t*>(A + (m0 + row) * lda + k0 + cv * VEC);
    }
    if (TRANS_B) {
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + 256 * it;
        const int row = idx / VPR, cv = idx % VPR;
        rb[it] = *reinterpret_cast<const vec_t*>(B + (n0 + row) * ldb + k0 + cv * VEC);
      }
    } else {
      constexpr int VPRB = 128 / VEC;     // vectors per k-row of the B tile
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + 256 * it;
        const int kk = idx / VPRB, cv = idx % VPRB;
        rb[it] = *reinterpret_cast<const vec_t*>(B + (k0 + kk) * ldb + n0 + cv * VEC);
      }
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + 256 * it;
      const int row = idx / VPR, cv = idx % VPR;
      *reinterpret_cast<vec_t*>(&As[buf][row * LDS_LD + cv * VEC]) = ra[it];
    }
    if (TRANS_B) {
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + 256 * it;
        const int row = idx / VPR, cv = idx % VPR;
        *reinterpret_cast<vec_t*>(&Bs[buf][row * LDS_LD + cv * VEC]) = rb[it];
      }
    } else {
      constexpr int VPRB = 128 / VEC;
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + 256 * it;
        const int kk = idx / VPRB, cv = idx % VPRB;
#pragma unroll
        for (int e = 0; e < VEC; ++e) Bs[buf][(cv * VEC + e) * LDS_LD + kk] = rb[it][e];
      }
    }
  };

  const int64_t nk = k / BK;
  load_global(0);
  store_lds(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load_global((kt + 1) * BK);
    // fragments: lane reads its KS consecutive k values (permuted k order)
    T af[4][KS], bf[4][KS];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const T* pa = &As[buf][(wm * 64 + i * 16 + lrow) * LDS_LD + lk * KS];
      const T* pb = &Bs[buf][(wn * 64 + i * 16 + lrow) * LDS_LD + lk * KS];
#pragma unroll
      for (int v = 0; v < KS / VEC; ++v) {
        vec_t x = *reinterpret_cast<const vec_t*>(pa + v * VEC);
        vec_t y = *reinterpret_cast<const vec_t*>(pb + v * VEC);
#pragma unroll
        for (int e = 0; e < VEC; ++e) { af[i][v * VEC + e] = x[e]; bf[i][v * VEC + e] = y[e]; }
      }
    }
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = M::mma(af[i][s], bf[j][s], acc[i][j]);
    if (kt + 1 < nk) {
      store_lds(buf ^ 1);
    }
    __syncthreads();
  }

  // epilogue: lane holds rows 4*(lane/16)+r, column lane%16 of each 16x16 tile
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 64 + i * 16 + lk * 4 + r;
        const int64_t col = n0 + wn * 64 + j * 16 + lrow;
        T* pc = C + row * ldc + col;
        const T prod = acc[i][j][r];
        *pc = (beta == T(0)) ? alpha * prod : fma(alpha, prod, beta * *pc);
      }
}

template <typename T>
int gemm_launch(int64_t m, int64_t n, int64_t k, const T* A, int64_t lda, const T* B, int64_t ldb,
                bool trans_b, T alpha, T beta, T* C, int64_t ldc, bool lower, hipStream_t st) {
  constexpr int BK = Mfma<T>::BK;
  GK_REQUIRE(m % 128 == 0 && n % 128 == 0 && k % BK == 0, GK_ERR_ARG,
             "gk_gemm: m, n must be multiples of 128 and k of %d (got %lld %lld %lld)", BK,
             (long long)m, (long long)n, (long long)k);
  GK_REQUIRE(lda % 4 == 0 && ldb % 4 == 0 && ldc >= n, GK_ERR_ARG, "gk_gemm: bad leading dimensions");
  GK_REQUIRE(!lower || m == n, GK_ERR_ARG, "gk_gemm: lower_only needs m == n");
  if (m == 0 || n == 0) return GK_OK;
  const int64_t tm = m / 128, tn = n / 128;
  const int64_t tiles = lower ? tm * (tm + 1) / 2 : tm * tn;
  if (k == 0) {
    GK_REQUIRE(beta == T(1), GK_ERR_ARG, "gk_gemm: k == 0 only with beta == 1");
    return GK_OK;
  }
  dim3 grid((unsigned)tiles), block(256);
  if (trans_b) {
    if (lower) hipLaunchKernelGGL((gemm_kernel<T, true, true>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles);
    else hipLaunchKernelGGL((gemm_kernel<T, true, false>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles);
  } else {
    if (lower) hipLaunchKernelGGL((gemm_kernel<T, false, true>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles);
    else hipLaunchKernelGGL((gemm_kernel<T, false, false>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles);
  }
  GK_HIP(hipGetLastError());
  return GK_OK;
}

// ------------------------------------------------------- diagonal blocks
// Cholesky of one NB x NB block in place (lower), then its inverse into Dinv.
// One workgroup of 1024 threads; the block stays in L2.
__global__ __launch_bounds__(1024) void potrf_diag_kernel(float* A, int64_t lda, int64_t k0,
                                                          float* Dinv, int32_t* info) {
  float* D = A + k0 * lda + k0;
  const int tid = threadIdx.x;
  __shared__ float piv;
  __shared__ float col[NB];
  for (int j = 0; j < NB; ++j) {
    if (tid == 0) {
      float d = D[j * lda + j];
      if (!(d > 0.0f)) {
        atomicCAS(info, 0, (int32_t)(k0 + j + 1));
        d = 1.0f;
      }
      piv = sqrtf(d);
      D[j * lda + j] = piv;
    }
    __syncthreads();
    const float ip = 1.0f / piv;
    for (int i = j + 1 + tid; i < NB; i += 1024) {
      float v = D[i * lda + j] * ip;
      D[i * lda + j] = v;
      col[i] = v;
    }
    __syncthreads();
    // trailing lower update: rows i > j, cols j < c <= i
    const int rem = NB - j - 1;
    const int64_t total = (int64_t)rem * (rem + 1) / 2;
    for (int64_t e = tid; e < total; e += 1024) {
      int ii = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while ((int64_t)(ii + 1) * (ii + 2) / 2 <= e) ++ii;
      while ((int64_t)ii * (ii + 1) / 2 > e) --ii;
      const int cc = (int)(e - (int64_t)ii * (ii + 1) / 2);
      const int i = j + 1 + ii, c = j + 1 + cc;
      D[i * lda + c] -= col[i] * col[c];
    }
    __syncthreads();
  }
  // inverse of the lower factor: thread c computes column c of L^-1
  float* X = Dinv + k0 * NB;     // NB x NB row-major block
  for (int c = tid; c < NB; c += 1024) {
    for (int i = 0; i < NB; ++i) {
      if (i < c) { X[i * NB + c] = 0.0f; continue; }
      float s = (i == c) ? 1.0f : 0.0f;
      for (int kk = c; kk < i; ++kk) s -= D[i * lda + kk] * X[kk * NB + c];
      X[i * NB + c] = s / D[i * lda + i];
    }
  }
}

// LU without pivoting of one NB x NB FP64 block, then inverses of the unit
// lower factor (Dinv) and of the upper factor (Dinv + lda*NB).
__global__ __launch_bounds__(1024) void getrf_diag_kernel(double* A, int64_t lda, int64_t k0,
                                                          double* Dinv, int64_t ld_total, int32_t* info) {
  double* D = A + k0 * lda + k0;
  const int tid = threadIdx.x;
  __shared__ double rowp[NB];
  __shared__ double colp[NB];
  for (int j = 0; j < NB; ++j) {
    __syncthreads();
    double p = D[j * lda + j];
    if (!(fabs(p) > 0.0)) {
      if (tid == 0) atomicCAS(info, 0, (int32_t)(k0 + j + 1));
      p = 1.0;
    }
    const double ip = 1.0 / p;
    for (int i = j + 1 + tid; i < NB; i += 1024) {
      const double v = D[i * lda + j] * ip;
      colp[i] = v;
      rowp[i] = D[j * lda + i];
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < NB; i += 1024) D[i * lda + j] = colp[i];
    const int rem = NB - j - 1;
    for (int64_t e = tid; e < (int64_t)rem * rem; e += 1024) {
      const int i = j + 1 + (int)(e / rem), c = j + 1 + (int)(e % rem);
      D[i * lda + c] -= colp[i] * rowp[c];
    }
  }
  __syncthreads();
  double* Li = Dinv + k0 * NB;               // unit lower inverse
  double* Ui = Dinv + ld_total * NB + k0 * NB; // upper inverse
  for (int c = tid; c < NB; c += 1024) {
    for (int i = 0; i < NB; ++i) {            // L^-1 column c (forward)
      if (i < c) { Li[i * NB + c] = 0.0; continue; }
      double s = (i == c) ? 1.0 : 0.0;
      for (int kk = c; kk < i; ++kk) s -= D[i * lda + kk] * Li[kk * NB + c];
      Li[i * NB + c] = s;
    }
    for (int i = NB - 1; i >= 0; --i) {       // U^-1 column c (backward)
      if (i > c) { Ui[i * NB + c] = 0.0; continue; }
      double s = (i == c) ? 1.0 : 0.0;
      for (int kk = i + 1; kk <= c; ++kk) s -= D[i * lda + kk] * Ui[kk * NB + c];
      Ui[i * NB + c] = s / D[i * lda + i];
    }
  }
}

// Identity padding of rows/cols [n, lda).
template <typename T>
__global__ void pad_identity_kernel(T* A, int64_t lda, int64_t n) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = lda - n;
  if (w <= 0) return;
  // region 1: rows [n, lda) x all cols ; region 2: rows [0, n) x cols [n, lda)
  const int64_t r1 = w * lda;
  if (idx < r1) {
    const int64_t i = n + idx / lda, j = idx % lda;
    A[i * lda + j] = (i == j) ? T(1) : T(0);
  } else if (idx < r1 + n * w) {
    const int64_t e = idx - r1;
    const int64_t i = e / w, j = n + e % w;
    A[i * lda + j] = T(0);
  }
}

template <typename T>
__global__ void copy_block_kernel(const T* src, int64_t lds, T* dst, int64_t ldd, int64_t rows, int64_t cols) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int64_t i = idx / cols, j = idx % cols;
  dst[i * ldd + j] = src[i * lds + j];
}

// ------------------------------------------------------ triangular solves
// b[r] -= sum_c M[r, c0 + c] x[c] for r in [r0, r1), c < NB.  8 threads per row.
template <typename T>
__global__ __launch_bounds__(256) void gemv_panel_kernel(const T* M, int64_t ld, int64_t r0, int64_t r1,
                                                         int64_t c0, const T* x, T* b) {
  __shared__ T xs[NB];
  xs[threadIdx.x] = x[threadIdx.x];
  __syncthreads();
  const int64_t r = r0 + (int64_t)blockIdx.x * 32 + (threadIdx.x >> 3);
  const int sub = threadIdx.x & 7;
  T s = T(0);
  if (r < r1) {
    const T* row = M + r * ld + c0 + sub * 32;
#pragma unroll 8
    for (int c = 0; c < 32; ++c) s = fma(row[c], xs[sub * 32 + c], s);
  }
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 4);
  if (r < r1 && sub == 0) b[r] -= s;
}

// b[c] -= sum_r M[r0 + r, c] x[r] for c in [0, c1), r < NB.
template <typename T>
__global__ __launch_bounds__(256) void gemv_panelT_kernel(const T* M, int64_t ld, int64_t r0, int64_t c1,
                                                          const T* x, T* b) {
  __shared__ T xs[NB];
  xs[threadIdx.x] = x[threadIdx.x];
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= c1) return;
  T s = T(0);
  for (int r = 0; r < NB; ++r) s = fma(M[(r0 + r) * ld + c], xs[r], s);
  b[c] -= s;
}

// x_k = Dinv_k b_k (TRANS: Dinv_k^T b_k), in place on b + k0.
template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void diag_apply_kernel(const T* Dk, T* b) {
  __shared__ T bs[NB];
  bs[threadIdx.x] = b[threadIdx.x];
  __syncthreads();
  const int i = threadIdx.x;
  T s = T(0);
  if (TRANS) { for (int c = 0; c < NB; ++c) s = fma(Dk[c * NB + i], bs[c], s); }
  else { for (int c = 0; c < NB; ++c) s = fma(Dk[i * NB + c], bs[c], s); }
  b[i] = s;
}

}  // namespace

// ------------------------------------------------------------------ API
extern "C" int gk_gemm_f32(gk_handle h, int64_t m, int64_t n, int64_t k, const float* A, int64_t lda,
                           const float* B, int64_t ldb, int32_t trans_b, float beta, float* C, int64_t ldc,
                           int32_t lower_only, gk_stream stream) {
  (void)h;
  return gemm_launch<float>(m, n, k, A, lda, B, ldb, trans_b != 0, -1.0f, beta, C, ldc, lower_only != 0,
                            (hipStream_t)stream);
}

extern "C" int gk_gemm_f64(gk_handle h, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                           const double* B, int64_t ldb, int32_t trans_b, double beta, double* C,
                           int64_t ldc, int32_t lower_only, gk_stream stream) {
  (void)h;
  return gemm_launch<double>(m, n, k, A, lda, B, ldb, trans_b != 0, -1.0, beta, C, ldc, lower_only != 0,
                             (hipStream_t)stream);
}

extern "C" int gk_cholesky_f32(gk_handle h, float* A, int64_t n, int64_t lda, float* dinv, int32_t* info,
                               gk_stream stream) {
  hipStream_t st = (hipStream_t)stream;
  GK_REQUIRE(A && dinv, GK_ERR_ARG, "gk_cholesky_f32: A and dinv required");
  GK_REQUIRE(lda % NB == 0 && n <= lda && n >= 0, GK_ERR_ARG,
             "gk_cholesky_f32: lda must be a multiple of %d and >= n", NB);
  const int64_t nb = lda / NB;
  void* ws = nullptr;
  // workspace: info (256 B) + W panel (lda x NB floats)
  GK_TRY(gk_workspace(h, WS_FACTOR, 256 + (size_t)lda * NB * sizeof(float), &ws));
  int32_t* dinfo = static_cast<int32_t*>(ws);
  float* W = reinterpret_cast<float*>(static_cast<char*>(ws) + 256);
  GK_HIP(hipMemsetAsync(dinfo, 0, sizeof(int32_t), st));
  if (n < lda) {
    const int64_t tot = (lda - n) * lda + n * (lda - n);
    hipLaunchKernelGGL(pad_identity_kernel<float>, dim3((unsigned)gk_ceil_div(tot, 256)), dim3(256), 0, st,
                       A, lda, n);
  }
  for (int64_t kb = 0; kb < nb; ++kb) {
    const int64_t k0 = kb * NB;
    hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(1024), 0, st, A, lda, k0, dinv, dinfo);
    const int64_t rem = lda - k0 - NB;
    if (rem <= 0) break;
    // W = A21 Linv^T ;  A21 <- W ; A22 -= W W^T (lower)
    GK_TRY(gemm_launch<float>(rem, NB, NB, A + (k0 + NB) * lda + k0, lda, dinv + k0 * NB, NB, true, 1.0f,
                              0.0f, W, NB, false, st));
    hipLaunchKernelGGL(copy_block_kernel<float>, dim3((unsigned)gk_ceil_div(rem * NB, 256)), dim3(256), 0,
                       st, W, (int64_t)NB, A + (k0 + NB) * lda + k0, lda, rem, (int64_t)NB);
    GK_TRY(gemm_launch<float>(rem, rem, NB, W, NB, W, NB, true, -1.0f, 1.0f, A + (k0 + NB) * lda + k0 + NB,
                              lda, true, st));
  }
  GK_HIP(hipGetLastError());
  int32_t hinfo = 0;
  GK_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  if (info) *info = hinfo;
  return hinfo == 0 ? GK_OK : GK_ERR_NOT_SPD;
}

extern "C" int gk_cholesky_solve_f32(gk_handle h, const float* L, const float* dinv, int64_t lda, float* b,
                                     gk_stream stream) {
  (void)h;
  hipStream_t st = (hipStream_t)stream;
  GK_REQUIRE(L && dinv && b && lda % NB == 0, GK_ERR_ARG, "gk_cholesky_solve_f32: bad arguments");
  const int64_t nb = lda / NB;
  // forward: L y = b
  for (int64_t kb = 0; kb < nb; ++kb) {
    const int64_t k0 = kb * NB;
    hipLaunchKernelGGL((diag_apply_kernel<float, false>), dim3(1), dim3(NB), 0, st, dinv + k0 * NB, b + k0);
    const int64_t r0 = k0 + NB;
    if (r0 < lda)
      hipLaunchKernelGGL(gemv_panel_kernel<float>, dim3((unsigned)gk_ceil_div(lda - r0, 32)), dim3(256), 0,
                         st, L, lda, r0, lda, k0, b + k0, b);
  }
  // backward: L^T x = y
  for (int64_t kb = nb - 1; kb >= 0; --kb) {
    const int64_t k0 = kb * NB;
    hipLaunchKernelGGL((diag_apply_kernel<float, true>), dim3(1), dim3(NB), 0, st, dinv + k0 * NB, b + k0);
    if (k0 > 0)
      hipLaunchKernelGGL(gemv_panelT_kernel<float>, dim3((unsigned)gk_ceil_div(k0, 256)), dim3(256), 0, st,
                         L, lda, k0, k0, b + k0, b);
  }
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_lu_f64(gk_handle h, double* A, int64_t n, int64_t lda, double* dinv, int32_t* info,
                         gk_stream stream) {
  hipStream_t st = (hipStream_t)stream;
  GK_REQUIRE(A && dinv, GK_ERR_ARG, "gk_lu_f64: A and dinv required");
  GK_REQUIRE(lda % NB == 0 && n <= lda && n >= 0, GK_ERR_ARG,
             "gk_lu_f64: lda must be a multiple of %d and >= n", NB);
  const int64_t nb = lda / NB;
  void* ws = nullptr;
  GK_TRY(gk_workspace(h, WS_FACTOR, 256 + 2 * (size_t)lda * NB * sizeof(double), &ws));
  int32_t* dinfo = static_cast<int32_t*>(ws);
  double* W1 = reinterpret_cast<double*>(static_cast<char*>(ws) + 256);   // U12: NB x rem
  double* W2 = W1 + (size_t)lda * NB;                                        // L21: rem x NB
  GK_HIP(hipMemsetAsync(dinfo, 0, sizeof(int32_t), st));
  if (n < lda) {
    const int64_t tot = (lda - n) * lda + n * (lda - n);
    hipLaunchKernelGGL(pad_identity_kernel<double>, dim3((unsigned)gk_ceil_div(tot, 256)), dim3(256), 0, st,
                       A, lda, n);
  }
  for (int64_t kb = 0; kb < nb; ++kb) {
    const int64_t k0 = kb * NB;
    hipLaunchKernelGGL(getrf_diag_kernel, dim3(1), dim3(1024), 0, st, A, lda, k0, dinv, lda, dinfo);
    const int64_t rem = lda - k0 - NB;
    if (rem <= 0) break;
    const double* Li = dinv + k0 * NB;
    const double* Ui = dinv + lda * NB + k0 * NB;
    // U12 = Linv A12 (NB x rem), L21 = A21 Uinv (rem x NB)
    GK_TRY(gemm_launch<double>(NB, rem, NB, Li, NB, A + k0 * lda + k0 + NB, lda, false, 1.0, 0.0, W1, rem,
                               false, st));
    GK_TRY(gemm_launch<double>(rem, NB, NB, A + (k0 + NB) * lda + k0, lda, Ui, NB, false, 1.0, 0.0, W2, NB,
                               false, st));
    hipLaunchKernelGGL(copy_block_kernel<double>, dim3((unsigned)gk_ceil_div(rem * NB, 256)), dim3(256), 0,
                       st, W1, rem, A + k0 * lda + k0 + NB, lda, (int64_t)NB, rem);
    hipLaunchKernelGGL(copy_block_kernel<double>, dim3((unsigned)gk_ceil_div(rem * NB, 256)), dim3(256), 0,
                       st, W2, (int64_t)NB, A + (k0 + NB) * lda + k0, lda, rem, (int64_t)NB);
    GK_TRY(gemm_launch<double>(rem, rem, NB, W2, NB, W1, rem, false, -1.0, 1.0, A + (k0 + NB) * lda + k0 + NB,
                               lda, false, st));
  }
  GK_HIP(hipGetLastError());
  int32_t hinfo = 0;
  GK_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  if (info) *info = hinfo;
  return hinfo == 0 ? GK_OK : GK_ERR_SINGULAR;
}

extern "C" int gk_lu_solve_f64(gk_handle h, const double* LU, const double* dinv, int64_t lda, double* b,
                               gk_stream stream) {
  (void)h;
  hipStream_t st = (hipStream_t)stream;
  GK_REQUIRE(LU && dinv && b && lda % NB == 0, GK_ERR_ARG, "gk_lu_solve_f64: bad arguments");
  const int64_t nb = lda / NB;
  for (int64_t kb = 0; kb < nb; ++kb) {          // L y = b (unit lower)
    const int64_t k0 = kb * NB;
    hipLaunchKernelGGL((diag_apply_kernel<double, false>), dim3(1), dim3(NB), 0, st, dinv + k0 * NB, b + k0);
    const int64_t r0 = k0 + NB;
    if (r0 < lda)
      hipLaunchKernelGGL(gemv_panel_kernel<double>, dim3((unsigned)gk_ceil_div(lda - r0, 32)), dim3(256), 0,
                         st, LU, lda, r0, lda, k0, b + k0, b);
  }
  for (int64_t kb = nb - 1; kb >= 0; --kb) {     // U x = y
    const int64_t k0 = kb * NB;
    hipLaunchKernelGGL((diag_apply_kernel<double, false>), dim3(1), dim3(NB), 0, st,
                       dinv + lda * NB + k0 * NB, b + k0);
    if (k0 > 0)
      hipLaunchKernelGGL(gemv_panel_kernel<double>, dim3((unsigned)gk_ceil_div(k0, 32)), dim3(256), 0, st,
                         LU, lda, (int64_t)0, k0, k0, b + k0, b);
  }
  GK_HIP(hipGetLastError());
  return GK_OK;
}
