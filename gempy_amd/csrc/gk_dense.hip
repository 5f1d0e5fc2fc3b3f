// Dense blocked factorizations for the co-kriging solve on CDNA4 MFMA.
//
//  * GEMM C = beta C + alpha A op(B) on v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64:
//    128x128 block tile, 4 waves of 64x64, LDS double buffer.  The K index is
//    permuted so that each lane reads its BK/4 consecutive k values with
//    128-bit LDS loads; MFMA only needs A and B to agree on it.
//  * Cholesky FP64: recursive on halves of 256-blocks, with a recursive TRSM,
//    so most flops are large-K GEMMs; Cholesky FP32 and LU without pivoting
//    (FP64): right-looking with 256-wide panels.  Each 256 diagonal block is factored
//    and inverted as two LDS-resident 128-wide halves joined by 128^3 GEMMs;
//    a panel solve is a GEMM with the inverted diagonal block.
//  * Blocked triangular solves of the LU with the inverted diagonal blocks
//    (the Cholesky solves are in gk_trsv.hip).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include "gk_internal.h"

namespace {

constexpr int NB = GK_BLOCK;   // 256

// K depth of one LDS stage of the GEMM (elements)
#ifndef GK_BK_F32
#define GK_BK_F32 16
#endif
#ifndef GK_BK_F64
#define GK_BK_F64 16
#endif
// GEMM register stages of the global -> LDS copy (1 or 2; FP32 only) and
// whether fragments are read from LDS one 16-byte vector at a time.  Measured
// on MI355X (K = 256 rank updates): BK 16 with chunked fragments beats BK 32
// with whole-stage fragments; a second register stage ran slower with 8 waves
// (91 against 99 TF/s) and does not fit the 4-wave tile's register budget.
#ifndef GK_GEMM_RS_F32
#define GK_GEMM_RS_F32 1
#endif
#ifndef GK_GEMM_CHUNKED
#define GK_GEMM_CHUNKED 1
#endif
// minimum waves per SIMD the GEMM is compiled for (0 = compiler's choice)
#ifndef GK_GEMM_WPE_F32
#define GK_GEMM_WPE_F32 3
#endif
#ifndef GK_GEMM_WPE_F64
#define GK_GEMM_WPE_F64 2
#endif

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
  using acc_t = f4v;
  static constexpr int BK = GK_BK_F32;
  static constexpr int WPE = GK_GEMM_WPE_F32;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<double> {
  using acc_t = d4v;
  static constexpr int BK = GK_BK_F64;
  static constexpr int WPE = GK_GEMM_WPE_F64;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// ------------------------------------------------------------------ GEMM
// Block tile TM x TM, WM x WN waves, each wave FM x FN MFMA tiles of 16 x 16.
// Measured on MI355X for FP32 rank-256 updates (m = 34816, TF/s): 128 x 128
// with 8 waves of 64 x 32 (2 x 4), 3 waves per SIMD and tiles grouped by 16
// rows 103.6; the same without grouping 98.8; 4 waves of 64 x 64 82.7;
// 256 x 256 tiles with 8 waves 71-72, with 4 waves (accumulators in AGPRs)
// 49-53 (one workgroup per CU leaves the stage barriers uncovered); a second
// register stage or BK 32 were slower.
// TM x TN tiles (TN defaults to TM) of WM x WN waves.
template <int TM_, int WM_, int WN_, int TN_ = TM_>
struct Tile {
  static constexpr int TM = TM_, TN = TN_, WM = WM_, WN = WN_, NT = 64 * WM * WN;
  static constexpr int FM = TM / WM / 16, FN = TN / WN / 16;
};
#ifndef GK_GEMM_GROUP
#define GK_GEMM_GROUP 16
#endif
#ifndef GK_GEMM_WM_F32
#define GK_GEMM_WM_F32 2
#endif
#ifndef GK_GEMM_WN_F32
#define GK_GEMM_WN_F32 4
#endif
#ifndef GK_GEMM_WM_F64
#define GK_GEMM_WM_F64 2
#endif
#ifndef GK_GEMM_WN_F64
#define GK_GEMM_WN_F64 2
#endif
template <typename T> struct GemmTile;
template <> struct GemmTile<float> { typedef Tile<128, GK_GEMM_WM_F32, GK_GEMM_WN_F32> type; };
template <> struct GemmTile<double> { typedef Tile<128, GK_GEMM_WM_F64, GK_GEMM_WN_F64> type; };

// Requirements (checked on the host): m, n multiples of TM, k multiple of BK,
// leading dimensions multiples of 4 (16-byte rows), 16-byte aligned bases.
template <typename T, bool TRANS_B, bool LOWER, typename TL, int WPE>
__global__ __launch_bounds__(TL::NT) __attribute__((amdgpu_waves_per_eu(WPE, 0))) void gemm_kernel(int64_t m, int64_t n, int64_t k, const T* __restrict__ A,
                                                   int64_t lda, const T* __restrict__ B, int64_t ldb,
                                                   T alpha, T beta, T* C, int64_t ldc, int64_t tiles_n,
                                                   int64_t n_tiles, int64_t sA, int64_t sB, int64_t sC) {
  using M = Mfma<T>;
  // batched launches: product blockIdx.y at A + y sA, B + y sB, C + y sC
  A += blockIdx.y * sA;
  B += blockIdx.y * sB;
  C += blockIdx.y * sC;
  constexpr int TM = TL::TM, TN = TL::TN, NT = TL::NT, FM = TL::FM, FN = TL::FN;
  static_assert(!LOWER || TN == TM, "lower-triangle products take square tiles");
  constexpr int BK = M::BK;
  constexpr int KS = BK / 4;              // k values per lane per stage
  constexpr int PAD = 16 / sizeof(T);     // keep rows 16-byte aligned
  constexpr int LDS_LD = BK + PAD;
  constexpr int VEC = 16 / sizeof(T);     // elements per 16-byte vector
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  __shared__ __attribute__((aligned(16))) T As[2][TM * LDS_LD];
  __shared__ __attribute__((aligned(16))) T Bs[2][TN * LDS_LD];

  // XCD-aware tile order: consecutive hardware block ids go to different
  // XCDs, so give each XCD a contiguous range of logical tiles.
  const int64_t bid = gk_xcd_tile(blockIdx.x, gridDim.x);
  // With GK_GEMM_GROUP = G > 0 the tiles go column by column within groups of
  // G tile rows, so a B tile is reused by the G rows while their A rows stay
  // in L2.
  int64_t tm, tn;
  if (LOWER) {
    tm = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while ((tm + 1) * (tm + 2) / 2 <= bid) ++tm;
    while (tm * (tm + 1) / 2 > bid) --tm;
    tn = bid - tm * (tm + 1) / 2;
    if (GK_GEMM_GROUP > 0) {            // group of rows [r0, r0 + S): a rectangle, then a triangle
      const int64_t G = GK_GEMM_GROUP, r0 = (tm / G) * G, S = min(G, tiles_n - r0);
      const int64_t local = bid - r0 * (r0 + 1) / 2;
      if (local < r0 * S) {
        tn = local / S;
        tm = r0 + local % S;
      } else {
        int64_t l2 = local - r0 * S, j = 0, acc = 0;
        while (acc + (S - j) <= l2) { acc += S - j; ++j; }
        tn = r0 + j;
        tm = tn + (l2 - acc);
      }
    }
  } else if (GK_GEMM_GROUP > 0) {
    const int64_t G = GK_GEMM_GROUP, tiles_m = n_tiles / tiles_n;
    const int64_t first = bid / (G * tiles_n) * G, S = min(tiles_m - first, G);
    const int64_t loc = bid - first * tiles_n;
    tm = first + loc % S;
    tn = loc / S;
  } else {
    tm = bid / tiles_n;
    tn = bid % tiles_n;
  }
  const int64_t m0 = tm * TM, n0 = tn * TN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave / TL::WN, wn = wave % TL::WN;
  const int lrow = lane & 15, lk = lane >> 4;

  typename M::acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = typename M::acc_t{0, 0, 0, 0};

  // global -> register staging: TM x BK tile = TM*BK/VEC vectors, NV per
  // thread (NVB for the TN x BK tile of B)
  constexpr int VPR = BK / VEC;           // vectors per row
  constexpr int NV = TM * BK / VEC / NT, NVB = TN * BK / VEC / NT;
  static_assert(NV >= 1 && NV * NT * VEC == TM * BK, "BK must make the tile a whole number of vectors per thread");
  static_assert(NVB >= 1 && NVB * NT * VEC == TN * BK, "BK must make the tile a whole number of vectors per thread");
  constexpr int RS = sizeof(T) == 4 ? GK_GEMM_RS_F32 : 1;
  vec_t ra[RS][NV], rb[RS][NVB];
  auto load_global = [&](int64_t k0, vec_t (&xa)[NV], vec_t (&xb)[NVB]) {
#pragma unroll
    for (int it = 0; it < NV; ++it) {
      const int idx = tid + NT * it;
      const int row = idx / VPR, cv = idx % VPR;
      xa[it] = *reinterpret_cast<const vec_t*>(A + (m0 + row) * lda + k0 + cv * VEC);
    }
    if (TRANS_B) {
#pragma unroll
      for (int it = 0; it < NVB; ++it) {
        const int idx = tid + NT * it;
        const int row = idx / VPR, cv = idx % VPR;
        xb[it] = *reinterpret_cast<const vec_t*>(B + (n0 + row) * ldb + k0 + cv * VEC);
      }
    } else {
      constexpr int VPRB = TN / VEC;      // vectors per k-row of the B tile
#pragma unroll
      for (int it = 0; it < NVB; ++it) {
        const int idx = tid + NT * it;
        const int kk = idx / VPRB, cv = idx % VPRB;
        xb[it] = *reinterpret_cast<const vec_t*>(B + (k0 + kk) * ldb + n0 + cv * VEC);
      }
    }
  };
  auto store_lds = [&](int buf, const vec_t (&xa)[NV], const vec_t (&xb)[NVB]) {
#pragma unroll
    for (int it = 0; it < NV; ++it) {
      const int idx = tid + NT * it;
      const int row = idx / VPR, cv = idx % VPR;
      *reinterpret_cast<vec_t*>(&As[buf][row * LDS_LD + cv * VEC]) = xa[it];
    }
    if (TRANS_B) {
#pragma unroll
      for (int it = 0; it < NVB; ++it) {
        const int idx = tid + NT * it;
        const int row = idx / VPR, cv = idx % VPR;
        *reinterpret_cast<vec_t*>(&Bs[buf][row * LDS_LD + cv * VEC]) = xb[it];
      }
    } else {
      constexpr int VPRB = TN / VEC;
#pragma unroll
      for (int it = 0; it < NVB; ++it) {
        const int idx = tid + NT * it;
        const int kk = idx / VPRB, cv = idx % VPRB;
#pragma unroll
        for (int e = 0; e < VEC; ++e) Bs[buf][(cv * VEC + e) * LDS_LD + kk] = xb[it][e];
      }
    }
  };
  auto multiply = [&](int buf) {
    if (GK_GEMM_CHUNKED) {        // fragments one 16-byte vector of k at a time
#pragma unroll
      for (int v = 0; v < KS / VEC; ++v) {
        vec_t xa[FM], xb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
          xa[i] = *reinterpret_cast<const vec_t*>(&As[buf][(wm * FM * 16 + i * 16 + lrow) * LDS_LD + lk * KS + v * VEC]);
#pragma unroll
        for (int j = 0; j < FN; ++j)
          xb[j] = *reinterpret_cast<const vec_t*>(&Bs[buf][(wn * FN * 16 + j * 16 + lrow) * LDS_LD + lk * KS + v * VEC]);
#pragma unroll
        for (int e = 0; e < VEC; ++e)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = M::mma(xa[i][e], xb[j][e], acc[i][j]);
      }
    } else {                      // all KS k values of the stage at once
      T af[FM][KS], bf[FN][KS];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const T* pa = &As[buf][(wm * FM * 16 + i * 16 + lrow) * LDS_LD + lk * KS];
#pragma unroll
        for (int v = 0; v < KS / VEC; ++v) {
          vec_t x = *reinterpret_cast<const vec_t*>(pa + v * VEC);
#pragma unroll
          for (int e = 0; e < VEC; ++e) af[i][v * VEC + e] = x[e];
        }
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const T* pb = &Bs[buf][(wn * FN * 16 + j * 16 + lrow) * LDS_LD + lk * KS];
#pragma unroll
        for (int v = 0; v < KS / VEC; ++v) {
          vec_t y = *reinterpret_cast<const vec_t*>(pb + v * VEC);
#pragma unroll
          for (int e = 0; e < VEC; ++e) bf[j][v * VEC + e] = y[e];
        }
      }
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = M::mma(af[i][s], bf[j][s], acc[i][j]);
    }
  };

  const int64_t nk = k / BK;
  if (RS == 2) {
    // stage kt lives in LDS buffer kt & 1 and, before that, in register set kt & 1:
    // the loads of stage kt+2 are in flight while stage kt is multiplied
    load_global(0, ra[0], rb[0]);
    if (nk > 1) load_global(BK, ra[RS - 1], rb[RS - 1]);
    store_lds(0, ra[0], rb[0]);
    __syncthreads();
    for (int64_t kt = 0; kt < nk; kt += 2) {
      if (kt + 2 < nk) load_global((kt + 2) * BK, ra[0], rb[0]);
      multiply(0);
      if (kt + 1 < nk) store_lds(1, ra[RS - 1], rb[RS - 1]);
      __syncthreads();
      if (kt + 1 >= nk) break;
      if (kt + 3 < nk) load_global((kt + 3) * BK, ra[RS - 1], rb[RS - 1]);
      multiply(1);
      if (kt + 2 < nk) store_lds(0, ra[0], rb[0]);
      __syncthreads();
    }
  } else {
    load_global(0, ra[0], rb[0]);
    store_lds(0, ra[0], rb[0]);
    __syncthreads();
    for (int64_t kt = 0; kt < nk; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nk) load_global((kt + 1) * BK, ra[0], rb[0]);
      multiply(buf);
      if (kt + 1 < nk) store_lds(buf ^ 1, ra[0], rb[0]);
      __syncthreads();
    }
  }

  // epilogue: lane holds column lane%16 of each 16x16 tile; its rows are
  // 4*(lane/16)+r for the f32 instruction and lane/16+4*r for the f64 one
  // (both checked with an in-bounds lane-layout probe on gfx950).
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = sizeof(T) == 8 ? lk + 4 * r : lk * 4 + r;
        const int64_t row = m0 + wm * FM * 16 + i * 16 + rr;
        const int64_t col = n0 + wn * FN * 16 + j * 16 + lrow;
        T* pc = C + row * ldc + col;
        const T prod = acc[i][j][r];
        *pc = (beta == T(0)) ? alpha * prod : fma(alpha, prod, beta * *pc);
      }
}

template <typename T, typename TL, int WPE>
void gemm_dispatch(dim3 grid, bool trans_b, bool lower, hipStream_t st, int64_t m, int64_t n, int64_t k, const T* A,
                   int64_t lda, const T* B, int64_t ldb, T alpha, T beta, T* C, int64_t ldc, int64_t tn,
                   int64_t tiles, int64_t sA, int64_t sB, int64_t sC) {
  dim3 block(TL::NT);
  if constexpr (TL::TN == TL::TM) {      // lower-triangle products: square tiles only
    if (lower) {
      if (trans_b) hipLaunchKernelGGL((gemm_kernel<T, true, true, TL, WPE>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles, sA, sB, sC);
      else hipLaunchKernelGGL((gemm_kernel<T, false, true, TL, WPE>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles, sA, sB, sC);
      return;
    }
  }
  if (trans_b) hipLaunchKernelGGL((gemm_kernel<T, true, false, TL, WPE>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles, sA, sB, sC);
  else hipLaunchKernelGGL((gemm_kernel<T, false, false, TL, WPE>), grid, block, 0, st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn, tiles, sA, sB, sC);
}

// C = alpha A op(B) + beta C; `batch` > 1 runs that many products, the i-th at
// A + i sA, B + i sB, C + i sC (one launch, products on blockIdx.y).
template <typename T>
int gemm_launch(int64_t m, int64_t n, int64_t k, const T* A, int64_t lda, const T* B, int64_t ldb,
                bool trans_b, T alpha, T beta, T* C, int64_t ldc, bool lower, hipStream_t st, int64_t batch = 1,
                int64_t sA = 0, int64_t sB = 0, int64_t sC = 0, bool row_tiles = false) {
  constexpr int BK = Mfma<T>::BK;
  GK_REQUIRE(m % 128 == 0 && n % 128 == 0 && k % BK == 0, GK_ERR_ARG,
             "gk_gemm: m, n must be multiples of 128 and k of %d (got %lld %lld %lld)", BK,
             (long long)m, (long long)n, (long long)k);
  GK_REQUIRE(lda % 4 == 0 && ldb % 4 == 0 && ldc >= n, GK_ERR_ARG, "gk_gemm: bad leading dimensions");
  GK_REQUIRE(!lower || m == n, GK_ERR_ARG, "gk_gemm: lower_only needs m == n");
  GK_REQUIRE(batch >= 1 && batch <= 65535 && sA % 4 == 0 && sB % 4 == 0, GK_ERR_ARG, "gk_gemm: bad batch");
  if (m == 0 || n == 0) return GK_OK;
  if (k == 0) {
    GK_REQUIRE(beta == T(1), GK_ERR_ARG, "gk_gemm: k == 0 only with beta == 1");
    return GK_OK;
  }
  hipEvent_t pa;
  const bool prof = gk_prof_begin(st, &pa);
  // Tall narrow FP32 products (the panel solve and the next panel's update:
  // m in the thousands, n = 256) and the 128^3 products of the diagonal
  // blocks have too few 128 x 128 tiles to fill the chip (one, for the
  // latter); 64 x 64 tiles of 4 waves give 4x as many workgroups.
  // row_tiles (n == 128): one tile spans all of C's columns, so a workgroup
  // reads every A row it writes before writing it (in place, C == A, is safe).
  GK_REQUIRE(!row_tiles || (n == 128 && !lower), GK_ERR_ARG, "gk_gemm: row tiles need n == 128");
  const bool narrow = !row_tiles && sizeof(T) == 4 && !lower && batch == 1 &&
                      ((n <= 256 && m >= 2048) || m * n <= 256 * 256);
  const int64_t TMe = narrow ? 64 : 128;
  const int64_t tm = m / TMe, tn = n / TMe;
  const int64_t tiles = lower ? tm * (tm + 1) / 2 : tm * tn;
  bool launched = false;
  if constexpr (sizeof(T) == 4) {
    if (row_tiles) {
      // FP32 row tiles of 64 x 128 on 4 waves: 30 KB of LDS and at most 102
      // VGPRs, so a workgroup fits on a CU beside two workgroups of the
      // split-fp16 trailing update (64 KB of LDS and 200 VGPRs per wave
      // each) and the factor panel's products run while that update holds
      // the GPU.
      gemm_dispatch<T, Tile<64, 1, 4, 128>, 5>(dim3((unsigned)(m / 64), (unsigned)batch), trans_b, false, st, m, n,
                                               k, A, lda, B, ldb, alpha, beta, C, ldc, 1, m / 64, sA, sB, sC);
      launched = true;
    }
  }
  if (launched) {
  } else if (narrow)
    gemm_dispatch<T, Tile<64, 2, 2>, Mfma<T>::WPE>(dim3((unsigned)tiles, 1), trans_b, false, st, m, n, k, A, lda, B,
                                                   ldb, alpha, beta, C, ldc, tn, tiles, 0, 0, 0);
  else
    gemm_dispatch<T, typename GemmTile<T>::type, Mfma<T>::WPE>(dim3((unsigned)tiles, (unsigned)batch), trans_b, lower,
                                                               st, m, n, k, A, lda, B, ldb, alpha, beta, C, ldc, tn,
                                                               tiles, sA, sB, sC);
  if (prof)
    gk_prof_end(st, pa, 4 * (sizeof(T) == 8) + 2 * trans_b + lower,
                (double)batch * (lower ? (double)m * (double)(m + 1) * (double)k : 2.0 * (double)m * (double)n * (double)k));
  GK_HIP(hipGetLastError());
  return GK_OK;
}

// ----------------------------------------------- split-fp16 rank-256 update
// C -= A B^T with K = 256 on v_mfma_f32_16x16x32_f16, for factor panels whose
// entries lie within [-16, 16] (rows of an equilibrated Cholesky factor: |L_ij|
// <= sqrt(H_ii + shift) ~ 1).  Each FP32 operand x is split once per panel
// into hi = fp16(x) and lo = fp16((x - hi) 2^11); then
//   a b ~ 2^-11 ((2^11 hi_a) hi_b + hi_a lo_b + lo_a hi_b)
// (2^11 hi_a is exact in fp16 below |x| = 16 x 2^11 / 2^11 ... i.e. |x| < 32),
// three MFMAs into one FP32 accumulator.  The dropped lo_a lo_b term and the
// two roundings leave ~2^-21 relative per product, against 2^-24 for FP32:
// the factor's backward error is within 2x of the FP32 update's on the
// co-kriging blocks, and fp16 MFMA runs at 16x the FP32 MFMA rate on gfx950.
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

constexpr int H3_K = 256;       // depth of the update (panel width)
constexpr int H3_BK = 32;       // k per LDS stage (one MFMA)
constexpr int H3_NK = H3_K / H3_BK;
constexpr int H3_TILE = 128 * H3_BK;          // halves of one array of one stage of a 128-row block
constexpr int64_t H3_ROWBLK = 2 * 128 * H3_K; // halves of a 128-row block (both arrays, all stages)

// Split panel layout ("tiled"): for 128-row block b and k-chunk c (32 wide),
// the hi part (128 x 32, row-major) then the lo part, so that one stage of a
// tile's operand is one contiguous 16 KB block, copied to LDS as it lies.
// Element (r, k): ((r / 128) * H3_NK + k / 32) * 2 H3_TILE + [lo] H3_TILE
// + (r % 128) * 32 + k % 32.  A row offset R (a multiple of 128) is R * 2 H3_K
// halves.
__global__ void split_h3_kernel(const float* __restrict__ W, int64_t ldw, int64_t n8, _Float16* __restrict__ P) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const f4v* w = reinterpret_cast<const f4v*>(W + (i / (H3_K / 8)) * ldw + (i % (H3_K / 8)) * 8);
  const f4v a = w[0];
  const f4v b = w[1];
  h8v hi, lo;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    hi[e] = (_Float16)a[e];
    hi[e + 4] = (_Float16)b[e];
    lo[e] = (_Float16)((a[e] - (float)hi[e]) * 2048.0f);
    lo[e + 4] = (_Float16)((b[e] - (float)hi[e + 4]) * 2048.0f);
  }
  const int64_t r = i / (H3_K / 8);
  const int kv = (int)(i % (H3_K / 8)), c = kv >> 2, kk = (kv & 3) * 8;
  _Float16* base = P + ((r >> 7) * H3_NK + c) * 2 * H3_TILE + (r & 127) * H3_BK + kk;
  *reinterpret_cast<h8v*>(base) = hi;
  *reinterpret_cast<h8v*>(base + H3_TILE) = lo;
}

#ifndef H3_RS
#define H3_RS 2                 // stages in flight in registers
#endif

// Tile (tm >= tn) of the lower triangle of nt x nt tiles for launch index b,
// in supertiles of H3_SUPER x H3_SUPER tiles (lower triangle of supertiles,
// row by row; inside one, column by column; the diagonal supertiles are
// triangles): an XCD's concurrently resident workgroups (its contiguous range
// of indices) then share a few A and B row blocks, which stay in its L2,
// instead of streaming a different B block for every tile of a tile row.
constexpr int H3_SUPER = 8;
__device__ __forceinline__ void h3_lower_tile(int64_t b, int64_t nt, int64_t* tm, int64_t* tn) {
  const int64_t G = H3_SUPER, S = (nt + G - 1) / G;
  // tiles before supertile row I (all rows above it full height G):
  // G^2 I (I - 1) / 2 + I G (G + 1) / 2
  auto before = [&](int64_t I) { return G * G * I * (I - 1) / 2 + I * G * (G + 1) / 2; };
  int64_t I = (int64_t)(sqrt(2.0 * (double)b) / (double)G);
  if (I > S - 1) I = S - 1;
  while (I > 0 && before(I) > b) --I;
  while (I + 1 < S && before(I + 1) <= b) ++I;
  const int64_t h = min(G, nt - I * G);              // tile rows in supertile row I
  int64_t l = b - before(I);
  if (l < I * G * h) {                                // rectangle supertile (I, J), J < I
    const int64_t J = l / (G * h), r = l % (G * h);
    *tn = J * G + r / h;
    *tm = I * G + r % h;
  } else {                                            // diagonal supertile: triangle of h rows
    l -= I * G * h;
    int64_t c = 0;                                    // column c holds h - c tiles
    while (l >= h - c) { l -= h - c; ++c; }
    *tn = I * G + c;
    *tm = I * G + c + l;
  }
}

// 128 x 128 tile, 4 waves of 64 x 64 (4 x 4 fragments of 16 x 16).  A, B:
// tiled split panels at their first row; NP = 2 adds the product of a second
// pair (A2, B2): a rank-512 update in one pass over C.
template <bool LOWER, int NP>
__device__ __forceinline__ void syrk_h3_tile(const _Float16* __restrict__ A, const _Float16* __restrict__ B,
                                             const _Float16* __restrict__ A2, const _Float16* __restrict__ B2,
                                             float* C, int64_t ldc, int64_t tm, int64_t tn,
                                             _Float16 (*S)[4 * H3_TILE]) {
  constexpr int NK = NP * H3_NK;              // stages
  const int64_t m0 = tm * 128, n0 = tn * 128;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lrow = lane & 15, lk = lane >> 4;
  const h8v* srcA = reinterpret_cast<const h8v*>(A + tm * H3_ROWBLK);
  const h8v* srcB = reinterpret_cast<const h8v*>(B + tn * H3_ROWBLK);
  const h8v* srcA2 = reinterpret_cast<const h8v*>((NP == 2 ? A2 : A) + tm * H3_ROWBLK);
  const h8v* srcB2 = reinterpret_cast<const h8v*>((NP == 2 ? B2 : B) + tn * H3_ROWBLK);

  // The accumulators start at -2^11 C (exact scaling), so the epilogue only
  // stores: C - 2^-11 sum = -2^-11 (-2^11 C + sum).  The C loads are issued
  // first and complete under the operand pipeline's start.  Lane layout of a
  // 16 x 16 fragment (v_mfma_f32_16x16x32_f16, checked with an in-bounds
  // probe): column lane % 16, rows 4 (lane / 16) + r.
  f4v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 64 + i * 16 + lk * 4 + r;
        const int64_t col = n0 + wn * 64 + j * 16 + lrow;
        acc[i][j][r] = -2048.0f * C[row * ldc + col];
      }

  // one stage = A's 16 KB block then B's: 2048 vectors of 8 halves, 8 per
  // thread, contiguous in memory and in LDS; H3_RS register sets keep H3_RS
  // stages in flight
  h8v st_reg[H3_RS][8];
  auto load_global = [&](int set, int c) {
    const h8v* sa = c < H3_NK ? srcA : srcA2;
    const h8v* sb = c < H3_NK ? srcB : srcB2;
    const int cc = c % H3_NK;
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int v = tid + 256 * it;           // 0..1023: A, 1024..2047: B
      st_reg[set][it] = it < 4 ? sa[cc * (2 * H3_TILE / 8) + v] : sb[cc * (2 * H3_TILE / 8) + v - 1024];
    }
  };
  auto store_lds = [&](int buf, int set) {
#pragma unroll
    for (int it = 0; it < 8; ++it) reinterpret_cast<h8v*>(S[buf])[tid + 256 * it] = st_reg[set][it];
  };
  auto multiply = [&](int buf) {
    const _Float16* SA = S[buf];
    const _Float16* SB = S[buf] + 2 * H3_TILE;
    h8v ah[4], al[4], as[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = (wm * 64 + i * 16 + lrow) * H3_BK + lk * 8;
      ah[i] = *reinterpret_cast<const h8v*>(SA + r);
      al[i] = *reinterpret_cast<const h8v*>(SA + H3_TILE + r);
      as[i] = ah[i] * (_Float16)2048.0f;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = (wn * 64 + j * 16 + lrow) * H3_BK + lk * 8;
      bh[j] = *reinterpret_cast<const h8v*>(SB + r);
      bl[j] = *reinterpret_cast<const h8v*>(SB + H3_TILE + r);
    }
    // term by term over all 16 fragments: an accumulator's next product is
    // issued 16 MFMAs after its previous one
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(as[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
  };

  // stage kt: registers set kt % H3_RS, then LDS buffer kt & 1 (loop fully
  // unrolled so the register sets are static)
#pragma unroll
  for (int q = 0; q < H3_RS; ++q) load_global(q, q);
  store_lds(0, 0);
  __syncthreads();
#pragma unroll
  for (int kt = 0; kt < NK; ++kt) {
    if (kt + H3_RS < NK) load_global(kt % H3_RS, kt + H3_RS);
    multiply(kt & 1);
    if (kt + 1 < NK) store_lds((kt + 1) & 1, (kt + 1) % H3_RS);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm * 64 + i * 16 + lk * 4 + r;
        const int64_t col = n0 + wn * 64 + j * 16 + lrow;
        C[row * ldc + col] = -0.00048828125f * acc[i][j][r];
      }
}

// Grid of g workgroups over t tiles (g <= t): workgroup b runs on XCD b % 8
// (round-robin dispatch) and takes tiles from its XCD's contiguous range
// (x t/8 + min(x, t%8), t/8 + [x < t%8] long), so the tiles resident at once
// on an XCD are neighbours and share operand blocks in its L2.  g < t keeps
// the workgroups resident (a persistent grid) and leaves the rest of the chip
// to other streams; the range is then handed out in order: each workgroup's
// first tile is its index l in the XCD, the next ones come from the XCD's
// counter ctr[x] (tile nx + ticket), as a dispatcher would hand them out.
// The workgroup that draws the XCD's last ticket (every workgroup draws one
// past the range) sets the counter back to 0 for the next launch.  Without
// counters the workgroups step through the range by their count.  Every tile
// is visited once: XCD x holds nx = g/8 + [x < g%8] >= 1 workgroups whenever
// g >= 8 or g == t.
template <bool LOWER, int NP>
__global__ __launch_bounds__(256, 2) void syrk_h3_kernel(const _Float16* __restrict__ A,
                                                         const _Float16* __restrict__ B,
                                                         const _Float16* __restrict__ A2,
                                                         const _Float16* __restrict__ B2, float* C, int64_t ldc,
                                                         int64_t tiles_n, int64_t tiles, int32_t* ctr) {
  __shared__ __attribute__((aligned(16))) _Float16 S[2][4 * H3_TILE];   // [buf][A hi | A lo | B hi | B lo]
  __shared__ int64_t next;
  const int64_t g = gridDim.x, x = blockIdx.x % 8, l = blockIdx.x / 8;
  const int64_t nx = g / 8 + (x < g % 8 ? 1 : 0);
  const int64_t start = x * (tiles / 8) + min(x, tiles % 8), cnt = tiles / 8 + (x < tiles % 8 ? 1 : 0);
  const bool dyn = ctr != nullptr && g < tiles;
  const int64_t last_ticket = (cnt > nx ? cnt - nx : 0) + nx - 1;
  int64_t j = l;
  while (true) {
    if (j < cnt) {
      int64_t tm, tn;
      if (LOWER) {
        h3_lower_tile(start + j, tiles_n, &tm, &tn);
      } else {
        tm = (start + j) / tiles_n;
        tn = (start + j) % tiles_n;
      }
      syrk_h3_tile<LOWER, NP>(A, B, A2, B2, C, ldc, tm, tn, S);
    }
    if (!dyn) {
      j += nx;
      if (j >= cnt) break;
      continue;
    }
    if (threadIdx.x == 0) {
      const int64_t v = atomicAdd(&ctr[x], 1);
      if (v == last_ticket) atomicExch(&ctr[x], 0);
      next = nx + v;
    }
    __syncthreads();
    j = next;
    if (j >= cnt) break;
  }
}

// Split an FP32 panel (rows x 256, dense, rows a multiple of 128) into the
// tiled fp16 layout.
int split_h3(const float* W, int64_t rows, _Float16* P, hipStream_t st, int64_t ldw = H3_K) {
  GK_REQUIRE(rows % 128 == 0 && ldw >= H3_K && ldw % 4 == 0, GK_ERR_ARG,
             "split_h3: rows must be a multiple of 128");
  const int64_t n8 = rows * H3_K / 8;
  if (n8 > 0)
    hipLaunchKernelGGL(split_h3_kernel, dim3((unsigned)gk_ceil_div(n8, 256)), dim3(256), 0, st, W, ldw, n8, P);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

// Workgroups of the persistent trailing update: two per CU (64 KB of LDS
// each) less GK_SYRK_RESERVE (default 16) slots, a multiple of 8 and >= 8.
static int64_t syrk_h3_cap() {
  static const int64_t cap = [] {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
    const char* e = std::getenv("GK_SYRK_RESERVE");
    const int64_t reserve = e && *e ? std::atoll(e) : 16;
    return std::max<int64_t>(8, (2 * (int64_t)cus - reserve) / 8 * 8);
  }();
  return cap;
}

// C[m x n] -= A B^T (K = 256) from tiled split panels (A: m rows, B: n rows),
// and -= A2 B2^T as well when A2 is given (K = 512 in one pass over C);
// lower: only the tiles on or below the diagonal.
int syrk_h3_launch(int64_t m, int64_t n, const _Float16* A, const _Float16* B, float* C, int64_t ldc, bool lower,
                   hipStream_t st, int32_t* ctr = nullptr, const _Float16* A2 = nullptr,
                   const _Float16* B2 = nullptr) {
  GK_REQUIRE(m % 128 == 0 && n % 128 == 0 && ldc >= n && (!lower || m == n), GK_ERR_ARG,
             "syrk_h3: m, n must be multiples of 128 (m == n for the lower update)");
  if (m == 0 || n == 0) return GK_OK;
  hipEvent_t pa;
  const bool prof = gk_prof_begin(st, &pa);
  const int64_t tm = m / 128, tn = n / 128;
  const int64_t tiles = lower ? tm * (tm + 1) / 2 : tm * tn;
  // the trailing (lower) update runs on a persistent grid of at most
  // syrk_h3_cap() workgroups: the panel chain beside it keeps CUs to run on
  const int64_t grid = lower ? std::min(tiles, syrk_h3_cap()) : tiles;
  const int np = A2 ? 2 : 1;
  GK_REQUIRE(!A2 == !B2, GK_ERR_ARG, "syrk_h3: A2 and B2 go together");
  int32_t* no_ctr = nullptr;
  if (lower && np == 2)
    hipLaunchKernelGGL((syrk_h3_kernel<true, 2>), dim3((unsigned)grid), dim3(256), 0, st, A, B, A2, B2, C, ldc, tn,
                       tiles, ctr);
  else if (lower)
    hipLaunchKernelGGL((syrk_h3_kernel<true, 1>), dim3((unsigned)grid), dim3(256), 0, st, A, B, A, B, C, ldc, tn,
                       tiles, ctr);
  else if (np == 2)
    hipLaunchKernelGGL((syrk_h3_kernel<false, 2>), dim3((unsigned)grid), dim3(256), 0, st, A, B, A2, B2, C, ldc, tn,
                       tiles, no_ctr);
  else
    hipLaunchKernelGGL((syrk_h3_kernel<false, 1>), dim3((unsigned)grid), dim3(256), 0, st, A, B, A, B, C, ldc, tn,
                       tiles, no_ctr);
  const double K = (double)H3_K * np;
  if (prof)
    gk_prof_end(st, pa, 10 + (lower ? 1 : 0), lower ? (double)m * (double)(m + 1) * K : 2.0 * (double)m * (double)n * K);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

// ------------------------------------------------------- diagonal blocks
// A 256-wide diagonal block is factored as two 128-wide halves: each half is
// factored and inverted by one workgroup with the block resident in LDS, and
// the halves are joined with 128^3 MFMA GEMMs (trsm/syrk and the off-diagonal
// quadrant of the inverse).
constexpr int DB = 128;

// LU without pivoting of a DB x DB FP64 block in LDS, then the inverses of its
// unit-lower factor (Li) and of its upper factor (Ui), each computed with the
// factor and the inverse packed in the same LDS buffer.
__global__ __launch_bounds__(1024) void getrf128_kernel(double* A, int64_t lda, double* Li, double* Ui,
                                                        int64_t ldx, int32_t* info, int64_t gbase) {
  __shared__ double S[DB * (DB + 1)];
  const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5;
  constexpr int LD = DB + 1;
  for (int e = tid; e < DB * DB; e += 1024) S[(e / DB) * LD + e % DB] = A[(e / DB) * lda + e % DB];
  __syncthreads();
  for (int j = 0; j < DB; ++j) {
    double p = S[j * LD + j];
    const bool bad = !(fabs(p) > 0.0);
    if (bad) p = 1.0;
    const double ip = 1.0 / p;
    __syncthreads();
    if (tid == 0 && bad) atomicCAS(info, 0, (int32_t)(gbase + j + 1));
    if (tid > j && tid < DB) S[tid * LD + j] *= ip;
    __syncthreads();
    for (int i = j + 1 + ty; i < DB; i += 32)
      for (int c = j + 1 + tx; c < DB; c += 32) S[i * LD + c] -= S[i * LD + j] * S[j * LD + c];
    __syncthreads();
  }
  for (int e = tid; e < DB * DB; e += 1024) A[(e / DB) * lda + e % DB] = S[(e / DB) * LD + e % DB];
  __syncthreads();
  // --- unit-lower inverse: strict lower L packed at [0, 8128), X packed lower at [8128, ...)
  double* Lp = S;
  double* Xp = S + DB * (DB - 1) / 2;
  auto psl = [](int i, int c) { return i * (i - 1) / 2 + c; };
  auto pl = [](int i, int c) { return i * (i + 1) / 2 + c; };
  for (int e = tid; e < DB * DB; e += 1024) {
    const int i = e / DB, c = e % DB;
    if (c < i) Lp[psl(i, c)] = A[i * lda + c];
  }
  __syncthreads();
  for (int e = tid; e < DB * DB; e += 1024) {
    const int i = e / DB, c = e % DB;
    if (c <= i) Xp[pl(i, c)] = (i == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < DB; ++k) {
    for (int i = k + 1 + ty; i < DB; i += 32) {
      const double l = Lp[psl(i, k)];
      for (int c = tx; c <= k; c += 32) Xp[pl(i, c)] -= l * Xp[pl(k, c)];
    }
    __syncthreads();
  }
  for (int e = tid; e < DB * DB; e += 1024) {
    const int i = e / DB, c = e % DB;
    Li[i * ldx + c] = (c <= i) ? Xp[pl(i, c)] : 0.0;
  }
  __syncthreads();
  // --- upper inverse: U packed upper at [0, 8256), Y packed upper at [8256, ...)
  double* Up = S;
  double* Yp = S + DB * (DB + 1) / 2;
  auto pu = [](int i, int c) { return i * (2 * DB - i + 1) / 2 + (c - i); };
  for (int e = tid; e < DB * DB; e += 1024) {
    const int i = e / DB, c = e % DB;
    if (c >= i) { Up[pu(i, c)] = A[i * lda + c]; Yp[pu(i, c)] = (i == c) ? 1.0 : 0.0; }
  }
  __syncthreads();
  for (int k = DB - 1; k >= 0; --k) {
    const double ip = 1.0 / Up[pu(k, k)];
    if (tid >= k && tid < DB) Yp[pu(k, tid)] *= ip;
    __syncthreads();
    for (int i = ty; i < k; i += 32) {
      const double u = Up[pu(i, k)];
      for (int c = k + tx; c < DB; c += 32) Yp[pu(i, c)] -= u * Yp[pu(k, c)];
    }
    __syncthreads();
  }
  for (int e = tid; e < DB * DB; e += 1024) {
    const int i = e / DB, c = e % DB;
    Ui[i * ldx + c] = (c >= i) ? Yp[pu(i, c)] : 0.0;
  }
}

// Blocked Cholesky + inverse of a DB x DB block (one workgroup of NT threads),
// two barriers per PB columns.  Phase 1 (wave 0, registers, lane values
// broadcast through scalar registers): factor the PB-column panel, then the PB
// rows of X = L^-1,
//   X[i][c] = (delta_ic - acc[i][c] - sum_{jb<=k<i} L[i][k] X[k][c]) / L[i][i],
// where acc[i][c] = sum_{k<jb} L[i][k] X[k][c] was applied eagerly.  Phase 2
// (all threads): the rank-PB trailing update of L and the eager update of the
// later rows of X by the new panel and rows, in one sweep per row.
#ifndef GK_POTRF_PB_F32
#define GK_POTRF_PB_F32 16
#endif
#ifndef GK_POTRF_PB_F64
#define GK_POTRF_PB_F64 4
#endif
// LDS: L and X = L^-1 as packed lower triangles, then two panel buffers P[b]
// holding the PB columns of the current / next panel column-major (P[b][t][i]
// = L[i][jb + t]): wave 0 loads its panel rows without per-row branches (16%
// less time in phase 1 than from the packed rows) and the sweep reads the
// panel with consecutive addresses.  FP32 72.5 KB, FP64 137 KB.
template <typename T, int PB>
struct BlkStore {
  T* S;
  static constexpr int SIZE = DB * (DB + 1) + 2 * PB * DB + PB;
  __device__ static int pl(int i, int c) { return i * (i + 1) / 2 + c; }
  __device__ T& L(int i, int c) { return S[pl(i, c)]; }
  __device__ T& X(int i, int c) { return S[DB * (DB + 1) / 2 + pl(i, c)]; }
  __device__ T& P(int b, int t, int i) { return S[DB * (DB + 1) + (b * PB + t) * DB + i]; }
  __device__ T& ipv(int j) { return S[DB * (DB + 1) + 2 * PB * DB + j]; }   // 1 / pivot of the current panel
};
// FP32 diagonal-block kernel: threads.  It runs on the side stream beside the
// trailing update, whose workgroups (512 threads, 80 VGPRs, 40 KB LDS) sit
// three to a CU: with 512 threads, < 96 VGPRs and < 80 KB of LDS it fits on a
// CU where one of them has finished, while a 1024-thread version needs an
// empty CU and waits for the end of the update.
#ifndef GK_POTRF_NT_F32
#define GK_POTRF_NT_F32 512
#endif

// Value of v in lane l (l uniform) as a scalar-register broadcast.
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Pivot d > 0 -> (sqrt(d), 1/sqrt(d)).  FP32: the hardware reciprocal square
// root (within 1 ulp), the factor being an FP32 preconditioner; FP64: IEEE.
__device__ __forceinline__ void pivot_pair(float d, float& piv, float& ip) {
  ip = __builtin_amdgcn_rsqf(d);
  piv = d * ip;
}
__device__ __forceinline__ void pivot_pair(double d, double& piv, double& ip) {
  piv = sqrt(d);
  ip = 1.0 / piv;
}

template <typename T, int PB, int NT>
__device__ void potrf_blocked(BlkStore<T, PB> B, T* A, int64_t lda, T* X, int64_t ldx, int32_t* info, int64_t gbase) {
  constexpr int RS = NT / 32;                 // rows of the trailing update in flight
  const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5, lane = tid & 63, wave = tid >> 6;
  // all loads of the block first (in flight together, 16 bytes each), then
  // the LDS stores
  constexpr int VL = 16 / sizeof(T);          // elements per 16-byte load
  constexpr int PER = DB * DB / (NT * VL);
  static_assert(PER * NT * VL == DB * DB, "the block must be a whole number of vectors per thread");
  typedef T vt __attribute__((ext_vector_type(VL)));
  vt vin[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = (tid + q * NT) * VL, i = e / DB, c = e % DB;
    if (c <= i) vin[q] = *reinterpret_cast<const vt*>(A + i * lda + c);
    else
#pragma unroll
      for (int u = 0; u < VL; ++u) vin[q][u] = T(0);
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int e = (tid + q * NT) * VL, i = e / DB, c = e % DB;
#pragma unroll
    for (int u = 0; u < VL; ++u) {
      if (c + u <= i) {
        B.L(i, c + u) = vin[q][u];
        B.X(i, c + u) = T(0);
        if (c + u < PB) B.P(0, c + u, i) = vin[q][u];
      }
    }
  }
  __syncthreads();
  int bad_col = 0;                            // first non-positive pivot + 1 (wave 0)
  // FP32: the rows of X are computed by the last two waves from the stored
  // panel (64 columns each) while the other waves apply the panel to the
  // trailing L tiles; the X tiles follow beside the next panel's factor.
  constexpr bool SPLIT_X = sizeof(T) == 4 && PB == 16 && NT >= 128;
  constexpr int XW = NT / 64 - 1;             // the last of the two waves of the X rows
  // ---- rows jb .. jb+PB-1 of X from the panel's diagonal block (lane l holds
  //      row jb + l in a) and the reciprocal pivots: lane owns columns lane
  //      and lane + 64
  // pb >= 0: the panel's entries come from the stored panel P[pb] (LDS
  // broadcasts: a wave other than the one that factored it), else from a.
  auto x_rows = [&](int jb, const T (&a)[PB], const T (&ipv)[PB], int pb) {
    T xa[PB], xb[PB];
    const int ca = lane, cb = lane + 64;
#pragma unroll
    for (int r = 0; r < PB; ++r) {
      const int i = jb + r;
      xa[r] = ca < jb ? B.X(i, ca) : (ca == i ? T(1) : T(0));  // B.X holds -acc[i][c] for c < jb
      xb[r] = cb < jb ? B.X(i, cb) : (cb == i ? T(1) : T(0));
    }
    // right-looking: row k is final once the rows above it are applied, and
    // then goes into every later row (each row takes the same products in the
    // same order as a row-by-row substitution; the dependent chain is 16
    // steps, not 120)
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      xa[k] *= ipv[k];
      xb[k] *= ipv[k];
#pragma unroll
      for (int r = k + 1; r < PB; ++r) {
        const T lrk = pb >= 0 ? B.P(pb, k, jb + r) : lane_value(a[k], r);   // L[jb+r][jb+k]
        xa[r] -= lrk * xa[k];
        xb[r] -= lrk * xb[k];
      }
    }
#pragma unroll
    for (int r = 0; r < PB; ++r) {
      const int i = jb + r;
      if (ca <= i) B.X(i, ca) = xa[r];
      if (cb <= i) B.X(i, cb) = xb[r];
    }
  };
  // the same for one column per lane (c = lane + coff), the panel's entries
  // from the stored panel P[pb] (LDS broadcasts): two waves share the rows
  auto x_rows1 = [&](int jb, int pb, int coff) {
    T x[PB], ipv[PB];
    const int c = lane + coff;
#pragma unroll
    for (int r = 0; r < PB; ++r) {
      const int i = jb + r;
      x[r] = c < jb ? B.X(i, c) : (c == i ? T(1) : T(0));
      ipv[r] = B.ipv(r);
    }
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      x[k] *= ipv[k];
#pragma unroll
      for (int r = k + 1; r < PB; ++r) x[r] -= B.P(pb, k, jb + r) * x[k];   // L[jb+r][jb+k]
    }
#pragma unroll
    for (int r = 0; r < PB; ++r)
      if (c <= jb + r) B.X(jb + r, c) = x[r];
  };
  // ---- phase 2 on MFMA (16 x 16 x 4 f32), one 16 x 16 tile per wave at a time
  //      (waves w, w + nw, ...):
  //      L tiles (ti >= tj >= c0/16):  L[i][c] -= sum_t L[i][jb+t] L[c][jb+t]
  //      X tiles (ti >= c0/16 > tj):   X[i][c] -= sum_t L[i][jb+t] X[jb+t][c]
  // Operands from the column-major panel buffer P[pb] (and the packed X rows
  // of the panel); a lane holds C[4 (lane/16) + r][lane % 16] of a tile.
  auto mfma_tiles = [&](auto jb, int pb, int c0, bool is_l, int w, int nw) {
    const int a = c0 / 16, nt = DB / 16 - a;             // first trailing tile row, tile rows left
    const int count = is_l ? nt * (nt + 1) / 2 : nt * a;
    const int lr = lane & 15, lq = lane >> 4;
    for (int tile = w; tile < count; tile += nw) {
      int ti, tj;
      if (is_l) {                                         // lower triangle of the trailing tiles
        int t = tile, r = 0;
        while (t > r) { t -= r + 1; ++r; }
        ti = a + r;
        tj = a + t;
      } else {
        ti = a + tile / a;
        tj = tile % a;
      }
      const int i0 = ti * 16, cb = tj * 16;
      f4v acc;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + lq * 4 + r, c = cb + lr;
        acc[r] = is_l ? (c <= i ? B.L(i, c) : 0.0f) : B.X(i, c);
      }
#pragma unroll
      for (int ks = 0; ks < PB / 4; ++ks) {
        const int t = ks * 4 + lq;
        const float av = -B.P(pb, t, i0 + lr);            // A[row][k] = L[i0+lr][jb+t]
        float bv;
        if (is_l) bv = B.P(pb, t, cb + lr);               // B[k][col] = L[cb+lr][jb+t]
        else bv = (cb + lr <= jb + t) ? B.X(jb + t, cb + lr) : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + lq * 4 + r, c = cb + lr;
        if (is_l) {
          if (c <= i) {
            B.L(i, c) = acc[r];
            if (c < c0 + PB) B.P(pb ^ 1, c - c0, i) = acc[r];
          }
        } else {
          B.X(i, c) = acc[r];
        }
      }
    }
  };
  for (int jb = 0, pb = 0; jb < DB; jb += PB, pb ^= 1) {
    if constexpr (SPLIT_X) {
      // the previous step's X tiles beside this step's panel (they touch
      // only X, which the panel does not read)
      if (wave > 0 && jb > 0) mfma_tiles(jb - PB, pb ^ 1, jb, false, wave - 1, NT / 64 - 1);
    }
    if (wave == 0) {
      // ---- panel: lane l holds rows jb + l (a) and jb + 64 + l (b)
      const int ra = jb + lane, rb = jb + 64 + lane;
      const int rac = min(ra, DB - 1), rbc = min(rb, DB - 1);
      T a[PB], b[PB], ipv[PB];
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        const T va = B.P(pb, c, rac), vb = B.P(pb, c, rbc);
        a[c] = (ra < DB && jb + c <= ra) ? va : T(0);
        b[c] = rb < DB ? vb : T(0);
      }
      // column j below the diagonal (aj, 0 on and above it) goes into the later
      // columns unmasked; the rows jb + 64 + l exist only while jb < 64
      auto factor = [&](auto with_b) {
#pragma unroll
        for (int j = 0; j < PB; ++j) {
          T d = lane_value(a[j], j);              // L[jb+j][jb+j], held by lane j
          if (!(d > T(0))) {
            if (bad_col == 0) bad_col = jb + j + 1;
            d = T(1);
          }
          T piv, ip;
          pivot_pair(d, piv, ip);
          ipv[j] = ip;
          const T aj = lane > j ? a[j] * ip : T(0);
          if constexpr (decltype(with_b)::value) b[j] *= ip;
#pragma unroll
          for (int c = j + 1; c < PB; ++c) {
            const T lcj = lane_value(aj, c);      // L[jb+c][jb+j]
            a[c] -= aj * lcj;
            if constexpr (decltype(with_b)::value) b[c] -= b[j] * lcj;
          }
          a[j] = lane == j ? piv : (lane > j ? aj : a[j]);
        }
      };
      if (jb < 64) factor(std::true_type{});
      else factor(std::false_type{});
#pragma unroll
      for (int c = 0; c < PB; ++c) {
        if (ra < DB && jb + c <= ra) B.L(ra, jb + c) = a[c];
        if (rb < DB) B.L(rb, jb + c) = b[c];
        if (ra < DB) B.P(pb, c, ra) = a[c];
        if (rb < DB) B.P(pb, c, rb) = b[c];
      }
      if constexpr (SPLIT_X) {
        if (lane < PB) B.ipv(lane) = ipv[lane];
      } else {
        x_rows(jb, a, ipv, -1);
      }
    }
    __syncthreads();
    const int c0 = jb + PB;
    if constexpr (SPLIT_X) {
      if (wave >= XW - 1) {
        x_rows1(jb, pb, wave == XW ? 64 : 0);
      } else {
        mfma_tiles(jb, pb, c0, true, wave, XW - 1);
      }
      __syncthreads();
      continue;                               // (the last step has no X tiles)
    }
    // ---- rows i >= c0: L[i][c] -= sum_t L[i][jb+t] L[c][jb+t]   (c0 <= c <= i;
    //      the next panel's columns also into P[pb^1])
    //                    -acc[i][c] -= sum_t L[i][jb+t] X[jb+t][c] (c < c0, X lower)
    for (int i = c0 + ty; i < DB; i += RS) {
      T li[PB];
#pragma unroll
      for (int t = 0; t < PB; ++t) li[t] = B.P(pb, t, i);
      for (int c = tx; c <= i; c += 32) {
        T sacc = T(0);
        if (c >= c0) {
#pragma unroll
          for (int t = 0; t < PB; ++t) sacc = fma(li[t], B.P(pb, t, c), sacc);
          const T v = B.L(i, c) - sacc;
          B.L(i, c) = v;
          if (c < c0 + PB) B.P(pb ^ 1, c - c0, i) = v;
        } else {
#pragma unroll
          for (int t = 0; t < PB; ++t) {
            const T xv = B.X(jb + t, c);
            sacc = fma(li[t], jb + t >= c ? xv : T(0), sacc);
          }
          B.X(i, c) -= sacc;
        }
      }
    }
    __syncthreads();
  }
  if (wave == 0 && lane == 0 && bad_col != 0) atomicCAS(info, 0, (int32_t)(gbase + bad_col));
  for (int e = tid * VL; e < DB * DB; e += NT * VL) {
    const int i = e / DB, c = e % DB;
    vt l, x;
#pragma unroll
    for (int u = 0; u < VL; ++u) {
      l[u] = c + u <= i ? B.L(i, c + u) : T(0);
      x[u] = c + u <= i ? B.X(i, c + u) : T(0);
    }
    if (c + VL - 1 <= i) {
      *reinterpret_cast<vt*>(A + i * lda + c) = l;
    } else {
#pragma unroll
      for (int u = 0; u < VL; ++u)
        if (c + u <= i) A[i * lda + c + u] = l[u];
    }
    *reinterpret_cast<vt*>(X + i * ldx + c) = x;
  }
}

// FP32: 72.5 KB of LDS, so the kernel fits on a CU beside workgroups of the
// concurrent trailing update (the look-ahead factorization runs it on a
// second stream).
__global__ __launch_bounds__(GK_POTRF_NT_F32) void potrf128b_f32_kernel(float* A, int64_t lda, float* X, int64_t ldx,
                                                                        int32_t* info, int64_t gbase) {
  __shared__ float S[BlkStore<float, GK_POTRF_PB_F32>::SIZE];
  potrf_blocked<float, GK_POTRF_PB_F32, GK_POTRF_NT_F32>(BlkStore<float, GK_POTRF_PB_F32>{S}, A, lda, X, ldx, info,
                                                         gbase);
}

__global__ __launch_bounds__(1024) void potrf128b_f64_kernel(double* A, int64_t lda, double* X, int64_t ldx,
                                                             int32_t* info, int64_t gbase) {
  __shared__ double S[BlkStore<double, GK_POTRF_PB_F64>::SIZE];
  potrf_blocked<double, GK_POTRF_PB_F64, 1024>(BlkStore<double, GK_POTRF_PB_F64>{S}, A, lda, X, ldx, info, gbase);
}

template <typename T> struct Potrf128;
template <> struct Potrf128<float> {
  static void launch(float* A, int64_t lda, float* X, int64_t ldx, int32_t* info, int64_t g, hipStream_t st) {
    hipLaunchKernelGGL(potrf128b_f32_kernel, dim3(1), dim3(GK_POTRF_NT_F32), 0, st, A, lda, X, ldx, info, g);
  }
};
template <> struct Potrf128<double> {
  static void launch(double* A, int64_t lda, double* X, int64_t ldx, int32_t* info, int64_t g, hipStream_t st) {
    hipLaunchKernelGGL(potrf128b_f64_kernel, dim3(1), dim3(1024), 0, st, A, lda, X, ldx, info, g);
  }
};

template <typename T>
__global__ void zero_block_kernel(T* X, int64_t ldx, int rows, int cols) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < rows * cols) X[(int64_t)(e / cols) * ldx + e % cols] = T(0);
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

// Factor + invert one 256-wide diagonal block from two 128-wide halves.
// Cholesky: L = [L11 0; L21 L22], L^-1 = [X11 0; -X22 L21 X11  X22].  L21 is
// formed in T1; with copy_l21 false the caller copies it into A later (no
// later step of the factorization reads it from A).  The upper-right quadrant
// of Dinv is zeroed once per factorization by the caller.
template <typename T>
int diag256_chol(T* A, int64_t lda, int64_t k0, T* dinv, int32_t* dinfo, T* T1, T* T2, hipStream_t st,
                 bool copy_l21 = true) {
  T* A11 = A + k0 * lda + k0;
  T* A21 = A + (k0 + DB) * lda + k0;
  T* A22 = A21 + DB;
  T* D = dinv + k0 * NB;                 // 256 x 256, ld NB
  Potrf128<T>::launch(A11, lda, D, NB, dinfo, k0, st);
  GK_TRY(gemm_launch<T>(DB, DB, DB, A21, lda, D, NB, true, T(1), T(0), T1, DB, false, st));        // L21
  if (copy_l21)
    hipLaunchKernelGGL(copy_block_kernel<T>, dim3(DB * DB / 256), dim3(256), 0, st, T1, (int64_t)DB, A21,
                       lda, (int64_t)DB, (int64_t)DB);
  GK_TRY(gemm_launch<T>(DB, DB, DB, T1, DB, T1, DB, true, T(-1), T(1), A22, lda, false, st));      // syrk
  Potrf128<T>::launch(A22, lda, D + DB * NB + DB, NB, dinfo, k0 + DB, st);
  GK_TRY(gemm_launch<T>(DB, DB, DB, T1, DB, D, NB, false, T(1), T(0), T2, DB, false, st));         // L21 X11
  GK_TRY(gemm_launch<T>(DB, DB, DB, D + DB * NB + DB, NB, T2, DB, false, T(-1), T(0), D + DB * NB, NB,
                        false, st));
  return GK_OK;
}

// LU: L^-1 = [Li11 0; -Li22 L21 Li11  Li22], U^-1 = [Ui11  -Ui11 U12 Ui22; 0 Ui22].
int diag256_lu(double* A, int64_t lda, int64_t k0, double* dinv, int32_t* dinfo, double* T, hipStream_t st) {
  double* A11 = A + k0 * lda + k0;
  double* A12 = A11 + DB;
  double* A21 = A + (k0 + DB) * lda + k0;
  double* A22 = A21 + DB;
  double* Li = dinv + k0 * NB;
  double* Ui = dinv + lda * NB + k0 * NB;
  double *T1 = T, *T2 = T + DB * DB, *T3 = T + 2 * DB * DB, *T4 = T + 3 * DB * DB;
  hipLaunchKernelGGL(getrf128_kernel, dim3(1), dim3(1024), 0, st, A11, lda, Li, Ui, (int64_t)NB, dinfo, k0);
  GK_TRY(gemm_launch<double>(DB, DB, DB, Li, NB, A12, lda, false, 1.0, 0.0, T1, DB, false, st));    // U12
  GK_TRY(gemm_launch<double>(DB, DB, DB, A21, lda, Ui, NB, false, 1.0, 0.0, T2, DB, false, st));    // L21
  hipLaunchKernelGGL(copy_block_kernel<double>, dim3(DB * DB / 256), dim3(256), 0, st, T1, (int64_t)DB, A12,
                     lda, (int64_t)DB, (int64_t)DB);
  hipLaunchKernelGGL(copy_block_kernel<double>, dim3(DB * DB / 256), dim3(256), 0, st, T2, (int64_t)DB, A21,
                     lda, (int64_t)DB, (int64_t)DB);
  GK_TRY(gemm_launch<double>(DB, DB, DB, T2, DB, T1, DB, false, -1.0, 1.0, A22, lda, false, st));
  hipLaunchKernelGGL(getrf128_kernel, dim3(1), dim3(1024), 0, st, A22, lda, Li + DB * NB + DB,
                     Ui + DB * NB + DB, (int64_t)NB, dinfo, k0 + DB);
  GK_TRY(gemm_launch<double>(DB, DB, DB, T2, DB, Li, NB, false, 1.0, 0.0, T3, DB, false, st));      // L21 Li11
  GK_TRY(gemm_launch<double>(DB, DB, DB, Li + DB * NB + DB, NB, T3, DB, false, -1.0, 0.0, Li + DB * NB, NB,
                             false, st));
  GK_TRY(gemm_launch<double>(DB, DB, DB, T1, DB, Ui + DB * NB + DB, NB, false, 1.0, 0.0, T4, DB, false, st));
  GK_TRY(gemm_launch<double>(DB, DB, DB, Ui, NB, T4, DB, false, -1.0, 0.0, Ui + DB, NB, false, st));
  hipLaunchKernelGGL(zero_block_kernel<double>, dim3(DB * DB / 256), dim3(256), 0, st, Li + DB, (int64_t)NB, DB, DB);
  hipLaunchKernelGGL(zero_block_kernel<double>, dim3(DB * DB / 256), dim3(256), 0, st, Ui + DB * NB, (int64_t)NB,
                     DB, DB);
  return GK_OK;
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

// Recursive Cholesky (lower, row-major, in place) of the trailing principal
// block of nblk 256-blocks at k0:
//   chol(A11); A21 <- A21 L11^-T (recursive TRSM); A22 -= A21 A21^T; chol(A22).
// Almost all flops land in GEMMs whose K is half the current order, where the
// MFMA GEMM runs far closer to its peak than with 256-deep rank updates.
template <typename T>
struct CholRec {
  gk_handle h;
  T* A;
  int64_t lda;
  T* dinv;
  int32_t* dinfo;
  T* W;            // lda x NB panel buffer
  T* Tb;           // 2 DB x DB scratch of the diagonal-block kernel
  hipStream_t st;
  int64_t done = 0;      // diagonal blocks factored so far
  bool failed = false;

  // X <- X L^-T for X = A[r0:r0+m, c0:c0+nbk*NB], L = A[c0:, c0:] (nbk blocks)
  int trsm(int64_t r0, int64_t m, int64_t c0, int64_t nbk) {
    if (nbk == 1) {
      GK_TRY(gemm_launch<T>(m, NB, NB, A + r0 * lda + c0, lda, dinv + c0 * NB, NB, true, T(1), T(0), W, NB, false,
                            st));
      hipLaunchKernelGGL(copy_block_kernel<T>, dim3((unsigned)gk_ceil_div(m * NB, 256)), dim3(256), 0, st, W,
                         (int64_t)NB, A + r0 * lda + c0, lda, m, (int64_t)NB);
      return GK_OK;
    }
    const int64_t a = nbk / 2, b = nbk - a;
    const int64_t ca = c0, cb = c0 + a * NB;
    GK_TRY(trsm(r0, m, ca, a));
    // Xb -= Xa Lb^T, Lb = A[cb:cb+b NB, ca:ca+a NB]
    GK_TRY(gemm_launch<T>(m, b * NB, a * NB, A + r0 * lda + ca, lda, A + cb * lda + ca, lda, true, T(-1), T(1),
                          A + r0 * lda + cb, lda, false, st));
    return trsm(r0, m, cb, b);
  }

  int chol(int64_t k0, int64_t nblk) {
    if (failed) return GK_OK;
    if (nblk == 1) {
      GK_TRY(diag256_chol<T>(A, lda, k0, dinv, dinfo, Tb, Tb + DB * DB, st));
      if (++done % 8 == 0) {          // stop early once a pivot has failed
        int32_t hinfo = 0;
        GK_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        GK_HIP(hipStreamSynchronize(st));
        failed = hinfo != 0;
      }
      return GK_OK;
    }
    const int64_t n1 = nblk / 2, n2 = nblk - n1;
    const int64_t k1 = k0 + n1 * NB;
    GK_TRY(chol(k0, n1));
    if (failed) return GK_OK;
    GK_TRY(trsm(k1, n2 * NB, k0, n1));
    GK_TRY(gemm_launch<T>(n2 * NB, n2 * NB, n1 * NB, A + k1 * lda + k0, lda, A + k1 * lda + k0, lda, true, T(-1),
                          T(1), A + k1 * lda + k1, lda, true, st));
    return chol(k1, n2);
  }
};

// One FP32 panel of the look-ahead factorization, on the side stream.  The
// part the next trailing update waits for: factor and invert diagonal block kb
// and solve the panel, W = A21 Dinv_kb^T.
// With P given, the panel is also split into its tiled fp16 hi / lo form for
// the split-fp16 updates.
static int chol32_panel(float* A, int64_t lda, int64_t kb, float* dinv, int32_t* dinfo, float* W, float* Tb,
                        hipStream_t s, _Float16* P = nullptr) {
  const int64_t k0 = kb * NB, rem = lda - k0 - NB;
  GK_TRY(diag256_chol<float>(A, lda, k0, dinv, dinfo, Tb, Tb + DB * DB, s, false));
  if (rem > 0) {
    GK_TRY(gemm_launch<float>(rem, NB, NB, A + (k0 + NB) * lda + k0, lda, dinv + k0 * NB, NB, true, 1.0f, 0.0f, W,
                              NB, false, s));
    if (P) GK_TRY(split_h3(W, rem, P, s));
  }
  return GK_OK;
}

// The split16 panel, as a blocked solve that leaves the inverse's
// off-diagonal quadrant for later: with the 256-block halves 1, 2 and the
// rows below them b,
//   X11 = L11^-1;  [L21; W1] = [A21; Ab1] X11^T  (one product);
//   [A22; Ab2] -= [L21; W1] L21^T;  X22 = L22^-1;  W2 = Ab2 X22^T,
// so W = [W1 W2] = Ab Dinv^T with Dinv = [X11 0; -X22 L21 X11  X22], whose
// lower-left quadrant (needed only by the triangular solves) is formed for
// all panels at the end (dinv_lower_left).  The products run in place in A
// on tiles spanning their 128 columns, and W is split from A: six kernels in
// sequence and no copies.
static int chol32_panel_split(float* A, int64_t lda, int64_t kb, float* dinv, int32_t* dinfo, hipStream_t s,
                              _Float16* P) {
  const int64_t k0 = kb * NB, rem = lda - k0 - NB;
  float* D = dinv + k0 * NB;                                     // 256 x 256, ld NB
  float* C1 = A + (k0 + DB) * lda + k0;                          // [A21; Ab1]
  Potrf128<float>::launch(A + k0 * lda + k0, lda, D, NB, dinfo, k0, s);
  GK_TRY(gemm_launch<float>(DB + rem, DB, DB, C1, lda, D, NB, true, 1.0f, 0.0f, C1, lda, false, s, 1, 0, 0, 0,
                            true));                              // [L21; W1], in place
  GK_TRY(gemm_launch<float>(DB + rem, DB, DB, C1, lda, C1, lda, true, -1.0f, 1.0f, C1 + DB, lda, false, s));
  Potrf128<float>::launch(C1 + DB, lda, D + DB * NB + DB, NB, dinfo, k0 + DB, s);
  if (rem > 0) {
    float* C2 = A + (k0 + NB) * lda + k0 + DB;                   // Ab2
    GK_TRY(gemm_launch<float>(rem, DB, DB, C2, lda, D + DB * NB + DB, NB, true, 1.0f, 0.0f, C2, lda, false, s, 1, 0, 0,
                              0, true));                         // W2, in place
    GK_TRY(split_h3(A + (k0 + NB) * lda + k0, rem, P, s, lda));
  }
  return GK_OK;
}

// Lower-left quadrants of the diagonal blocks' inverses, all panels at once:
// Dinv21 = -X22 (L21 X11), two batched 128^3 products per panel (T: nb x DB x
// DB scratch).
static int dinv_lower_left(const float* A, int64_t lda, float* dinv, float* T, hipStream_t s) {
  const int64_t nb = lda / NB, sD = (int64_t)NB * NB;
  if (nb == 0) return GK_OK;
  GK_TRY(gemm_launch<float>(DB, DB, DB, A + DB * lda, lda, dinv, NB, false, 1.0f, 0.0f, T, DB, false, s, nb,
                            (int64_t)NB * lda + NB, sD, (int64_t)DB * DB));
  return gemm_launch<float>(DB, DB, DB, dinv + DB * NB + DB, NB, T, DB, false, -1.0f, 0.0f, dinv + DB * NB, NB, false,
                            s, nb, sD, (int64_t)DB * DB, sD);
}

// Chain planes of the triangular solves (planes of lda x 256, D =
// GK_CHAIN_DEPTH), for all block rows kb at once from the finished factor:
// plane D+1-d: G_d = Dinv_kb L_{kb,kb-d} (kb >= d); plane 2D+1-d: H_d =
// L_{kb+d,kb} Dinv_kb (kb + d < nb).  One batched launch per plane; every
// product is 256 x 256 x 256 and the block addresses advance by a fixed stride
// per block row (NB x NB in dinv, NB rows + NB columns in A).
template <typename T>
static int chain_planes(const T* A, int64_t lda, T* dinv, hipStream_t s) {
  constexpr int D = GK_CHAIN_DEPTH;
  const int64_t nb = lda / NB, sD = (int64_t)NB * NB, sL = (int64_t)NB * lda + NB;
  for (int d = 1; d <= D; ++d) {
    const int64_t cnt = nb - d;
    if (cnt <= 0) continue;
    T* G = dinv + (D + 1 - d) * lda * NB;
    T* H = dinv + (2 * D + 1 - d) * lda * NB;
    GK_TRY(gemm_launch<T>(NB, NB, NB, dinv + d * sD, NB, A + d * NB * lda, lda, false, T(1), T(0), G + d * sD, NB,
                          false, s, cnt, sD, sL, sD));
    GK_TRY(gemm_launch<T>(NB, NB, NB, A + d * NB * lda, lda, dinv, NB, false, T(1), T(0), H, NB, false, s, cnt, sL,
                          sD, sD));
  }
  return GK_OK;
}

// The rest of the panel, off the sequential path: W back into A21 (columns no
// other kernel touches until the end) and the diagonal block's L21 (from T1,
// before the next panel reuses it).
static int chol32_panel_post(float* A, int64_t lda, int64_t kb, const float* W, const float* T1, hipStream_t s) {
  const int64_t k0 = kb * NB, rem = lda - k0 - NB;
  hipLaunchKernelGGL(copy_block_kernel<float>, dim3(DB * DB / 256), dim3(256), 0, s, T1, (int64_t)DB,
                     A + (k0 + DB) * lda + k0, lda, (int64_t)DB, (int64_t)DB);
  if (rem > 0)
    hipLaunchKernelGGL(copy_block_kernel<float>, dim3((unsigned)gk_ceil_div(rem * NB, 256)), dim3(256), 0, s, W,
                       (int64_t)NB, A + (k0 + NB) * lda + k0, lda, rem, (int64_t)NB);
  return GK_OK;
}

// Paired rank-512 trailing updates of the split16 factorization (on unless
// GK_FACTOR_PAIRS=0).
static bool factor_pairs() {
  static const bool v = [] { const char* e = std::getenv("GK_FACTOR_PAIRS"); return !(e && *e == '0'); }();
  return v;
}

// split16 (FP32 only): rank-256 updates with the split-fp16 MFMA product; for
// matrices whose factor entries stay within [-16, 16] (unit-diagonal ones).
template <typename T>
static int cholesky_impl(gk_handle h, T* A, int64_t n, int64_t lda, T* dinv, int32_t* info, hipStream_t st,
                         bool split16 = false) {
  GK_REQUIRE(A && dinv, GK_ERR_ARG, "gk_cholesky: A and dinv required");
  GK_REQUIRE(lda % NB == 0 && n <= lda && n >= 0, GK_ERR_ARG,
             "gk_cholesky: lda must be a multiple of %d and >= n", NB);
  const int64_t nb = lda / NB;
  split16 = split16 && sizeof(T) == 4;
  void* ws = nullptr;
  const size_t split_bytes = split16 ? 4 * 2 * (size_t)lda * NB * sizeof(_Float16) : 0;
  GK_TRY(gk_workspace(h, WS_FACTOR, 256 + (2 * (size_t)lda * NB + 2 * DB * DB) * sizeof(T) + split_bytes, &ws));
  int32_t* dinfo = static_cast<int32_t*>(ws);
  int32_t* tile_ctr = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + 64);   // 8 XCD tile counters
  T* W = reinterpret_cast<T*>(static_cast<char*>(ws) + 256);       // two panel buffers (FP32 look-ahead)
  T* Tb = W + 2 * (size_t)lda * NB;
  // four tiled split panels of lda x NB x 2 halves
  _Float16* Hs = split16 ? reinterpret_cast<_Float16*>(Tb + 2 * DB * DB) : nullptr;
  GK_HIP(hipMemsetAsync(ws, 0, 256, st));                              // info word and tile counters
  GK_HIP(hipMemsetAsync(dinv, 0, (size_t)lda * NB * sizeof(T), st));   // Dinv plane: upper-right quadrants stay 0
  if (n < lda) {
    const int64_t tot = (lda - n) * lda + n * (lda - n);
    hipLaunchKernelGGL(pad_identity_kernel<T>, dim3((unsigned)gk_ceil_div(tot, 256)), dim3(256), 0, st, A, lda, n);
  }
  if (sizeof(T) == 8) {
    if (nb > 0) {
      CholRec<T> R{h, A, lda, dinv, dinfo, W, Tb, st};
      GK_TRY(R.chol(0, nb));
    }
  } else {
    // FP32: right-looking order.  Each rank-256 update rounds the trailing
    // block relative to the Schur complement as it shrinks, which a one-shot
    // large-K update does not; on the kernel matrices of this method the
    // recursive order needs a ~60x larger diagonal shift to factor in FP32.
    // Look-ahead: the caller's stream first updates the next panel's columns
    // and signals; the side stream factors that panel (diagonal block, panel
    // solve) and signals, then copies it back while the caller's stream
    // applies the rest of the trailing update.  W alternates
    // between two buffers: the side stream writes panel kb+2 only after the
    // caller's stream has passed the update that last read it.
    // split16: the panels go in pairs (kb - 1, kb), kb odd: the trailing block
    // receives both in one rank-512 update after kb, which halves the passes
    // over it.  The next two column blocks take the pair's product first, so
    // that the side stream factors both panels of the next pair (the second
    // after applying the first to it) beside that update.  Four split
    // buffers: the pair being applied and the two panels factored meanwhile.
    if constexpr (sizeof(T) == 4) {
      hipStream_t s2;
      hipEvent_t* ev;
      GK_TRY(gk_side_stream(h, &s2, &ev));
      T* Wb[2] = {W, W + (size_t)lda * NB};
      _Float16* Hb[4];
      for (int b = 0; b < 4; ++b) Hb[b] = Hs ? Hs + 2 * b * (size_t)lda * NB : nullptr;
      auto sp = [&](int64_t kb) { return Hb[kb % 4]; };
      const int64_t RB = NB * 2 * H3_K;         // halves of NB rows of a split panel
      const bool pairs = split16 && factor_pairs();
      // panel kb on the side stream; ev[1] marks it factored (split16: its
      // split form in sp(kb)); without split16 the copies back into A follow
      auto panel = [&](int64_t kb) -> int {
        T* V = Wb[kb & 1];
        if (split16) GK_TRY(chol32_panel_split(A, lda, kb, dinv, dinfo, s2, sp(kb)));
        else GK_TRY(chol32_panel(A, lda, kb, dinv, dinfo, V, Tb, s2));
        GK_HIP(hipEventRecord(ev[1], s2));
        if (split16) return GK_OK;
        return chol32_panel_post(A, lda, kb, V, Tb, s2);
      };
      GK_HIP(hipEventRecord(ev[0], st));
      GK_HIP(hipStreamWaitEvent(s2, ev[0], 0));
      if (nb > 0) GK_TRY(panel(0));
      for (int64_t kb = 0; kb + 1 < nb; ++kb) {
        const int64_t k0 = kb * NB, rem = lda - k0 - NB;
        float* Cn = A + (k0 + NB) * lda + k0 + NB;          // column block kb + 1, rows below k0 + NB
        if (pairs && !(kb & 1)) {
          // kb opens a pair.  Column block kb + 1 holds every earlier panel's
          // product (the last pair's update of two column blocks ran before
          // panel kb was factored); the side stream applies panel kb to it
          // and factors it, beside the trailing update of the last pair.
          GK_TRY(syrk_h3_launch(rem, NB, sp(kb), sp(kb), Cn, lda, false, s2));
          GK_TRY(panel(kb + 1));
        } else if (pairs) {
          // kb closes the pair (kb - 1, kb): its product with the next two
          // column blocks (rank 512; the block above the diagonal of the
          // second is computed and never read), then the next panel on the
          // side stream, and the rest of the trailing block (column blocks
          // from kb + 3) on the caller's stream.
          GK_HIP(hipStreamWaitEvent(st, ev[1], 0));         // panel kb factored
          GK_TRY(syrk_h3_launch(rem, std::min<int64_t>(2 * NB, rem), sp(kb), sp(kb), Cn, lda, false, st, nullptr,
                                sp(kb - 1) + RB, sp(kb - 1) + RB));
          GK_HIP(hipEventRecord(ev[0], st));
          GK_HIP(hipStreamWaitEvent(s2, ev[0], 0));
          GK_TRY(panel(kb + 1));
          if (rem > 2 * NB)
            GK_TRY(syrk_h3_launch(rem - 2 * NB, rem - 2 * NB, sp(kb) + 2 * RB, sp(kb) + 2 * RB,
                                  A + (k0 + 3 * NB) * lda + k0 + 3 * NB, lda, true, st, tile_ctr,
                                  sp(kb - 1) + 3 * RB, sp(kb - 1) + 3 * RB));
        } else {
          T* Wk = Wb[kb & 1];
          GK_HIP(hipStreamWaitEvent(st, ev[1], 0));         // panel kb factored, Wk = its rows below
          // next panel's columns: A[k0+NB:, k0+NB : k0+2NB] -= Wk Wk[0:NB]^T
          if (split16)
            GK_TRY(syrk_h3_launch(rem, NB, sp(kb), sp(kb), Cn, lda, false, st));
          else
            GK_TRY(gemm_launch<T>(rem, NB, NB, Wk, NB, Wk, NB, true, T(-1), T(1), A + (k0 + NB) * lda + k0 + NB, lda,
                                  false, st));
          GK_HIP(hipEventRecord(ev[0], st));
          GK_HIP(hipStreamWaitEvent(s2, ev[0], 0));
          GK_TRY(panel(kb + 1));
          if (rem > NB) {                                   // the rest of the trailing update (lower)
            if (split16)
              GK_TRY(syrk_h3_launch(rem - NB, rem - NB, sp(kb) + RB, sp(kb) + RB,
                                    A + (k0 + 2 * NB) * lda + k0 + 2 * NB, lda, true, st, tile_ctr));
            else
              GK_TRY(gemm_launch<T>(rem - NB, rem - NB, NB, Wk + NB * NB, NB, Wk + NB * NB, NB, true, T(-1), T(1),
                                    A + (k0 + 2 * NB) * lda + k0 + 2 * NB, lda, true, st));
          }
        }
        if (kb % 16 == 15) {                                // stop early once a pivot has failed
          int32_t hinfo = 0;
          GK_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s2));
          GK_HIP(hipStreamSynchronize(s2));
          if (hinfo != 0) break;
        }
      }
      GK_HIP(hipEventRecord(ev[2], s2));
      GK_HIP(hipStreamWaitEvent(st, ev[2], 0));
      if (split16) GK_TRY(dinv_lower_left(A, lda, dinv, Wb[0], st));
    }
  }
  GK_TRY(chain_planes<T>(A, lda, dinv, st));   // chain planes of the triangular solves
  GK_HIP(hipGetLastError());
  int32_t hinfo = 0;
  GK_HIP(hipMemcpyAsync(&hinfo, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  if (info) *info = hinfo;
  return hinfo == 0 ? GK_OK : GK_ERR_NOT_SPD;
}

extern "C" int gk_cholesky_f32_split16(gk_handle h, float* A, int64_t n, int64_t lda, float* dinv, int32_t* info,
                                       gk_stream stream) {
  return cholesky_impl<float>(h, A, n, lda, dinv, info, (hipStream_t)stream, true);
}

extern "C" int gk_test_cholesky_f32_split16(gk_handle h, float* A, int64_t n, int64_t lda, float* dinv,
                                            int32_t* info, gk_stream stream) {
  GK_REQUIRE(h && A && dinv, GK_ERR_ARG, "gk_test_cholesky_f32_split16: NULL argument");
  return cholesky_impl<float>(h, A, n, lda, dinv, info, (hipStream_t)stream, true);
}

extern "C" int gk_test_syrk_split16(gk_handle h, int64_t m, int64_t n, const float* A, const float* B,
                                    const float* A2, const float* B2, float* C, int64_t ldc, int32_t lower,
                                    gk_stream stream) {
  GK_REQUIRE(h && A && B && C && !A2 == !B2, GK_ERR_ARG, "gk_test_syrk_split16: NULL argument");
  GK_REQUIRE(m > 0 && n > 0 && m % 128 == 0 && n % 128 == 0 && ldc >= n && (!lower || m == n), GK_ERR_ARG,
             "gk_test_syrk_split16: m, n must be positive multiples of 128 (m == n when lower)");
  hipStream_t st = (hipStream_t)stream;
  void* ws = nullptr;
  const int np = A2 ? 2 : 1;
  GK_TRY(gk_workspace(h, WS_FACTOR, 256 + 2 * np * (size_t)(m + n) * NB * sizeof(_Float16), &ws));
  int32_t* ctr = static_cast<int32_t*>(ws);
  _Float16* Ap = reinterpret_cast<_Float16*>(static_cast<char*>(ws) + 256);
  _Float16* Bp = Ap + 2 * (size_t)m * NB;
  _Float16* Ap2 = Bp + 2 * (size_t)n * NB;
  _Float16* Bp2 = Ap2 + 2 * (size_t)m * NB;
  GK_HIP(hipMemsetAsync(ctr, 0, 256, st));
  GK_TRY(split_h3(A, m, Ap, st));
  GK_TRY(split_h3(B, n, Bp, st));
  if (np == 2) {
    GK_TRY(split_h3(A2, m, Ap2, st));
    GK_TRY(split_h3(B2, n, Bp2, st));
  }
  return syrk_h3_launch(m, n, Ap, Bp, C, ldc, lower != 0, st, ctr, np == 2 ? Ap2 : nullptr, np == 2 ? Bp2 : nullptr);
}

extern "C" int gk_cholesky_f32(gk_handle h, float* A, int64_t n, int64_t lda, float* dinv, int32_t* info,
                               gk_stream stream) {
  return cholesky_impl<float>(h, A, n, lda, dinv, info, (hipStream_t)stream);
}
extern "C" int gk_cholesky_f64(gk_handle h, double* A, int64_t n, int64_t lda, double* dinv, int32_t* info,
                               gk_stream stream) {
  return cholesky_impl<double>(h, A, n, lda, dinv, info, (hipStream_t)stream);
}
// The public solves read the sync-free kernels' error word before returning
// (one stream synchronisation): a wait that reached its spin limit leaves b
// unusable and is reported here instead of by a later, unrelated solve.
extern "C" int gk_cholesky_solve_f32(gk_handle h, const float* L, const float* dinv, int64_t lda, float* b,
                                     gk_stream stream) {
  GK_TRY(gk_trsv_chol_impl<float>(h, L, dinv, lda, b, (hipStream_t)stream));
  return gk_trsv_check(h, (hipStream_t)stream);
}
extern "C" int gk_cholesky_solve_f64(gk_handle h, const double* L, const double* dinv, int64_t lda, double* b,
                                     gk_stream stream) {
  GK_TRY(gk_trsv_chol_impl<double>(h, L, dinv, lda, b, (hipStream_t)stream));
  return gk_trsv_check(h, (hipStream_t)stream);
}

extern "C" int gk_lu_f64(gk_handle h, double* A, int64_t n, int64_t lda, double* dinv, int32_t* info,
                         gk_stream stream) {
  hipStream_t st = (hipStream_t)stream;
  GK_REQUIRE(A && dinv, GK_ERR_ARG, "gk_lu_f64: A and dinv required");
  GK_REQUIRE(lda % NB == 0 && n <= lda && n >= 0, GK_ERR_ARG,
             "gk_lu_f64: lda must be a multiple of %d and >= n", NB);
  const int64_t nb = lda / NB;
  void* ws = nullptr;
  GK_TRY(gk_workspace(h, WS_FACTOR, 256 + (2 * (size_t)lda * NB + 4 * DB * DB) * sizeof(double), &ws));
  int32_t* dinfo = static_cast<int32_t*>(ws);
  double* W1 = reinterpret_cast<double*>(static_cast<char*>(ws) + 256);   // U12: NB x rem
  double* W2 = W1 + (size_t)lda * NB;                                        // L21: rem x NB
  double* T = W2 + (size_t)lda * NB;                                         // 4 x DB x DB
  GK_HIP(hipMemsetAsync(dinfo, 0, sizeof(int32_t), st));
  if (n < lda) {
    const int64_t tot = (lda - n) * lda + n * (lda - n);
    hipLaunchKernelGGL(pad_identity_kernel<double>, dim3((unsigned)gk_ceil_div(tot, 256)), dim3(256), 0, st,
                       A, lda, n);
  }
  for (int64_t kb = 0; kb < nb; ++kb) {
    const int64_t k0 = kb * NB;
    GK_TRY(diag256_lu(A, lda, k0, dinv, dinfo, T, st));
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
