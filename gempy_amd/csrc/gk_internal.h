// Internal definitions shared by the gk HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <vector>
#include "../../include/gk.h"

void gk_set_error(const char* fmt, ...);

#define GK_HIP(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      gk_set_error("%s:%d %s -> %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return GK_ERR_HIP;                                                           \
    }                                                                              \
  } while (0)

#define GK_REQUIRE(cond, code, ...)                                                \
  do {                                                                             \
    if (!(cond)) {                                                                 \
      gk_set_error(__VA_ARGS__);                                                   \
      return code;                                                                 \
    }                                                                              \
  } while (0)

#define GK_TRY(x)                                                                  \
  do {                                                                             \
    int rc_ = (x);                                                                 \
    if (rc_ != GK_OK) return rc_;                                                  \
  } while (0)

// Per-device context: growable device workspaces owned by the handle, one per
// purpose so that nested calls (the solver calling the matvec calling the
// evaluator) never alias each other's buffers.
enum gk_ws_slot {
  WS_EVAL_SRC = 0,   // packed sources of one evaluation
  WS_EVAL_PART,      // per-chunk partial sums of a split evaluation
  WS_MATVEC,         // data points and field values of the matrix-free operator
  WS_FACTOR,         // panel buffers of a factorization
  WS_SOLVER,         // Krylov basis, preconditioner blocks
  WS_MATRIX,         // assembled matrices of gk_solve
  WS_MATRIX2,
  WS_ADJOINT,        // vectors of the adjoint
  WS_SMALL,          // small scratch (bounding boxes, counts)
  WS_TRSV,           // work tables, readiness flags and partial sums of the triangular solves
  WS_ASSEMBLE,       // reference-point tables of the assembly
  WS_MESH,           // case tables, edge flags and scans of the isosurface extraction
  WS_LANCZOS,        // Lanczos basis of the condition-number estimate
  WS_KNN,            // cells and sorted points of the grid neighbour search
  WS_COUNT
};
// Independent sets of tables, flags and buffers of the sync-free triangular
// solves in WS_TRSV: up to this many solves run at once, each on its own
// stream with its own set.
constexpr int GK_TRSV_CTX = 3;

struct gk_context {
  int device = 0;
  void* ws[WS_COUNT] = {};
  size_t ws_bytes[WS_COUNT] = {};
  void* solver_state = nullptr;            // factorization kept by the last gk_solve
  void (*solver_free)(void*) = nullptr;
  // state of the sync-free triangular solves (GK_TRSV_CTX sets in WS_TRSV)
  int trsv_nb = 0, trsv_slots = 0, trsv_nfwd = 0, trsv_nbwd = 0;
  unsigned trsv_epoch[GK_TRSV_CTX] = {};
  unsigned long long trsv_tickets[GK_TRSV_CTX] = {};
  // side stream (default priority) and events of the look-ahead factorization
  hipStream_t side = nullptr;
  hipEvent_t side_ev[3] = {};
  // auxiliary stream (default priority) for set-up work that overlaps a factorization
  hipStream_t aux = nullptr;
  hipEvent_t aux_ev[2] = {};
  // device address of gk_eval_noalloc's file-scope scratch array (this device's copy)
  char* eval_scratch = nullptr;
  // where the marching-cubes case tables were last uploaded (inside WS_MESH)
  void* mesh_tables = nullptr;
};

// Takes the device address of the evaluation scratch array for h's device
// (called by gk_create with that device current).
int gk_eval_scratch_init(gk_handle h);
// Neighbours per data point of the FP32 form's exact local inverse (the
// learned network keeps GK_SPAI_K); gk_learned.hip.
#define GK_LOCAL_K GK_LOCAL_NEIGHBOURS
extern "C" int gk_knn_local(gk_handle h, const gk_system* sys, const double* box, int32_t* idx, float* hk,
                            gk_stream stream);
extern "C" int gk_local_inverse_local(gk_handle h, const gk_system* sys, const int32_t* idx, const double* dscale,
                                      float* Y, gk_stream stream);
extern "C" int gk_spai_apply_k(const gk_system* sys, int32_t k, const int32_t* idx, const float* Y, const double* u,
                               double* y, hipStream_t st);

// FP32 Cholesky whose rank-256 updates run on the split-fp16 MFMA product
// (gk_dense.hip); for unit-diagonal matrices (factor entries within [-16, 16]).
extern "C" int gk_cholesky_f32_split16(gk_handle h, float* A, int64_t n, int64_t lda, float* dinv, int32_t* info,
                                       gk_stream stream);

// Depth of the chain step of the triangular solves: it multiplies the planes
// Dinv_k L_{k,k-d} (forward) and L_{k+d,k} Dinv_k (backward) for d = 1..depth;
// the factorization leaves 1 + 2 depth planes (GK_CHOL_PLANES).
constexpr int GK_CHAIN_DEPTH = 2;
static_assert(1 + 2 * GK_CHAIN_DEPTH == GK_CHOL_PLANES, "plane count of the Cholesky output");

// The handle's side stream and its three events, created on first use.
int gk_side_stream(gk_handle h, hipStream_t* s2, hipEvent_t** ev);
// The handle's auxiliary stream and its two events, created on first use.
int gk_aux_stream(gk_handle h, hipStream_t* s3, hipEvent_t** ev);

int gk_workspace(gk_handle h, int slot, size_t bytes, void** out);
// L^-T L^-1 b in place (one launch per direction) with the solve set ctx
// (0 .. GK_TRSV_CTX-1; solves running at once on different streams take
// different sets); gk_trsv_check reports a readiness wait that reached its
// spin limit since the last check.
template <typename T>
int gk_trsv_chol_impl(gk_handle h, const T* L, const T* dinv, int64_t lda, T* b, hipStream_t st, int ctx = 0);
int gk_trsv_check(gk_handle h, hipStream_t st);

// Optional per-launch timing of the dense kernels (gk_profile): HIP events are
// recorded around each launch of an instrumented kernel while enabled.  Kinds:
// GEMM instantiations 4*(double) + 2*trans_b + lower (0..7), triangular solves
// 8 (FP32) and 9 (FP64); `work` is the launch's algorithmic flops (GEMM) or
// bytes of the factor read (triangular solves).
struct gk_prof_rec {
  hipEvent_t a, b;
  int kind;
  double work;
};
bool gk_prof_begin(hipStream_t st, hipEvent_t* a);
void gk_prof_end(hipStream_t st, hipEvent_t a, int kind, double work);

static inline int64_t gk_ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
static inline int64_t gk_round_up(int64_t a, int64_t b) { return gk_ceil_div(a, b) * b; }

// Logical tile of hardware block bid of a grid of nblk: blocks are dispatched
// to the 8 XCDs round-robin (XCD bid % 8), so XCD x runs ceil((nblk - x) / 8)
// of them; it is given the contiguous range of logical tiles that starts at
// x q + min(x, r) (q, r = nblk / 8, nblk % 8), so tiles that share operands
// share that XCD's L2.  A bijection for every nblk.
__device__ __forceinline__ int64_t gk_xcd_tile(int64_t bid, int64_t nblk) {
  const int64_t q = nblk / 8, r = nblk % 8, x = bid % 8;
  return x * q + (x < r ? x : r) + bid / 8;
}

// ------------------------------------------------------------------ covariance
// All device code works in range-normalised coordinates (x / a), so the cubic
// covariance is a polynomial in t = rho^2 and rho.
//   C/c_o            = 1 - 7 t + rho t (35/4 - 7/2 t + 3/4 t^2)
//   (C'/r) a^2/c_o   = -14 + rho (105/4 + t (-35/2 + 21/4 t))
//   (C''-C'/r) a^2/c_o = 105/4 rho (1 - t)^2
template <typename T>
__device__ __forceinline__ T gk_cov_poly(T t, T rho) {
  T q = fma(t, T(0.75), T(-3.5));
  q = fma(q, t, T(8.75));
  q = fma(q, rho, T(-7.0));
  return fma(q, t, T(1.0));
}

template <typename T>
__device__ __forceinline__ T gk_dcov_poly(T t, T rho) {
  T p = fma(t, T(5.25), T(-17.5));
  p = fma(p, t, T(26.25));
  return fma(p, rho, T(-14.0));
}

// 1/sqrt(t) to working precision; t is clamped away from 0 (rho = t * y = 0 there).
__device__ __forceinline__ float gk_rsqrt(float t) {
  return __builtin_amdgcn_rsqf(fmaxf(t, 1e-30f));
}
__device__ __forceinline__ double gk_rsqrt(double t) {
  t = fmax(t, 1e-280);
  double y = __builtin_amdgcn_rsq(t);
  // one Newton step: y <- y (3 - t y^2) / 2  (relative error squares)
  double r = fma(-t * y, y, 1.0);
  return fma(0.5 * y, r, y);
}

// Hessians (xx, yy, zz, xy, xz, yz per point, FP64) of the system's field for
// `weights`, or of sum_s w_s C(|x - s|) over explicit sources, at n points;
// part: scratch (>= 48 n bytes) for the source-split partial sums.
int gk_hessian_system(gk_handle h, const gk_system* sys, const double* weights, int64_t n, const double* xyz,
                      double* hess6, double* part, size_t part_cap, hipStream_t st);
int gk_hessian_sources(gk_handle h, double range, double c_o, const double* center, int64_t n_src,
                       const double* src_xyz, const double* src_w, int64_t n, const double* xyz, double* hess6,
                       double* part, size_t part_cap, hipStream_t st);
