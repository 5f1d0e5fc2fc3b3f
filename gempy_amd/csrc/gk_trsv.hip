// Sync-free blocked triangular solves with a Cholesky factor: one launch for
// L y = b (y into a workspace vector) and one for L^T x = y (x into b).  No
// launch writes a vector that it also reads before the matching flag (the
// forward c steps read b, the backward ones y): a copy of the old line read
// on one XCD could otherwise stay in that XCD's L2 and be read there after
// the new value's flag.
//
// L is row-major (ld = lda, a multiple of 256); its 256 x 256 diagonal blocks
// are not read.  The factorization leaves 1 + 2 D planes of 256 x 256 blocks
// (lda x 256 each, block k at row k0 = 256 k; D = GK_CHAIN_DEPTH = 2):
//   plane 0: Dinv_k = L_kk^-1 (lower triangular, zeros stored above the
//            diagonal: the solves read whole blocks)
//   planes D+1-d (d = 1..D):  G_{k,d} = Dinv_k L_{k,k-d}
//   planes 2D+1-d (d = 1..D): H_{k,d} = L_{k+d,k} Dinv_k
// Forward, block row k:
//   c_k = Dinv_k (b_k - sum_{j <= k-D-1} L_kj y_j)                ("c step")
//   y_k = c_k - sum_{d=1..D} G_{k,d} y_{k-d}                      ("chain step")
// and backward x_k = c'_k - sum_d H_{k,d}^T x_{k+d} with c'_k =
// Dinv_k^T (y_k - sum_{j >= k+D+1} L_jk^T x_j).  Only the chain step lies on
// the sequential path: it multiplies G_{k,D} (loaded into registers before
// its waits) as soon as y_{k-D} is published, loads each next block while the
// next y is pending, and takes c_k, computed ahead.  The sums over far blocks
// are split into "far chunks" whose size grows with the distance from the
// diagonal (d in [3,5), [5,9), then 8 blocks each); the c step adds their
// published partials (the nearest one last).  Measured at N = 35k: the solves
// are bound by the rate at which the far chunks stream L (each published y_j
// releases up to 33 MB of block reads, ~2.7 TB/s achieved), not by the chain;
// a depth of 3 with 4-block chunks ran slower.  A workgroup covers a
// 256 x 256 block with 8 waves of 32 rows (two waves per SIMD, so a wave can
// keep its 32 prefetched rows in registers; FP64 takes a block in two
// 128-column halves and prefetches the first).  Every item is one row-block
// (forward) or column-block (backward) product over 1..8 blocks followed by a
// short epilogue, so all item types share one loop: a CU runs only a few items
// per launch, and a separate unrolled copy per item type would run from a cold
// instruction cache on the sequential path.
//
// Ordering and progress: every workgroup takes a ticket (atomic counter) when it
// starts; the ticket, not blockIdx, selects its work from a host-built table in
// which every item depends only on items with smaller tickets.  A workgroup that
// waits therefore waits on workgroups that have already started, so the grid
// always drains.  Readiness flags hold the launch's epoch (no clearing between
// launches).  Each wait is bounded: past the spin limit the workgroup raises an
// error word, publishes its flags without computing, and every later wait sees
// the error and returns at once.
#include <vector>
#include <algorithm>
#include "gk_internal.h"

namespace {

constexpr int NB = GK_BLOCK;            // 256
constexpr unsigned SPIN_LIMIT = 1u << 21;

// Rows of a 256-block per wave (RPW), waves per workgroup (NW), columns per
// lane and pass (VW; a pass covers 64 VW columns): 8 waves of 32 rows, 2
// waves per SIMD, so a wave may hold one pass of its rows (128 VGPRs) loaded
// ahead of its waits: FP32 4 columns per lane (one pass per block), FP64 2
// (two passes).
template <typename T> struct Shape;
template <> struct Shape<float> { static constexpr int RPW = 32, NW = 8, NT = 512, WPE = 2, VW = 4; };
template <> struct Shape<double> { static constexpr int RPW = 32, NW = 8, NT = 512, WPE = 2, VW = 2; };

// distance range [lo, hi) of far chunk s
constexpr int NEAR = GK_CHAIN_DEPTH + 1;     // the chain step covers distances 1 .. NEAR-1
inline void chunk_range(int s, int* lo, int* hi) {
  static_assert(NEAR == 3, "chunk ranges start at the first distance past the chain");
  if (s == 0) { *lo = 3; *hi = 5; }
  else if (s == 1) { *lo = 5; *hi = 9; }
  else { *lo = 9 + 8 * (s - 2); *hi = 17 + 8 * (s - 2); }
}

template <typename T> struct Vec { typedef T type __attribute__((ext_vector_type(Shape<T>::VW))); };

struct Item { int k, s, jlo, jhi; };    // s >= 0: far chunk s of row k; CHAIN / CSTEP: the row's steps
constexpr int CHAIN = -1, CSTEP = -2;

// Relaxed polling (an acquire load would invalidate the caches on every
// poll), then one acquire fence by the polling thread once the flag is seen;
// a workgroup barrier carries the ordering to the other waves.
__device__ __forceinline__ bool wait_epoch(const unsigned* flag, unsigned epoch, int* err) {
  for (unsigned it = 0; it < SPIN_LIMIT; ++it) {
    if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      return true;
    }
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return false;
}

// thread 0 waits, the workgroup joins; returns false on error
__device__ __forceinline__ bool group_wait(const unsigned* flag, unsigned epoch, int* err, int* s_ok) {
  if (threadIdx.x == 0) *s_ok = wait_epoch(flag, epoch, err) ? 1 : 0;
  __syncthreads();
  return *s_ok != 0;
}

// wait for flags[0..n) in parallel (one per thread, strided), then join
__device__ __forceinline__ bool group_wait_all(const unsigned* flags, int n, unsigned epoch, int* err, int* s_ok) {
  if (threadIdx.x == 0) *s_ok = 1;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    if (!wait_epoch(flags + i, epoch, err)) *s_ok = 0;
  __syncthreads();
  return *s_ok != 0;
}

// Every wave first waits for its own global stores to complete (the
// workgroup barrier alone does not: it orders LDS, and the release of thread 0
// covers only its own wave's outstanding stores); after the barrier the
// release store of thread 0 writes back the L2 and publishes the flag.
__device__ __forceinline__ void group_publish(unsigned* flag, unsigned epoch) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// RPW rows of a wave, lane l holding columns VW l .. VW l + VW - 1 of a pass
template <typename T>
struct Rows {
  static constexpr int R = Shape<T>::RPW;
  typename Vec<T>::type v[R];
};

template <typename T>
__device__ __forceinline__ void loadr(Rows<T>& R, const T* M, int64_t ld, int64_t r0, int l) {
  typedef typename Vec<T>::type vec;
  const T* base = M + r0 * ld + Shape<T>::VW * l;
  int64_t off = 0;
#pragma unroll
  for (int r = 0; r < Rows<T>::R; ++r) {
    R.v[r] = __builtin_nontemporal_load(reinterpret_cast<const vec*>(base + off));
    off += ld;
    // FP64: each row offset from the previous one (the 32 row addresses
    // would otherwise be kept live beside the rows and the row partials)
    if constexpr (sizeof(T) == 8) asm volatile("" : "+v"(off));
  }
}

// Completes the outstanding loads into R before the caller's flag wait: the
// registers are asm operands, so the loads cannot be sunk past the wait.
template <typename T>
__device__ __forceinline__ void pin(Rows<T>& R) {
#pragma unroll
  for (int q = 0; q < Rows<T>::R; q += 8)
    asm volatile("s_waitcnt vmcnt(0)"
                 : "+v"(R.v[q]), "+v"(R.v[q + 1]), "+v"(R.v[q + 2]), "+v"(R.v[q + 3]), "+v"(R.v[q + 4]),
                   "+v"(R.v[q + 5]), "+v"(R.v[q + 6]), "+v"(R.v[q + 7])
                 :
                 : "memory");
}

template <typename T>
__device__ __forceinline__ void pin1(T& v) {
  asm volatile("s_waitcnt vmcnt(0)" : "+v"(v) : : "memory");
}

// Sum over the 16 lanes of a DPP row, in every lane of the row: xor 1 and
// xor 2 (quad permutes), then the mirrored quad within each half row, then
// the mirrored half row.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffll), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename T>
__device__ __forceinline__ T row16_sum(T v) {
  v += dpp_mov<0xB1>(v);      // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);      // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);     // row_half_mirror
  v += dpp_mov<0x140>(v);     // row_mirror
  return v;
}

// Lane partials p[r] (row r of the wave) summed over the 64 lanes: lane
// l < RPW returns the sum of row l.
template <typename T>
__device__ __forceinline__ T lane_sums(const T (&p)[Shape<T>::RPW], int l) {
  constexpr int RPW = Shape<T>::RPW;
  T v1 = T(0), v2 = T(0);
  const int i = l & 15;
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const T s = row16_sum(p[r]);
    if (r < 16) v1 = i == r ? s : v1;
    else v2 = i == r - 16 ? s : v2;
  }
  v1 += __shfl_xor(v1, 16);
  v1 += __shfl_xor(v1, 32);
  if constexpr (RPW > 16) {
    v2 += __shfl_xor(v2, 16);
    v2 += __shfl_xor(v2, 32);
    return l < 16 ? v1 : v2;
  }
  return v1;
}

// p[r] += R[r][.] . x (the lane's VW columns)
template <typename T>
__device__ __forceinline__ void rowacc(const Rows<T>& R, typename Vec<T>::type xv, T (&p)[Shape<T>::RPW]) {
#pragma unroll
  for (int r = 0; r < Rows<T>::R; ++r) {
    T s = p[r];
#pragma unroll
    for (int e = 0; e < Shape<T>::VW; ++e) s = fma(R.v[r][e], xv[e], s);
    p[r] = s;
  }
}

// acc (the lane's VW columns) += sum_r R[r][.] x[r], x the wave's RPW entries
template <typename T>
__device__ __forceinline__ void colacc(const Rows<T>& R, const T* x, typename Vec<T>::type& acc) {
#pragma unroll
  for (int r = 0; r < Rows<T>::R; ++r) {
    const T xr = x[r];
#pragma unroll
    for (int e = 0; e < Shape<T>::VW; ++e) acc[e] = fma(R.v[r][e], xr, acc[e]);
  }
}

// Half of a wave's rows (16), for the software-pipelined far chunks: one
// half is multiplied while the next half's loads are in flight.
template <typename T>
struct Half {
  typename Vec<T>::type v[16];
};

template <typename T>
__device__ __forceinline__ void loadh(Half<T>& H, const T* base, int64_t ld) {
  typedef typename Vec<T>::type vec;
  int64_t off = 0;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    H.v[r] = __builtin_nontemporal_load(reinterpret_cast<const vec*>(base + off));
    off += ld;
    asm volatile("" : "+v"(off));       // each row offset from the previous one (the base stays global)
  }
}

template <typename T>
__device__ __forceinline__ void pinh(Half<T>& H) {
#pragma unroll
  for (int q = 0; q < 16; q += 8)
    asm volatile("s_waitcnt vmcnt(0)"
                 : "+v"(H.v[q]), "+v"(H.v[q + 1]), "+v"(H.v[q + 2]), "+v"(H.v[q + 3]), "+v"(H.v[q + 4]),
                   "+v"(H.v[q + 5]), "+v"(H.v[q + 6]), "+v"(H.v[q + 7])
                 :
                 : "memory");
}

// Completes the products into p[0..15] before the next loads are issued (the
// loads overwrite the half just consumed; hoisting them above the products
// would need a second copy of it).
template <typename T>
__device__ __forceinline__ void fence16(T* p) {
#pragma unroll
  for (int q = 0; q < 16; q += 8)
    asm volatile("" : "+v"(p[q]), "+v"(p[q + 1]), "+v"(p[q + 2]), "+v"(p[q + 3]), "+v"(p[q + 4]), "+v"(p[q + 5]),
                 "+v"(p[q + 6]), "+v"(p[q + 7]) : : "memory");
}

// p[r] += H[r][.] . x for the 16 rows of a half
template <typename T>
__device__ __forceinline__ void rowacc_h(const Half<T>& H, typename Vec<T>::type xv, T* p) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    T s = p[r];
#pragma unroll
    for (int e = 0; e < Shape<T>::VW; ++e) s = fma(H.v[r][e], xv[e], s);
    p[r] = s;
  }
}

// acc (the lane's columns) += sum_r H[r][.] x[r], x the half's 16 entries
template <typename T>
__device__ __forceinline__ void colacc_h(const Half<T>& H, const T* x, typename Vec<T>::type& acc) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const T xr = x[r];
#pragma unroll
    for (int e = 0; e < Shape<T>::VW; ++e) acc[e] = fma(H.v[r][e], xr, acc[e]);
  }
}

struct TrsvCtl {
  const Item* items;
  const int* rowbase;       // far-chunk slot of (k, s) = rowbase[k] + s
  unsigned* yflag;          // [nb] solution block published
  unsigned* cflag;          // [nb] c step published
  unsigned* pflag;          // [slots]
  unsigned long long* ticket;
  unsigned long long ticket_base;
  unsigned epoch;
  int n_items;
  int nb;
  int* err;
};

template <typename T>
__device__ __forceinline__ bool take_ticket(const TrsvCtl& ctl, Item* it, int* s_id) {
  if (threadIdx.x == 0) *s_id = (int)(atomicAdd(ctl.ticket, 1ull) - ctl.ticket_base);
  __syncthreads();
  if (*s_id < 0 || *s_id >= ctl.n_items) {                    // ticket counter out of step
    if (threadIdx.x == 0) __hip_atomic_store(ctl.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
  }
  *it = ctl.items[*s_id];
  return true;
}

// Far chunks stream their blocks: all x blocks of the chunk are staged in
// LDS once (no barrier inside the loop), and the wave's rows are taken in
// halves of 16, the loads of the next half in flight while the current one is
// multiplied.  The first block's first pass is loaded before the wait.
constexpr int MAXCHUNK = 8;              // blocks per far chunk at most

template <typename T>
__device__ __forceinline__ void far_fwd(const Item& it, const T* __restrict__ L, int64_t lda, const T* b, T* part,
                                        const TrsvCtl& ctl, T* xall, int* s_ok) {
  typedef typename Vec<T>::type vec;
  constexpr int RPW = Shape<T>::RPW, NT = Shape<T>::NT, VW = Shape<T>::VW, PW = 64 * VW, PS = NB / PW;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row = RPW * w + (l & (RPW - 1));
  const int64_t k0 = (int64_t)it.k * NB;
  const T* M = L + k0 * lda + (int64_t)it.jlo * NB + VW * l;
  const int nblk = it.jhi - it.jlo, U = nblk * PS * 2;
  auto base = [&](int u) -> const T* {          // unit u: block j, pass h, half q
    const int j = u / (2 * PS), h = (u >> 1) % PS, q = u & 1;
    return M + (int64_t)j * NB + h * PW + (int64_t)(RPW * w + 16 * q) * lda;
  };
  Half<T> A, B;
  loadh(A, base(0), lda);
  loadh(B, base(1), lda);
  pinh(A);
  pinh(B);
  const bool ok = group_wait(ctl.yflag + it.jhi - 1, ctl.epoch, ctl.err, s_ok);
  T sum = T(0);
  if (ok) {
    const T* x = b + (int64_t)it.jlo * NB;
    for (int e = threadIdx.x; e < nblk * NB; e += NT) xall[e] = x[e];
    __syncthreads();
    T p[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) p[r] = T(0);
#pragma unroll 1
    for (int u = 0; u < U; u += 2) {
      const int j = u / (2 * PS), h = (u >> 1) % PS;
      const vec xv = *reinterpret_cast<const vec*>(xall + j * NB + h * PW + VW * l);
      rowacc_h<T>(A, xv, p);
      fence16(p);
      if (u + 2 < U) loadh(A, base(u + 2), lda);
      rowacc_h<T>(B, xv, p + 16);
      fence16(p + 16);
      if (u + 3 < U) loadh(B, base(u + 3), lda);
    }
    sum = lane_sums<T>(p, l);
  }
  if (ok && l < RPW) part[(int64_t)(ctl.rowbase[it.k] + it.s) * NB + row] = sum;
  group_publish(ctl.pflag + ctl.rowbase[it.k] + it.s, ctl.epoch);
}

template <typename T>
__device__ __forceinline__ void far_bwd(const Item& it, const T* __restrict__ L, int64_t lda, T* b, T* part,
                                        const TrsvCtl& ctl, T* xall, T (*red)[NB], int* s_ok) {
  typedef typename Vec<T>::type vec;
  constexpr int RPW = Shape<T>::RPW, NW = Shape<T>::NW, NT = Shape<T>::NT, VW = Shape<T>::VW, PW = 64 * VW,
                PS = NB / PW;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t k0 = (int64_t)it.k * NB;
  const T* M = L + (int64_t)it.jlo * NB * lda + k0 + VW * l;
  const int nblk = it.jhi - it.jlo, U = nblk * PS * 2;
  auto base = [&](int u) -> const T* {          // unit u: block j, pass h, half q
    const int j = u / (2 * PS), h = (u >> 1) % PS, q = u & 1;
    return M + ((int64_t)j * NB + RPW * w + 16 * q) * lda + h * PW;
  };
  Half<T> A, B;
  loadh(A, base(0), lda);
  loadh(B, base(1), lda);
  pinh(A);
  pinh(B);
  const bool ok = group_wait(ctl.yflag + it.jlo, ctl.epoch, ctl.err, s_ok);
  T sum = T(0);
  if (ok) {
    const T* x = b + (int64_t)it.jlo * NB;
    for (int e = threadIdx.x; e < nblk * NB; e += NT) xall[e] = x[e];
    __syncthreads();
    vec acc[PS];
#pragma unroll
    for (int h = 0; h < PS; ++h)
#pragma unroll
      for (int e = 0; e < VW; ++e) acc[h][e] = T(0);
#pragma unroll 1
    for (int u = 0; u < U; u += 2) {
      const int j = u / (2 * PS), h = (u >> 1) % PS;
      const T* xw = xall + j * NB + RPW * w;
      vec a = {};
#pragma unroll
      for (int e = 0; e < VW; ++e) a[e] = T(0);
      colacc_h<T>(A, xw, a);
      asm volatile("" : "+v"(a) : : "memory");
      if (u + 2 < U) loadh(A, base(u + 2), lda);
      colacc_h<T>(B, xw + 16, a);
      asm volatile("" : "+v"(a) : : "memory");
      if (u + 3 < U) loadh(B, base(u + 3), lda);
      if constexpr (PS == 1) {
        acc[0] += a;
      } else {
        if (h == 0) acc[0] += a;
        else acc[PS - 1] += a;
      }
    }
#pragma unroll
    for (int h = 0; h < PS; ++h) *reinterpret_cast<vec*>(&red[w][h * PW + VW * l]) = acc[h];
    __syncthreads();
    if (threadIdx.x < NB) {
#pragma unroll
      for (int q = 0; q < NW; ++q) sum += red[q][threadIdx.x];
    }
  }
  if (ok && threadIdx.x < NB) part[(int64_t)(ctl.rowbase[it.k] + it.s) * NB + threadIdx.x] = sum;
  group_publish(ctl.pflag + ctl.rowbase[it.k] + it.s, ctl.epoch);
}

// Forward items: row-block products sum_j M_j x_j; the lanes keep per-row
// partials over all blocks and passes and reduce once.  PF: the first pass
// of the first block is loaded into registers before the item's waits.
template <typename T, bool PF>
__global__ __launch_bounds__(Shape<T>::NT) __attribute__((amdgpu_waves_per_eu(Shape<T>::WPE, Shape<T>::WPE)))
void trsv_fwd_kernel(const T* __restrict__ L, int64_t lda, const T* __restrict__ planes, const T* b, T* y_out,
                     T* cbuf, T* part, TrsvCtl ctl) {
  __shared__ int s_id, s_ok, s_ok2;
  typedef typename Vec<T>::type vec;
  __shared__ __attribute__((aligned(32))) T xs[NB * MAXCHUNK];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  constexpr int RPW = Shape<T>::RPW, NT = Shape<T>::NT, PW = 64 * Shape<T>::VW;   // PW: columns per pass
  const int row = RPW * w + (l & (RPW - 1));     // row of the 256-block reduced into lanes l < RPW
  Item it;
  if (!take_ticket<T>(ctl, &it, &s_id)) return;
  if (it.s >= 0) {
    far_fwd<T>(it, L, lda, y_out, part, ctl, xs, &s_ok);
    return;
  }
  const int64_t k0 = (int64_t)it.k * NB;
  // c step: Dinv_k t (t staged in LDS);  chain: sum_d G_{k,d} y_{k-d} (the
  // blocks that exist), one wait per block
  const int64_t plane = (int64_t)ctl.nb * NB * NB;
  const T* M;
  int64_t ld, bstride = 0;
  int nblk;
  const T* x = nullptr;
  const unsigned* wflag = nullptr;
  if (it.s == CSTEP) {
    M = planes + k0 * NB; ld = NB; nblk = 1;
  } else {
    nblk = min(it.k, NEAR - 1);
    M = planes + (NEAR - nblk) * plane + k0 * NB; ld = NB; bstride = plane;
    x = y_out + k0 - (int64_t)nblk * NB; wflag = ctl.yflag + it.k - nblk;
  }
  Rows<T> R;
  if constexpr (PF) {
    if (nblk > 0) { loadr(R, M, ld, RPW * w, l); pin(R); }
  }
  bool ok = true;
  T y = T(0);
  if (it.s == CSTEP) {                                        // t = b_k - far partials (the nearest last)
    const int f0 = ctl.rowbase[it.k], nfar = ctl.rowbase[it.k + 1] - f0;
    T t = b[k0 + row];
    if (nfar > 1) {
      ok = group_wait_all(ctl.pflag + f0 + 1, nfar - 1, ctl.epoch, ctl.err, &s_ok2);
      if (ok)
        for (int s = 1; s < nfar; ++s) t -= part[(int64_t)(f0 + s) * NB + row];
    }
    if (ok && nfar > 0) {
      ok = group_wait(ctl.pflag + f0, ctl.epoch, ctl.err, &s_ok);
      if (ok) t -= part[(int64_t)f0 * NB + row];
    }
    if (l < RPW) xs[row] = t;
  } else if (nblk == 0) {
    ok = group_wait(ctl.cflag + it.k, ctl.epoch, ctl.err, &s_ok);
    if (ok) y = cbuf[k0 + row];
  }
  T sum = T(0);
  if (ok) {
    T p[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) p[r] = T(0);
#pragma unroll 1
    for (int j = 0; j < nblk; ++j) {
      if (!PF || j > 0) loadr(R, M + j * bstride, ld, RPW * w, l);
      if (wflag) {           // chain: c_k (ready links ahead) before the last block's wait
        if (j == nblk - 1) {
          ok = group_wait(ctl.cflag + it.k, ctl.epoch, ctl.err, &s_ok);
          if (ok) { y = cbuf[k0 + row]; pin1(y); }
        }
        if (ok) ok = group_wait(wflag + j, ctl.epoch, ctl.err, &s_ok2);
        if (!ok) break;
      }
      if (x) {
        for (int e = threadIdx.x; e < NB; e += NT) xs[e] = x[(int64_t)j * NB + e];
      }
      __syncthreads();
      rowacc<T>(R, *reinterpret_cast<const vec*>(xs + Shape<T>::VW * l), p);
#pragma unroll
      for (int h = 1; h < NB / PW; ++h) {
        asm volatile("" ::: "memory");           // one pass of rows live at a time
        loadr(R, M + j * bstride + h * PW, ld, RPW * w, l);
        rowacc<T>(R, *reinterpret_cast<const vec*>(xs + h * PW + Shape<T>::VW * l), p);
      }
      __syncthreads();
    }
    sum = lane_sums<T>(p, l);
  }
  if (it.s == CSTEP) {
    if (ok && l < RPW) cbuf[k0 + row] = sum;
    group_publish(ctl.cflag + it.k, ctl.epoch);
  } else {
    if (ok && l < RPW) y_out[k0 + row] = y - sum;
    group_publish(ctl.yflag + it.k, ctl.epoch);
  }
}

// Backward items are column-block products sum_j M_j^T x_j sharing one loop,
// as in the forward kernel.
template <typename T, bool PF>
__global__ __launch_bounds__(Shape<T>::NT) __attribute__((amdgpu_waves_per_eu(Shape<T>::WPE, Shape<T>::WPE)))
void trsv_bwd_kernel(const T* __restrict__ L, int64_t lda, const T* __restrict__ planes, T* b, const T* y_in,
                     T* cbuf, T* part, TrsvCtl ctl) {
  typedef typename Vec<T>::type vec;
  __shared__ int s_id, s_ok, s_ok2;
  constexpr int RPW = Shape<T>::RPW, NW = Shape<T>::NW, NT = Shape<T>::NT, VW = Shape<T>::VW, PW = 64 * VW;
  __shared__ __attribute__((aligned(32))) T xs[NB * MAXCHUNK];
  __shared__ __attribute__((aligned(32))) T red[NW][NB];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  Item it;
  if (!take_ticket<T>(ctl, &it, &s_id)) return;
  if (it.s >= 0) {
    far_bwd<T>(it, L, lda, b, part, ctl, xs, red, &s_ok);
    return;
  }
  const int64_t k0 = (int64_t)it.k * NB;
  // c step: Dinv_k^T t (t staged in LDS);  chain: sum_d H_{k,d}^T x_{k+d}
  // (the blocks that exist), one wait per block
  const int64_t plane = (int64_t)ctl.nb * NB * NB;
  const T* M;
  int64_t ld, bstride = 0, xstride = NB;
  int nblk;
  const T* x = nullptr;
  const unsigned* wflag = nullptr;
  if (it.s == CSTEP) {
    M = planes + k0 * NB; ld = NB; nblk = 1;
  } else {
    nblk = min(ctl.nb - 1 - it.k, NEAR - 1);
    M = planes + (2 * NEAR - 1 - nblk) * plane + k0 * NB; ld = NB; bstride = plane;
    x = b + k0 + (int64_t)nblk * NB; xstride = -NB; wflag = ctl.yflag + it.k + nblk;
  }
  Rows<T> R;
  if constexpr (PF) {
    if (nblk > 0) { loadr(R, M, ld, RPW * w, l); pin(R); }
  }
  bool ok = true;
  T cv = T(0);
  if (it.s == CSTEP) {                                        // t = y_k - far partials (the nearest last)
    const int f0 = ctl.rowbase[it.k], nfar = ctl.rowbase[it.k + 1] - f0;
    T t = threadIdx.x < NB ? y_in[k0 + threadIdx.x] : T(0);
    if (nfar > 1) {
      ok = group_wait_all(ctl.pflag + f0 + 1, nfar - 1, ctl.epoch, ctl.err, &s_ok2);
      if (ok && threadIdx.x < NB)
        for (int s = 1; s < nfar; ++s) t -= part[(int64_t)(f0 + s) * NB + threadIdx.x];
    }
    if (ok && nfar > 0) {
      ok = group_wait(ctl.pflag + f0, ctl.epoch, ctl.err, &s_ok);
      if (ok && threadIdx.x < NB) t -= part[(int64_t)f0 * NB + threadIdx.x];
    }
    if (threadIdx.x < NB) xs[threadIdx.x] = t;
  } else if (nblk == 0) {
    ok = group_wait(ctl.cflag + it.k, ctl.epoch, ctl.err, &s_ok);
    if (ok && threadIdx.x < NB) cv = cbuf[k0 + threadIdx.x];
  }
  vec acc[NB / PW];
#pragma unroll
  for (int h = 0; h < NB / PW; ++h)
#pragma unroll
    for (int e = 0; e < VW; ++e) acc[h][e] = T(0);
  T sum = T(0);
  if (ok) {
#pragma unroll 1
    for (int j = 0; j < nblk; ++j) {
      if (!PF || j > 0) loadr(R, M + j * bstride, ld, RPW * w, l);
      if (wflag) {           // chain: c'_k (ready links ahead) before the last block's wait
        if (j == nblk - 1) {
          ok = group_wait(ctl.cflag + it.k, ctl.epoch, ctl.err, &s_ok);
          if (ok && threadIdx.x < NB) { cv = cbuf[k0 + threadIdx.x]; pin1(cv); }
        }
        if (ok) ok = group_wait(wflag - j, ctl.epoch, ctl.err, &s_ok2);
        if (!ok) break;
      }
      if (x) {
        for (int e = threadIdx.x; e < NB; e += NT) xs[e] = x[(int64_t)j * xstride + e];
      }
      __syncthreads();
      colacc<T>(R, xs + RPW * w, acc[0]);
#pragma unroll
      for (int h = 1; h < NB / PW; ++h) {
        asm volatile("" ::: "memory");           // one pass of rows live at a time
        loadr(R, M + j * bstride + h * PW, ld, RPW * w, l);
        colacc<T>(R, xs + RPW * w, acc[h]);
      }
      __syncthreads();
    }
    // column c = threadIdx.x (< 256) gets the sum of the waves' partials
#pragma unroll
    for (int h = 0; h < NB / PW; ++h) *reinterpret_cast<vec*>(&red[w][h * PW + VW * l]) = acc[h];
    __syncthreads();
    if (threadIdx.x < NB) {
#pragma unroll
      for (int q = 0; q < NW; ++q) sum += red[q][threadIdx.x];
    }
  }
  if (it.s == CSTEP) {
    if (ok && threadIdx.x < NB) cbuf[k0 + threadIdx.x] = sum;
    group_publish(ctl.cflag + it.k, ctl.epoch);
  } else {
    if (ok && threadIdx.x < NB) b[k0 + threadIdx.x] = cv - sum;
    group_publish(ctl.yflag + it.k, ctl.epoch);
  }
}

// Host-side work tables.  Forward: the c steps of rows 0..NEAR-1 (no far
// chunks), then for each block j: its chain step, the far chunks whose nearest
// block is j (row k = j + lo_s, blocks at distances [lo_s, hi_s)), and the c
// step of row j + NEAR (all its far chunks are then queued).  Backward mirrors it with
// k = j - lo_s.  Every wait of an item refers to an item queued before it.
int nfar_of(int k) {              // far chunks of a row with k blocks on its far side
  int s = 0, lo, hi;
  for (;;) { chunk_range(s, &lo, &hi); if (lo > k) return s; ++s; }
}

void build_tables(int nb, std::vector<Item>& fwd, std::vector<int>& fbase, std::vector<Item>& bwd,
                  std::vector<int>& bbase) {
  fbase.assign(nb + 1, 0);
  bbase.assign(nb + 1, 0);
  for (int k = 0; k < nb; ++k) {
    fbase[k + 1] = fbase[k] + nfar_of(k);
    bbase[k + 1] = bbase[k] + nfar_of(nb - 1 - k);
  }
  fwd.clear();
  bwd.clear();
  for (int k = 0; k < std::min(NEAR, nb); ++k) fwd.push_back({k, CSTEP, 0, 0});
  for (int j = 0; j < nb; ++j) {
    fwd.push_back({j, CHAIN, 0, 0});
    for (int s = 0;; ++s) {
      int lo, hi;
      chunk_range(s, &lo, &hi);
      const int k = j + lo;
      if (k >= nb) break;
      fwd.push_back({k, s, std::max(0, k - hi + 1), k - lo + 1});
    }
    if (j + NEAR < nb) fwd.push_back({j + NEAR, CSTEP, 0, 0});
  }
  for (int k = nb - 1; k >= std::max(nb - NEAR, 0); --k) bwd.push_back({k, CSTEP, 0, 0});
  for (int j = nb - 1; j >= 0; --j) {
    bwd.push_back({j, CHAIN, 0, 0});
    for (int s = 0;; ++s) {
      int lo, hi;
      chunk_range(s, &lo, &hi);
      const int k = j - lo;
      if (k < 0) break;
      bwd.push_back({k, s, k + lo, std::min(nb, k + hi)});
    }
    if (j - NEAR >= 0) bwd.push_back({j - NEAR, CSTEP, 0, 0});
  }
}

}  // namespace

// Workspace: [fwd items][bwd items][fwd base][bwd base][yflag][cflag][pflag][ticket][err][c][y][partials]
struct TrsvLayout {
  size_t o_fi, o_bi, o_fb, o_bb, o_yf, o_cf, o_pf, o_tk, o_err, o_c, o_y, o_part, bytes;
  int nfwd, nbwd, slots;
};

static TrsvLayout trsv_layout(int nb, int nfwd, int nbwd, int slots) {
  TrsvLayout L{};
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += gk_round_up((int64_t)b, 256); return r; };
  L.o_fi = take(sizeof(Item) * nfwd);
  L.o_bi = take(sizeof(Item) * nbwd);
  L.o_fb = take(sizeof(int) * (nb + 1));
  L.o_bb = take(sizeof(int) * (nb + 1));
  L.o_yf = take(sizeof(unsigned) * nb);
  L.o_cf = take(sizeof(unsigned) * nb);
  L.o_pf = take(sizeof(unsigned) * std::max(slots, 1));
  L.o_tk = take(sizeof(unsigned long long));
  L.o_err = take(sizeof(int));
  L.o_c = take(sizeof(double) * NB * (size_t)nb);
  L.o_y = take(sizeof(double) * NB * (size_t)nb);
  L.o_part = take(sizeof(double) * NB * (size_t)std::max(slots, 1));
  L.bytes = o;
  L.nfwd = nfwd; L.nbwd = nbwd; L.slots = slots;
  return L;
}

// Tables, flags and buffers of solve set ctx: WS_TRSV holds GK_TRSV_CTX
// copies of the layout back to back.
static int trsv_prepare(gk_handle h, int nb, hipStream_t st, char** base, TrsvLayout* lay, int ctx) {
  std::vector<Item> fwd, bwd;
  std::vector<int> fb, bb;
  const bool rebuild = h->trsv_nb != nb || h->ws[WS_TRSV] == nullptr;
  if (rebuild) {
    build_tables(nb, fwd, fb, bwd, bb);
    for (const std::vector<Item>* tab : {&fwd, &bwd})
      for (const Item& q : *tab)
        GK_REQUIRE(q.s < 0 || q.jhi - q.jlo <= MAXCHUNK, GK_ERR_ARG, "triangular solve: far chunk too large");
  }
  const int slots = rebuild ? std::max(fb[nb], bb[nb]) : h->trsv_slots;
  const int nfwd = rebuild ? (int)fwd.size() : h->trsv_nfwd;
  const int nbwd = rebuild ? (int)bwd.size() : h->trsv_nbwd;
  *lay = trsv_layout(nb, nfwd, nbwd, slots);
  void* p = nullptr;
  const void* before = h->ws[WS_TRSV];
  GK_TRY(gk_workspace(h, WS_TRSV, lay->bytes * GK_TRSV_CTX, &p));
  if (rebuild || p != before) {
    if (!rebuild) build_tables(nb, fwd, fb, bwd, bb);
    for (int c = 0; c < GK_TRSV_CTX; ++c) {
      char* b = static_cast<char*>(p) + c * lay->bytes;
      GK_HIP(hipMemcpyAsync(b + lay->o_fi, fwd.data(), sizeof(Item) * fwd.size(), hipMemcpyHostToDevice, st));
      GK_HIP(hipMemcpyAsync(b + lay->o_bi, bwd.data(), sizeof(Item) * bwd.size(), hipMemcpyHostToDevice, st));
      GK_HIP(hipMemcpyAsync(b + lay->o_fb, fb.data(), sizeof(int) * fb.size(), hipMemcpyHostToDevice, st));
      GK_HIP(hipMemcpyAsync(b + lay->o_bb, bb.data(), sizeof(int) * bb.size(), hipMemcpyHostToDevice, st));
      GK_HIP(hipMemsetAsync(b + lay->o_yf, 0, lay->o_c - lay->o_yf, st));   // flags, ticket, error
      h->trsv_epoch[c] = 0;
      h->trsv_tickets[c] = 0;
    }
    GK_HIP(hipStreamSynchronize(st));          // host tables go out of scope
    h->trsv_nb = nb;
    h->trsv_slots = slots;
    h->trsv_nfwd = nfwd;
    h->trsv_nbwd = nbwd;
  }
  *base = static_cast<char*>(p) + ctx * lay->bytes;
  return GK_OK;
}

template <typename T>
int gk_trsv_chol_impl(gk_handle h, const T* L, const T* planes, int64_t lda, T* b, hipStream_t st, int ctx) {
  GK_REQUIRE(h && L && planes && b && lda > 0 && lda % NB == 0, GK_ERR_ARG, "triangular solve: bad arguments");
  GK_REQUIRE(ctx >= 0 && ctx < GK_TRSV_CTX, GK_ERR_ARG, "triangular solve: solve set %d out of range", ctx);
  constexpr bool PF = true;                     // first pass of the first block before the waits
  const int nb = (int)(lda / NB);
  char* base = nullptr;
  TrsvLayout lay;
  GK_TRY(trsv_prepare(h, nb, st, &base, &lay, ctx));
  TrsvCtl c;
  c.yflag = reinterpret_cast<unsigned*>(base + lay.o_yf);
  c.cflag = reinterpret_cast<unsigned*>(base + lay.o_cf);
  c.pflag = reinterpret_cast<unsigned*>(base + lay.o_pf);
  c.ticket = reinterpret_cast<unsigned long long*>(base + lay.o_tk);
  c.err = reinterpret_cast<int*>(base + lay.o_err);
  c.nb = nb;
  T* cbuf = reinterpret_cast<T*>(base + lay.o_c);
  T* ybuf = reinterpret_cast<T*>(base + lay.o_y);
  T* part = reinterpret_cast<T*>(base + lay.o_part);
  // forward
  c.items = reinterpret_cast<const Item*>(base + lay.o_fi);
  c.rowbase = reinterpret_cast<const int*>(base + lay.o_fb);
  c.ticket_base = h->trsv_tickets[ctx];
  c.epoch = ++h->trsv_epoch[ctx];
  c.n_items = lay.nfwd;
  hipEvent_t pa;
  const bool prof = gk_prof_begin(st, &pa);
  hipLaunchKernelGGL((trsv_fwd_kernel<T, PF>), dim3((unsigned)lay.nfwd), dim3(Shape<T>::NT), 0, st, L, lda, planes, b, ybuf,
                     cbuf, part, c);
  h->trsv_tickets[ctx] += (unsigned long long)lay.nfwd;
  // backward
  c.items = reinterpret_cast<const Item*>(base + lay.o_bi);
  c.rowbase = reinterpret_cast<const int*>(base + lay.o_bb);
  c.ticket_base = h->trsv_tickets[ctx];
  c.epoch = ++h->trsv_epoch[ctx];
  c.n_items = lay.nbwd;
  hipLaunchKernelGGL((trsv_bwd_kernel<T, PF>), dim3((unsigned)lay.nbwd), dim3(Shape<T>::NT), 0, st, L, lda, planes, b, ybuf,
                     cbuf, part, c);
  h->trsv_tickets[ctx] += (unsigned long long)lay.nbwd;
  if (prof) {    // each direction reads the blocks at distance >= NEAR, Dinv and the chain planes
    const double far = nb > NEAR ? (double)(nb - NEAR) * (nb - NEAR + 1) / 2.0 : 0.0;
    double nearb = nb;
    for (int d = 1; d < NEAR; ++d) nearb += std::max(nb - d, 0);
    gk_prof_end(st, pa, sizeof(T) == 8 ? 9 : 8, 2.0 * sizeof(T) * (far + nearb) * NB * NB);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    h->trsv_nb = 0;              // counters no longer in step: rebuild on the next call
    GK_REQUIRE(false, GK_ERR_HIP, "triangular solve launch: %s", hipGetErrorString(e));
  }
  return GK_OK;
}

template int gk_trsv_chol_impl<float>(gk_handle, const float*, const float*, int64_t, float*, hipStream_t, int);
template int gk_trsv_chol_impl<double>(gk_handle, const double*, const double*, int64_t, double*, hipStream_t, int);

// Error words of the sync-free solves (spin limit reached) of every solve
// set; the caller's stream has waited for the solves of the other sets.
int gk_trsv_check(gk_handle h, hipStream_t st) {
  if (!h->ws[WS_TRSV] || h->trsv_nb == 0) return GK_OK;
  TrsvLayout lay = trsv_layout(h->trsv_nb, h->trsv_nfwd, h->trsv_nbwd, h->trsv_slots);
  char* base = static_cast<char*>(h->ws[WS_TRSV]);
  int errs[GK_TRSV_CTX] = {};
  for (int c = 0; c < GK_TRSV_CTX; ++c)
    GK_HIP(hipMemcpyAsync(&errs[c], base + c * lay.bytes + lay.o_err, sizeof(int), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  int err = 0;
  for (int c = 0; c < GK_TRSV_CTX; ++c) err = err ? err : errs[c];
  if (err) {
    h->trsv_nb = 0;              // reset tables, flags and counters on the next call
    GK_REQUIRE(false, GK_ERR_HIP, err == 1 ? "triangular solve: a readiness wait reached its spin limit"
                                           : "triangular solve: work ticket out of range");
  }
  return GK_OK;
}

// Test hook: store `code` into the error word of the handle's triangular-solve
// workspace, as a readiness wait reaching its spin limit would (the kernels
// then skip their waits and drain).  Needs a solve to have run on the handle.
extern "C" int gk_test_trsv_set_error(gk_handle h, int32_t code, gk_stream stream) {
  GK_REQUIRE(h, GK_ERR_ARG, "gk_test_trsv_set_error: NULL handle");
  GK_REQUIRE(h->ws[WS_TRSV] && h->trsv_nb != 0, GK_ERR_ARG,
             "gk_test_trsv_set_error: no triangular-solve workspace yet");
  hipStream_t st = (hipStream_t)stream;
  TrsvLayout lay = trsv_layout(h->trsv_nb, h->trsv_nfwd, h->trsv_nbwd, h->trsv_slots);
  int v = code;
  GK_HIP(hipMemcpyAsync(static_cast<char*>(h->ws[WS_TRSV]) + lay.o_err, &v, sizeof(int), hipMemcpyHostToDevice, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}
