// Sync-free blocked triangular solves with a Cholesky factor: one launch for
// L y = b and one for L^T x = y (in place on b).
//
// L is row-major (ld = lda, a multiple of 256); its 256 x 256 diagonal blocks
// are not read.  The factorization leaves three planes of 256 x 256 blocks
// (lda x 256 each, block k at row k0 = 256 k):
//   plane 0: Dinv_k = L_kk^-1 (lower triangular)
//   plane 1: G_k = Dinv_k L_{k,k-1}         (k >= 1)
//   plane 2: H_k = L_{k+1,k} Dinv_k         (k <= nb-2)
// Forward, block row k:
//   c_k = Dinv_k (b_k - sum_{j <= k-2} L_kj y_j)      ("c step")
//   y_k = c_k - G_k y_{k-1}                           ("chain step")
// and backward x_k = c'_k - H_k^T x_{k+1} with c'_k = Dinv_k^T (y_k - sum_{j >= k+2}
// L_jk^T x_j).  Only the chain step lies on the sequential path, and its block
// G_k (H_k) is loaded into registers before the wait, so a link of the chain
// costs one flag hand-off and a register-resident 256 x 256 product instead
// of two HBM block reads by one CU (FP32; the FP64 variant loads after the
// wait, its rows do not fit in registers).  The sums over far blocks are
// split into "far chunks" whose size grows with the distance from the
// diagonal (d in [2,3), [3,5), [5,9), then 8 blocks each); the c step adds
// their published partials (the nearest one last).  A workgroup covers a
// 256 x 256 block: FP32 8 waves of 32 rows (two waves per SIMD, so a wave can
// keep its 32 prefetched rows in registers), FP64 16 waves of 16 rows.
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

// Rows of a 256-block per wave: FP32 32 (8 waves, 2 per SIMD, so a wave may
// hold 32 prefetched rows = 128 VGPRs), FP64 16 (16 waves).
template <typename T> struct Shape;
template <> struct Shape<float> { static constexpr int RPW = 32, NW = 8, NT = 512, WPE = 2; };
template <> struct Shape<double> { static constexpr int RPW = 16, NW = 16, NT = 1024, WPE = 4; };

// distance range [lo, hi) of far chunk s
inline void chunk_range(int s, int* lo, int* hi) {
  if (s == 0) { *lo = 2; *hi = 3; }
  else if (s == 1) { *lo = 3; *hi = 5; }
  else if (s == 2) { *lo = 5; *hi = 9; }
  else { *lo = 9 + 8 * (s - 3); *hi = 17 + 8 * (s - 3); }
}

template <typename T> struct V4;
template <> struct V4<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct V4<double> { typedef double type __attribute__((ext_vector_type(4))); };

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

// The barrier completes every wave's stores at workgroup scope; the release
// store of thread 0 then makes them visible device-wide before the flag.
__device__ __forceinline__ void group_publish(unsigned* flag, unsigned epoch) {
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// RPW rows of a wave, lane l holding columns 4l..4l+3 of each
template <typename T>
struct Rows {
  static constexpr int R = Shape<T>::RPW;
  typename V4<T>::type v[R];
};

template <typename T>
__device__ __forceinline__ void loadr(Rows<T>& R, const T* M, int64_t ld, int64_t r0, int64_t c0, int l) {
  typedef typename V4<T>::type v4;
  const T* base = M + r0 * ld + c0 + 4 * l;
#pragma unroll
  for (int r = 0; r < Rows<T>::R; ++r)
    R.v[r] = __builtin_nontemporal_load(reinterpret_cast<const v4*>(base + r * ld));
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

// Row dots of 16 rows R.v[q..q+15]: lane l gets row q + (l & 15).
template <typename T, bool MASK_UPPER>
__device__ __forceinline__ T rowdot16(const Rows<T>& R, int q, typename V4<T>::type xv, int rloc0, int l) {
  typedef typename V4<T>::type v4;
  T acc[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    v4 a = R.v[q + r];
    if (MASK_UPPER) {
      const int row = rloc0 + q + r, c = 4 * l;
      a.x = c <= row ? a.x : T(0);
      a.y = c + 1 <= row ? a.y : T(0);
      a.z = c + 2 <= row ? a.z : T(0);
      a.w = c + 3 <= row ? a.w : T(0);
    }
    T s = a.x * xv.x;
    s = fma(a.y, xv.y, s);
    s = fma(a.z, xv.z, s);
    s = fma(a.w, xv.w, s);
    acc[r] = s;
  }
#pragma unroll
  for (int width = 8; width >= 1; width >>= 1) {        // transpose-reduce over lanes
    const bool upper = (l & width) != 0;
#pragma unroll
    for (int i = 0; i < width; ++i) {
      const T send = upper ? acc[i] : acc[i + width];
      const T keep = upper ? acc[i + width] : acc[i];
      acc[i] = keep + __shfl_xor(send, width);
    }
  }
  T t = acc[0];
  t += __shfl_xor(t, 16);
  t += __shfl_xor(t, 32);
  return t;
}

// Row dots: lane l < RPW returns sum_c R[l][c] x[c] (x: 256 entries);
// MASK_UPPER keeps c <= rloc0 + row (the lower triangle of a diagonal block).
template <typename T, bool MASK_UPPER>
__device__ __forceinline__ T rowdot(const Rows<T>& R, const T* x, int rloc0, int l) {
  typedef typename V4<T>::type v4;
  const v4 xv = *reinterpret_cast<const v4*>(x + 4 * l);
  T t = rowdot16<T, MASK_UPPER>(R, 0, xv, rloc0, l);
#pragma unroll
  for (int q = 16; q < Rows<T>::R; q += 16) {
    const T u = rowdot16<T, MASK_UPPER>(R, q, xv, rloc0, l);
    if ((l >> 4) == q / 16) t = u;
  }
  return t;
}

// Column dots of a wave: acc (columns 4l..4l+3) += sum_r R[r][.] x[r], x the
// wave's RPW entries; MASK_LOWER keeps rloc0 + r >= column.
template <typename T, bool MASK_LOWER>
__device__ __forceinline__ void coldot(const Rows<T>& R, const T* x, int rloc0, int l, typename V4<T>::type& acc) {
  typedef typename V4<T>::type v4;
#pragma unroll
  for (int r = 0; r < Rows<T>::R; ++r) {
    v4 a = R.v[r];
    const T xr = x[r];
    if (MASK_LOWER) {
      const int row = rloc0 + r, c = 4 * l;
      a.x = row >= c ? a.x : T(0);
      a.y = row >= c + 1 ? a.y : T(0);
      a.z = row >= c + 2 ? a.z : T(0);
      a.w = row >= c + 3 ? a.w : T(0);
    }
    acc.x = fma(a.x, xr, acc.x);
    acc.y = fma(a.y, xr, acc.y);
    acc.z = fma(a.z, xr, acc.z);
    acc.w = fma(a.w, xr, acc.w);
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

// PF: load the step's matrix rows before its waits (FP32)
template <typename T, bool PF>
__global__ __launch_bounds__(Shape<T>::NT) __attribute__((amdgpu_waves_per_eu(Shape<T>::WPE, Shape<T>::WPE)))
void trsv_fwd_kernel(const T* __restrict__ L, int64_t lda, const T* __restrict__ planes, T* b, T* cbuf, T* part,
                     TrsvCtl ctl) {
  __shared__ int s_id, s_ok, s_ok2;
  __shared__ __attribute__((aligned(32))) T tv[NB];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  constexpr int RPW = Shape<T>::RPW, NW = Shape<T>::NW, NT = Shape<T>::NT;
  const int row = RPW * w + (l & (RPW - 1));     // row of the 256-block reduced into lanes l < RPW
  Item it;
  if (!take_ticket<T>(ctl, &it, &s_id)) return;
  const int64_t k0 = (int64_t)it.k * NB;
  const T* Dinv = planes + k0 * NB;
  const T* G = planes + (int64_t)ctl.nb * NB * NB + k0 * NB;
  Rows<T> R;
  if (it.s >= 0) {                                            // far chunk: sum_j L_kj y_j
    const int slot = ctl.rowbase[it.k] + it.s;
    const bool pre = PF && it.jhi - it.jlo == 1;
    if constexpr (PF) {
      if (pre) { loadr(R, L, lda, k0 + RPW * w, (int64_t)it.jlo * NB, l); pin(R); }
    }
    if (group_wait(ctl.yflag + it.jhi - 1, ctl.epoch, ctl.err, &s_ok)) {
      T sum = T(0);
      if (pre) {
        sum = rowdot<T, false>(R, b + (int64_t)it.jlo * NB, 0, l);
      } else {
        #pragma unroll 1
        for (int j = it.jlo; j < it.jhi; ++j) {
          loadr(R, L, lda, k0 + RPW * w, (int64_t)j * NB, l);
          sum += rowdot<T, false>(R, b + (int64_t)j * NB, 0, l);
        }
      }
      if (l < RPW) part[(int64_t)slot * NB + row] = sum;
    }
    group_publish(ctl.pflag + slot, ctl.epoch);
    return;
  }
  if (it.s == CSTEP) {                                        // c_k = Dinv_k (b_k - far partials)
    if constexpr (PF) { loadr(R, Dinv, NB, RPW * w, 0, l); pin(R); }
    const int f0 = ctl.rowbase[it.k], nfar = ctl.rowbase[it.k + 1] - f0;
    T t = b[k0 + row];
    bool ok = true;
    if (nfar > 1) {                                           // the far partials first, the nearest last
      ok = group_wait_all(ctl.pflag + f0 + 1, nfar - 1, ctl.epoch, ctl.err, &s_ok2);
      if (ok)
        for (int s = 1; s < nfar; ++s) t -= part[(int64_t)(f0 + s) * NB + row];
    }
    if (ok && nfar > 0) {
      ok = group_wait(ctl.pflag + f0, ctl.epoch, ctl.err, &s_ok);
      if (ok) t -= part[(int64_t)f0 * NB + row];
    }
    if (ok) {
      if (l < RPW) tv[row] = t;
      __syncthreads();
      if (!PF) loadr(R, Dinv, NB, RPW * w, 0, l);
      const T c = rowdot<T, true>(R, tv, RPW * w, l);
      if (l < RPW) cbuf[k0 + row] = c;
    }
    group_publish(ctl.cflag + it.k, ctl.epoch);
    return;
  }
  // chain step: y_k = c_k - G_k y_{k-1}
  if constexpr (PF) {
    if (it.k > 0) { loadr(R, G, NB, RPW * w, 0, l); pin(R); }
  }
  bool ok = group_wait(ctl.cflag + it.k, ctl.epoch, ctl.err, &s_ok);
  T y = T(0);
  if (ok) { y = cbuf[k0 + row]; pin1(y); }         // c_k lands before the wait for y_{k-1}
  if (ok && it.k > 0) ok = group_wait(ctl.yflag + it.k - 1, ctl.epoch, ctl.err, &s_ok2);
  if (ok) {
    if (it.k > 0) {
      if (!PF) loadr(R, G, NB, RPW * w, 0, l);
      y -= rowdot<T, false>(R, b + k0 - NB, 0, l);
    }
    if (l < RPW) b[k0 + row] = y;
  }
  group_publish(ctl.yflag + it.k, ctl.epoch);
}

template <typename T, bool PF>
__global__ __launch_bounds__(Shape<T>::NT) __attribute__((amdgpu_waves_per_eu(Shape<T>::WPE, Shape<T>::WPE)))
void trsv_bwd_kernel(const T* __restrict__ L, int64_t lda, const T* __restrict__ planes, T* b, T* cbuf, T* part,
                     TrsvCtl ctl) {
  typedef typename V4<T>::type v4;
  __shared__ int s_id, s_ok, s_ok2;
  __shared__ __attribute__((aligned(32))) T tv[NB];
  constexpr int RPW = Shape<T>::RPW, NW = Shape<T>::NW, NT = Shape<T>::NT;
  __shared__ __attribute__((aligned(32))) T red[NW][NB];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  Item it;
  if (!take_ticket<T>(ctl, &it, &s_id)) return;
  const int64_t k0 = (int64_t)it.k * NB;
  const T* Dinv = planes + k0 * NB;
  const T* H = planes + 2 * (int64_t)ctl.nb * NB * NB + k0 * NB;
  // column c = threadIdx.x (< 256) gets the sum of the 16 waves' partials
  auto reduce_waves = [&](v4 a) -> T {
    *reinterpret_cast<v4*>(&red[w][4 * l]) = a;
    __syncthreads();
    T s = T(0);
    if (threadIdx.x < NB) {
#pragma unroll
      for (int q = 0; q < NW; ++q) s += red[q][threadIdx.x];
    }
    __syncthreads();
    return s;
  };
  Rows<T> R;
  v4 acc = {T(0), T(0), T(0), T(0)};
  if (it.s >= 0) {                                            // far chunk: sum_j L_jk^T x_j
    const int slot = ctl.rowbase[it.k] + it.s;
    const bool pre = PF && it.jhi - it.jlo == 1;
    if constexpr (PF) {
      if (pre) { loadr(R, L, lda, (int64_t)it.jlo * NB + RPW * w, k0, l); pin(R); }
    }
    if (group_wait(ctl.yflag + it.jlo, ctl.epoch, ctl.err, &s_ok)) {
      #pragma unroll 1
      for (int j = it.jlo; j < it.jhi; ++j) {
        if (!pre) loadr(R, L, lda, (int64_t)j * NB + RPW * w, k0, l);
        coldot<T, false>(R, b + (int64_t)j * NB + RPW * w, 0, l, acc);
      }
      const T sum = reduce_waves(acc);
      if (threadIdx.x < NB) part[(int64_t)slot * NB + threadIdx.x] = sum;
    }
    group_publish(ctl.pflag + slot, ctl.epoch);
    return;
  }
  if (it.s == CSTEP) {                                        // c'_k = Dinv_k^T (y_k - far partials)
    if constexpr (PF) { loadr(R, Dinv, NB, RPW * w, 0, l); pin(R); }
    const int f0 = ctl.rowbase[it.k], nfar = ctl.rowbase[it.k + 1] - f0;
    T t = threadIdx.x < NB ? b[k0 + threadIdx.x] : T(0);
    bool ok = true;
    if (nfar > 1) {
      ok = group_wait_all(ctl.pflag + f0 + 1, nfar - 1, ctl.epoch, ctl.err, &s_ok2);
      if (ok && threadIdx.x < NB)
        for (int s = 1; s < nfar; ++s) t -= part[(int64_t)(f0 + s) * NB + threadIdx.x];
    }
    if (ok && nfar > 0) {
      ok = group_wait(ctl.pflag + f0, ctl.epoch, ctl.err, &s_ok);
      if (ok && threadIdx.x < NB) t -= part[(int64_t)f0 * NB + threadIdx.x];
    }
    if (ok) {
      if (threadIdx.x < NB) tv[threadIdx.x] = t;
      __syncthreads();
      if (!PF) loadr(R, Dinv, NB, RPW * w, 0, l);
      coldot<T, true>(R, tv + RPW * w, RPW * w, l, acc);
      const T c = reduce_waves(acc);
      if (threadIdx.x < NB) cbuf[k0 + threadIdx.x] = c;
    }
    group_publish(ctl.cflag + it.k, ctl.epoch);
    return;
  }
  // chain step: x_k = c'_k - H_k^T x_{k+1}
  const bool last = it.k + 1 >= ctl.nb;
  if constexpr (PF) {
    if (!last) { loadr(R, H, NB, RPW * w, 0, l); pin(R); }
  }
  bool ok = group_wait(ctl.cflag + it.k, ctl.epoch, ctl.err, &s_ok);
  T cv = T(0);
  if (ok && threadIdx.x < NB) { cv = cbuf[k0 + threadIdx.x]; pin1(cv); }
  if (ok && !last) ok = group_wait(ctl.yflag + it.k + 1, ctl.epoch, ctl.err, &s_ok2);
  if (ok) {
    T x = T(0);
    if (!last) {
      if (!PF) loadr(R, H, NB, RPW * w, 0, l);
      coldot<T, false>(R, b + k0 + NB + RPW * w, 0, l, acc);
      x = reduce_waves(acc);
    }
    if (threadIdx.x < NB) b[k0 + threadIdx.x] = cv - x;
  }
  group_publish(ctl.yflag + it.k, ctl.epoch);
}

// Host-side work tables.  Forward: c steps of rows 0 and 1, then for each block
// j: its chain step, the far chunks whose nearest block is j (row k = j + lo_s,
// blocks at distances [lo_s, hi_s)), and the c step of row j + 2 (all its far
// chunks are then queued).  Backward mirrors it with k = j - lo_s.
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
  fwd.push_back({0, CSTEP, 0, 0});
  if (nb > 1) fwd.push_back({1, CSTEP, 0, 0});
  for (int j = 0; j < nb; ++j) {
    fwd.push_back({j, CHAIN, 0, 0});
    for (int s = 0;; ++s) {
      int lo, hi;
      chunk_range(s, &lo, &hi);
      const int k = j + lo;
      if (k >= nb) break;
      fwd.push_back({k, s, std::max(0, k - hi + 1), k - lo + 1});
    }
    if (j + 2 < nb) fwd.push_back({j + 2, CSTEP, 0, 0});
  }
  bwd.push_back({nb - 1, CSTEP, 0, 0});
  if (nb > 1) bwd.push_back({nb - 2, CSTEP, 0, 0});
  for (int j = nb - 1; j >= 0; --j) {
    bwd.push_back({j, CHAIN, 0, 0});
    for (int s = 0;; ++s) {
      int lo, hi;
      chunk_range(s, &lo, &hi);
      const int k = j - lo;
      if (k < 0) break;
      bwd.push_back({k, s, k + lo, std::min(nb, k + hi)});
    }
    if (j - 2 >= 0) bwd.push_back({j - 2, CSTEP, 0, 0});
  }
}

}  // namespace

// Workspace: [fwd items][bwd items][fwd base][bwd base][yflag][cflag][pflag][ticket][err][c][partials]
struct TrsvLayout {
  size_t o_fi, o_bi, o_fb, o_bb, o_yf, o_cf, o_pf, o_tk, o_err, o_c, o_part, bytes;
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
  L.o_part = take(sizeof(double) * NB * (size_t)std::max(slots, 1));
  L.bytes = o;
  L.nfwd = nfwd; L.nbwd = nbwd; L.slots = slots;
  return L;
}

static int trsv_prepare(gk_handle h, int nb, hipStream_t st, char** base, TrsvLayout* lay) {
  std::vector<Item> fwd, bwd;
  std::vector<int> fb, bb;
  const bool rebuild = h->trsv_nb != nb || h->ws[WS_TRSV] == nullptr;
  if (rebuild) build_tables(nb, fwd, fb, bwd, bb);
  const int slots = rebuild ? std::max(fb[nb], bb[nb]) : h->trsv_slots;
  const int nfwd = rebuild ? (int)fwd.size() : h->trsv_nfwd;
  const int nbwd = rebuild ? (int)bwd.size() : h->trsv_nbwd;
  *lay = trsv_layout(nb, nfwd, nbwd, slots);
  void* p = nullptr;
  const void* before = h->ws[WS_TRSV];
  GK_TRY(gk_workspace(h, WS_TRSV, lay->bytes, &p));
  *base = static_cast<char*>(p);
  if (rebuild || p != before) {
    if (!rebuild) build_tables(nb, fwd, fb, bwd, bb);
    GK_HIP(hipMemcpyAsync(*base + lay->o_fi, fwd.data(), sizeof(Item) * fwd.size(), hipMemcpyHostToDevice, st));
    GK_HIP(hipMemcpyAsync(*base + lay->o_bi, bwd.data(), sizeof(Item) * bwd.size(), hipMemcpyHostToDevice, st));
    GK_HIP(hipMemcpyAsync(*base + lay->o_fb, fb.data(), sizeof(int) * fb.size(), hipMemcpyHostToDevice, st));
    GK_HIP(hipMemcpyAsync(*base + lay->o_bb, bb.data(), sizeof(int) * bb.size(), hipMemcpyHostToDevice, st));
    GK_HIP(hipMemsetAsync(*base + lay->o_yf, 0, lay->o_c - lay->o_yf, st));   // flags, ticket, error
    GK_HIP(hipStreamSynchronize(st));          // host tables go out of scope
    h->trsv_nb = nb;
    h->trsv_slots = slots;
    h->trsv_nfwd = nfwd;
    h->trsv_nbwd = nbwd;
    h->trsv_epoch = 0;
    h->trsv_tickets = 0;
  }
  return GK_OK;
}

template <typename T>
int gk_trsv_chol_impl(gk_handle h, const T* L, const T* planes, int64_t lda, T* b, hipStream_t st) {
  GK_REQUIRE(h && L && planes && b && lda > 0 && lda % NB == 0, GK_ERR_ARG, "triangular solve: bad arguments");
  constexpr bool PF = sizeof(T) == 4;          // FP32 block rows fit in registers
  const int nb = (int)(lda / NB);
  char* base = nullptr;
  TrsvLayout lay;
  GK_TRY(trsv_prepare(h, nb, st, &base, &lay));
  TrsvCtl c;
  c.yflag = reinterpret_cast<unsigned*>(base + lay.o_yf);
  c.cflag = reinterpret_cast<unsigned*>(base + lay.o_cf);
  c.pflag = reinterpret_cast<unsigned*>(base + lay.o_pf);
  c.ticket = reinterpret_cast<unsigned long long*>(base + lay.o_tk);
  c.err = reinterpret_cast<int*>(base + lay.o_err);
  c.nb = nb;
  T* cbuf = reinterpret_cast<T*>(base + lay.o_c);
  T* part = reinterpret_cast<T*>(base + lay.o_part);
  // forward
  c.items = reinterpret_cast<const Item*>(base + lay.o_fi);
  c.rowbase = reinterpret_cast<const int*>(base + lay.o_fb);
  c.ticket_base = h->trsv_tickets;
  c.epoch = ++h->trsv_epoch;
  c.n_items = lay.nfwd;
  hipEvent_t pa;
  const bool prof = gk_prof_begin(st, &pa);
  hipLaunchKernelGGL((trsv_fwd_kernel<T, PF>), dim3((unsigned)lay.nfwd), dim3(Shape<T>::NT), 0, st, L, lda, planes, b, cbuf,
                     part, c);
  h->trsv_tickets += (unsigned long long)lay.nfwd;
  // backward
  c.items = reinterpret_cast<const Item*>(base + lay.o_bi);
  c.rowbase = reinterpret_cast<const int*>(base + lay.o_bb);
  c.ticket_base = h->trsv_tickets;
  c.epoch = ++h->trsv_epoch;
  c.n_items = lay.nbwd;
  hipLaunchKernelGGL((trsv_bwd_kernel<T, PF>), dim3((unsigned)lay.nbwd), dim3(Shape<T>::NT), 0, st, L, lda, planes, b, cbuf,
                     part, c);
  h->trsv_tickets += (unsigned long long)lay.nbwd;
  if (prof)      // each direction reads the strict lower block triangle and two planes of diagonal blocks
    gk_prof_end(st, pa, sizeof(T) == 8 ? 9 : 8,
                2.0 * sizeof(T) * ((double)nb * (nb - 1) / 2.0 * NB * NB + 2.0 * nb * NB * NB));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    h->trsv_nb = 0;              // counters no longer in step: rebuild on the next call
    GK_REQUIRE(false, GK_ERR_HIP, "triangular solve launch: %s", hipGetErrorString(e));
  }
  return GK_OK;
}

template int gk_trsv_chol_impl<float>(gk_handle, const float*, const float*, int64_t, float*, hipStream_t);
template int gk_trsv_chol_impl<double>(gk_handle, const double*, const double*, int64_t, double*, hipStream_t);

// Error word of the sync-free solves (spin limit reached); cleared by reading.
int gk_trsv_check(gk_handle h, hipStream_t st) {
  if (!h->ws[WS_TRSV] || h->trsv_nb == 0) return GK_OK;
  TrsvLayout lay = trsv_layout(h->trsv_nb, h->trsv_nfwd, h->trsv_nbwd, h->trsv_slots);
  char* base = static_cast<char*>(h->ws[WS_TRSV]);
  int err = 0;
  GK_HIP(hipMemcpyAsync(&err, base + lay.o_err, sizeof(int), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  if (err) {
    h->trsv_nb = 0;              // reset tables, flags and counters on the next call
    GK_REQUIRE(false, GK_ERR_HIP, err == 1 ? "triangular solve: a readiness wait reached its spin limit"
                                           : "triangular solve: work ticket out of range");
  }
  return GK_OK;
}
