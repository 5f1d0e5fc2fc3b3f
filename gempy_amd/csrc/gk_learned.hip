// Learned sparse approximate inverse (SPAI) of the equilibrated covariance
// block H = D A_cc D, for the matrix-free GMRES of the largest systems.
//
// Pattern: the data points (orientations, then rest interface points) each own
// a 3-dim block of unknowns (an orientation: its x, y, z gradient unknowns; an
// interface point: its one increment unknown in slot 0, slots 1-2 dummy).  Row
// block P of M has a 3x3 block for each of P's K nearest data points Q.  A small
// MLP maps pair features of (P, Q) to that 3x3 block; it is trained to minimise
// the local SPAI residual || M_P H_loc(P) - E_P ||_F^2 (H_loc = H restricted to
// the unknowns of P's neighbourhood), whose minimiser is the neighbourhood
// inverse row (the local cardinal-function preconditioner).
#include <algorithm>
#include "gk_internal.h"
#include "gk_entry.h"

extern "C" int gk_source_box(gk_handle h, const gk_system* sys, double* box, gk_stream stream);

using namespace gk_detail;

namespace {

constexpr int K = GK_SPAI_K;          // neighbours per point (including itself)
constexpr int F = GK_SPAI_FEATURES;   // features per pair
constexpr int H1 = GK_SPAI_HIDDEN;
constexpr int H2 = GK_SPAI_HIDDEN;
constexpr int O = 9;
constexpr int NPARAM = H1 * F + H1 + H2 * H1 + H2 + O * H2 + O;
constexpr int LD = 3 * K;             // local block dimension

__device__ __forceinline__ void point_xyz(const gk_system& s, int64_t p, double* o) {
  const double* q = p < s.n_ori ? s.ori_xyz + 3 * p : s.rest_xyz + 3 * (p - s.n_ori);
  o[0] = q[0]; o[1] = q[1]; o[2] = q[2];
}

// unknown index of dimension d of data point p (-1 for the dummy slots)
__device__ __forceinline__ int64_t unknown_of(const gk_system& s, int64_t p, int d) {
  if (p < s.n_ori) return (int64_t)d * s.n_ori + p;
  return d == 0 ? 3 * s.n_ori + (p - s.n_ori) : -1;
}

// ---------------------------------------------------------------- neighbours
template <int K>
__global__ __launch_bounds__(256) void knn_kernel(gk_system sys, int64_t n, int32_t* idx, float* hk) {
  __shared__ float sx[256], sy[256], sz[256];
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const double ia = 1.0 / sys.range;
  float px = 0, py = 0, pz = 0;
  if (p < n) {
    double x[3];
    point_xyz(sys, p, x);
    px = (float)((x[0] - sys.center[0]) * ia);
    py = (float)((x[1] - sys.center[1]) * ia);
    pz = (float)((x[2] - sys.center[2]) * ia);
  }
  float bd[K];
  int32_t bi[K];
#pragma unroll
  for (int j = 0; j < K; ++j) { bd[j] = 3.4e38f; bi[j] = -1; }
  for (int64_t base = 0; base < n; base += 256) {
    __syncthreads();
    const int64_t q = base + threadIdx.x;
    if (q < n) {
      double x[3];
      point_xyz(sys, q, x);
      sx[threadIdx.x] = (float)((x[0] - sys.center[0]) * ia);
      sy[threadIdx.x] = (float)((x[1] - sys.center[1]) * ia);
      sz[threadIdx.x] = (float)((x[2] - sys.center[2]) * ia);
    }
    __syncthreads();
    const int cnt = (int)std::min<int64_t>(256, n - base);
    for (int j = 0; j < cnt; ++j) {
      const float dx = sx[j] - px, dy = sy[j] - py, dz = sz[j] - pz;
      float d = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      if (d < bd[K - 1]) {          // insertion into the sorted list (ties keep the lower index)
        int32_t id = (int32_t)(base + j);
#pragma unroll
        for (int t = 0; t < K; ++t) {
          if (d < bd[t]) {
            const float td = bd[t];
            const int32_t ti = bi[t];
            bd[t] = d; bi[t] = id;
            d = td; id = ti;
          }
        }
      }
    }
  }
  if (p < n) {
#pragma unroll
    for (int j = 0; j < K; ++j) idx[p * K + j] = bi[j];
    hk[p] = sqrtf(bd[K - 1]);
  }
}

#ifndef GK_LI_GRID
#define GK_LI_GRID 128          // workgroups of the FP32 form's local-inverse kernel
#endif

// ------------------------------------------------------- grid neighbours
// The exact K nearest data points of every data point (itself first) through a
// uniform grid of cells (~2 points per cell on average): a query scans cubes of
// cells of growing radius r (only the new shell each time) until its K-th
// distance is within r cell sizes, which every point closer than that lies
// inside.  Ties are broken by the lower index, as in the brute-force search.
struct KnnGrid {
  float lo[3];
  float inv_c, c;
  int n[3];
};

__device__ __forceinline__ void norm_point(const gk_system& s, int64_t p, double ia, float* o) {
  double x[3];
  point_xyz(s, p, x);
  o[0] = (float)((x[0] - s.center[0]) * ia);
  o[1] = (float)((x[1] - s.center[1]) * ia);
  o[2] = (float)((x[2] - s.center[2]) * ia);
}

__device__ __forceinline__ int cell_coord(float v, float lo, float inv_c, int n) {
  const int c = (int)floorf((v - lo) * inv_c);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

__global__ void knn_cells_kernel(gk_system sys, int64_t n, KnnGrid g, int32_t* cell, int32_t* count) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float x[3];
  norm_point(sys, p, 1.0 / sys.range, x);
  const int cx = cell_coord(x[0], g.lo[0], g.inv_c, g.n[0]), cy = cell_coord(x[1], g.lo[1], g.inv_c, g.n[1]),
            cz = cell_coord(x[2], g.lo[2], g.inv_c, g.n[2]);
  const int c = (cz * g.n[1] + cy) * g.n[0] + cx;
  cell[p] = c;
  atomicAdd(count + c, 1);
}

// exclusive scan of count[0..nc) into start[0..nc] (one workgroup of 1024)
__global__ __launch_bounds__(1024) void knn_scan_kernel(const int32_t* count, int32_t nc, int32_t* start) {
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  const int per = (nc + 1023) / 1024;
  const int b = t * per, e = min(nc, b + per);
  int32_t sum = 0;
  for (int i = b; i < e; ++i) sum += count[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int32_t v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int32_t run = part[t] - sum;
  for (int i = b; i < e; ++i) {
    start[i] = run;
    run += count[i];
  }
  if (t == 1023) start[nc] = part[1023];
}

__global__ void knn_scatter_kernel(gk_system sys, int64_t n, const int32_t* cell, const int32_t* start,
                                   int32_t* fill, float4* spos) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int c = cell[p];
  const int slot = start[c] + atomicAdd(fill + c, 1);
  float x[3];
  norm_point(sys, p, 1.0 / sys.range, x);
  spos[slot] = make_float4(x[0], x[1], x[2], __int_as_float((int)p));
}

template <int K>
__global__ __launch_bounds__(256) void knn_grid_kernel(gk_system sys, int64_t n, KnnGrid g, const int32_t* start,
                                                       const float4* spos, int32_t* idx, float* hk) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float x[3];
  norm_point(sys, p, 1.0 / sys.range, x);
  const int cx = cell_coord(x[0], g.lo[0], g.inv_c, g.n[0]), cy = cell_coord(x[1], g.lo[1], g.inv_c, g.n[1]),
            cz = cell_coord(x[2], g.lo[2], g.inv_c, g.n[2]);
  float bd[K];
  int32_t bi[K];
#pragma unroll
  for (int j = 0; j < K; ++j) { bd[j] = 3.4e38f; bi[j] = 0x7fffffff; }
  const int rmax = max(g.n[0], max(g.n[1], g.n[2]));
  for (int r = 0; r <= rmax; ++r) {
    for (int dz = -r; dz <= r; ++dz) {
      const int z = cz + dz;
      if (z < 0 || z >= g.n[2]) continue;
      for (int dy = -r; dy <= r; ++dy) {
        const int y = cy + dy;
        if (y < 0 || y >= g.n[1]) continue;
        const bool inner = abs(dz) < r && abs(dy) < r;
        for (int dx = -r; dx <= r; dx += (inner ? 2 * r : 1)) {   // only the shell of the cube
          const int xx = cx + dx;
          if (r > 0 && inner && dx != -r && dx != r) continue;
          if (xx < 0 || xx >= g.n[0]) continue;
          const int c = (z * g.n[1] + y) * g.n[0] + xx;
          for (int s = start[c], e = start[c + 1]; s < e; ++s) {
            const float4 q = spos[s];
            const float ddx = q.x - x[0], ddy = q.y - x[1], ddz = q.z - x[2];
            float d = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
            int32_t id = __float_as_int(q.w);
            if (d < bd[K - 1] || (d == bd[K - 1] && id < bi[K - 1])) {
#pragma unroll
              for (int t = 0; t < K; ++t) {
                if (d < bd[t] || (d == bd[t] && id < bi[t])) {
                  const float td = bd[t];
                  const int32_t ti = bi[t];
                  bd[t] = d; bi[t] = id;
                  d = td; id = ti;
                }
              }
            }
          }
          if (r == 0) break;
        }
      }
    }
    // every point within r cell sizes of the query has been seen (a margin
    // for the rounding of the cell coordinates)
    const float reach = ((float)r - 1e-3f) * g.c;
    if (bi[K - 1] != 0x7fffffff && bd[K - 1] <= reach * reach) break;
  }
#pragma unroll
  for (int j = 0; j < K; ++j) idx[p * K + j] = bi[j];
  hk[p] = sqrtf(bd[K - 1]);
}

// ------------------------------------------------------------------ features
// f = [ (q-p)/h_P (3), |q-p|/h_P, P is orientation, Q is orientation, P == Q, 10 h_P ]
__global__ void features_kernel(gk_system sys, int64_t nq, const int32_t* pts, const int32_t* idx,
                                const float* hk, float* feat) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq * K) return;
  const int64_t b = t / K;
  const int j = (int)(t % K);
  const int64_t p = pts ? pts[b] : b;
  const int64_t q = idx[p * K + j];
  double xp[3], xq[3];
  point_xyz(sys, p, xp);
  point_xyz(sys, q, xq);
  const double ia = 1.0 / sys.range;
  const double h = std::max((double)hk[p], 1e-12);
  const double d0 = (xq[0] - xp[0]) * ia / h, d1 = (xq[1] - xp[1]) * ia / h, d2 = (xq[2] - xp[2]) * ia / h;
  float* f = feat + t * F;
  f[0] = (float)d0; f[1] = (float)d1; f[2] = (float)d2;
  f[3] = (float)sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  f[4] = p < sys.n_ori ? 1.0f : 0.0f;
  f[5] = q < sys.n_ori ? 1.0f : 0.0f;
  f[6] = p == q ? 1.0f : 0.0f;
  f[7] = (float)(10.0 * h);
}

// ------------------------------------------------------------- local blocks
// H_loc of point pts[b]: (3K x 3K) equilibrated covariance among the unknowns of
// its neighbourhood; dummy slots are identity rows/columns.
__global__ __launch_bounds__(256) void local_blocks_kernel(Geo G, int64_t nq, const int32_t* pts, const int32_t* idx,
                                                           const double* dscale, float* Hloc) {
  const int64_t b = blockIdx.x;
  const int64_t p = pts ? pts[b] : b;
  __shared__ int64_t un[LD];
  if (threadIdx.x < LD) {
    const int j = threadIdx.x / 3, d = threadIdx.x % 3;
    un[threadIdx.x] = unknown_of(G.s, idx[p * K + j], d);
  }
  __syncthreads();
  float* Hb = Hloc + b * LD * LD;
  for (int e = threadIdx.x; e < LD * LD; e += 256) {
    const int r = e / LD, c = e % LD;
    const int64_t ur = un[r], uc = un[c];
    double v;
    if (ur < 0 || uc < 0) v = (r == c) ? 1.0 : 0.0;
    else v = entry(G, ur, uc) * dscale[ur] * dscale[uc];
    Hb[e] = (float)v;
  }
}

// Exact local sparse approximate inverse: for point p, the rows of
// H_loc^-1 that belong to p's own unknowns, as K 3x3 blocks Y[p][j] (the same
// layout the network predicts).  H_loc (FP64, equilibrated, dummy slots
// identity) is symmetric positive definite; its unknowns are ordered with p's
// own three last, so with H_loc = L L^T the wanted columns H_loc^-1 e_s are
// L^-T z with z = L^-1 e_s nonzero only in its last three entries (the inverse
// of L's last 3x3 block).  Cholesky in LDS (LD^3/6 multiply-adds, a third of a
// Gauss-Jordan elimination's), then one wave solves L^T x = z for the three
// columns with broadcasts instead of barriers.
template <int K>
__global__ __launch_bounds__(256) void local_inverse_kernel(Geo G, int64_t n, const int32_t* idx,
                                                            const double* dscale, float* Y) {
  constexpr int LD = 3 * K;
  __shared__ double A[LD][LD + 1];
  __shared__ int64_t un[LD];
  __shared__ int orig[LD];                // compact row -> slot row (3 slot + component)
  __shared__ int self, nc;
  const int tid = threadIdx.x;
  // a fixed grid walks the points: the kernel holds at most gridDim.x CUs'
  // worth of LDS, so a factorization running beside it keeps the rest
  for (int64_t p = blockIdx.x; p < n; p += gridDim.x) {
  __syncthreads();
  if (tid == 0) {
    self = 0;
    for (int j = 0; j < K; ++j)
      if (idx[p * K + j] == p) { self = j; break; }
  }
  __syncthreads();
  // slot s < K - 1: neighbour s (s + 1 past self); slot K - 1: p itself.
  // A rest point has one unknown of its three slot rows; the others are
  // identity rows, which decouple exactly, so the matrix is factored on the
  // real unknowns only (the three rows of p itself are always kept, last).
  if (tid < LD) {
    const int sl = tid / 3, d = tid % 3;
    const int j = sl == K - 1 ? self : (sl < self ? sl : sl + 1);
    un[tid] = unknown_of(G.s, idx[p * K + j], d);
  }
  __syncthreads();
  if (tid == 0) {
    int c = 0;
    for (int r = 0; r < LD; ++r)
      if (un[r] >= 0 || r >= LD - 3) orig[c++] = r;
    nc = c;
  }
  __syncthreads();
  const int m_all = nc;
  for (int e = tid; e < m_all * m_all; e += 256) {
    const int r = e / m_all, c = e % m_all;
    if (c > r) continue;
    const int64_t ur = un[orig[r]], uc = un[orig[c]];
    A[r][c] = (ur < 0 || uc < 0) ? (r == c ? 1.0 : 0.0) : entry(G, ur, uc) * dscale[ur] * dscale[uc];
  }
  __syncthreads();
  for (int k = 0; k < m_all; ++k) {
    if (tid == 0) A[k][k] = sqrt(A[k][k]);
    __syncthreads();
    const double ip = 1.0 / A[k][k];
    for (int r = k + 1 + tid; r < m_all; r += 256) A[r][k] *= ip;
    __syncthreads();
    // trailing lower triangle: rows k+1+i (16 thread groups), columns k+1+t <= row
    const int m = m_all - k - 1;
    for (int i = tid >> 4; i < m; i += 16) {
      const double ark = A[k + 1 + i][k];
      for (int t = tid & 15; t <= i; t += 16) A[k + 1 + i][k + 1 + t] -= ark * A[k + 1 + t][k];
    }
    __syncthreads();
  }
  if (tid < 64) {
    // z = L^-1 e_{n-3+q}: only z[n-3 .. n-1] are nonzero (forward substitution
    // on L's last 3x3 block), then x = L^-T z; lane l owns rows l, l + 64
    constexpr int RPL = (LD + 63) / 64;
    double x[RPL][3];
    const int b0 = m_all - 3;
    const double l00 = A[b0][b0], l10 = A[b0 + 1][b0], l11 = A[b0 + 1][b0 + 1];
    const double l20 = A[b0 + 2][b0], l21 = A[b0 + 2][b0 + 1], l22 = A[b0 + 2][b0 + 2];
    double z[3][3];                      // z[q][t]: column q, entry b0 + t
    for (int q = 0; q < 3; ++q) {
      const double e0 = q == 0, e1 = q == 1, e2 = q == 2;
      z[q][0] = e0 / l00;
      z[q][1] = (e1 - l10 * z[q][0]) / l11;
      z[q][2] = (e2 - l20 * z[q][0] - l21 * z[q][1]) / l22;
    }
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int r = tid + 64 * u;
#pragma unroll
      for (int q = 0; q < 3; ++q) x[u][q] = r >= b0 && r < m_all ? z[q][r - b0] : 0.0;
    }
    for (int k = m_all - 1; k >= 0; --k) {
      const double ikk = 1.0 / A[k][k];
      const int owner = k & 63, uo = k >> 6;
      double xk[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double v = 0.0;
#pragma unroll
        for (int u = 0; u < RPL; ++u)
          if (u == uo) v = x[u][q];
        v = __shfl(v, owner) * ikk;          // x_k = (z_k - sum_{i>k} L_ik x_i) / L_kk
        xk[q] = v;
      }
#pragma unroll
      for (int u = 0; u < RPL; ++u) {
        const int r = tid + 64 * u;
        if (r == k) {
#pragma unroll
          for (int q = 0; q < 3; ++q) x[u][q] = xk[q];
        } else if (r < k) {
          const double lkr = A[k][r];
#pragma unroll
          for (int q = 0; q < 3; ++q) x[u][q] -= lkr * xk[q];
        }
      }
    }
    // Y[p][j][3 r + c] = (H_loc^-1)[3 self + r][3 j + c] = x_r[slot(j), c];
    // the rows of identity unknowns (not factored) hold 0
#pragma unroll
    for (int u = 0; u < RPL; ++u) {
      const int row = tid + 64 * u;
      if (row >= m_all) continue;
      const int orow = orig[row];
      const int sl = orow / 3, c = orow % 3;
      const int j = sl == K - 1 ? self : (sl < self ? sl : sl + 1);
#pragma unroll
      for (int q = 0; q < 3; ++q) Y[((int64_t)p * K + j) * 9 + 3 * q + c] = (float)x[u][q];
    }
  }
  // zeros for the slot rows left out (identity unknowns of neighbours)
  for (int row = tid; row < LD - 3; row += 256) {
    if (un[row] >= 0) continue;
    const int sl = row / 3, c = row % 3;
    const int j = sl < self ? sl : sl + 1;
#pragma unroll
    for (int q = 0; q < 3; ++q) Y[((int64_t)p * K + j) * 9 + 3 * q + c] = 0.0f;
  }
  }
}

// ---------------------------------------------------------------------- MLP
struct MlpPtrs {
  const float *W1, *b1, *W2, *b2, *W3, *b3;
};
__device__ __forceinline__ MlpPtrs mlp_ptrs(const float* P) {
  MlpPtrs m;
  m.W1 = P; m.b1 = m.W1 + H1 * F; m.W2 = m.b1 + H1; m.b2 = m.W2 + H2 * H1; m.W3 = m.b2 + H2; m.b3 = m.W3 + O * H2;
  return m;
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(int64_t n, const float* params, const float* x, float* y,
                                                          float* acts) {
  __shared__ float P[NPARAM];
  for (int i = threadIdx.x; i < NPARAM; i += 256) P[i] = params[i];
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const MlpPtrs m = mlp_ptrs(P);
  float in[F], h1[H1], h2[H2];
#pragma unroll
  for (int f = 0; f < F; ++f) in[f] = x[t * F + f];
  for (int i = 0; i < H1; ++i) {
    float s = m.b1[i];
#pragma unroll
    for (int f = 0; f < F; ++f) s = fmaf(m.W1[i * F + f], in[f], s);
    h1[i] = tanhf(s);
  }
  for (int i = 0; i < H2; ++i) {
    float s = m.b2[i];
    for (int k = 0; k < H1; ++k) s = fmaf(m.W2[i * H1 + k], h1[k], s);
    h2[i] = tanhf(s);
  }
  // the self block is the identity plus the network output, so that an
  // untrained network is the diagonal (Jacobi) preconditioner
  const bool self = in[6] > 0.5f;
  for (int o = 0; o < O; ++o) {
    float s = m.b3[o];
    for (int k = 0; k < H2; ++k) s = fmaf(m.W3[o * H2 + k], h2[k], s);
    y[t * O + o] = s + ((self && (o % 4 == 0)) ? 1.0f : 0.0f);
  }
  if (acts) {
    for (int i = 0; i < H1; ++i) acts[t * (H1 + H2) + i] = h1[i];
    for (int i = 0; i < H2; ++i) acts[t * (H1 + H2) + H1 + i] = h2[i];
  }
}

// Parameter gradients of sum_t dy_t . y_t; LDS accumulation per block, then one
// partial row per block (summed in block order by mlp_reduce_kernel).
__global__ __launch_bounds__(256) void mlp_backward_kernel(int64_t n, const float* params, const float* x,
                                                           const float* acts, const float* dy, float* partial) {
  __shared__ float P[NPARAM];
  __shared__ float Gs[NPARAM];
  for (int i = threadIdx.x; i < NPARAM; i += 256) { P[i] = params[i]; Gs[i] = 0.0f; }
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) {
    const MlpPtrs m = mlp_ptrs(P);
    float* gW1 = Gs; float* gb1 = gW1 + H1 * F; float* gW2 = gb1 + H1; float* gb2 = gW2 + H2 * H1;
    float* gW3 = gb2 + H2; float* gb3 = gW3 + O * H2;
    float in[F], h1[H1], h2[H2], d[O], d2[H2], d1[H1];
    for (int f = 0; f < F; ++f) in[f] = x[t * F + f];
    for (int i = 0; i < H1; ++i) h1[i] = acts[t * (H1 + H2) + i];
    for (int i = 0; i < H2; ++i) h2[i] = acts[t * (H1 + H2) + H1 + i];
    for (int o = 0; o < O; ++o) d[o] = dy[t * O + o];
    for (int k = 0; k < H2; ++k) d2[k] = 0.0f;
    for (int o = 0; o < O; ++o) {
      if (d[o] == 0.0f) continue;
      atomicAdd(&gb3[o], d[o]);
      for (int k = 0; k < H2; ++k) {
        atomicAdd(&gW3[o * H2 + k], d[o] * h2[k]);
        d2[k] = fmaf(m.W3[o * H2 + k], d[o], d2[k]);
      }
    }
    for (int k = 0; k < H1; ++k) d1[k] = 0.0f;
    for (int i = 0; i < H2; ++i) {
      const float g = d2[i] * (1.0f - h2[i] * h2[i]);
      atomicAdd(&gb2[i], g);
      for (int k = 0; k < H1; ++k) {
        atomicAdd(&gW2[i * H1 + k], g * h1[k]);
        d1[k] = fmaf(m.W2[i * H1 + k], g, d1[k]);
      }
    }
    for (int i = 0; i < H1; ++i) {
      const float g = d1[i] * (1.0f - h1[i] * h1[i]);
      atomicAdd(&gb1[i], g);
      for (int f = 0; f < F; ++f) atomicAdd(&gW1[i * F + f], g * in[f]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NPARAM; i += 256) partial[(int64_t)blockIdx.x * NPARAM + i] = Gs[i];
}

__global__ void mlp_reduce_kernel(const float* partial, int64_t nblk, float* grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  float s = 0.0f;
  for (int64_t b = 0; b < nblk; ++b) s += partial[b * NPARAM + i];
  grad[i] = s;
}

// ---------------------------------------------------------------- SPAI loss
// For point pts[b] with outputs Y[b*K + j][9] (3x3 blocks, row-major):
// R = M_P H_loc - E_P on the valid rows of P, loss_b = ||R||^2,
// dY = 2 R H_loc^T scattered back to the blocks (zero on the dummy rows).
__global__ __launch_bounds__(64) void spai_loss_kernel(gk_system sys, int64_t nq, const int32_t* pts,
                                                       const int32_t* idx, const float* Hloc, const float* Y,
                                                       float* dY, float* loss) {
  __shared__ float H[LD][LD + 1];
  __shared__ float M[3][LD];
  __shared__ float R[3][LD];
  __shared__ float red[64];
  const int64_t b = blockIdx.x;
  const int64_t p = pts ? pts[b] : b;
  const int rows = p < sys.n_ori ? 3 : 1;
  for (int e = threadIdx.x; e < LD * LD; e += 64) H[e / LD][e % LD] = Hloc[b * LD * LD + e];
  for (int e = threadIdx.x; e < 3 * LD; e += 64) {
    const int r = e / LD, c = e % LD;
    const int j = c / 3, cc = c % 3;
    M[r][c] = Y[(b * K + j) * O + 3 * r + cc];
  }
  __syncthreads();
  int self = 0;
  for (int j = 0; j < K; ++j) if (idx[p * K + j] == p) { self = j; break; }
  float acc = 0.0f;
  for (int e = threadIdx.x; e < 3 * LD; e += 64) {
    const int r = e / LD, c = e % LD;
    float s = 0.0f;
    for (int k = 0; k < LD; ++k) s = fmaf(M[r][k], H[k][c], s);
    if (c == 3 * self + r) s -= 1.0f;
    if (r >= rows) s = 0.0f;
    R[r][c] = s;
    acc = fmaf(s, s, acc);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[b] = red[0];
  for (int e = threadIdx.x; e < 3 * LD; e += 64) {
    const int r = e / LD, c = e % LD;
    float s = 0.0f;
    for (int k = 0; k < LD; ++k) s = fmaf(R[r][k], H[c][k], s);
    const int j = c / 3, cc = c % 3;
    dY[(b * K + j) * O + 3 * r + cc] = 2.0f * s;
  }
}

__global__ void adam_kernel(float* params, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2, float scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = grad[i] * scale;
  m[i] = b1 * m[i] + (1.0f - b1) * g;
  v[i] = b2 * v[i] + (1.0f - b2) * g * g;
  params[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

// -------------------------------------------------------------- SPAI apply
// y = M u on the covariance unknowns (u, y FP64 of length Nc; blocks FP32).
template <int K>
__global__ void spai_apply_kernel(gk_system sys, int64_t n, const int32_t* idx, const float* Y, const double* u,
                                  double* y) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < K; ++j) {
    const int64_t q = idx[p * K + j];
    double xq[3];
    for (int c = 0; c < 3; ++c) {
      const int64_t uc = unknown_of(sys, q, c);
      xq[c] = uc >= 0 ? u[uc] : 0.0;
    }
    const float* B = Y + (p * K + j) * O;
    for (int r = 0; r < 3; ++r)
      acc[r] += (double)B[3 * r] * xq[0] + (double)B[3 * r + 1] * xq[1] + (double)B[3 * r + 2] * xq[2];
  }
  for (int r = 0; r < 3; ++r) {
    const int64_t ur = unknown_of(sys, p, r);
    if (ur >= 0) y[ur] = acc[r];
  }
}

}  // namespace

// ------------------------------------------------------------------- API
extern "C" int gk_spai_nparams(void) { return NPARAM; }

extern "C" int gk_knn(gk_handle h, const gk_system* sys, int32_t* idx, float* hk, gk_stream stream) {
  GK_REQUIRE(h && sys && idx && hk, GK_ERR_ARG, "gk_knn: NULL argument");
  const int64_t n = sys->n_ori + sys->n_rest;
  GK_REQUIRE(n >= K, GK_ERR_ARG, "gk_knn: needs at least %d data points (got %lld)", K, (long long)n);
  hipLaunchKernelGGL(knn_kernel<K>, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, *sys, n,
                     idx, hk);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

// The neighbourhoods, local inverses and products of the FP32 form's exact
// local inverse, with GK_LOCAL_K neighbours (the network keeps GK_SPAI_K).
// box: bounding box of the data points in the system's coordinates
// (xmin, ymin, zmin, xmax, ymax, zmax), as gk_source_box returns it.
extern "C" int gk_knn_local(gk_handle h, const gk_system* sys, const double* box, int32_t* idx, float* hk,
                            gk_stream stream) {
  GK_REQUIRE(h && sys && box && idx && hk, GK_ERR_ARG, "gk_knn_local: NULL argument");
  const int64_t n = sys->n_ori + sys->n_rest;
  GK_REQUIRE(n >= GK_LOCAL_K, GK_ERR_ARG, "gk_knn_local: needs at least %d data points (got %lld)", GK_LOCAL_K,
             (long long)n);
  hipStream_t st = (hipStream_t)stream;
  // grid in the normalised frame (x - centre) / range, cells of ~2 points
  KnnGrid g{};
  double ext[3], vol = 1.0;
  for (int d = 0; d < 3; ++d) {
    const double a = (box[d] - sys->center[d]) / sys->range, b = (box[3 + d] - sys->center[d]) / sys->range;
    g.lo[d] = (float)a - 1e-6f;
    ext[d] = std::max(b - a, 1e-9) + 2e-6;
  }
  const double emax = std::max(ext[0], std::max(ext[1], ext[2]));
  for (int d = 0; d < 3; ++d) vol *= std::max(ext[d], emax * 1e-3);
  double c = std::cbrt(vol * 2.0 / (double)n);
  int64_t ncell = 0;
  for (int it = 0; it < 64; ++it) {
    ncell = 1;
    for (int d = 0; d < 3; ++d) ncell *= std::max<int64_t>(1, (int64_t)std::ceil(ext[d] / c));
    if (ncell <= 4 * n + 64) break;
    c *= 1.25;
  }
  for (int d = 0; d < 3; ++d) g.n[d] = (int)std::max<int64_t>(1, (int64_t)std::ceil(ext[d] / c));
  g.c = (float)c;
  g.inv_c = (float)(1.0 / c);
  const int32_t nc = g.n[0] * g.n[1] * g.n[2];
  void* ws = nullptr;
  const size_t need = gk_round_up((size_t)n * 4, 256) + 3 * gk_round_up(((size_t)nc + 1) * 4, 256) +
                      (size_t)n * sizeof(float4);
  GK_TRY(gk_workspace(h, WS_KNN, need, &ws));
  char* b = static_cast<char*>(ws);
  int32_t* cell = reinterpret_cast<int32_t*>(b);
  b += gk_round_up((size_t)n * 4, 256);
  int32_t* count = reinterpret_cast<int32_t*>(b);
  b += gk_round_up(((size_t)nc + 1) * 4, 256);
  int32_t* start = reinterpret_cast<int32_t*>(b);
  b += gk_round_up(((size_t)nc + 1) * 4, 256);
  int32_t* fill = reinterpret_cast<int32_t*>(b);
  b += gk_round_up(((size_t)nc + 1) * 4, 256);
  float4* spos = reinterpret_cast<float4*>(b);
  GK_HIP(hipMemsetAsync(count, 0, (size_t)nc * 4, st));
  GK_HIP(hipMemsetAsync(fill, 0, (size_t)nc * 4, st));
  const unsigned nb = (unsigned)gk_ceil_div(n, 256);
  hipLaunchKernelGGL(knn_cells_kernel, dim3(nb), dim3(256), 0, st, *sys, n, g, cell, count);
  hipLaunchKernelGGL(knn_scan_kernel, dim3(1), dim3(1024), 0, st, count, nc, start);
  hipLaunchKernelGGL(knn_scatter_kernel, dim3(nb), dim3(256), 0, st, *sys, n, cell, start, fill, spos);
  hipLaunchKernelGGL(knn_grid_kernel<GK_LOCAL_K>, dim3(nb), dim3(256), 0, st, *sys, n, g, start, spos, idx, hk);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_test_knn_local(gk_handle h, const gk_system* sys, int32_t* idx, float* hk, gk_stream stream) {
  GK_REQUIRE(h && sys && idx && hk, GK_ERR_ARG, "gk_test_knn_local: NULL argument");
  double box[6];
  GK_TRY(gk_source_box(h, sys, box, stream));
  return gk_knn_local(h, sys, box, idx, hk, stream);
}

extern "C" int gk_local_inverse_local(gk_handle h, const gk_system* sys, const int32_t* idx, const double* dscale,
                                      float* Y, gk_stream stream) {
  GK_REQUIRE(h && sys && idx && dscale && Y, GK_ERR_ARG, "gk_local_inverse_local: bad arguments");
  const int64_t n = sys->n_ori + sys->n_rest;
  if (n == 0) return GK_OK;
  // a fixed grid (the work spreads over the factorization it runs beside)
  hipLaunchKernelGGL(local_inverse_kernel<GK_LOCAL_K>, dim3((unsigned)std::min<int64_t>(n, GK_LI_GRID)), dim3(256), 0,
                     (hipStream_t)stream, make_geo(sys), n, idx, dscale, Y);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_spai_apply_k(const gk_system* sys, int32_t k, const int32_t* idx, const float* Y, const double* u,
                               double* y, hipStream_t st) {
  const int64_t n = sys->n_ori + sys->n_rest;
  if (n == 0) return GK_OK;
  GK_REQUIRE(k == K || k == GK_LOCAL_K, GK_ERR_ARG, "gk_spai_apply: %d neighbours not supported", k);
  if (k == GK_LOCAL_K)
    hipLaunchKernelGGL(spai_apply_kernel<GK_LOCAL_K>, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, st, *sys, n,
                       idx, Y, u, y);
  else
    hipLaunchKernelGGL(spai_apply_kernel<K>, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, st, *sys, n, idx, Y,
                       u, y);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_spai_features(gk_handle h, const gk_system* sys, int64_t nq, const int32_t* pts, const int32_t* idx,
                                const float* hk, float* feat, gk_stream stream) {
  GK_REQUIRE(h && sys && idx && hk && feat && nq >= 0, GK_ERR_ARG, "gk_spai_features: bad arguments");
  if (nq == 0) return GK_OK;
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)gk_ceil_div(nq * K, 256)), dim3(256), 0, (hipStream_t)stream,
                     *sys, nq, pts, idx, hk, feat);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_spai_local_blocks(gk_handle h, const gk_system* sys, int64_t nq, const int32_t* pts,
                                    const int32_t* idx, const double* dscale, float* Hloc, gk_stream stream) {
  GK_REQUIRE(h && sys && pts && idx && dscale && Hloc && nq >= 0, GK_ERR_ARG, "gk_spai_local_blocks: bad arguments");
  if (nq == 0) return GK_OK;
  hipLaunchKernelGGL(local_blocks_kernel, dim3((unsigned)nq), dim3(256), 0, (hipStream_t)stream, make_geo(sys), nq,
                     pts, idx, dscale, Hloc);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_spai_local_inverse(gk_handle h, const gk_system* sys, const int32_t* idx, const double* dscale,
                                     float* Y, gk_stream stream) {
  GK_REQUIRE(h && sys && idx && dscale && Y, GK_ERR_ARG, "gk_spai_local_inverse: bad arguments");
  const int64_t n = sys->n_ori + sys->n_rest;
  if (n == 0) return GK_OK;
  hipLaunchKernelGGL(local_inverse_kernel<K>, dim3((unsigned)n), dim3(256), 0, (hipStream_t)stream, make_geo(sys), n,
                     idx, dscale, Y);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_mlp_forward(gk_handle h, int64_t n, const float* params, const float* x, float* y, float* acts,
                              gk_stream stream) {
  GK_REQUIRE(h && params && x && y && n >= 0, GK_ERR_ARG, "gk_mlp_forward: bad arguments");
  if (n == 0) return GK_OK;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                     params, x, y, acts);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_mlp_backward(gk_handle h, int64_t n, const float* params, const float* x, const float* acts,
                               const float* dy, float* grad, gk_stream stream) {
  GK_REQUIRE(h && params && x && acts && dy && grad && n > 0, GK_ERR_ARG, "gk_mlp_backward: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb = gk_ceil_div(n, 256);
  void* ws = nullptr;
  GK_TRY(gk_workspace(h, WS_EVAL_PART, (size_t)nb * NPARAM * sizeof(float) + 256, &ws));
  float* partial = static_cast<float*>(ws);
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((unsigned)nb), dim3(256), 0, st, n, params, x, acts, dy, partial);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((unsigned)gk_ceil_div(NPARAM, 256)), dim3(256), 0, st, partial, nb, grad);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_spai_loss(gk_handle h, const gk_system* sys, int64_t nq, const int32_t* pts, const int32_t* idx,
                            const float* Hloc, const float* Y, float* dY, float* loss, gk_stream stream) {
  GK_REQUIRE(h && sys && idx && Hloc && Y && dY && loss && nq >= 0, GK_ERR_ARG, "gk_spai_loss: bad arguments");
  if (nq == 0) return GK_OK;
  hipLaunchKernelGGL(spai_loss_kernel, dim3((unsigned)nq), dim3(64), 0, (hipStream_t)stream, *sys, nq, pts, idx, Hloc,
                     Y, dY, loss);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_adam(gk_handle h, float* params, const float* grad, float* m, float* v, int32_t n, float lr,
                       int32_t step, float grad_scale, gk_stream stream) {
  GK_REQUIRE(h && params && grad && m && v && n > 0 && step >= 1, GK_ERR_ARG, "gk_adam: bad arguments");
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, params, grad,
                     m, v, n, lr, b1, b2, 1e-8f, c1, c2, grad_scale);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_spai_apply_impl(const gk_system* sys, const int32_t* idx, const float* Y, const double* u, double* y,
                                  hipStream_t st) {
  const int64_t n = sys->n_ori + sys->n_rest;
  if (n == 0) return GK_OK;
  hipLaunchKernelGGL(spai_apply_kernel<K>, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, st, *sys, n, idx, Y, u,
                     y);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_spai_apply(gk_handle h, const gk_system* sys, const int32_t* idx, const float* Y, const double* u,
                             double* y, gk_stream stream) {
  GK_REQUIRE(h && sys && idx && Y && u && y, GK_ERR_ARG, "gk_spai_apply: NULL argument");
  return gk_spai_apply_impl(sys, idx, Y, u, y, (hipStream_t)stream);
}
