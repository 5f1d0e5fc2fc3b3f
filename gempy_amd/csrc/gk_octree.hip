// Octree refinement of the evaluation grid (the engine's octree levels:
// _number_octree_levels, octree_error_threshold at fault.approved.txt:210-214;
// outputs per level at test_example_models_I.py:20; grid sizing
// resolution = base * 2^(levels-1) at regular_grid.py:36-43).
//
// A level holds a list of active voxels (linear index (ix*ny + iy)*nz + iz at
// that level's resolution).  Each voxel is probed at its centre and its 8
// corners; it is refined when any id field differs among those 9 points, and
// its 8 children (at twice the resolution) become the next level's voxels,
// written in voxel order (deterministic compaction: per-block counts, host
// exclusive scan, per-block LDS scan).
#include "gk_internal.h"

namespace {

__global__ void octree_points_kernel(double x0, double y0, double z0, double hx, double hy, double hz,
                                     int64_t ny, int64_t nz, const int64_t* vox, int64_t n, double* xyz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * 9) return;
  const int64_t v = vox[t / 9];
  const int k = (int)(t % 9);
  const int64_t iz = v % nz, iy = (v / nz) % ny, ix = v / (nz * ny);
  double fx, fy, fz;
  if (k == 0) { fx = 0.5; fy = 0.5; fz = 0.5; }
  else { const int c = k - 1; fx = (c >> 2) & 1; fy = (c >> 1) & 1; fz = c & 1; }
  xyz[3 * t + 0] = x0 + (ix + fx) * hx;
  xyz[3 * t + 1] = y0 + (iy + fy) * hy;
  xyz[3 * t + 2] = z0 + (iz + fz) * hz;
}

__global__ void octree_mark_kernel(const int32_t* ids, int32_t nf, int64_t n, uint8_t* refine, int32_t* counts) {
  __shared__ int32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int r = 0;
  if (v < n) {
    for (int f = 0; f < nf && !r; ++f) {
      const int32_t* p = ids + (int64_t)f * n * 9 + v * 9;
      for (int k = 1; k < 9; ++k) if (p[k] != p[0]) { r = 1; break; }
    }
    refine[v] = (uint8_t)r;
  }
  if (r) atomicAdd(&cnt, 1);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = cnt;
}

__global__ void octree_children_kernel(const uint8_t* refine, const int64_t* vox, int64_t n, const int64_t* offsets,
                                       int64_t ny, int64_t nz, int64_t* children) {
  __shared__ int32_t scan[256];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (v < n) ? refine[v] : 0;
  scan[threadIdx.x] = r;
  __syncthreads();
  for (int s = 1; s < 256; s <<= 1) {        // inclusive Hillis-Steele scan
    const int32_t add = threadIdx.x >= s ? scan[threadIdx.x - s] : 0;
    __syncthreads();
    scan[threadIdx.x] += add;
    __syncthreads();
  }
  if (!r) return;
  const int64_t pos = offsets[blockIdx.x] + scan[threadIdx.x] - 1;
  const int64_t w = vox[v];
  const int64_t iz = w % nz, iy = (w / nz) % ny, ix = w / (nz * ny);
  const int64_t NY = 2 * ny, NZ = 2 * nz;
  for (int c = 0; c < 8; ++c) {
    const int64_t cx = 2 * ix + ((c >> 2) & 1), cy = 2 * iy + ((c >> 1) & 1), cz = 2 * iz + (c & 1);
    children[8 * pos + c] = (cx * NY + cy) * NZ + cz;
  }
}

__global__ void octree_centres_kernel(double x0, double y0, double z0, double hx, double hy, double hz, int64_t ny,
                                      int64_t nz, const int64_t* vox, int64_t n, double* xyz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t v = vox[t];
  const int64_t iz = v % nz, iy = (v / nz) % ny, ix = v / (nz * ny);
  xyz[3 * t + 0] = x0 + (ix + 0.5) * hx;
  xyz[3 * t + 1] = y0 + (iy + 0.5) * hy;
  xyz[3 * t + 2] = z0 + (iz + 0.5) * hz;
}

}  // namespace

extern "C" int gk_octree_points(gk_handle h, const double* extent, const int64_t* res, const int64_t* voxels,
                                int64_t n, int32_t centres_only, double* xyz, gk_stream stream) {
  GK_REQUIRE(h && extent && res && (n == 0 || (voxels && xyz)) && n >= 0, GK_ERR_ARG,
             "gk_octree_points: bad arguments");
  GK_REQUIRE(res[0] > 0 && res[1] > 0 && res[2] > 0, GK_ERR_ARG, "gk_octree_points: bad resolution");
  if (n == 0) return GK_OK;
  hipStream_t st = (hipStream_t)stream;
  const double hx = (extent[1] - extent[0]) / res[0], hy = (extent[3] - extent[2]) / res[1],
               hz = (extent[5] - extent[4]) / res[2];
  if (centres_only)
    hipLaunchKernelGGL(octree_centres_kernel, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, st, extent[0],
                       extent[2], extent[4], hx, hy, hz, res[1], res[2], voxels, n, xyz);
  else
    hipLaunchKernelGGL(octree_points_kernel, dim3((unsigned)gk_ceil_div(n * 9, 256)), dim3(256), 0, st, extent[0],
                       extent[2], extent[4], hx, hy, hz, res[1], res[2], voxels, n, xyz);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_octree_refine(gk_handle h, int32_t n_fields, const int32_t* ids, int64_t n, const int64_t* voxels,
                                const int64_t* res, uint8_t* refine, int64_t* children, int64_t* n_children,
                                gk_stream stream) {
  GK_REQUIRE(h && n_children && res && n >= 0 && n_fields >= 1, GK_ERR_ARG, "gk_octree_refine: bad arguments");
  *n_children = 0;
  if (n == 0) return GK_OK;
  GK_REQUIRE(ids && voxels && refine && children, GK_ERR_ARG, "gk_octree_refine: NULL array");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nb = gk_ceil_div(n, 256);
  void* ws = nullptr;
  GK_TRY(gk_workspace(h, WS_SMALL, (size_t)nb * (sizeof(int32_t) + sizeof(int64_t)) + 512, &ws));
  int32_t* counts = static_cast<int32_t*>(ws);
  int64_t* offsets = reinterpret_cast<int64_t*>(static_cast<char*>(ws) + gk_round_up(nb * 4, 256));
  hipLaunchKernelGGL(octree_mark_kernel, dim3((unsigned)nb), dim3(256), 0, st, ids, n_fields, n, refine, counts);
  std::vector<int32_t> hc(nb);
  std::vector<int64_t> ho(nb);
  GK_HIP(hipMemcpyAsync(hc.data(), counts, nb * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  int64_t total = 0;
  for (int64_t b = 0; b < nb; ++b) { ho[b] = total; total += hc[b]; }
  GK_HIP(hipMemcpyAsync(offsets, ho.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (total > 0)
    hipLaunchKernelGGL(octree_children_kernel, dim3((unsigned)nb), dim3(256), 0, st, refine, voxels, n, offsets,
                       res[1], res[2], children);
  GK_HIP(hipGetLastError());
  GK_HIP(hipStreamSynchronize(st));
  *n_children = 8 * total;
  return GK_OK;
}
