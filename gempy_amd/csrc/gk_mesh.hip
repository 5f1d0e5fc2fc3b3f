// Isosurface extraction (marching cubes) of a lattice field: the consumer of
// the evaluated scalar field that turns each surface's threshold into a
// triangle mesh (the reference calls skimage's marching cubes per element,
// gempy/modules/mesh_extranction/marching_cubes.py:59-101).
//
// One vertex per lattice edge whose end values straddle the level (a corner
// is "above" when its value > level), placed by linear interpolation along
// the edge; vertices are shared by all cubes that use the edge.  The
// triangulation of the 256 corner cases is built here from first principles:
// on every cube face the crossing points are joined by segments that cut off
// the face's "above" corners (so two cubes sharing a face always agree: the
// mesh is watertight), segments are directed along the face boundary taken
// counter-clockwise as seen from outside the cube (entering the "above" run
// -> leaving it), they chain into closed loops across the faces, and every
// loop is triangulated as a fan.  Triangles wind so that their normals
// (right-hand rule) point towards increasing field values.
//
// Passes: classify cubes (case byte, triangle count, used-edge flags) ->
// exclusive scans (vertex ids, face offsets) -> vertex positions -> faces.
#include <algorithm>
#include <cmath>
#include <vector>
#include "gk_internal.h"

namespace {

constexpr int MC_MAX_TRI = 5;    // triangles per cube case, at most (the construction yields <= 5)

struct McTables {
  uint8_t ntri[256];
  uint16_t edges[256];                 // bit e: local edge e crosses
  int8_t tri[256][MC_MAX_TRI][3];      // local edges of each triangle
};

// corner c = x + 2y + 4z; local edge e = 4 * axis + k, k the two other
// coordinates (in increasing axis order) as bits
void edge_corners(int e, int* c0, int* c1) {
  const int axis = e / 4, k = e % 4;
  int o[3] = {0, 0, 0}, j = 0;
  for (int a = 0; a < 3; ++a)
    if (a != axis) o[a] = (k >> j++) & 1;
  *c0 = o[0] + 2 * o[1] + 4 * o[2];
  o[axis] = 1;
  *c1 = o[0] + 2 * o[1] + 4 * o[2];
}

int edge_of(int ca, int cb) {
  for (int e = 0; e < 12; ++e) {
    int c0, c1;
    edge_corners(e, &c0, &c1);
    if ((c0 == ca && c1 == cb) || (c0 == cb && c1 == ca)) return e;
  }
  return -1;
}

McTables build_tables() {
  McTables T{};
  // the six faces: corners in counter-clockwise order seen from outside
  int faces[6][4];
  for (int f = 0; f < 6; ++f) {
    const int axis = f / 2, side = f % 2;
    const double n[3] = {axis == 0 ? (side ? 1.0 : -1.0) : 0.0, axis == 1 ? (side ? 1.0 : -1.0) : 0.0,
                         axis == 2 ? (side ? 1.0 : -1.0) : 0.0};
    std::vector<int> cs;
    for (int c = 0; c < 8; ++c)
      if (((c >> axis) & 1) == side) cs.push_back(c);
    const int u = (axis + 1) % 3, v = (axis + 2) % 3;        // (u, v, axis) is right-handed
    std::sort(cs.begin(), cs.end(), [&](int a, int b) {
      const double au = ((a >> u) & 1) - 0.5, av = ((a >> v) & 1) - 0.5;
      const double bu = ((b >> u) & 1) - 0.5, bv = ((b >> v) & 1) - 0.5;
      return std::atan2(av, au) < std::atan2(bv, bu);        // counter-clockwise about +axis
    });
    if (n[axis] < 0) std::reverse(cs.begin(), cs.end());     // counter-clockwise about the outward normal
    for (int i = 0; i < 4; ++i) faces[f][i] = cs[i];
  }
  for (int cs = 0; cs < 256; ++cs) {
    int next[12];
    for (int& e : next) e = -1;
    uint16_t used = 0;
    for (int f = 0; f < 6; ++f) {
      int cross[4], dir[4], nc = 0;     // crossing edges in boundary order, +1 entering the "above" run
      for (int i = 0; i < 4; ++i) {
        const int a = faces[f][i], b = faces[f][(i + 1) % 4];
        const int ia = (cs >> a) & 1, ib = (cs >> b) & 1;
        if (ia != ib) {
          cross[nc] = edge_of(a, b);
          dir[nc] = ib ? 1 : -1;
          ++nc;
        }
      }
      // every entering crossing joins the next leaving one (cuts off the
      // "above" corners between them)
      for (int i = 0; i < nc; ++i) {
        if (dir[i] != 1) continue;
        for (int j = 1; j < nc; ++j) {
          const int k = (i + j) % nc;
          if (dir[k] == -1) {
            next[cross[i]] = cross[k];
            used |= (uint16_t)((1u << cross[i]) | (1u << cross[k]));
            break;
          }
        }
      }
    }
    T.edges[cs] = used;
    int nt = 0;
    bool seen[12] = {};
    for (int s = 0; s < 12; ++s) {
      if (next[s] < 0 || seen[s]) continue;
      std::vector<int> loop;
      for (int e = s; !seen[e]; e = next[e]) {
        seen[e] = true;
        loop.push_back(e);
      }
      // the loop runs clockwise about the direction of increasing values:
      // the fan is wound the other way round
      for (size_t i = 1; i + 1 < loop.size() && nt < MC_MAX_TRI; ++i) {
        T.tri[cs][nt][0] = (int8_t)loop[0];
        T.tri[cs][nt][1] = (int8_t)loop[i + 1];
        T.tri[cs][nt][2] = (int8_t)loop[i];
        ++nt;
      }
    }
    T.ntri[cs] = (uint8_t)nt;
  }
  return T;
}

struct McArgs {
  const double* f;
  const uint8_t* mask;
  int64_t nx, ny, nz;
  double level;
  const McTables* T;
};

__device__ __forceinline__ int64_t pidx(const McArgs& a, int64_t x, int64_t y, int64_t z) {
  return (x * a.ny + y) * a.nz + z;
}

// global edge of local edge e of cube (x, y, z): 3 * point + axis
__device__ __forceinline__ int64_t gedge(const McArgs& a, int64_t x, int64_t y, int64_t z, int e) {
  const int axis = e >> 2, k = e & 3;
  int o[3] = {0, 0, 0}, j = 0;
  for (int d = 0; d < 3; ++d)
    if (d != axis) o[d] = (k >> j++) & 1;
  return 3 * pidx(a, x + o[0], y + o[1], z + o[2]) + axis;
}

__global__ void mc_classify_kernel(McArgs a, uint8_t* cases, uint32_t* ntri, uint32_t* eflag) {
  const int64_t ncx = a.nx - 1, ncy = a.ny - 1, ncz = a.nz - 1;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncx * ncy * ncz) return;
  const int64_t z = c % ncz, y = (c / ncz) % ncy, x = c / (ncz * ncy);
  int cs = 0;
  // a cube is meshed when the mask is set at its far corner (+1, +1, +1)
  bool ok = !a.mask || a.mask[pidx(a, x + 1, y + 1, z + 1)];
  if (ok) {
    for (int k = 0; k < 8; ++k) {
      const double v = a.f[pidx(a, x + (k & 1), y + ((k >> 1) & 1), z + ((k >> 2) & 1))];
      if (v != v) { ok = false; break; }
      cs |= (v > a.level ? 1 : 0) << k;
    }
  }
  if (!ok) cs = 0;
  cases[c] = (uint8_t)cs;
  ntri[c] = a.T->ntri[cs];
  const uint16_t used = a.T->edges[cs];
  for (int e = 0; e < 12; ++e)
    if (used & (1u << e)) eflag[gedge(a, x, y, z, e)] = 1u;
}

__global__ void mc_vertices_kernel(McArgs a, const uint32_t* eflag, const uint32_t* vid, double ox, double oy,
                                   double oz, double hx, double hy, double hz, double* verts) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 3 * a.nx * a.ny * a.nz || !eflag[g]) return;
  const int axis = (int)(g % 3);
  const int64_t p = g / 3;
  const int64_t z = p % a.nz, y = (p / a.nz) % a.ny, x = p / (a.nz * a.ny);
  const int64_t q = axis == 0 ? pidx(a, x + 1, y, z) : (axis == 1 ? pidx(a, x, y + 1, z) : pidx(a, x, y, z + 1));
  const double v0 = a.f[p], v1 = a.f[q];
  const double t = (a.level - v0) / (v1 - v0);
  double P[3] = {(double)x, (double)y, (double)z};
  P[axis] += t;
  double* o = verts + 3 * (int64_t)vid[g];
  o[0] = ox + P[0] * hx;
  o[1] = oy + P[1] * hy;
  o[2] = oz + P[2] * hz;
}

__global__ void mc_faces_kernel(McArgs a, const uint8_t* cases, const uint32_t* foff, const uint32_t* vid,
                                int32_t* faces) {
  const int64_t ncx = a.nx - 1, ncy = a.ny - 1, ncz = a.nz - 1;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncx * ncy * ncz) return;
  const int cs = cases[c];
  const int nt = a.T->ntri[cs];
  if (nt == 0) return;
  const int64_t z = c % ncz, y = (c / ncz) % ncy, x = c / (ncz * ncy);
  int32_t* o = faces + 3 * (int64_t)foff[c];
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) o[3 * t + k] = (int32_t)vid[gedge(a, x, y, z, a.T->tri[cs][t][k])];
}

// Exclusive scan of n uint32 in chunks of 1024 (256 threads x 4), the chunk
// totals scanned by one workgroup; *total receives the sum.
__global__ __launch_bounds__(256) void scan_chunks_kernel(const uint32_t* in, uint32_t* out, int64_t n,
                                                          uint32_t* sums) {
  __shared__ uint32_t s[256];
  const int64_t base = (int64_t)blockIdx.x * 1024 + 4 * threadIdx.x;
  uint32_t v[4], acc = 0;
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    acc += v[k];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - acc;
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) sums[blockIdx.x] = s[255];
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(uint32_t* sums, int64_t nb, uint32_t* total) {
  __shared__ uint32_t s[1024];
  const int64_t per = (nb + 1023) / 1024;
  const int64_t b0 = threadIdx.x * per;
  uint32_t acc = 0;
  for (int64_t i = b0; i < b0 + per && i < nb; ++i) acc += sums[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - acc;
  for (int64_t i = b0; i < b0 + per && i < nb; ++i) {
    const uint32_t v = sums[i];
    sums[i] = run;
    run += v;
  }
  if (threadIdx.x == 1023) *total = s[1023];
}

__global__ void scan_add_kernel(uint32_t* out, int64_t n, const uint32_t* sums) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / 1024];
}

void exclusive_scan(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* sums, uint32_t* total, hipStream_t st) {
  const int64_t nb = std::max<int64_t>(gk_ceil_div(n, 1024), 1);
  hipLaunchKernelGGL(scan_chunks_kernel, dim3((unsigned)nb), dim3(256), 0, st, in, out, n, sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, st, sums, nb, total);
  hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, st, out, n, sums);
}

}  // namespace

extern "C" int gk_marching_cubes(gk_handle h, const double* field, const int64_t* res, const uint8_t* mask,
                                 double level, const double* origin, const double* spacing, double* verts,
                                 int64_t max_verts, int32_t* faces, int64_t max_faces, int64_t* n_verts,
                                 int64_t* n_faces, gk_stream stream) {
  GK_REQUIRE(h && field && res && origin && spacing && n_verts && n_faces, GK_ERR_ARG,
             "gk_marching_cubes: NULL argument");
  const int64_t nx = res[0], ny = res[1], nz = res[2];
  GK_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, GK_ERR_ARG, "gk_marching_cubes: bad resolution");
  const int64_t np_ = nx * ny * nz;
  GK_REQUIRE(3 * np_ < (int64_t)UINT32_MAX, GK_ERR_ARG, "gk_marching_cubes: lattice too large");
  *n_verts = 0;
  *n_faces = 0;
  if (nx < 2 || ny < 2 || nz < 2) return GK_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nc = (nx - 1) * (ny - 1) * (nz - 1);
  const int64_t ne = 3 * np_;
  // workspace: tables | eflag ne | vid ne | cases nc | ntri nc | foff nc | sums | totals
  size_t need = 0;
  auto take = [&](size_t bytes) { size_t o = need; need += gk_round_up(bytes, 256); return o; };
  const size_t o_T = take(sizeof(McTables)), o_ef = take(ne * 4), o_vid = take(ne * 4), o_cs = take(nc),
               o_nt = take(nc * 4), o_fo = take(nc * 4),
               o_sum = take((size_t)(std::max(gk_ceil_div(ne, 1024), gk_ceil_div(nc, 1024)) + 1) * 4),
               o_tot = take(16);
  void* p = nullptr;
  GK_TRY(gk_workspace(h, WS_MESH, need, &p));
  char* b = static_cast<char*>(p);
  if (h->mesh_tables != b + o_T) {        // tables uploaded once per workspace
    static const McTables host = build_tables();
    GK_HIP(hipMemcpyAsync(b + o_T, &host, sizeof(McTables), hipMemcpyHostToDevice, st));
    h->mesh_tables = b + o_T;
  }
  McArgs a{field, mask, nx, ny, nz, level, reinterpret_cast<const McTables*>(b + o_T)};
  uint32_t* ef = reinterpret_cast<uint32_t*>(b + o_ef);
  uint32_t* vid = reinterpret_cast<uint32_t*>(b + o_vid);
  uint8_t* cs = reinterpret_cast<uint8_t*>(b + o_cs);
  uint32_t* nt = reinterpret_cast<uint32_t*>(b + o_nt);
  uint32_t* fo = reinterpret_cast<uint32_t*>(b + o_fo);
  uint32_t* sums = reinterpret_cast<uint32_t*>(b + o_sum);
  uint32_t* tot = reinterpret_cast<uint32_t*>(b + o_tot);
  GK_HIP(hipMemsetAsync(ef, 0, ne * 4, st));
  hipLaunchKernelGGL(mc_classify_kernel, dim3((unsigned)gk_ceil_div(nc, 256)), dim3(256), 0, st, a, cs, nt, ef);
  exclusive_scan(ef, vid, ne, sums, tot, st);
  exclusive_scan(nt, fo, nc, sums, tot + 1, st);
  uint32_t counts[2] = {0, 0};
  GK_HIP(hipMemcpyAsync(counts, tot, sizeof(counts), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  *n_verts = counts[0];
  *n_faces = counts[1];
  if (!verts || !faces) return GK_OK;
  GK_REQUIRE(counts[0] <= max_verts && counts[1] <= max_faces, GK_ERR_ARG,
             "gk_marching_cubes: %u vertices / %u faces exceed the output capacity (%lld / %lld)", counts[0],
             counts[1], (long long)max_verts, (long long)max_faces);
  hipLaunchKernelGGL(mc_vertices_kernel, dim3((unsigned)gk_ceil_div(ne, 256)), dim3(256), 0, st, a, ef, vid,
                     origin[0], origin[1], origin[2], spacing[0], spacing[1], spacing[2], verts);
  hipLaunchKernelGGL(mc_faces_kernel, dim3((unsigned)gk_ceil_div(nc, 256)), dim3(256), 0, st, a, cs, fo, vid, faces);
  GK_HIP(hipGetLastError());
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}
