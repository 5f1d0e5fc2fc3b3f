// Gravity forward on a centred grid (the reference's geophysics input:
// tz = calculate_gravity_gradient(centered_grid), densities per unit, result
// Solutions.gravity; test/test_modules/test_geophysics/test_gravity.py:67-89).
//
// tz of every kernel cell is the vertical attraction of a homogeneous
// rectangular prism (unit density) at the device, from the closed form over
// its 8 corners (Nagy 1966; Plouff 1976); the gravity of device d is
// sum_k tz[k] * density(id[d][k]), the ids being the lith block at the
// device's kernel points.  Both are HBM-light: tz is computed once per kernel
// (n cells), the contraction streams the ids once.
#include <cmath>
#include "gk_internal.h"

namespace {

__global__ void tz_kernel(int64_t n, const double* c, const double* lo, const double* hi, double G, double* tz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x[2] = {c[3 * i] - lo[3 * i], c[3 * i] + hi[3 * i]};
  const double y[2] = {c[3 * i + 1] - lo[3 * i + 1], c[3 * i + 1] + hi[3 * i + 1]};
  const double z[2] = {c[3 * i + 2] - lo[3 * i + 2], c[3 * i + 2] + hi[3 * i + 2]};
  // corner order: x slowest, z fastest; signs (1, -1, -1, 1, -1, 1, 1, -1)
  const double mu[8] = {1, -1, -1, 1, -1, 1, 1, -1};
  double s = 0.0;
  for (int k = 0; k < 8; ++k) {
    const double X = x[(k >> 2) & 1], Y = y[(k >> 1) & 1], Z = z[k & 1];
    const double R = sqrt(X * X + Y * Y + Z * Z);
    const double q = X * Y, den = Z * R;
    const double at = den != 0.0 ? atan(q / den) : (q > 0.0 ? 1.5707963267948966 : (q < 0.0 ? -1.5707963267948966 : 0.0));
    s -= mu[k] * (X * log(Y + R) + Y * log(X + R) - Z * at);
  }
  tz[i] = G * s;
}

struct Dens {
  int32_t n;
  double v[GK_MAX_THRESHOLDS + 1];
};

__global__ __launch_bounds__(256) void gravity_kernel(int64_t nk, const double* tz, const int32_t* ids, Dens d,
                                                      double* out) {
  __shared__ double red[256];
  const int64_t dev = blockIdx.x;
  const int32_t* id = ids + dev * nk;
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < nk; k += 256) {
    const int32_t u = id[k] - 1;
    acc += tz[k] * ((u >= 0 && u < d.n) ? d.v[u] : 0.0);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[dev] = red[0];
}

}  // namespace

extern "C" int gk_gravity_tz(gk_handle h, int64_t n, const double* centres, const double* half_lo,
                             const double* half_hi, double G, double* tz, gk_stream stream) {
  GK_REQUIRE(h && n >= 0 && (n == 0 || (centres && half_lo && half_hi && tz)), GK_ERR_ARG,
             "gk_gravity_tz: bad arguments");
  if (n == 0) return GK_OK;
  hipLaunchKernelGGL(tz_kernel, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, n, centres,
                     half_lo, half_hi, G, tz);
  GK_HIP(hipGetLastError());
  return GK_OK;
}

extern "C" int gk_gravity(gk_handle h, int64_t n_devices, int64_t n_kernel, const double* tz, const int32_t* ids,
                          int32_t n_densities, const double* densities, double* out, gk_stream stream) {
  GK_REQUIRE(h && n_devices >= 0 && n_kernel >= 0 && densities && n_densities >= 0 &&
             n_densities <= GK_MAX_THRESHOLDS + 1, GK_ERR_ARG, "gk_gravity: bad arguments");
  if (n_devices == 0) return GK_OK;
  GK_REQUIRE(out && (n_kernel == 0 || (tz && ids)), GK_ERR_ARG, "gk_gravity: NULL array");
  Dens d{};
  d.n = n_densities;
  for (int i = 0; i < n_densities; ++i) d.v[i] = densities[i];
  hipLaunchKernelGGL(gravity_kernel, dim3((unsigned)n_devices), dim3(256), 0, (hipStream_t)stream, n_kernel, tz, ids,
                     d, out);
  GK_HIP(hipGetLastError());
  return GK_OK;
}
