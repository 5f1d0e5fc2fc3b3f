// Finite faults: the fault block's drift values tapered by an implicit
// function (the reference's FiniteFaultData(implicit_function,
// implicit_function_transform, pivot), gempy/core/data/__init__.py:20; the
// ellipsoid of gempy/modules/custom_implicit_functions/ellipsoid_implicit_function.py:21-35).
//
// values[i] = (values[i] - 1) * f(x'_i) with x' = R^T (x - pivot) + pivot (the
// inverse of the transform's rotation about the pivot) and
//   f(x) = 1 - sigmoid(k1 k2 k3 (sum_d ((x_d - c_d) / (r_d k_d))^2 - 1)),
// so the block's two sides differ by f: the offset fades out with the
// ellipsoid, and f == 1 reproduces the infinite fault up to a constant.
#include <cmath>
#include "gk_internal.h"

namespace {

struct Taper {
  double R[9];      // row-major rotation (x' = R^T (x - pivot) + pivot)
  double pivot[3], center[3], radii[3], k[3];
  double origin[3], spacing[3];
  int64_t res[3];
};

__global__ void taper_kernel(int64_t n, const double* xyz, Taper t, double* values) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[3];
  if (xyz) {
    x[0] = xyz[3 * i]; x[1] = xyz[3 * i + 1]; x[2] = xyz[3 * i + 2];
  } else {                                  // lattice point i (z fastest)
    const int64_t iz = i % t.res[2], iy = (i / t.res[2]) % t.res[1], ix = i / (t.res[2] * t.res[1]);
    x[0] = t.origin[0] + ix * t.spacing[0];
    x[1] = t.origin[1] + iy * t.spacing[1];
    x[2] = t.origin[2] + iz * t.spacing[2];
  }
  const double d[3] = {x[0] - t.pivot[0], x[1] - t.pivot[1], x[2] - t.pivot[2]};
  double s = 0.0;
  for (int r = 0; r < 3; ++r) {
    // row r of R^T = column r of R
    const double xr = t.R[r] * d[0] + t.R[3 + r] * d[1] + t.R[6 + r] * d[2] + t.pivot[r];
    const double u = (xr - t.center[r]) / (t.radii[r] * t.k[r]);
    s += u * u;
  }
  const double z = (s - 1.0) * (t.k[0] * t.k[1] * t.k[2]);
  const double f = 1.0 - 1.0 / (1.0 + exp(-z));
  values[i] = (values[i] - 1.0) * f;
}

}  // namespace

extern "C" int gk_finite_fault_values(gk_handle h, int64_t n, const double* xyz, const double* lattice,
                                      const double* rotation, const double* pivot, const double* center,
                                      const double* radii, const double* k, double* values, gk_stream stream) {
  GK_REQUIRE(h && n >= 0 && rotation && pivot && center && radii && k && (n == 0 || values), GK_ERR_ARG,
             "gk_finite_fault_values: NULL argument");
  GK_REQUIRE(xyz || lattice || n == 0, GK_ERR_ARG, "gk_finite_fault_values: points or a lattice required");
  Taper t{};
  for (int i = 0; i < 9; ++i) t.R[i] = rotation[i];
  for (int d = 0; d < 3; ++d) {
    GK_REQUIRE(radii[d] > 0.0 && k[d] > 0.0, GK_ERR_ARG, "gk_finite_fault_values: radii and slopes must be positive");
    t.pivot[d] = pivot[d]; t.center[d] = center[d]; t.radii[d] = radii[d]; t.k[d] = k[d];
  }
  if (!xyz) {
    for (int d = 0; d < 3; ++d) {
      t.origin[d] = lattice[d];
      t.spacing[d] = lattice[3 + d];
      t.res[d] = (int64_t)lattice[6 + d];
    }
    GK_REQUIRE(t.res[0] * t.res[1] * t.res[2] == n, GK_ERR_ARG, "gk_finite_fault_values: lattice size != n");
  }
  if (n == 0) return GK_OK;
  hipLaunchKernelGGL(taper_kernel, dim3((unsigned)gk_ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, n, xyz, t,
                     values);
  GK_HIP(hipGetLastError());
  return GK_OK;
}
