// Library state, error reporting and the public wrappers of libgkhip.so.
#include <cstring>
#include <cstdlib>
#include "gk_internal.h"

static thread_local char g_err[2048] = "";

void gk_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* gk_last_error(void) { return g_err; }
extern "C" int gk_version(void) { return 100; }

// Workspace slot of at least `bytes` (grown by 1/8 beyond the request).
// Growing a slot synchronizes the device before freeing the old buffer (work
// queued on any stream may still read it), so a slot that grows in the middle
// of a pipelined solve would stall its side streams.  Entry points take their
// slots before their first launch, and repeated solves of one size never grow
// a slot after the first.
int gk_workspace(gk_handle h, int slot, size_t bytes, void** out) {
  GK_REQUIRE(h && slot >= 0 && slot < WS_COUNT, GK_ERR_ARG, "gk: bad handle or workspace slot");
  if (bytes > h->ws_bytes[slot]) {
    if (h->ws[slot]) {
      GK_HIP(hipDeviceSynchronize());
      GK_HIP(hipFree(h->ws[slot]));
      h->ws[slot] = nullptr;
      h->ws_bytes[slot] = 0;
    }
    size_t nb = bytes + bytes / 8 + 4096;
    GK_HIP(hipMalloc(&h->ws[slot], nb));
    h->ws_bytes[slot] = nb;
  }
  *out = h->ws[slot];
  return GK_OK;
}

extern "C" int gk_create(int device, gk_handle* out) {
  GK_REQUIRE(out != nullptr, GK_ERR_ARG, "gk_create: out is NULL");
  int n = 0;
  GK_HIP(hipGetDeviceCount(&n));
  GK_REQUIRE(device >= 0 && device < n, GK_ERR_ARG, "gk_create: device %d not present (%d devices)",
             device, n);
  GK_HIP(hipSetDevice(device));
  gk_context* c = new gk_context();
  c->device = device;
  const int rc = gk_eval_scratch_init(c);
  if (rc != GK_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return GK_OK;
}

extern "C" void gk_destroy(gk_handle h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  if (h->solver_state && h->solver_free) h->solver_free(h->solver_state);
  for (int i = 0; i < WS_COUNT; ++i)
    if (h->ws[i]) (void)hipFree(h->ws[i]);
  for (hipEvent_t& e : h->side_ev)
    if (e) (void)hipEventDestroy(e);
  if (h->side) (void)hipStreamDestroy(h->side);
  for (hipEvent_t& e : h->aux_ev)
    if (e) (void)hipEventDestroy(e);
  if (h->aux) (void)hipStreamDestroy(h->aux);
  delete h;
}

int gk_aux_stream(gk_handle h, hipStream_t* s3, hipEvent_t** ev) {
  if (!h->aux) {
    GK_HIP(hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
    for (hipEvent_t& e : h->aux_ev) GK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  *s3 = h->aux;
  *ev = h->aux_ev;
  return GK_OK;
}

int gk_side_stream(gk_handle h, hipStream_t* s2, hipEvent_t** ev) {
  if (!h->side) {
    GK_HIP(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    for (hipEvent_t& e : h->side_ev) GK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  *s2 = h->side;
  *ev = h->side_ev;
  return GK_OK;
}

// implemented in the other translation units
extern "C" int gk_eval_impl(gk_handle h, const gk_system* sys, const double* weights, gk_eval_desc* d,
                            hipStream_t st, bool force_mask, const double* src_box, bool reuse_sources);
extern "C" int gk_matvec_impl(gk_handle h, const gk_system* sys, int32_t precision, const double* x,
                              double* y, hipStream_t st, void* scratch, const double* src_box);
extern "C" size_t gk_matvec_scratch_doubles(const gk_system* sys);
extern "C" int gk_assemble_impl(gk_handle h, const gk_system* sys, int32_t dtype, void* A, int64_t lda,
                                double* rhs, double* dscale, hipStream_t st, int lower_only);
extern "C" int gk_assemble_drift_impl(const gk_system* sys, double* U, hipStream_t st);

// Bounding box of all sources (orientation, rest and reference points), read
// back once per evaluation to decide whether the r < a mask can be dropped.
__global__ void source_box_kernel(gk_system sys, double* box) {
  __shared__ double lo[3][256], hi[3][256];
  const int t = threadIdx.x;
  double l[3] = {1e300, 1e300, 1e300}, u[3] = {-1e300, -1e300, -1e300};
  const int64_t n = sys.n_ori + sys.n_rest + sys.n_surf;
  for (int64_t i = t; i < n; i += 256) {
    const double* p = i < sys.n_ori ? sys.ori_xyz + 3 * i
                      : (i < sys.n_ori + sys.n_rest ? sys.rest_xyz + 3 * (i - sys.n_ori)
                                                    : sys.ref_xyz + 3 * (i - sys.n_ori - sys.n_rest));
    for (int k = 0; k < 3; ++k) { l[k] = fmin(l[k], p[k]); u[k] = fmax(u[k], p[k]); }
  }
  for (int k = 0; k < 3; ++k) { lo[k][t] = l[k]; hi[k][t] = u[k]; }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < 3; ++k) {
        lo[k][t] = fmin(lo[k][t], lo[k][t + w]);
        hi[k][t] = fmax(hi[k][t], hi[k][t + w]);
      }
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 3; ++k) { box[k] = lo[k][0]; box[3 + k] = hi[k][0]; }
}

extern "C" int gk_eval(gk_handle h, const gk_system* sys, const double* weights, gk_eval_desc* desc,
                       gk_stream stream) {
  GK_REQUIRE(h != nullptr, GK_ERR_ARG, "gk_eval: handle is NULL");
  hipStream_t st = (hipStream_t)stream;
  double box[6];
  const double* pbox = nullptr;
  if (desc && desc->kind == GK_POINTS_LATTICE && sys && sys->n_ori + sys->n_rest + sys->n_surf > 0) {
    void* ws = nullptr;
    GK_TRY(gk_workspace(h, WS_SMALL, 256, &ws));
    hipLaunchKernelGGL(source_box_kernel, dim3(1), dim3(256), 0, st, *sys, static_cast<double*>(ws));
    GK_HIP(hipMemcpyAsync(box, ws, sizeof(box), hipMemcpyDeviceToHost, st));
    GK_HIP(hipStreamSynchronize(st));
    pbox = box;
  }
  return gk_eval_impl(h, sys, weights, desc, st, false, pbox, false);
}

// Operator product; with `src_box` (the bounding box of the system's points,
// from gk_source_box) the r < a mask is dropped when no pair can reach the range.
extern "C" int gk_matvec_boxed(gk_handle h, const gk_system* sys, int32_t precision, const double* x, double* y,
                               gk_stream stream, const double* src_box) {
  GK_REQUIRE(h && sys && x && y, GK_ERR_ARG, "gk_matvec: NULL argument");
  GK_REQUIRE(precision == GK_F32 || precision == GK_F64, GK_ERR_ARG, "gk_matvec: bad precision");
  void* scratch = nullptr;
  GK_TRY(gk_workspace(h, WS_MATVEC, gk_matvec_scratch_doubles(sys) * sizeof(double), &scratch));
  return gk_matvec_impl(h, sys, precision, x, y, (hipStream_t)stream, scratch, src_box);
}

extern "C" int gk_matvec(gk_handle h, const gk_system* sys, int32_t precision, const double* x, double* y,
                         gk_stream stream) {
  return gk_matvec_boxed(h, sys, precision, x, y, stream, nullptr);
}

// Bounding box of the system's points ([lo xyz, hi xyz]), synchronously.
extern "C" int gk_source_box(gk_handle h, const gk_system* sys, double* box, gk_stream stream) {
  GK_REQUIRE(h && sys && box, GK_ERR_ARG, "gk_source_box: NULL argument");
  GK_REQUIRE(sys->n_ori + sys->n_rest + sys->n_surf > 0, GK_ERR_ARG, "gk_source_box: empty system");
  hipStream_t st = (hipStream_t)stream;
  void* ws = nullptr;
  GK_TRY(gk_workspace(h, WS_SMALL, 256, &ws));
  hipLaunchKernelGGL(source_box_kernel, dim3(1), dim3(256), 0, st, *sys, static_cast<double*>(ws));
  GK_HIP(hipMemcpyAsync(box, ws, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  GK_HIP(hipStreamSynchronize(st));
  return GK_OK;
}

extern "C" int gk_assemble(gk_handle h, const gk_system* sys, int32_t dtype, void* A, int64_t lda,
                           double* rhs, double* dscale, gk_stream stream) {
  GK_REQUIRE(h && sys, GK_ERR_ARG, "gk_assemble: NULL argument");
  return gk_assemble_impl(h, sys, dtype, A, lda, rhs, dscale, (hipStream_t)stream, 0);
}

extern "C" int gk_assemble_drift(gk_handle h, const gk_system* sys, double* U, gk_stream stream) {
  GK_REQUIRE(h && sys && U, GK_ERR_ARG, "gk_assemble_drift: NULL argument");
  return gk_assemble_drift_impl(sys, U, (hipStream_t)stream);
}

// ------------------------------------------------------------- profiling
namespace {
bool g_prof_on = false;
std::vector<gk_prof_rec> g_prof_recs;
std::vector<hipEvent_t> g_prof_pool;
size_t g_prof_used = 0;

hipEvent_t prof_event() {
  if (g_prof_used == g_prof_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_prof_pool.push_back(e);
  }
  return g_prof_pool[g_prof_used++];
}
}  // namespace

bool gk_prof_begin(hipStream_t st, hipEvent_t* a) {
  if (!g_prof_on) return false;
  *a = prof_event();
  if (!*a || hipEventRecord(*a, st) != hipSuccess) return false;
  return true;
}

void gk_prof_end(hipStream_t st, hipEvent_t a, int kind, double work) {
  hipEvent_t b = prof_event();
  if (!b || hipEventRecord(b, st) != hipSuccess) return;
  g_prof_recs.push_back({a, b, kind, work});
}

extern "C" int gk_profile(gk_handle h, int32_t enable) {
  GK_REQUIRE(h, GK_ERR_ARG, "gk_profile: NULL handle");
  GK_HIP(hipDeviceSynchronize());
  g_prof_on = enable != 0;
  if (enable) {
    g_prof_recs.clear();
    g_prof_used = 0;
  }
  return GK_OK;
}

extern "C" int gk_profile_read(gk_handle h, int32_t kind, int64_t* launches, double* ms, double* work) {
  GK_REQUIRE(h && launches && ms && work, GK_ERR_ARG, "gk_profile_read: NULL argument");
  GK_HIP(hipDeviceSynchronize());
  int64_t n = 0;
  double t = 0.0, w = 0.0;
  for (const gk_prof_rec& r : g_prof_recs) {
    if (r.kind != kind) continue;
    float e = 0.0f;
    GK_HIP(hipEventElapsedTime(&e, r.a, r.b));
    ++n;
    t += e;
    w += r.work;
  }
  *launches = n;
  *ms = t;
  *work = w;
  return GK_OK;
}
