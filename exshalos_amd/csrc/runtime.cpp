// Context runtime: error messages, device buffers, HIP-event timing and the
// rocFFT plan cache.
#include <rocfft/rocfft.h>

#include <cstdarg>
#include <cstring>
#include <mutex>

#include <cstdlib>

#include "common.hpp"
#include "kspace.hpp"

namespace exh {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

void* ctx_buf(exh_ctx* c, Slot s, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (c->cap[s] < bytes) {
    EXH_REQUIRE(!c->no_alloc, EXH_ERR_INTERNAL,
                "an allocation-free call needs a device buffer (slot %d)", (int)s);
    if (c->buf[s]) {
      EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
      EXH_CHECK_HIP(hipFree(c->buf[s]));
      c->buf[s] = nullptr;
      c->cap[s] = 0;
    }
    size_t want = (bytes + 4095) & ~(size_t)4095;
    EXH_CHECK_HIP(hipMalloc(&c->buf[s], want));
    c->cap[s] = want;
    c->buf_gen[s] = ++c->n_allocs;
  }
  return c->buf[s];
}

void* ctx_pinned(exh_ctx* c, size_t bytes) {
  if (c->pinned_cap < bytes) {
    EXH_REQUIRE(!c->no_alloc, EXH_ERR_INTERNAL,
                "an allocation-free call needs pinned host memory");
    if (c->pinned) EXH_CHECK_HIP(hipHostFree(c->pinned));
    c->pinned = nullptr;
    size_t want = (bytes + 4095) & ~(size_t)4095;
    EXH_CHECK_HIP(hipHostMalloc(&c->pinned, want, hipHostMallocDefault));
    c->pinned_cap = want;
  }
  return c->pinned;
}

TimedScope::TimedScope(exh_ctx* ctx, const char* name) : c(ctx) {
  if (!c->timing) return;
  Timing t;
  t.name = name;
  EXH_CHECK_HIP(hipEventCreate(&t.start));
  EXH_CHECK_HIP(hipEventCreate(&t.stop));
  EXH_CHECK_HIP(hipEventRecord(t.start, c->stream));
  c->events.push_back(t);
  idx = (int)c->events.size() - 1;
}

TimedScope::~TimedScope() {
  if (idx >= 0) (void)hipEventRecord(c->events[idx].stop, c->stream);
}

// ---------------------------------------------------------------------------
// rocFFT
// ---------------------------------------------------------------------------
static std::once_flag g_fft_once;

#define EXH_CHECK_FFT(expr)                                                        \
  do {                                                                             \
    rocfft_status _s = (expr);                                                     \
    if (_s != rocfft_status_success) {                                             \
      ::exh::set_error("%s:%d: %s failed (rocfft status %d)", __FILE__, __LINE__, \
                       #expr, (int)_s);                                            \
      throw ::exh::Error{EXH_ERR_FFT};                                             \
    }                                                                              \
  } while (0)

// dims 3: the [n0][n1][n2] real transform; dims 1: `batch` 1D real
// transforms of length n2 (n0, n1 unused).
static FftPlan& get_plan(exh_ctx* c, int n0, int n1, int n2, int prec, int inverse,
                         long long batch, int dims) {
  std::call_once(g_fft_once, [] { rocfft_setup(); });
  if (dims == 1) n0 = n1 = 0;
  FftKey key{n0, n1, n2, prec, inverse, batch, dims};
  auto it = c->plans.find(key);
  if (it != c->plans.end()) return it->second;
  FftPlan p;
  // rocFFT lengths are fastest-varying first
  size_t len[3] = {(size_t)n2, (size_t)n1, (size_t)n0};
  rocfft_plan_description desc = nullptr;
  EXH_CHECK_FFT(rocfft_plan_description_create(&desc));
  if (inverse)
    EXH_CHECK_FFT(rocfft_plan_description_set_data_layout(
        desc, rocfft_array_type_hermitian_interleaved, rocfft_array_type_real, nullptr,
        nullptr, 0, nullptr, 0, 0, nullptr, 0));
  else
    EXH_CHECK_FFT(rocfft_plan_description_set_data_layout(
        desc, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved, nullptr,
        nullptr, 0, nullptr, 0, 0, nullptr, 0));
  rocfft_plan plan = nullptr;
  EXH_CHECK_FFT(rocfft_plan_create(
      &plan, rocfft_placement_notinplace,
      inverse ? rocfft_transform_type_real_inverse : rocfft_transform_type_real_forward,
      prec == 4 ? rocfft_precision_single : rocfft_precision_double, (size_t)dims, len,
      (size_t)batch,
      desc));
  rocfft_plan_description_destroy(desc);
  size_t ws = 0;
  EXH_CHECK_FFT(rocfft_plan_get_work_buffer_size(plan, &ws));
  rocfft_execution_info info = nullptr;
  EXH_CHECK_FFT(rocfft_execution_info_create(&info));
  if (ws) {
    EXH_CHECK_HIP(hipMalloc(&p.work, ws));
    EXH_CHECK_FFT(rocfft_execution_info_set_work_buffer(info, p.work, ws));
  }
  p.plan = plan;
  p.info = info;
  p.work_bytes = ws;
  return c->plans.emplace(key, p).first->second;
}

static void run_plan(exh_ctx* c, FftPlan& p, const void* in, void* out) {
  EXH_CHECK_FFT(rocfft_execution_info_set_stream((rocfft_execution_info)p.info, c->stream));
  void* ib[1] = {const_cast<void*>(in)};
  void* ob[1] = {out};
  EXH_CHECK_FFT(rocfft_execute((rocfft_plan)p.plan, ib, ob, (rocfft_execution_info)p.info));
}

// The strided axes go through the LDS-staged passes of fft.hip when both
// have a supported length (EXH_FFT_AXIS=0 keeps rocFFT's 3D plans).
static bool use_axis_passes(int n0, int n1) {
  static const bool on = [] {
    const char* e = std::getenv("EXH_FFT_AXIS");
    return !(e && std::atoi(e) == 0);
  }();
  return on && fft_axis_supported(n0) && fft_axis_supported(n1);
}

// Real transforms of `rows` rows along the contiguous axis: the LDS row
// kernels of fft.hip where they cover n2, else a batched rocFFT plan.
static void z_real(exh_ctx* c, int n2, int prec, bool inverse, long long rows, const void* in,
                   void* out) {
  if (fft_rows_supported(n2))
    fft_rows_pass(c, prec, inverse, n2, in, out, rows);
  else
    run_plan(c, get_plan(c, 0, 0, n2, prec, inverse ? 1 : 0, rows, 1), in, out);
}

bool fft_plan_free(int n0, int n1, int n2) {
  return use_axis_passes(n0, n1) && fft_rows_supported(n2);
}

void fft_r2c(exh_ctx* c, int n0, int n1, int n2, int prec, int batch, const void* in, void* out) {
  TimedScope ts(c, "fft_r2c");
  if (!use_axis_passes(n0, n1)) {
    run_plan(c, get_plan(c, n0, n1, n2, prec, 0, batch, 3), in, out);
    return;
  }
  const long long nz2 = n2 / 2 + 1, plane = (long long)n1 * nz2;
  z_real(c, n2, prec, false, (long long)batch * n0 * n1, in, out);
  fft_axis_pass(c, prec, false, n1, out, nz2, nz2, plane, (long long)batch * n0);
  fft_axis_pass(c, prec, false, n0, out, plane, plane, n0 * plane, batch);
}

void fft_c2r(exh_ctx* c, int n0, int n1, int n2, int prec, int batch, void* in, void* out) {
  TimedScope ts(c, "fft_c2r");
  if (!use_axis_passes(n0, n1)) {
    run_plan(c, get_plan(c, n0, n1, n2, prec, 1, batch, 3), in, out);
    return;
  }
  const long long nz2 = n2 / 2 + 1, plane = (long long)n1 * nz2;
  fft_axis_pass(c, prec, true, n0, in, plane, plane, n0 * plane, batch);
  fft_axis_pass(c, prec, true, n1, in, nz2, nz2, plane, (long long)batch * n0);
  z_real(c, n2, prec, true, (long long)batch * n0 * n1, in, out);
}

// EXH_FFT_POISSON=0: the Poisson factor in its own kernel instead of the
// first inverse pass.
static bool fuse_poisson() {
  static const bool on = [] {
    const char* e = std::getenv("EXH_FFT_POISSON");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

bool fft_c2r_poisson(exh_ctx* c, int prec, const void* src, void* work, void* out,
                     const PoissonParams& p) {
  if (!use_axis_passes(p.n0, p.n1) || !fuse_poisson()) return false;
  TimedScope ts(c, "fft_c2r");
  const long long nz2 = p.n2 / 2 + 1, plane = (long long)p.n1 * nz2;
  fft_axis0_inverse_poisson(c, prec, src, work, p);
  fft_axis_pass(c, prec, true, p.n1, work, nz2, nz2, plane, p.n0);
  z_real(c, p.n2, prec, true, (long long)p.n0 * p.n1, work, out);
  return true;
}

void fft_release(exh_ctx* c) {
  for (auto& kv : c->plans) {
    if (kv.second.plan) rocfft_plan_destroy((rocfft_plan)kv.second.plan);
    if (kv.second.info) rocfft_execution_info_destroy((rocfft_execution_info)kv.second.info);
    if (kv.second.work) (void)hipFree(kv.second.work);
  }
  c->plans.clear();
  fft_axis_release(c);
}

}  // namespace exh
