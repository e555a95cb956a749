// Shared definitions for the gfx950 HIP library behind the C-ABI in
// include/exhip.h: error reporting, the per-context state (stream, rocFFT
// plans, reusable device buffers) and small device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "exhip.h"

namespace exh {

// ---------------------------------------------------------------------------
// errors: every entry point returns an int status and keeps a thread-local
// message readable through exh_last_error().
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);

struct Error {
  int code;
};

#define EXH_CHECK_HIP(expr)                                                       \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::exh::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,         \
                       hipGetErrorString(_e));                                    \
      throw ::exh::Error{EXH_ERR_HIP};                                            \
    }                                                                             \
  } while (0)

#define EXH_REQUIRE(cond, code, ...)     \
  do {                                   \
    if (!(cond)) {                       \
      ::exh::set_error(__VA_ARGS__);     \
      throw ::exh::Error{code};          \
    }                                    \
  } while (0)

#define EXH_LAUNCH_CHECK() EXH_CHECK_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// reusable device buffers owned by a context
// ---------------------------------------------------------------------------
enum Slot {
  BUF_CK0,      // half-complex grid (field modes, potentials)
  BUF_CK1,      // second half-complex grid
  BUF_CKF,      // half-complex work grid of the density field's transform
  BUF_R0,       // real grid (potential)
  BUF_R1,       // real grid (2LPT source / second potential)
  BUF_R2,       // real grid (the pipeline's density field)
  BUF_S,        // displacement field [ng][3]
  BUF_V,        // velocity field [ng][3]
  BUF_FLAG,     // int32 halo flags
  BUF_CLAIM,    // uint64 per-cell claims used by the halo rounds
  BUF_PEAK_IDX, // lexicographic peak indices (sort values)
  BUF_PEAK_CELL,// peak cell indices in lexicographic order
  BUF_PEAK_KEY, // peak densities
  BUF_SORT_K,   // sort ping-pong keys
  BUF_SORT_V,   // sort ping-pong values
  BUF_SORT_AUX, // sort scratch
  BUF_SORT_K2,  // second sort work buffer (keys)
  BUF_SORT_V2,  // second sort work buffer (values)
  BUF_CAND,     // halo candidate state
  BUF_HALO,     // halo outputs
  BUF_MISC,     // small scratch (counters, tables)
  BUF_SCAN,     // hipcub temporary storage
  BUF_GRIDK,    // transformed grids for P(k)
  BUF_PART,     // P(k) block partials
  BUF_TAB,      // P(k) spline knots
  BUF_SHELL,    // shell offset table of the halo finder
  BUF_HALOTAB,  // halo finder tables (barrier, mass function, spheres, taus)
  BUF_PTAB,     // power spectrum phase and window tables
  BUF_GCNT,     // tracer counts of the grid deposit
  BUF_SHELLS,   // bispectrum: real fields of every k shell and tracer
  BUF_SHELLK,   // bispectrum: combined modes, shell index per mode, shell work grids
  BUF_EFF,      // halo growth: delta with flagged cells set to -Mtot
  BUF_ADJ_TAB,  // adjoint: per-bin factors
  BUF_ADJ_S1,   // adjoint: first-order displacements [ng][3]
  BUF_ADJ_POS,  // adjoint: recomputed particle positions [ng][3]
  BUF_ADJ_GRID, // adjoint: recomputed density grid
  BUF_ADJ_G,    // adjoint: dL/dgrid, then dL/dphi
  BUF_ADJ_GS,   // adjoint: dL/dpos, then dL/dS1 [ng][3]
  BUF_ADJ_G6,   // adjoint: dL/dS_ab of the 2LPT source [6][ng]
  BUF_NSLOTS
};

struct FftKey {
  int n0, n1, n2, prec, inverse;
  long long batch;
  int dims;  // 3: full real transform; 1: along the contiguous axis only
  bool operator<(const FftKey& o) const {
    return std::tie(n0, n1, n2, prec, inverse, batch, dims) <
           std::tie(o.n0, o.n1, o.n2, o.prec, o.inverse, o.batch, o.dims);
  }
};

struct FftPlan {
  void* plan = nullptr;   // rocfft_plan
  void* info = nullptr;   // rocfft_execution_info
  void* work = nullptr;   // work buffer
  size_t work_bytes = 0;
};

struct Timing {
  std::string name;
  hipEvent_t start, stop;
};

// Lattice offsets grouped by r^2: shell s occupies offs[start[s] ..
// start[s+1]) in the reference's cube order.
struct ShellTab {
  const int* start = nullptr;
  const unsigned* offs = nullptr;
  int stab = 0;
};

// Device arrays of the last halo run (views into context buffers).
struct HaloState {
  int* hx = nullptr;
  int* hcount = nullptr;
  int* hr2 = nullptr;
  void* Mh = nullptr;
  void* posh = nullptr;
  void* velh = nullptr;
  long long* hidx = nullptr;
  int* flag = nullptr;
  ShellTab sh;
  int sh_tmax = -1;
  long long ng = 0;
  bool have_posh = false, have_velh = false;
};

}  // namespace exh

// Byte key of a cached table: appends the raw bytes of n objects.
template <typename U>
inline void key_append(std::vector<unsigned char>& key, const U* p, size_t n) {
  const unsigned char* b = reinterpret_cast<const unsigned char*>(p);
  key.insert(key.end(), b, b + n * sizeof(U));
}

struct exh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority) for work that overlaps the main
  // stream's, ordered by the two events; created on first use
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_spec = nullptr;
  // third stream (default priority) for the peak sort's small segments,
  // forked per level and joined at the end of the sort; created on first use
  hipStream_t side = nullptr;
  hipEvent_t ev_side_fork = nullptr, ev_side_join = nullptr;
  bool own_stream = false;
  // set for the calls that must not allocate: ctx_buf and ctx_pinned then
  // raise instead of allocating
  bool no_alloc = false;
  // per-context tunables (exh_ctx_set_tunable); 0 = the library default
  // (or its environment override)
  int tune_deposit_tile = 0;
  int tune_deposit_acc = 0;
  int tune_halo_claims = 0;
  void* buf[exh::BUF_NSLOTS] = {};
  size_t cap[exh::BUF_NSLOTS] = {};
  // allocation number of each slot's current buffer (0: none yet)
  unsigned long long buf_gen[exh::BUF_NSLOTS] = {};
  unsigned long long n_allocs = 0;
  // halo growth claims: the allocation they were cleared for and the last
  // round tag (epoch) used in them
  unsigned long long claim_gen = 0;
  unsigned claim_epoch = 0;
  int claim_mode = 0;  // 32 or 64: the width of the claims in the array
  std::map<exh::FftKey, exh::FftPlan> plans;
  std::map<int, void*> fft_tw;  // twiddle tables of the strided FFT passes
  // per-kernel HIP-event timing (enabled by exh_ctx_set_timing)
  bool timing = false;
  std::vector<exh::Timing> events;
  // results of the last halo run kept on the device
  exh::HaloState halo;
  long long nh = 0;
  int last_prec = 4;
  // pinned host staging for small results
  void* pinned = nullptr;
  size_t pinned_cap = 0;
  // host-built tables kept on the device while their inputs are unchanged
  std::vector<unsigned char> halotab_key;  // halo tables in BUF_HALOTAB
  int halotab_ncells = 0;
  double halotab_bmin = 0;
  std::vector<unsigned char> spline_key;  // P(k) spline in BUF_TAB
  // P(k) sums left on the device by a deferred run_power (copied out by
  // power_copy_out): offsets into BUF_PART and the layout
  struct PowerPending {
    size_t o_sum = 0, o_cnt = 0;
    int Nk = 0, NV = 0, NPs = 0, ls = 0;
  } power_pending;
  double spline_x0 = 0, spline_x1 = 0;    // its first and last knot
  long long shell_noffs = 0;
  // halo finder statistics of the last run
  long long stat_peaks = 0, stat_candidates = 0, stat_rounds = 0, stat_sort_levels = 0,
            stat_sort_doubling = 0, stat_sort_async_open = 0;
};

namespace exh {

void* ctx_buf(exh_ctx* c, Slot s, size_t bytes);
void* ctx_pinned(exh_ctx* c, size_t bytes);

// rocFFT wrappers (fft.cpp).  Layouts are C order [n0][n1][n2] real and
// [n0][n1][n2/2+1] complex; `batch` grids are contiguous.
void fft_r2c(exh_ctx* c, int n0, int n1, int n2, int prec, int batch, const void* in, void* out);
void fft_c2r(exh_ctx* c, int n0, int n1, int n2, int prec, int batch, void* in, void* out);
void fft_release(exh_ctx* c);
// true when a 3D transform of this shape runs on the strided and row passes
// alone: no rocFFT plan, whose work buffer calls with the same shape share,
// so two such transforms may run on different streams at once
bool fft_plan_free(int n0, int n1, int n2);
// strided complex passes of the 3D transforms (fft.hip)
bool fft_axis_supported(int n);
void fft_axis_pass(exh_ctx* c, int prec, bool inverse, int N, void* data, long long stride,
                   long long ncol, long long line_stride, long long nlines);
void fft_axis_release(exh_ctx* c);
// real transforms along the contiguous axis of length n (fft.hip): real rows
// of pitch n <-> half-complex rows of pitch n/2+1
bool fft_rows_supported(int n);
void fft_rows_pass(exh_ctx* c, int prec, bool inverse, int N, const void* in, void* out,
                   long long nrows);

// event timing around a launch sequence
struct TimedScope {
  exh_ctx* c;
  int idx = -1;
  TimedScope(exh_ctx* ctx, const char* name);
  ~TimedScope();
};

inline unsigned grid_for(long long n, int block) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 0x7fffffffLL) g = 0x7fffffffLL;
  return (unsigned)g;
}

// integer floor(sqrt(s)) for s >= 0, exact (the reference evaluates
// floor(sqrt((double)s)) with a correctly rounded sqrt).
__host__ __device__ inline int isqrt_floor(int s) {
  int t = (int)sqrt((double)s);
  while ((t + 1) * (t + 1) <= s) t++;
  while (t * t > s) t--;
  return t;
}

// periodic single wrap (exshalos_h.c:22-30)
__host__ __device__ inline int wrap1(int i, int j, int n) {
  int r = i + j;
  if (r >= n) r -= n;
  if (r < 0) r += n;
  return r;
}

// Division by a fixed divisor 1 <= d < 2^31 through a 64-bit reciprocal
// m = ceil(2^64 / d): for 0 <= x < 2^31 the high 64 bits of x m are
// exactly floor(x / d) (the error x (m - 2^64/d) / 2^64 < 2^-33 cannot
// carry a fraction of at most (d-1)/d across an integer).
struct FastDiv {
  unsigned d;
  unsigned long long m;
};

inline FastDiv fast_div(unsigned d) {
  FastDiv f;
  f.d = d;
  f.m = d > 1 ? ~0ull / d + 1 : 0;
  return f;
}

__device__ inline unsigned fdiv(unsigned x, const FastDiv& f) {
  return f.d == 1 ? x : (unsigned)__umul64hi((unsigned long long)x, f.m);
}

template <typename T>
struct vec3;
template <>
struct vec3<float> {
  using type = float3;
};
template <>
struct vec3<double> {
  using type = double3;
};

template <typename T>
struct cplx;
template <>
struct cplx<float> {
  using type = float2;
};
template <>
struct cplx<double> {
  using type = double2;
};

}  // namespace exh
