// extern "C" entry points of include/exhip.h.  Each call validates its
// arguments, dispatches on precision and converts exceptions into status
// codes; no entry point exits the process.
#include <cstring>

#include <cstdlib>

#include "common.hpp"
#include "host_tables.hpp"
#include "kernels.hpp"

namespace exh {
const char* last_error();
}

using namespace exh;

#define EXH_API_BEGIN try {
#define EXH_API_END                                   \
  }                                                   \
  catch (const exh::Error& e) {                       \
    return e.code;                                    \
  }                                                   \
  catch (const std::exception& e) {                   \
    exh::set_error("internal error: %s", e.what());   \
    return EXH_ERR_INTERNAL;                          \
  }                                                   \
  return EXH_OK;

static void check_ctx(exh_ctx* c) {
  EXH_REQUIRE(c != nullptr, EXH_ERR_ARG, "null context");
  EXH_CHECK_HIP(hipSetDevice(c->device));
}

static void check_box(const exh_box* b) {
  EXH_REQUIRE(b != nullptr, EXH_ERR_ARG, "null box");
  EXH_REQUIRE(b->precision == 4 || b->precision == 8, EXH_ERR_ARG, "precision must be 4 or 8");
  EXH_REQUIRE(b->nd[0] >= 4 && b->nd[1] >= 4 && b->nd[2] >= 4, EXH_ERR_ARG,
              "grid too small (at least 4 cells per side for the 7-point stencils)");
  EXH_REQUIRE(b->nd[0] % 2 == 0 && b->nd[1] % 2 == 0 && b->nd[2] % 2 == 0, EXH_ERR_ARG,
              "grid sizes must be even");
}

static void check_src(const exh_field_src* s) {
  EXH_REQUIRE(s != nullptr && s->data != nullptr, EXH_ERR_ARG, "null field source");
  EXH_REQUIRE(s->kind == EXH_SRC_MODES || s->kind == EXH_SRC_DELTA || s->kind == EXH_SRC_DELTAK,
              EXH_ERR_ARG, "unknown field source kind %d", s->kind);
  if (s->kind == EXH_SRC_MODES)
    EXH_REQUIRE(s->k && s->P && s->Nk >= 3, EXH_ERR_ARG, "modes need a P(k) table (Nk >= 3)");
}

static void check_halo_opts(const exh_halo_opts* o) {
  EXH_REQUIRE(o != nullptr, EXH_ERR_ARG, "null halo options");
  EXH_REQUIRE(o->k && o->P && o->Nk >= 3, EXH_ERR_ARG, "halo finder needs P(k) (Nk >= 3)");
  EXH_REQUIRE(o->sigma_radii == 0 || (o->sigma_radii >= 4 && o->sigma_radii <= 100000),
              EXH_ERR_ARG, "sigma_radii must be 0 or 4..100000 (got %d)", o->sigma_radii);
}

extern "C" {

const char* exh_version(void) { return "exshalos-amd 0.1 (gfx950)"; }

const char* exh_last_error(void) { return exh::last_error(); }

int exh_ctx_create(int device, exh_ctx** out) {
  EXH_API_BEGIN
  EXH_REQUIRE(out != nullptr, EXH_ERR_ARG, "null output");
  int n = 0;
  EXH_CHECK_HIP(hipGetDeviceCount(&n));
  EXH_REQUIRE(device >= 0 && device < n, EXH_ERR_ARG, "device %d not present (%d devices)", device, n);
  EXH_CHECK_HIP(hipSetDevice(device));
  exh_ctx* c = new exh_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    EXH_CHECK_HIP(e);
  }
  c->own_stream = true;
  *out = c;
  EXH_API_END
}

int exh_ctx_destroy(exh_ctx* c) {
  EXH_API_BEGIN
  if (!c) return EXH_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  fft_release(c);
  for (int s = 0; s < BUF_NSLOTS; s++)
    if (c->buf[s]) (void)hipFree(c->buf[s]);
  for (auto& t : c->events) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->aux) {
    (void)hipStreamSynchronize(c->aux);
    (void)hipStreamDestroy(c->aux);
  }
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  if (c->ev_side_fork) (void)hipEventDestroy(c->ev_side_fork);
  if (c->ev_side_join) (void)hipEventDestroy(c->ev_side_join);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_spec) (void)hipEventDestroy(c->ev_spec);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  EXH_API_END
}

int exh_ctx_set_stream(exh_ctx* c, void* stream) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (stream) {
    if (c->own_stream) EXH_CHECK_HIP(hipStreamDestroy(c->stream));
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else if (!c->own_stream) {
    EXH_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
  }
  EXH_API_END
}

int exh_ctx_synchronize(exh_ctx* c) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

int exh_ctx_set_timing(exh_ctx* c, int enable) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  for (auto& t : c->events) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  c->events.clear();
  c->timing = enable != 0;
  EXH_API_END
}

int exh_ctx_timings(exh_ctx* c, char* names, int names_len, double* ms, int max) {
  int n = 0;
  try {
    check_ctx(c);
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    std::string all;
    for (auto& t : c->events) {
      if (n >= max) break;
      float v = 0;
      EXH_CHECK_HIP(hipEventElapsedTime(&v, t.start, t.stop));
      if (ms) ms[n] = v;
      all += t.name;
      all += '\n';
      n++;
    }
    if (names && names_len > 0) {
      std::strncpy(names, all.c_str(), names_len - 1);
      names[names_len - 1] = 0;
    }
  } catch (const exh::Error& e) {
    return e.code;
  }
  return n;
}

int exh_ctx_set_tunable(exh_ctx* c, int key, int value) {
  EXH_API_BEGIN
  check_ctx(c);
  if (key == EXH_TUNE_DEPOSIT_TILE) {
    EXH_REQUIRE(value == 0 || value == 16 || value == 3216 || value == 32, EXH_ERR_ARG,
                "deposit tile must be 0, 16, 3216 or 32 (got %d)", value);
    c->tune_deposit_tile = value;
  } else if (key == EXH_TUNE_DEPOSIT_ACC) {
    EXH_REQUIRE(value == 0 || value == 1 || value == 4 || value == 8, EXH_ERR_ARG,
                "deposit accumulator must be 0, 1, 4 or 8 (got %d)", value);
    c->tune_deposit_acc = value;
  } else if (key == EXH_TUNE_HALO_CLAIMS) {
    EXH_REQUIRE(value == 0 || value == 64, EXH_ERR_ARG, "halo claims must be 0 or 64 (got %d)",
                value);
    c->tune_halo_claims = value;
  } else {
    EXH_REQUIRE(false, EXH_ERR_ARG, "unknown tunable %d", key);
  }
  EXH_API_END
}

int exh_ctx_halo_stats(exh_ctx* c, long long* out5) {
  EXH_API_BEGIN
  EXH_REQUIRE(c && out5, EXH_ERR_ARG, "null argument");
  out5[0] = c->stat_peaks;
  out5[1] = c->stat_candidates;
  out5[2] = c->stat_rounds;
  out5[3] = c->stat_sort_levels;
  out5[4] = c->stat_sort_doubling;
  EXH_API_END
}

int exh_box_init(exh_box* box, int ndx, int ndy, int ndz, double Lc_or_Mc, double Om0,
                 double redshift, double dc, int precision) {
  EXH_API_BEGIN
  EXH_REQUIRE(box != nullptr, EXH_ERR_ARG, "null box");
  EXH_REQUIRE(precision == 4 || precision == 8, EXH_ERR_ARG, "precision must be 4 or 8");
  EXH_REQUIRE(ndx > 0 && ndy > 0 && ndz > 0, EXH_ERR_ARG, "grid sizes must be positive");
  int rc = box_init(box, ndx, ndy, ndz, Lc_or_Mc, Om0, redshift, dc, precision);
  EXH_REQUIRE(rc == 0, EXH_ERR_ARG,
              "Lc/Mc = %g: a cell larger than 100 Mpc/h or lighter than 1e5 Msun/h is not "
              "supported",
              Lc_or_Mc);
  EXH_API_END
}

int exh_density_grid(exh_ctx* c, const exh_box* b, const exh_field_src* modes, double k_smooth,
                     void* delta, void* deltak) {
  EXH_API_BEGIN
  check_ctx(c);
  check_box(b);
  check_src(modes);
  EXH_REQUIRE(modes->kind == EXH_SRC_MODES, EXH_ERR_ARG, "density grid needs white-noise modes");
  EXH_REQUIRE(delta != nullptr, EXH_ERR_ARG, "null delta");
  if (b->precision == 4)
    launch_field<float>(c, b, modes, k_smooth, delta, deltak);
  else
    launch_field<double>(c, b, modes, k_smooth, delta, deltak);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

int exh_halo_tables(const exh_box* b, const exh_halo_opts* o, void* sig_grid, void* barrier,
                    int nB, double* Mknots, double* Int, int* Ncells) {
  EXH_API_BEGIN
  check_box(b);
  check_halo_opts(o);
  EXH_REQUIRE(sig_grid && barrier && Mknots && Int && nB >= 1, EXH_ERR_ARG, "null table");
  int nc = halo_tables(b, o, sig_grid, barrier, nB, Mknots, Int);
  EXH_REQUIRE(nc >= 0, EXH_ERR_RANGE, "sigma(M) evaluated outside its knots");
  if (Ncells) *Ncells = nc;
  EXH_API_END
}

int exh_find_halos(exh_ctx* c, const exh_box* b, const void* delta, const exh_halo_opts* o,
                   long long* nh) {
  EXH_API_BEGIN
  check_ctx(c);
  check_box(b);
  check_halo_opts(o);
  EXH_REQUIRE(delta != nullptr, EXH_ERR_ARG, "null delta");
  long long n = b->precision == 4 ? run_find_halos<float>(c, b, (const float*)delta, o)
                                  : run_find_halos<double>(c, b, (const double*)delta, o);
  if (b->precision == 4)
    run_halo_posh<float>(c, b, nullptr, nullptr, 1);
  else
    run_halo_posh<double>(c, b, nullptr, nullptr, 1);
  c->halo.have_posh = true;
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (nh) *nh = n;
  EXH_API_END
}

static void copy_out(exh_ctx* c, void* dst, const void* src, size_t bytes, int to_host) {
  if (!dst || bytes == 0) return;
  EXH_CHECK_HIP(hipMemcpyAsync(dst, src, bytes,
                               to_host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                               c->stream));
}

int exh_halos_get(exh_ctx* c, int* x, int* count, int* r2, void* Mh, int to_host) {
  EXH_API_BEGIN
  check_ctx(c);
  size_t n = (size_t)c->nh;
  size_t es = c->last_prec == 4 ? 4 : 8;
  copy_out(c, x, c->halo.hx, n * 12, to_host);
  copy_out(c, count, c->halo.hcount, n * 4, to_host);
  copy_out(c, r2, c->halo.hr2, n * 4, to_host);
  copy_out(c, Mh, c->halo.Mh, n * es, to_host);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

int exh_halos_posh(exh_ctx* c, void* posh, void* velh, int to_host) {
  EXH_API_BEGIN
  check_ctx(c);
  size_t n = (size_t)c->nh;
  size_t es = c->last_prec == 4 ? 4 : 8;
  EXH_REQUIRE(c->halo.have_posh || n == 0, EXH_ERR_ARG, "no halo positions computed");
  EXH_REQUIRE(!velh || c->halo.have_velh || n == 0, EXH_ERR_ARG, "no halo velocities computed");
  copy_out(c, posh, c->halo.posh, n * 3 * es, to_host);
  copy_out(c, velh, c->halo.velh, n * 3 * es, to_host);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

int exh_halos_flag(exh_ctx* c, long long* flag, int to_host) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_REQUIRE(flag != nullptr && c->halo.flag != nullptr, EXH_ERR_ARG, "no flag grid");
  exh_box b{};
  b.ng = c->halo.ng;
  long long* dst = flag;
  if (to_host) dst = (long long*)ctx_buf(c, BUF_R1, (size_t)b.ng * 8);
  run_halo_flag(c, &b, dst);
  if (to_host)
    EXH_CHECK_HIP(hipMemcpyAsync(flag, dst, (size_t)b.ng * 8, hipMemcpyDeviceToHost, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

int exh_lpt(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth, int do_2lpt,
            int out_pos, void* S, void* V) {
  EXH_API_BEGIN
  check_ctx(c);
  check_box(b);
  check_src(s);
  EXH_REQUIRE(S != nullptr, EXH_ERR_ARG, "null displacement output");
  EXH_REQUIRE(b->nd[0] == b->nd[1] && b->nd[1] == b->nd[2], EXH_ERR_UNSUPPORTED,
              "LPT needs a cubic grid (the reference's y stencil indexes rows with nd[0], "
              "lpt.c:66)");
  if (b->precision == 4)
    launch_lpt<float>(c, b, s, k_smooth, do_2lpt, out_pos, (float*)S, (float*)V);
  else
    launch_lpt<double>(c, b, s, k_smooth, do_2lpt, out_pos, (double*)S, (double*)V);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

int exh_pk_adjoint(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                   int do_2lpt, int window, int Nk, double k_min, double k_max,
                   int kbin_precision, const double* gP, void* g_noise, double* P_out,
                   long long* count_out) {
  EXH_API_BEGIN
  check_ctx(c);
  check_box(b);
  check_src(s);
  EXH_REQUIRE(gP != nullptr && g_noise != nullptr, EXH_ERR_ARG, "null gradient array");
  EXH_REQUIRE(Nk >= 1 && k_max > k_min, EXH_ERR_ARG, "bad binning");
  EXH_REQUIRE(kbin_precision == 4 || kbin_precision == 8, EXH_ERR_ARG,
              "kbin_precision must be 4 or 8");
  if (b->precision == 4)
    run_pk_adjoint<float>(c, b, s, k_smooth, do_2lpt, window, Nk, k_min, k_max, kbin_precision,
                          gP, (float*)g_noise, P_out, count_out);
  else
    run_pk_adjoint<double>(c, b, s, k_smooth, do_2lpt, window, Nk, k_min, k_max, kbin_precision,
                           gP, (double*)g_noise, P_out, count_out);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

}  // extern "C"

// Runs the caller's launches on another stream of the context until the
// scope ends.
struct StreamSwitch {
  exh_ctx* c;
  hipStream_t saved;
  StreamSwitch(exh_ctx* ctx, hipStream_t s) : c(ctx), saved(ctx->stream) { c->stream = s; }
  ~StreamSwitch() { c->stream = saved; }
};

static void ensure_aux(exh_ctx* c) {
  if (c->aux) return;
  EXH_CHECK_HIP(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
  EXH_CHECK_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  EXH_CHECK_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  EXH_CHECK_HIP(hipEventCreateWithFlags(&c->ev_spec, hipEventDisableTiming));
}

// Generate_Halos_Box_from_Pk (box.c:4-68).  The LPT depends only on the
// field's delta_k and the halo finder only on delta, so the LPT runs on the
// auxiliary stream while the finder (latency-bound, with a host round trip
// per sort level and growth round) runs on the main stream; the two share
// no buffer.  The halo positions wait for both.
// Mass assignment and power spectrum of the particle positions, chained
// after the LPT on the same stream (exh_halos_box_power).
struct SpectrumStage {
  bool on = false;
  int window = 2, interlacing = 1, Nk = 0, kbin = 8;
  double k_min = 0, k_max = 0;
  void* grid = nullptr;
  double* k_out = nullptr;
  double* P_out = nullptr;
  long long* count_out = nullptr;
};

template <typename T>
static long long pipeline(exh_ctx* c, const exh_box* b, const exh_field_src* modes,
                          const exh_halo_opts* o, double k_smooth, int do_2lpt, int out_vel,
                          void* delta_out, void* S_out, void* V_out,
                          const SpectrumStage& sp = SpectrumStage{}) {
  const size_t ng = (size_t)b->ng;
  T* delta = delta_out ? (T*)delta_out : (T*)ctx_buf(c, BUF_R2, ng * sizeof(T));
  // delta_k is kept for the LPT, which divides it by -k^2 as the reference
  // does with its deltak array (box.c:17-44)
  const size_t nhalf = (size_t)b->nd[0] * b->nd[1] * b->nz2;
  void* deltak = ctx_buf(c, BUF_CK1, nhalf * 2 * sizeof(T));
  // displacements in scratch (the halo positions average them); the last
  // LPT pass writes the particle positions into S_out
  T* S = (T*)ctx_buf(c, BUF_S, ng * 3 * sizeof(T));
  T* V = out_vel ? (V_out ? (T*)V_out : (T*)ctx_buf(c, BUF_V, ng * 3 * sizeof(T))) : nullptr;
  exh_field_src ks{};
  ks.kind = EXH_SRC_DELTAK;
  ks.data = deltak;
  ensure_aux(c);
  // The second stream's chain starts at the point EXH_LPT_OVERLAP names:
  // -2 once the field's modes (delta_k) are written, -1 after the field,
  // 0..2 at the halo finder's phases (see run_find_halos), 3 after it.  The
  // mass assignment and power spectrum follow the LPT on that stream, and
  // are queued at the point EXH_SPECTRUM_AT names (same numbering, never
  // before the LPT).  Defaults, as measured at nd = 512: the LPT alone
  // starts before the growth rounds, whose later rounds leave most of the
  // GPU idle; LPT + grid + P(k) start together with the finder (queued
  // later, the grid and P(k) slow the growth rounds more than they gain).
  static const int fork_env = [] {
    const char* e = std::getenv("EXH_LPT_OVERLAP");
    return e ? std::atoi(e) : -100;
  }();
  static const int spec_env = [] {
    const char* e = std::getenv("EXH_SPECTRUM_AT");
    return e ? std::atoi(e) : -100;
  }();
  int fork_at = fork_env != -100 ? fork_env : (sp.on ? 0 : 2);
  // the field's inverse transform and the LPT's may overlap only where
  // neither uses a rocFFT plan (their work buffer would be shared)
  if (fork_at == -2 && !fft_plan_free(b->nd[0], b->nd[1], b->nd[2])) fork_at = -1;
  const int spec_at = std::max(fork_at, spec_env != -100 ? spec_env : fork_at);
  bool forked = false, spec_queued = !sp.on;
  auto queue_spectrum = [&]() {
    spec_queued = true;
    EXH_CHECK_HIP(hipEventRecord(c->ev_spec, c->stream));
    EXH_CHECK_HIP(hipStreamWaitEvent(c->aux, c->ev_spec, 0));
    StreamSwitch sw(c, c->aux);
    const int nd = b->nd[0];
    run_grid<T>(c, (const T*)S_out, nullptr, nullptr, nullptr, (long long)ng, 1, nd, b->L[0],
                b->Om0, b->redshift, -1, sp.window, 0.0, 0.0, sp.interlacing, 1, (T*)sp.grid);
    run_power<T>(c, (const T*)sp.grid, 1, nd, b->L[0], sp.window, 0.0, sp.interlacing, sp.Nk,
                 sp.k_min, sp.k_max, 0, -1, sp.kbin, nullptr, nullptr, nullptr, true);
    EXH_CHECK_HIP(hipEventRecord(c->ev_join, c->stream));
  };
  auto fork_lpt = [&]() {
    forked = true;
    EXH_CHECK_HIP(hipEventRecord(c->ev_fork, c->stream));
    EXH_CHECK_HIP(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    {
      StreamSwitch sw(c, c->aux);
      launch_lpt<T>(c, b, &ks, k_smooth, do_2lpt, 0, S, V, (T*)S_out);
      EXH_CHECK_HIP(hipEventRecord(c->ev_join, c->stream));
    }
    if (!spec_queued && spec_at == fork_at) queue_spectrum();
  };
  auto at_point = [&](int p) {
    if (p == fork_at) fork_lpt();
    if (forked && !spec_queued && p == spec_at) queue_spectrum();
  };
  try {
    launch_field<T>(c, b, modes, 1e5, delta, deltak, [&]() { at_point(-2); });
    at_point(-1);
  } catch (...) {
    if (forked) (void)hipStreamSynchronize(c->aux);
    throw;
  }
  long long nh;
  try {
    nh = run_find_halos<T>(c, b, delta, o, at_point);
  } catch (...) {
    if (forked) (void)hipStreamSynchronize(c->aux);  // no LPT work outlives a failed call
    throw;
  }
  // the finder ended before those points (or they are > 2)
  if (!forked) fork_lpt();
  if (!spec_queued) queue_spectrum();
  EXH_CHECK_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  run_halo_posh<T>(c, b, S, V, 0);
  c->halo.have_posh = true;
  c->halo.have_velh = V != nullptr;
  if (sp.on) power_copy_out(c, sp.k_out, sp.P_out, sp.count_out);
  return nh;
}

extern "C" int exh_halos_box_from_pk(exh_ctx* c, const exh_box* b, const exh_field_src* modes,
                          const exh_halo_opts* o, double k_smooth, int do_2lpt, int out_vel,
                          void* delta, void* S, void* V, long long* nh) {
  EXH_API_BEGIN
  check_ctx(c);
  check_box(b);
  check_src(modes);
  check_halo_opts(o);
  EXH_REQUIRE(modes->kind == EXH_SRC_MODES, EXH_ERR_ARG, "pipeline needs white-noise modes");
  EXH_REQUIRE(b->nd[0] == b->nd[1] && b->nd[1] == b->nd[2], EXH_ERR_UNSUPPORTED,
              "the pipeline needs a cubic grid");
  long long n = b->precision == 4
                    ? pipeline<float>(c, b, modes, o, k_smooth, do_2lpt, out_vel, delta, S, V)
                    : pipeline<double>(c, b, modes, o, k_smooth, do_2lpt, out_vel, delta, S, V);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (nh) *nh = n;
  EXH_API_END
}

extern "C" int exh_halos_box_power(exh_ctx* c, const exh_box* b, const exh_field_src* modes,
                                   const exh_halo_opts* o, double k_smooth, int do_2lpt,
                                   void* pos, int window, int interlacing, int Nk, double k_min,
                                   double k_max, int kbin_precision, void* grid, double* k_out,
                                   double* P_out, long long* count_out, long long* nh) {
  EXH_API_BEGIN
  check_ctx(c);
  check_box(b);
  check_src(modes);
  check_halo_opts(o);
  EXH_REQUIRE(modes->kind == EXH_SRC_MODES, EXH_ERR_ARG, "pipeline needs white-noise modes");
  EXH_REQUIRE(b->nd[0] == b->nd[1] && b->nd[1] == b->nd[2], EXH_ERR_UNSUPPORTED,
              "the pipeline needs a cubic grid");
  EXH_REQUIRE(pos != nullptr && grid != nullptr, EXH_ERR_ARG, "null positions or grid");
  EXH_REQUIRE(window == EXH_WINDOW_NGP || window == EXH_WINDOW_CIC || window == EXH_WINDOW_TSC,
              EXH_ERR_UNSUPPORTED, "window %d: only NGP (1), CIC (2) and TSC (5)", window);
  EXH_REQUIRE(Nk >= 1 && k_max > k_min, EXH_ERR_ARG, "bad binning");
  EXH_REQUIRE(kbin_precision == 4 || kbin_precision == 8, EXH_ERR_ARG,
              "kbin_precision must be 4 or 8");
  SpectrumStage sp;
  sp.on = true;
  sp.window = window;
  sp.interlacing = interlacing;
  sp.Nk = Nk;
  sp.kbin = kbin_precision;
  sp.k_min = k_min;
  sp.k_max = k_max;
  sp.grid = grid;
  sp.k_out = k_out;
  sp.P_out = P_out;
  sp.count_out = count_out;
  long long n =
      b->precision == 4
          ? pipeline<float>(c, b, modes, o, k_smooth, do_2lpt, 0, nullptr, pos, nullptr, sp)
          : pipeline<double>(c, b, modes, o, k_smooth, do_2lpt, 0, nullptr, pos, nullptr, sp);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (nh) *nh = n;
  EXH_API_END
}

template <typename T>
static long long pipeline_grid(exh_ctx* c, const exh_box* b, const T* delta,
                               const exh_halo_opts* o, double k_smooth, int do_2lpt, int out_vel,
                               const T* S_in, const T* V_in, T* S_out, T* V_out) {
  const size_t ng = (size_t)b->ng;
  long long nh = run_find_halos<T>(c, b, delta, o);
  if (S_in) {
    run_halo_posh<T>(c, b, S_in, V_in, 0);
    c->halo.have_velh = V_in != nullptr;
  } else {
    exh_field_src s{};
    s.kind = EXH_SRC_DELTA;
    s.data = delta;
    T* S = (T*)ctx_buf(c, BUF_S, ng * 3 * sizeof(T));
    T* V = out_vel ? (V_out ? V_out : (T*)ctx_buf(c, BUF_V, ng * 3 * sizeof(T))) : nullptr;
    launch_lpt<T>(c, b, &s, k_smooth, do_2lpt, 0, S, V, S_out);
    run_halo_posh<T>(c, b, S, V, 0);
    c->halo.have_velh = V != nullptr;
  }
  c->halo.have_posh = true;
  return nh;
}

extern "C" {

int exh_halos_box_from_grid(exh_ctx* c, const exh_box* b, const void* delta,
                            const exh_halo_opts* o, double k_smooth, int do_2lpt, int out_vel,
                            const void* S_in, const void* V_in, void* S_out, void* V_out,
                            long long* nh) {
  EXH_API_BEGIN
  check_ctx(c);
  check_box(b);
  check_halo_opts(o);
  EXH_REQUIRE(delta != nullptr, EXH_ERR_ARG, "null delta");
  EXH_REQUIRE(S_in || (b->nd[0] == b->nd[1] && b->nd[1] == b->nd[2]), EXH_ERR_UNSUPPORTED,
              "LPT needs a cubic grid");
  long long n =
      b->precision == 4
          ? pipeline_grid<float>(c, b, (const float*)delta, o, k_smooth, do_2lpt, out_vel,
                                 (const float*)S_in, (const float*)V_in, (float*)S_out,
                                 (float*)V_out)
          : pipeline_grid<double>(c, b, (const double*)delta, o, k_smooth, do_2lpt, out_vel,
                                  (const double*)S_in, (const double*)V_in, (double*)S_out,
                                  (double*)V_out);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (nh) *nh = n;
  EXH_API_END
}

int exh_grid_compute(exh_ctx* c, const void* pos, const void* vel, const void* mass,
                     const int* type, long long np, int ntype, int nd, double L, double Om0,
                     double z, int direction, int window, double R, double R_times,
                     int interlacing, int folds, int precision, void* grid) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_REQUIRE(precision == 4 || precision == 8, EXH_ERR_ARG, "precision must be 4 or 8");
  EXH_REQUIRE(grid != nullptr && (pos != nullptr || np == 0), EXH_ERR_ARG, "null array");
  EXH_REQUIRE(np >= 0 && nd > 0 && L > 0, EXH_ERR_ARG, "bad sizes");
  EXH_REQUIRE(ntype == 1 || type != nullptr, EXH_ERR_ARG, "ntype > 1 needs a type array");
  EXH_REQUIRE(direction >= -1 && direction <= 2, EXH_ERR_ARG, "direction must be -1, 0, 1 or 2");
  if (precision == 4)
    run_grid<float>(c, (const float*)pos, (const float*)vel, (const float*)mass, type, np, ntype,
                    nd, L, Om0, z, direction, window, R, R_times, interlacing, folds,
                    (float*)grid);
  else
    run_grid<double>(c, (const double*)pos, (const double*)vel, (const double*)mass, type, np,
                     ntype, nd, L, Om0, z, direction, window, R, R_times, interlacing, folds,
                     (double*)grid);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  EXH_API_END
}

int exh_grid_compute_noalloc(void* stream, const void* pos, const void* vel, const void* mass,
                             const int* type, long long np, int ntype, int nd, double L,
                             double Om0, double z, int direction, int window, double R,
                             double R_times, int interlacing, int folds, int precision,
                             void* grid) {
  EXH_API_BEGIN
  EXH_REQUIRE(precision == 4 || precision == 8, EXH_ERR_ARG, "precision must be 4 or 8");
  EXH_REQUIRE(grid != nullptr && (pos != nullptr || np == 0), EXH_ERR_ARG, "null array");
  EXH_REQUIRE(np >= 0 && nd > 0 && L > 0, EXH_ERR_ARG, "bad sizes");
  EXH_REQUIRE(ntype >= 1 && ntype <= 64, EXH_ERR_ARG,
              "the allocation-free deposit takes 1 to 64 tracer types (got %d)", ntype);
  EXH_REQUIRE(ntype * (interlacing ? 2 : 1) <= NOALLOC_GRID_WORDS, EXH_ERR_ARG,
              "too many grids for the allocation-free deposit's scratch");
  EXH_REQUIRE(ntype == 1 || type != nullptr, EXH_ERR_ARG, "ntype > 1 needs a type array");
  EXH_REQUIRE(direction >= -1 && direction <= 2, EXH_ERR_ARG, "direction must be -1, 0, 1 or 2");
  int dev = 0;
  EXH_CHECK_HIP(hipGetDevice(&dev));
  void* scr = noalloc_grid_scratch(dev);
  // a context on the caller's stream that holds no buffers and may not
  // allocate any (ctx_buf raises)
  exh_ctx bare;
  bare.device = dev;
  bare.stream = (hipStream_t)stream;
  bare.no_alloc = true;
  if (precision == 4)
    run_grid<float>(&bare, (const float*)pos, (const float*)vel, (const float*)mass, type, np,
                    ntype, nd, L, Om0, z, direction, window, R, R_times, interlacing, folds,
                    (float*)grid, scr);
  else
    run_grid<double>(&bare, (const double*)pos, (const double*)vel, (const double*)mass, type, np,
                     ntype, nd, L, Om0, z, direction, window, R, R_times, interlacing, folds,
                     (double*)grid, scr);
  EXH_API_END
}

int exh_power_compute(exh_ctx* c, const void* grid, int ntype, int nd, double L, int window,
                      double R, int interlacing, int Nk, double k_min, double k_max, int l_max,
                      int direction, int precision, int kbin_precision, double* k_out,
                      double* P_out, long long* count_out) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_REQUIRE(precision == 4 || precision == 8, EXH_ERR_ARG, "precision must be 4 or 8");
  EXH_REQUIRE(kbin_precision == 4 || kbin_precision == 8, EXH_ERR_ARG,
              "kbin_precision must be 4 or 8");
  EXH_REQUIRE(grid != nullptr, EXH_ERR_ARG, "null grid");
  EXH_REQUIRE(k_max > k_min, EXH_ERR_ARG, "k_max must exceed k_min");
  EXH_REQUIRE(direction >= -1 && direction <= 2, EXH_ERR_ARG, "direction must be -1, 0, 1 or 2");
  EXH_REQUIRE(window != 3 || R > 0, EXH_ERR_ARG, "the sphere window needs R > 0");
  if (precision == 4)
    run_power<float>(c, (const float*)grid, ntype, nd, L, window, R, interlacing, Nk, k_min,
                     k_max, l_max, direction, kbin_precision, k_out, P_out, count_out);
  else
    run_power<double>(c, (const double*)grid, ntype, nd, L, window, R, interlacing, Nk, k_min,
                      k_max, l_max, direction, kbin_precision, k_out, P_out, count_out);
  EXH_API_END
}

}  // extern "C"

int exh_abundance_compute(exh_ctx* c, const void* Mh, long long nh, double Mmin, double Mmax,
                          int Nm, double Lc, int ndx, int ndy, int ndz, int precision,
                          double* Mmean_out, double* dn_out, double* dn_err_out) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_REQUIRE(precision == 4 || precision == 8, EXH_ERR_ARG, "precision must be 4 or 8");
  EXH_REQUIRE(Lc > 0 && ndx > 0 && ndy > 0 && ndz > 0, EXH_ERR_ARG, "bad box");
  if (!Mh) {  // the masses of the context's last halo catalogue
    EXH_REQUIRE(c->halo.Mh != nullptr || c->nh == 0, EXH_ERR_ARG, "no halo catalogue");
    EXH_REQUIRE(c->last_prec == precision, EXH_ERR_ARG,
                "the last catalogue was computed at precision %d", c->last_prec);
    Mh = c->halo.Mh;
    nh = c->nh;
  }
  if (precision == 4)
    run_abundance<float>(c, (const float*)Mh, nh, Mmin, Mmax, Nm, Lc, ndx, ndy, ndz, Mmean_out,
                         dn_out, dn_err_out);
  else
    run_abundance<double>(c, (const double*)Mh, nh, Mmin, Mmax, Nm, Lc, ndx, ndy, ndz, Mmean_out,
                          dn_out, dn_err_out);
  EXH_API_END
}

int exh_bispectrum_size(int Nk) {
  if (Nk < 1 || Nk > 4096) return EXH_ERR_ARG;
  std::vector<std::array<int, 3>> tri;
  bispectrum_configs(Nk, tri);
  return (int)tri.size();
}

int exh_bispectrum_compute(exh_ctx* c, const void* grid, int ntype, int nd, double L, int window,
                           double R, int interlacing, int Nk, double k_min, double k_max,
                           int precision, long long cap, double* kP, double* Pk, double* Nk_out,
                           double* kB, double* Bk, double* Ntri_out, int* ntri) {
  EXH_API_BEGIN
  check_ctx(c);
  EXH_REQUIRE(precision == 4 || precision == 8, EXH_ERR_ARG, "precision must be 4 or 8");
  EXH_REQUIRE(grid && kP && Pk && Nk_out && kB && Bk && Ntri_out && ntri, EXH_ERR_ARG,
              "null array");
  EXH_REQUIRE(window != 3 || R > 0, EXH_ERR_ARG, "the sphere window needs R > 0");
  *ntri = precision == 4
              ? run_bispectrum<float>(c, (const float*)grid, ntype, nd, L, window, R, interlacing,
                                      Nk, k_min, k_max, cap, kP, Pk, Nk_out, kB, Bk, Ntri_out)
              : run_bispectrum<double>(c, (const double*)grid, ntype, nd, L, window, R,
                                       interlacing, Nk, k_min, k_max, cap, kP, Pk, Nk_out, kB, Bk,
                                       Ntri_out);
  EXH_API_END
}
