/*
 * exhip.h -- C-ABI of the MI355X (gfx950) library for the ExSHalos
 * field -> 1LPT/2LPT -> excursion-set halos -> mass assignment -> P(k) path.
 *
 * The reference exposes this path through two CPython extension modules,
 * pyexshalos.lib.exshalos and pyexshalos.lib.spectrum (pyexshalos/lib/
 * __init__.py:1-4).  Each entry point below replaces one of their C
 * functions; the citation names the function it stands in for.  Bindings
 * (the Python shim in exshalos_amd/lib, or any ctypes/cffi/C caller) pass
 * plain pointers and sizes: device pointers for grids and particle arrays,
 * host pointers for the small tables and for outputs marked "host".
 *
 * Conventions
 *  - Every function returns EXH_OK (0) or a negative status; the message of
 *    the last failure on the calling thread is returned by exh_last_error().
 *    Nothing here exits the process (the reference calls exit(0) or lets GSL
 *    abort, e.g. exshalos_h.h:40-44, exshalos_h.c:88-91).
 *  - `precision` is 4 (fields, FFTs and displacements in float, as the
 *    reference's default build, include/exshalos/exshalos_h.h:17) or 8
 *    (double).  Device arrays typed "T" are float or double accordingly.
 *  - Grids are C order [nd0][nd1][nd2]; half-complex grids are
 *    [nd0][nd1][nd2/2+1] complex, interleaved (re, im).
 *  - A context owns one HIP stream (or borrows the caller's), the rocFFT
 *    plans and reusable device buffers; contexts are independent, there is
 *    no process-global state (the reference keeps its parameters in
 *    globals, exshalos_h.c:127-132).  Calls on one context are serialised on
 *    its stream; calls on different contexts may run concurrently.
 */
#ifndef EXHIP_H
#define EXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXH_OK 0
#define EXH_ERR_ARG (-1)         /* invalid argument or shape */
#define EXH_ERR_HIP (-2)         /* HIP runtime failure */
#define EXH_ERR_FFT (-3)         /* rocFFT failure */
#define EXH_ERR_RANGE (-4)       /* value outside a tabulated range (GSL aborts there) */
#define EXH_ERR_UNSUPPORTED (-5) /* option the library does not implement */
#define EXH_ERR_INTERNAL (-6)

#define EXH_WINDOW_NONE 0
#define EXH_WINDOW_NGP 1
#define EXH_WINDOW_CIC 2
#define EXH_WINDOW_SPHERE 3      /* gridmodule.c:57-84, spectrum_h.c:60-63 */
#define EXH_WINDOW_EXPONENTIAL 4 /* gridmodule.c:86-119, spectrum_h.c:64-67 */
#define EXH_WINDOW_TSC 5 /* not in the reference; deconvolved by sinc^6 */

/* Field sources accepted by the LPT and pipeline entry points. */
#define EXH_SRC_MODES 0  /* unit white-noise modes + P(k) (Compute_Den)  */
#define EXH_SRC_DELTA 1  /* real-space density grid (Compute_Denk)       */
#define EXH_SRC_DELTAK 2 /* half-complex density grid, unnormalised modes */

typedef struct exh_ctx exh_ctx;

/* Box and cosmology, derived as set_cosmology + set_box do
 * (exshalos_h.c:43-106).  With precision 4 every value is rounded to float
 * where the reference stores an fft_real. */
typedef struct {
  int nd[3];
  int nz2;
  int nmin;
  int precision;
  long long ng;
  double Om0, redshift, dc, rhoc, Hz, Omz, rhomz, Dv;
  double Lc, Mcell, L[3], Mtot, kl[3], Normx, Normk;
} exh_box;

/* A density field given in one of three forms (EXH_SRC_*). */
typedef struct {
  int kind;
  const void* data;  /* device: noise [nd0][nd1][nz2][2] T, delta [ng] T, or deltak */
  const double* k;   /* host: P(k) knots (EXH_SRC_MODES) */
  const double* P;   /* host */
  int Nk;
  int fixed;         /* fixed amplitudes (phase from the noise pair) */
  double phase;      /* constant phase rotation */
  double R_max;      /* 0 < R_max < 1e5: xi(r) of P(k) cut beyond R_max by FFTLog before
                        the spline (Inter_Power, density_grid.c:52-59; the knots must be
                        log-spaced); otherwise P(k) as given */
} exh_field_src;

/* Halo-finder options (find_halos / halos_box_from_* arguments). */
typedef struct {
  int Nmin;            /* minimum number of cells of a kept halo */
  double a, beta, alpha; /* ellipsoidal barrier; (1, 0, 0) = constant barrier */
  unsigned long seed;  /* seed of the mass draw (taus) */
  const double* k;     /* host: P(k) used for sigma(M) (find_halos.c:4-44) */
  const double* P;
  int Nk;
  int sigma_radii;     /* 0: the reference's sigma(R) table, Nk radii in the grid's float
                          type (find_halos.c:16-30, 493); > 0: that many radii (at least 4),
                          sigma^2 and the Tinker integral in double.  The mass draw splines
                          over these knots: with 64 radii draws near 600 cells (Lc = 4)
                          move by more than 50% from the reference's, with 256 by < 1e-3 */
} exh_halo_opts;

/* ---------------------------------------------------------------- basics */
const char* exh_version(void);
/* Message of the last failed call on this thread. */
const char* exh_last_error(void);

/* Create a context on HIP device `device` with its own non-blocking
 * stream.  No reference counterpart (the reference has no device state). */
int exh_ctx_create(int device, exh_ctx** out);
int exh_ctx_destroy(exh_ctx* ctx);
/* Run subsequent work on `stream` (a hipStream_t, e.g. torch's current
 * stream); NULL restores the context's own stream. */
int exh_ctx_set_stream(exh_ctx* ctx, void* stream);
int exh_ctx_synchronize(exh_ctx* ctx);
/* Per-kernel HIP-event timing: when enabled, each library stage records a
 * start/stop event pair on the context stream; exh_ctx_timings() returns
 * up to `max` (name, milliseconds) pairs of the calls since the last reset,
 * names separated by '\n' in `names`. */
int exh_ctx_set_timing(exh_ctx* ctx, int enable);
int exh_ctx_timings(exh_ctx* ctx, char* names, int names_len, double* ms, int max);
/* Per-context tunables (no reference counterpart): launch shapes the
 * autotuner (exshalos_amd/autotune.py) picks per grid size; 0 restores the
 * library's default.  EXH_TUNE_DEPOSIT_TILE: particles per window-deposit
 * workgroup, 16 (16^3), 3216 (32 x 16 x 16) or 32 (32 x 32 x 16);
 * EXH_TUNE_DEPOSIT_ACC: the window's sums, 1 (32-bit fixed point, float
 * grids without masses), 4 (float) or 8 (double).  Results agree to the
 * deposit's float rounding whatever the setting.  EXH_TUNE_HALO_CLAIMS: the
 * width of the halo growth rounds' per-cell claims, 0 (32 bits below 2^24
 * candidates, else 64) or 64; the catalogue is the same either way. */
#define EXH_TUNE_DEPOSIT_TILE 1
#define EXH_TUNE_DEPOSIT_ACC 2
#define EXH_TUNE_HALO_CLAIMS 3
int exh_ctx_set_tunable(exh_ctx* ctx, int key, int value);
/* Halo-finder statistics of the last run (5 values): peaks, candidates,
 * growth rounds, device-wide sort levels, sort levels that needed pointer
 * doubling. */
int exh_ctx_halo_stats(exh_ctx* ctx, long long* out5);

/* set_cosmology + set_box (exshalos_h.c:43-106).  Lc_or_Mc < 100 is a cell
 * size in Mpc/h, > 1e5 a cell mass in Msun/h, anything else EXH_ERR_ARG. */
int exh_box_init(exh_box* box, int ndx, int ndy, int ndz, double Lc_or_Mc, double Om0,
                 double redshift, double dc, int precision);

/* ----------------------------------------------------------------- field */
/* Gaussian density grid from unit white-noise modes: replaces
 * density_grid_compute (exshalos.c:46-121) / Compute_Den
 * (density_grid.c:71-217).  delta: device [ng] T, written as c2r(deltak)
 * * Normx.  deltak (optional, device [nd0][nd1][nz2][2] T) receives the
 * unnormalised modes. */
int exh_density_grid(exh_ctx* ctx, const exh_box* box, const exh_field_src* modes,
                     double k_smooth, void* delta, void* deltak);

/* ----------------------------------------------------------------- halos */
/* Halo tables of find_halos.c:4-93,200-212 computed on the host:
 * sig_grid[Ncells] (T), barrier[nB] (T; entries >= Ncells repeat the last),
 * Mknots[Nr], Int[Nr] (double) with Nr = opts->sigma_radii, or opts->Nk
 * when that is 0.  Returns Ncells through *Ncells. */
int exh_halo_tables(const exh_box* box, const exh_halo_opts* opts, void* sig_grid,
                    void* barrier, int nB, double* Mknots, double* Int, int* Ncells);

/* Excursion-set halos of a density grid: replaces find_halos
 * (exshalos.c:509-610) / Find_Halos (find_halos.c:479-563): peaks,
 * Lomuto quicksort with the reference's tie order, spherical growth against
 * the barrier, masses.  delta: device [ng] T.  The catalogue stays on the
 * device; *nh receives its size. */
int exh_find_halos(exh_ctx* ctx, const exh_box* box, const void* delta,
                   const exh_halo_opts* opts, long long* nh);

/* Copy the last catalogue out (each pointer may be NULL):
 * x [nh][3] int32 centre cells, count [nh] int32, r2 [nh] int32 (cells
 * with r^2 < r2 belong to the halo), Mh [nh] T.  to_host: destinations are
 * host memory (1) or device memory (0). */
int exh_halos_get(exh_ctx* ctx, int* x, int* count, int* r2, void* Mh, int to_host);
/* Halo positions (and velocities) of the last pipeline run
 * (Compute_Posh, lpt.c:245-255): posh/velh [nh][3] T. */
int exh_halos_posh(exh_ctx* ctx, void* posh, void* velh, int to_host);
/* Flag grid of the last halo run in the reference's encoding
 * (find_halos.c:349-376): ng free, ng+1 rejected, h member of halo h, -h
 * its centre.  flag: [ng] int64. */
int exh_halos_flag(exh_ctx* ctx, long long* flag, int to_host);

/* ------------------------------------------------------------------- LPT */
/* Displacements of the ng particles: replaces lpt_compute
 * (exshalos.c:613-707) / Compute_1LPT + Compute_2LPT + Compute_Pos
 * (lpt.c:4-242).  S: device [ng][3] T (displacements, or positions when
 * out_pos); V: device [ng][3] T or NULL. */
int exh_lpt(exh_ctx* ctx, const exh_box* box, const exh_field_src* src, double k_smooth,
            int do_2lpt, int out_pos, void* S, void* V);

/* Gradient of a loss of the monopole power spectrum of the LPT particles
 * with respect to the white-noise modes: the adjoint of modes -> delta_k
 * (Compute_Den, density_grid.c:94-198) -> 1LPT (+ 2LPT) displacements and
 * positions (lpt.c:4-242) -> unweighted, normalised mass assignment of the
 * ng particles on the box's own grid (Tracer_Grid, gridmodule.c:146-270;
 * window CIC or TSC, no interlacing) -> P(k) monopole in Nk bins
 * (Compute_gridk + Power_Spectrum, spectrum_h.c:83-167,
 * powermodule.c:4-108).  The reference has no adjoint: this entry point is
 * the transpose of the forward steps the library runs for exh_lpt,
 * exh_grid_compute and exh_power_compute, recomputing the positions instead
 * of storing them.  modes: EXH_SRC_MODES (unit noise; fixed and phase are
 * not supported).  gP: host [Nk] dL/dP_b.  g_noise: device
 * [nd0][nd1][nz2][2] T, overwritten with dL/dnoise (zero at the copied
 * modes of the kz = 0, nd/2 planes, whose noise the field ignores).  P_out,
 * count_out (host [Nk], optional): the forward P(k) and mode counts. */
int exh_pk_adjoint(exh_ctx* ctx, const exh_box* box, const exh_field_src* modes,
                   double k_smooth, int do_2lpt, int window, int Nk, double k_min,
                   double k_max, int kbin_precision, const double* gP, void* g_noise,
                   double* P_out, long long* count_out);

/* --------------------------------------------------------------- pipeline */
/* Halos and particles from a linear P(k): replaces halos_box_from_pk
 * (exshalos.c:710-940) / Generate_Halos_Box_from_Pk (box.c:4-68).
 * modes: EXH_SRC_MODES source.  delta (optional) receives the density grid;
 * S (optional) the particle positions [ng][3] (OUT_LPT); V (optional) the
 * particle velocities.  The catalogue, positions and masses stay on the
 * device (exh_halos_get / exh_halos_posh / exh_halos_flag). */
int exh_halos_box_from_pk(exh_ctx* ctx, const exh_box* box, const exh_field_src* modes,
                          const exh_halo_opts* opts, double k_smooth, int do_2lpt,
                          int out_vel, void* delta, void* S, void* V, long long* nh);

/* Halos from a given density grid: replaces halos_box_from_grid
 * (exshalos.c:943-1103) / Generate_Halos_Box_from_Grid (box.c:71-150).
 * With S_in (device [ng][3] T) the given displacements are attributed to
 * the halos (IN_disp); otherwise LPT is computed from the grid and written
 * to S_out / V_out when those are not NULL. */
int exh_halos_box_from_grid(exh_ctx* ctx, const exh_box* box, const void* delta,
                            const exh_halo_opts* opts, double k_smooth, int do_2lpt,
                            int out_vel, const void* S_in, const void* V_in, void* S_out,
                            void* V_out, long long* nh);

/* Halos, particle positions, their grid and its power spectrum in one call:
 * halos_box_from_pk (exshalos.c:710-940) with OUT_LPT, then grid_compute
 * (spectrum.c:46-108) of the positions (unweighted, real space, one
 * tracer) and power_compute (spectrum.c:111-197) of that grid (l_max = 0),
 * the chain pyexshalos users run as three calls.  It has no single
 * counterpart in the reference: it exists so that the LPT, the mass
 * assignment and the power spectrum run on a second stream while the halo
 * finder (latency-bound) runs.  Results equal the three calls': pos: device
 * [ng][3] T; grid: device [1+interlacing][nd^3] T; k_out, P_out, count_out:
 * host [Nk]; the catalogue stays on the device as after
 * exh_halos_box_from_pk. */
int exh_halos_box_power(exh_ctx* ctx, const exh_box* box, const exh_field_src* modes,
                        const exh_halo_opts* opts, double k_smooth, int do_2lpt, void* pos,
                        int window, int interlacing, int Nk, double k_min, double k_max,
                        int kbin_precision, void* grid, double* k_out, double* P_out,
                        long long* count_out, long long* nh);

/* -------------------------------------------------------------- spectrum */
/* Mass assignment: replaces grid_compute (spectrum.c:46-108) / Tracer_Grid
 * (gridmodule.c:146-375).  pos/vel: device [np][3] T; mass: device [np] T
 * or NULL; type: device [np] int32 or NULL (ntype == 1).  window: NGP, CIC,
 * SPHERE (radius R), EXPONENTIAL (scale R, cut at R * R_times) or TSC; R
 * and R_times are read only by the last two, as in the reference.  grid:
 * device [ntype][1+interlacing][nd^3] T, overwritten.  Unweighted grids
 * are normalised to rho/rho_bar - 1 by the deposited weight, weighted grids
 * are left as mass sums (gridmodule.c:311). */
int exh_grid_compute(exh_ctx* ctx, const void* pos, const void* vel, const void* mass,
                     const int* type, long long np, int ntype, int nd, double L,
                     double Om0, double z, int direction, int window, double R,
                     double R_times, int interlacing, int folds, int precision, void* grid);

/* Mass assignment that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as exh_grid_compute,
 * without a context: the work is queued on `stream` (a hipStream_t; NULL:
 * the null stream of the current device) and not waited for.  The
 * deposit's counters (particles per type, or deposited weight per grid)
 * live in a static 1 KiB device array, one per GPU, kept until the process
 * ends and cleared by every call; calls that overlap on one GPU share it
 * and race.  Limits: ntype <= 64 (EXH_ERR_ARG beyond). */
int exh_grid_compute_noalloc(void* stream, const void* pos, const void* vel, const void* mass,
                             const int* type, long long np, int ntype, int nd, double L,
                             double Om0, double z, int direction, int window, double R,
                             double R_times, int interlacing, int folds, int precision,
                             void* grid);

/* Power spectra of all tracer pairs: replaces power_compute
 * (spectrum.c:111-197) / Compute_gridk + Power_Spectrum
 * (spectrum_h.c:83-167, powermodule.c:4-108).  grid: device
 * [ntype][1+interlacing][nd^3] T (not modified).  Outputs (host):
 * k_out[Nk], P_out[NPs][l_max/2+1][Nk], count_out[Nk] with NPs =
 * ntype(ntype+1)/2.  window: the assignment to deconvolve (R: the
 * smoothing length of SPHERE and EXPONENTIAL).  kbin_precision selects the
 * arithmetic of the bin index: 4 follows the reference's float evaluation,
 * 8 double.  Bin sums are accumulated in double. */
int exh_power_compute(exh_ctx* ctx, const void* grid, int ntype, int nd, double L,
                      int window, double R, int interlacing, int Nk, double k_min,
                      double k_max, int l_max, int direction, int precision,
                      int kbin_precision, double* k_out, double* P_out, long long* count_out);

/* Differential halo abundance: replaces abundance_compute
 * (spectrum.c:535-600) / Measure_Abundance (abundance.c:4-36).  Mh: device
 * [nh] T, or NULL for the masses of the context's last halo catalogue.
 * Mmin or Mmax < 0 default to 0.9999 x the smallest and 1.0001 x the
 * largest mass.  Nm logarithmic bins of float width; a mass counts in bin
 * floor((ln M - ln Mmin) / dlnM) only for 0 < bin < Nm, as in the
 * reference.  Outputs (host) [Nm]: mean mass, dn/dlnM and its Poisson
 * error per volume (Lc ndx)(Lc ndy)(Lc ndz), in T arithmetic. */
int exh_abundance_compute(exh_ctx* ctx, const void* Mh, long long nh, double Mmin, double Mmax,
                          int Nm, double Lc, int ndx, int ndy, int ndz, int precision,
                          double* Mmean_out, double* dn_out, double* dn_err_out);

/* Number of bispectrum configurations for Nk shells (the order of
 * bimodule.c:49-317), or EXH_ERR_ARG. */
int exh_bispectrum_size(int Nk);

/* Bispectra of all tracer triples and the shell power spectra: replaces
 * bi_compute (spectrum.c:291-406) / Bi_Spectrum (bimodule.c:4-373).  grid:
 * device [ntype][1+interlacing][nd^3] T (not modified); window and R as in
 * exh_power_compute.  Every shell's real field is computed once (one c2r
 * per shell and tracer, plus the shell indicator) and kept on the device.
 * Outputs (host), Ntri = exh_bispectrum_size(Nk) <= cap: kP[Nk] mean |k|
 * per shell, Pk[ntype(ntype+1)/2][Nk], Nk_out[Nk] (sum of the indicator
 * squared / ng), kB[3][Ntri] mean |k| of the three shells, Bk[ntype^3][Ntri]
 * (sum of the triple product / sum of the indicators' triple product x
 * L^(3/2)), Ntri_out[Ntri] (that indicator sum / ng); *ntri = Ntri.  Sums
 * in double; outputs rounded to T as the reference's arrays. */
int exh_bispectrum_compute(exh_ctx* ctx, const void* grid, int ntype, int nd, double L,
                           int window, double R, int interlacing, int Nk, double k_min,
                           double k_max, int precision, long long cap, double* kP, double* Pk,
                           double* Nk_out, double* kB, double* Bk, double* Ntri_out, int* ntri);

#ifdef __cplusplus
}
#endif

#endif /* EXHIP_H */
