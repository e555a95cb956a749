/*
 * CPU oracle for the ExSHalos field -> LPT -> halos -> grid -> P(k) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithms, used as the checker for the HIP library.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product path never calls it.
 *
 * The file is compiled twice: with -DREAL=float (suffix _f32), which follows
 * the reference's single-precision build (fft_real = float, with its
 * double / long double islands kept where the reference has them), and with
 * -DREAL=double (suffix _f64), the FP64 oracle.  Expressions whose rounding
 * matters keep the reference's operand types and evaluation order; the code
 * structure is this project's own.  The 3D FFTs are not here: the Python
 * driver (oracle/pipeline.py) runs them with numpy.fft in float64.
 *
 * Pinning (see tests/test_oracle_*.py): natural cubic spline vs scipy,
 * sphere-count table vs brute force, Lomuto permutation vs a hand trace,
 * peak finder vs a numpy mask, growth vs the spike closed form and
 * invariants, LPT vs plane-wave closed forms, deposit vs mass conservation
 * and single-particle weights, binning vs brute-force mode enumeration,
 * and the taus generator behind the halo-mass draw vs GSL's published
 * known answer for gsl_rng_taus (seed 1, 10000th draw 2733957125).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif

/* -DORACLE_OMP builds the CPU baseline: the loops the reference runs under
 * OpenMP (the deposit, gridmodule.c:173-366; the deconvolution,
 * spectrum_h.c:116-157; the binning, powermodule.c:23-81; the
 * normalisations) run threaded, the deposit with atomic adds instead of
 * the reference's racy ones; everything the reference runs sequentially
 * stays sequential.  The parity oracle is built without it. */
#ifdef ORACLE_OMP
#include <omp.h>
#define ORA_ATOMIC _Pragma("omp atomic")
#define ORA_PAR_FOR _Pragma("omp parallel for schedule(static)")
#ifdef ORACLE_SHARED
/* Threads of the OpenMP build's parallel loops; returns the count in use. */
int ora_omp_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}
#endif
#else
#define ORA_ATOMIC
#define ORA_PAR_FOR
#endif

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#ifdef ORACLE_F32
#define SFX _f32
#else
#define SFX _f64
#endif
#define FN(name) CAT(name, SFX)

#define EXH_M_MAX 1e+16 /* exshalos_h.h:47 */
#define EXH_LC_MAX 1.0e+2
#define EXH_MC_MIN 1.0e+5

/* ------------------------------------------------------------------------ */
/* Derived parameters: exshalos_h.c:43-106                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  REAL Om0, redshift, dc, rhoc, Hz, Omz, rhomz, Dv;
  REAL Lc, Mcell, L[3], Mtot, kl[3], Normx, Normk;
  int nd[3], nz2, nmin;
  long long ng;
} FN(ora_params);

/* set_cosmology (exshalos_h.c:43-59) and set_box (exshalos_h.c:71-106).
 * `Lc_or_Mc` is a cell size when < 100 and a cell mass when > 1e5; the
 * reference decides that from the previous call's global box.Lc
 * (exshalos_h.c:78), a per-call parameter is used here.  Returns 0 on
 * success, -1 for a value in [100, 1e5]. */
int FN(ora_set_params)(FN(ora_params) * p, int ndx, int ndy, int ndz,
                       REAL Lc_or_Mc, REAL Om0, REAL redshift, REAL dc) {
  memset(p, 0, sizeof(*p));
  p->Om0 = Om0;
  p->redshift = redshift;
  p->rhoc = 2.775e+11;
  p->Hz = 100.0 * sqrt(p->Om0 * pow(1.0 + p->redshift, 3.0) + (1.0 - p->Om0));
  p->Omz = p->Om0 * pow(1.0 + p->redshift, 3.0) /
           (p->Om0 * pow(1.0 + p->redshift, 3.0) + (1.0 - p->Om0));
  p->rhomz = p->Om0 * p->rhoc;
  p->Dv = (18 * M_PI * M_PI + 82.0 * (p->Omz - 1.0) -
           39.0 * pow(p->Omz - 1.0, 2.0)) /
          p->Omz;
  if (dc <= 0.0)
    p->dc = 1.686 * pow(p->Omz, 0.0055);
  else
    p->dc = dc;

  p->nd[0] = ndx;
  p->nd[1] = ndy;
  p->nd[2] = ndz;
  p->ng = (long long)ndx * (long long)ndy * (long long)ndz;
  p->nz2 = ndz / 2 + 1;
  if (Lc_or_Mc < EXH_LC_MAX) {
    p->Lc = Lc_or_Mc;
    p->Mcell = p->rhomz * pow(p->Lc, 3.0);
  } else if (Lc_or_Mc > EXH_MC_MIN) {
    p->Mcell = Lc_or_Mc;
    p->Lc = pow(p->Mcell / p->rhomz, 1.0 / 3.0);
  } else {
    return -1;
  }
  for (int a = 0; a < 3; a++) p->L[a] = p->Lc * p->nd[a];
  p->Mtot = p->rhomz * p->L[0] * p->L[1] * p->L[2];
  for (int a = 0; a < 3; a++) p->kl[a] = 2.0 * M_PI / p->L[a];
  p->Normx = 1.0 / sqrt(p->L[0] * p->L[1] * p->L[2]);
  p->Normk = sqrt(p->L[0] * p->L[1] * p->L[2]) / (p->nd[0] * p->nd[1] * p->nd[2]);
  p->nmin = p->nd[0];
  if (p->nmin > p->nd[1]) p->nmin = p->nd[1];
  if (p->nmin > p->nd[2]) p->nmin = p->nd[2];
  return 0;
}

/* Periodic single-wrap addition (exshalos_h.c:22-30). */
static inline int wrap1(int i, int j, int n) {
  int r = i + j;
  if (r >= n) r -= n;
  if (r < 0) r += n;
  return r;
}

/* ------------------------------------------------------------------------ */
/* Natural cubic spline on double data (the reference uses GSL's cspline,    */
/* density_grid.c:62, find_halos.c:33-36, 88-89).  Coefficients follow the    */
/* textbook natural spline: second derivatives M_i from the tridiagonal       */
/* system h_{i-1}M_{i-1} + 2(h_{i-1}+h_i)M_i + h_i M_{i+1} = 6(dy_i/h_i -     */
/* dy_{i-1}/h_{i-1}), M_0 = M_{n-1} = 0, solved by the Thomas algorithm.     */
/* Evaluation outside [x_0, x_{n-1}] is an error in the reference (GSL        */
/* EDOM); here it is reported through *err.                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  const double *x, *y;
  double *m; /* second derivatives */
} ora_spline;

static int spline_init(ora_spline *s, const double *x, const double *y, int n) {
  s->n = n;
  s->x = x;
  s->y = y;
  s->m = (double *)calloc((size_t)n, sizeof(double));
  if (!s->m) return -1;
  if (n < 3) return 0;
  int ns = n - 2;
  double *cp = (double *)malloc((size_t)ns * sizeof(double));
  double *dp = (double *)malloc((size_t)ns * sizeof(double));
  for (int i = 0; i < ns; i++) {
    double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
    double diag = 2.0 * (h0 + h1);
    double rhs = 6.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
    double lower = h0; /* coefficient of M_i for row i+1 */
    if (i == 0) {
      cp[i] = h1 / diag;
      dp[i] = rhs / diag;
    } else {
      double den = diag - lower * cp[i - 1];
      cp[i] = h1 / den;
      dp[i] = (rhs - lower * dp[i - 1]) / den;
    }
  }
  s->m[ns] = dp[ns - 1];
  for (int i = ns - 2; i >= 0; i--) s->m[i + 1] = dp[i] - cp[i] * s->m[i + 2];
  free(cp);
  free(dp);
  return 0;
}

static double spline_eval(const ora_spline *s, double xv, int *err) {
  const double *x = s->x;
  int n = s->n;
  if (xv < x[0] || xv > x[n - 1]) {
    if (err) *err = 1;
    if (xv < x[0]) xv = x[0];
    if (xv > x[n - 1]) xv = x[n - 1];
  }
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (x[mid] > xv)
      hi = mid;
    else
      lo = mid;
  }
  double h = x[hi] - x[lo];
  double t = xv - x[lo];
  double b = (s->y[hi] - s->y[lo]) / h - h * (2.0 * s->m[lo] + s->m[hi]) / 6.0;
  double c = s->m[lo] / 2.0;
  double d = (s->m[hi] - s->m[lo]) / (6.0 * h);
  return s->y[lo] + t * (b + t * (c + t * d));
}

static void spline_free(ora_spline *s) { free(s->m); }

/* Exported for pinning against scipy.interpolate.CubicSpline(natural). */
int FN(ora_spline_eval_many)(const double *x, const double *y, int n,
                             const double *xq, double *out, int nq) {
  ora_spline s;
  if (spline_init(&s, x, y, n)) return -1;
  int err = 0;
  for (int i = 0; i < nq; i++) out[i] = spline_eval(&s, xq[i], &err);
  spline_free(&s);
  return err;
}

/* ------------------------------------------------------------------------ */
/* Gaussian field from white-noise modes: density_grid.c:71-205.             */
/* The reference draws the modes from GSL's taus + gsl_ran_gaussian in its   */
/* own loop order; here they are an input array noise[nd0][nd1][nz2][2] so    */
/* that the GPU and the oracle consume identical modes.  Modes filled by the  */
/* Hermitian copy rule (kz in {0, nd/2} planes) ignore their noise entry;     */
/* self-conjugate modes use sqrt(2)*A*noise.re with a zero imaginary part.    */
/* Output: deltak[nd0][nd1][nz2][2], unnormalised (the reference applies      */
/* Normx after the c2r transform, density_grid.c:213-216).                    */
/* ------------------------------------------------------------------------ */
/* R_max < 1e5 (the truncated-P(k) case of Inter_Power) also gives the DC   */
/* mode an amplitude: |k| is replaced by (kl0 kl1 kl2)^(1/3) / 4 and the      */
/* mode is self-conjugate, sqrt(2) A g (density_grid.c:158-161, 169-175).     */
int FN(ora_field_modes)(const FN(ora_params) * p, const double *Kin, const double *Pin,
                        int Nk, const REAL *noise, REAL k_smooth, int fixed,
                        REAL phase, REAL R_max, REAL *deltak) {
  ora_spline sp;
  if (spline_init(&sp, Kin, Pin, Nk)) return -1;
  const int n0 = p->nd[0], n1 = p->nd[1], n2 = p->nd[2], nz2 = p->nz2;
  int err = 0;
  for (int i = 0; i < n0; i++) {
    REAL kx = (i < n0 / 2) ? i * p->kl[0] : (i - n0) * p->kl[0];
    for (int j = 0; j < n1; j++) {
      REAL ky = (j < n1 / 2) ? j * p->kl[1] : (j - n1) * p->kl[1];
      for (int k = 0; k < nz2; k++) {
        REAL kz = k * p->kl[2];
        if (k == n2 / 2) kz = -n2 / 2.0 * p->kl[2];
        size_t id = ((size_t)i * n1 + j) * (size_t)nz2 + k;
        REAL *out = deltak + 2 * id;
        /* Hermitian partner copies in the kz = 0 and kz = nd/2 planes. */
        if ((k == 0 || k == n2 / 2) && (i > n0 / 2 || j > n1 / 2)) {
          int pi = -1, pj = -1;
          if (i > n0 / 2 && j > n1 / 2) {
            pi = n0 - i;
            pj = n1 - j;
          } else if (i == 0 || i == n0 / 2) {
            pi = i;
            pj = n1 - j;
          } else if (j == 0 || j == n1 / 2) {
            pi = n0 - i;
            pj = j;
          } else if (i > n0 / 2 && j < n1 / 2) {
            pi = n0 - i;
            pj = n1 - j;
          }
          if (pi >= 0) {
            const REAL *src = deltak + 2 * (((size_t)pi * n1 + pj) * (size_t)nz2 + k);
            out[0] = src[0];
            out[1] = -src[1];
            continue;
          }
        }
        REAL kmod = (REAL)sqrt(kx * kx + ky * ky + kz * kz);
        REAL A = 0.0;
        if (kmod <= k_smooth) {
          if (kmod > 0.0) {
            A = (REAL)sqrt(spline_eval(&sp, (double)kmod, &err) / 2.0);
          } else if (kmod == 0.0 && R_max < 100000.0) {
            /* the product of the three REAL frequencies is rounded to REAL,
             * pow runs in double, its result is rounded to REAL before the
             * division by 4.0 (in double), then stored as REAL */
            kmod = (REAL)pow(p->kl[0] * p->kl[1] * p->kl[2], 1.0 / 3.0) / 4.0;
            A = (REAL)sqrt(spline_eval(&sp, (double)kmod, &err) / 2.0);
          }
        }
        const REAL *g = noise + 2 * id;
        int selfconj = (i == 0 || i == n0 / 2) && (j == 0 || j == n1 / 2) &&
                       (k == 0 || k == n2 / 2);
        if (selfconj) {
          out[0] = fixed ? (REAL)(sqrt(2.0) * A) : (REAL)(sqrt(2.0) * A * g[0]);
          out[1] = 0.0;
        } else if (fixed) {
          /* fixed amplitude: the noise pair encodes the phase through
           * atan2 of the two components (the reference draws a flat phase,
           * density_grid.c:178-181). */
          REAL theta = (REAL)atan2((double)g[1], (double)g[0]);
          out[0] = A * cos(theta);
          out[1] = A * sin(theta);
        } else {
          out[0] = A * g[0];
          out[1] = A * g[1];
        }
        if (phase != 0.0) {
          /* density_grid.c:189-192 rotates with the already-updated real part */
          out[0] = out[0] * cos(phase) - out[1] * sin(phase);
          out[1] = out[0] * sin(phase) + out[1] * cos(phase);
        }
      }
    }
  }
  spline_free(&sp);
  return err ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* sigma(R), sigma^2(cells), Tinker integral: find_halos.c:4-93              */
/* ------------------------------------------------------------------------ */
/* exshalos_h.c:33-38.  Built by gcc -O2, the reference's sin and cos of
 * the same argument become one glibc sincos call (which may differ from sin
 * by an ulp), and pow(x, 2) a double multiply; both are written out here so
 * the result does not depend on the optimiser. */
static REAL tophat_W(REAL k, REAL R) {
  REAL r;
  double s, c, x = (double)(k * R);
  sincos(x, &s, &c);
  r = 3.0 / (x * x) * (s / (k * R) - c);
  return r;
}

static REAL calc_sigma2(const REAL *k, const REAL *P, int Nk, REAL R) {
  REAL acc = 0.0;
  for (int i = 0; i < Nk - 1; i++)
    acc += (k[i + 1] - k[i]) / 2.0 *
           (P[i] * pow(k[i] * tophat_W(k[i], R), 2) +
            P[i + 1] * pow(k[i + 1] * tophat_W(k[i + 1], R), 2));
  return acc / (2.0 * M_PI * M_PI);
}

static REAL tinker_fh(REAL sigma) { /* find_halos.c:47-71, model 1 */
  REAL B = 0.466, d = 2.06, e = 0.99, f = 0.48, g = 1.310, r;
  r = B * (pow(sigma / e, -d) + pow(sigma, -f)) * exp(-g / (sigma * sigma));
  return r;
}

/* The same three in double, for the FP64 sigma(R) table option. */
static double tophat_W_d(double k, double R) {
  double s, c, x = k * R;
  sincos(x, &s, &c);
  return 3.0 / (x * x) * (s / x - c);
}
static double calc_sigma2_d(const double *k, const double *P, int Nk, double R) {
  double acc = 0.0;
  for (int i = 0; i < Nk - 1; i++) {
    const double a = k[i] * tophat_W_d(k[i], R), b = k[i + 1] * tophat_W_d(k[i + 1], R);
    acc += (k[i + 1] - k[i]) / 2.0 * (P[i] * (a * a) + P[i + 1] * (b * b));
  }
  return acc / (2.0 * M_PI * M_PI);
}
static double tinker_fh_d(double sigma) {
  return 0.466 * (pow(sigma / 0.99, -2.06) + pow(sigma, -0.48)) * exp(-1.310 / (sigma * sigma));
}

static REAL barrier_of(REAL S, int do_eb, REAL dc, REAL a, REAL beta, REAL alpha) {
  REAL r; /* find_halos.c:200-212 */
  if (do_eb == 0)
    r = dc;
  else
    r = sqrt(a) * dc * (1.0 + beta * pow(S / (a * dc * dc), alpha));
  return r;
}

/* Tables used by the halo finder.
 * Sig_grid[0..Ncells-1]: sigma^2(i*Mcell) (find_halos.c:15-44, Nr = Nk).
 * Bt[0..nB-1]: barrier per cell count; entries i >= Ncells repeat
 *   Bt[Ncells-1] (the reference reads past the end of Sig_Grid there,
 *   find_halos.c:316).
 * Mspl_x/Int (Nr each): knots of the M -> integrated mass function spline.
 * Returns Ncells. */
/* sigma_radii > 0: the FP64 table option, that many radii with sigma^2
 * (a trapezoid over the input k) and the Tinker integral in double;
 * Mknots and Int then hold sigma_radii knots. */
int FN(ora_halo_tables)(const FN(ora_params) * p, const REAL *K, const REAL *P, int Nk,
                        int do_eb, REAL a, REAL beta, REAL alpha, REAL *Sig_grid,
                        REAL *Bt, int nB, double *Mknots, double *Int, int sigma_radii) {
  const int Nr = sigma_radii > 0 ? sigma_radii : Nk;
  const int Ncells = (int)floor(EXH_M_MAX / p->Mcell);
  double *Sig = (double *)malloc(sizeof(double) * Nr);
  double *Rd = (double *)malloc(sizeof(double) * Nr);
  if (sigma_radii > 0) {
    double *Kd = (double *)malloc(sizeof(double) * Nk), *Pd = (double *)malloc(sizeof(double) * Nk);
    for (int i = 0; i < Nk; i++) {
      Kd[i] = K[i];
      Pd[i] = P[i];
    }
    const double Rmin = pow(p->Mcell * 0.9 * 3.0 / (4.0 * M_PI * p->rhomz), 1.0 / 3.0);
    const double Rmax = pow(EXH_M_MAX * 3.0 / (4.0 * M_PI * p->rhomz), 1.0 / 3.0);
    for (int i = 0; i < Nr; i++) {
      Rd[i] = pow(10, log10(Rmin) + i * (log10(Rmax) - log10(Rmin)) / (Nr - 1));
      Mknots[i] = 4.0 / 3.0 * M_PI * ((double)p->rhomz * pow(Rd[i], 3.0));
      Sig[i] = sqrt(calc_sigma2_d(Kd, Pd, Nk, Rd[i]));
    }
    free(Kd);
    free(Pd);
  } else {
    REAL Rmin = (REAL)pow(p->Mcell * 0.9 * 3.0 / (4.0 * M_PI * p->rhomz), 1.0 / 3.0);
    REAL Rmax = (REAL)pow(EXH_M_MAX * 3.0 / (4.0 * M_PI * p->rhomz), 1.0 / 3.0);
    for (int i = 0; i < Nr; i++) {
      REAL Ri = pow(10, log10(Rmin) + i * (log10(Rmax) - log10(Rmin)) / (Nr - 1));
      Rd[i] = Ri;
      Mknots[i] = (double)4.0 / 3.0 * M_PI * ((REAL)p->rhomz * pow(Ri, 3));
      Sig[i] = (double)sqrt(calc_sigma2(K, P, Nk, Ri));
    }
  }
  ora_spline sp;
  spline_init(&sp, Mknots, Sig, Nr);
  int err = 0;
  Sig_grid[0] = 1e+30;
  for (int i = 1; i < Ncells; i++)
    Sig_grid[i] = pow(spline_eval(&sp, (double)i * p->Mcell, &err), 2.0);
  spline_free(&sp);

  Int[0] = 0.0;
  for (int i = 1; i < Nr; i++) {
    if (sigma_radii > 0)
      Int[i] = Int[i - 1] - (log(Sig[i]) - log(Sig[i - 1])) / 2.0 *
                                (tinker_fh_d(Sig[i]) / pow(Rd[i], 3.0) +
                                 tinker_fh_d(Sig[i - 1]) / pow(Rd[i - 1], 3.0));
    else
      Int[i] = Int[i - 1] - (double)(log(Sig[i]) - log(Sig[i - 1])) / 2.0 *
                                (tinker_fh(Sig[i]) / pow((REAL)Rd[i], 3.0) +
                                 tinker_fh(Sig[i - 1]) / pow((REAL)Rd[i - 1], 3.0));
    if (Int[i] == Int[i - 1]) Int[i] = Int[i - 1] * 1.000001;
  }
  for (int i = 0; i < nB; i++) {
    int c = i < Ncells ? i : Ncells - 1;
    Bt[i] = barrier_of(Sig_grid[c], do_eb, p->dc, a, beta, alpha);
  }
  free(Rd);
  free(Sig);
  return err ? -Ncells : Ncells;
}

/* ------------------------------------------------------------------------ */
/* Peaks: find_halos.c:96-156.  Strict 6-neighbour periodic maxima in        */
/* lexicographic (i, j, k) order.                                            */
/* ------------------------------------------------------------------------ */
long long FN(ora_find_peaks)(const int *nd, const REAL *delta, int *px, REAL *pden,
                             long long cap) {
  const int n0 = nd[0], n1 = nd[1], n2 = nd[2];
  long long np = 0;
  for (int i = 0; i < n0; i++)
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n2; k++) {
        size_t c = ((size_t)i * n1 + j) * n2 + k;
        REAL v = delta[c];
        if (v > delta[((size_t)wrap1(i, 1, n0) * n1 + j) * n2 + k] &&
            v > delta[((size_t)wrap1(i, -1, n0) * n1 + j) * n2 + k] &&
            v > delta[((size_t)i * n1 + wrap1(j, 1, n1)) * n2 + k] &&
            v > delta[((size_t)i * n1 + wrap1(j, -1, n1)) * n2 + k] &&
            v > delta[((size_t)i * n1 + j) * n2 + wrap1(k, 1, n2)] &&
            v > delta[((size_t)i * n1 + j) * n2 + wrap1(k, -1, n2)]) {
          if (np < cap) {
            px[3 * np] = i;
            px[3 * np + 1] = j;
            px[3 * np + 2] = k;
            pden[np] = v;
          }
          np++;
        }
      }
  return np;
}

/* Lomuto quicksort of the peaks, pivot = last element, "<=" to the left,
 * ascending and unstable (find_halos.c:159-197).  The permutation is
 * returned through `perm` (perm[new position] = old position) together with
 * the sorted keys, so that the tie order can be compared.  An explicit stack
 * replaces the recursion; the visiting order of sub-ranges does not change
 * the result because sub-ranges are disjoint. */
void FN(ora_lomuto_sort)(REAL *key, int *perm, long long n) {
  if (n < 2) return;
  long long cap = 128, top = 0;
  long long *st = (long long *)malloc(sizeof(long long) * 2 * cap);
  st[0] = 0;
  st[1] = n - 1;
  top = 1;
  while (top > 0) {
    top--;
    long long lo = st[2 * top], hi = st[2 * top + 1];
    if (lo >= hi) continue;
    REAL pivot = key[hi];
    long long i = lo - 1;
    for (long long j = lo; j < hi; j++) {
      if (key[j] <= pivot) {
        i++;
        REAL tk = key[i];
        key[i] = key[j];
        key[j] = tk;
        int tp = perm[i];
        perm[i] = perm[j];
        perm[j] = tp;
      }
    }
    {
      REAL tk = key[i + 1];
      key[i + 1] = key[hi];
      key[hi] = tk;
      int tp = perm[i + 1];
      perm[i + 1] = perm[hi];
      perm[hi] = tp;
    }
    long long pi = i + 1;
    if (top + 2 > cap) {
      cap *= 2;
      st = (long long *)realloc(st, sizeof(long long) * 2 * cap);
    }
    st[2 * top] = lo;
    st[2 * top + 1] = pi - 1;
    top++;
    st[2 * top] = pi + 1;
    st[2 * top + 1] = hi;
    top++;
  }
  free(st);
}

/* Visit the lattice offsets of the cube [-t, t]^3 in the reference's (i, j,
 * k) order.  Used by growth, shrink and flagging below. */
#define FOR_CUBE(t, di, dj, dk)             \
  for (int di = -(t); di <= (t); di++)      \
    for (int dj = -(t); dj <= (t); dj++)    \
      for (int dk = -(t); dk <= (t); dk++)

static inline size_t cell_of(const int *c, int di, int dj, int dk, const int *nd) {
  return ((size_t)wrap1(c[0], di, nd[0]) * nd[1] + wrap1(c[1], dj, nd[1])) * nd[2] +
         wrap1(c[2], dk, nd[2]);
}

/* Spherical growth against the barrier: find_halos.c:215-379.
 * flag[ng] (int64) must be initialised to ng.  Peaks must be sorted as by
 * ora_lomuto_sort; peaks are visited from index np-1 down to 1.
 * Bt: barrier per cell count (see ora_halo_tables), length nB; Bmin =
 * Bt[Ncells-1].  Output per kept halo h: hx[3h..3h+2] centre, hcount[h],
 * hr2[h] = the squared radius bound (cells with r^2 < hr2 belong to it).
 * Returns the number of halos, or a negative value on error. */
long long FN(ora_grow_halos)(const FN(ora_params) * p, long long np, const int *px,
                             const REAL *pden, const REAL *delta, const REAL *Bt,
                             int nB, int Ncells, int Nmin, long long *flag, int *hx,
                             int *hcount, int *hr2) {
  const int *nd = p->nd;
  const long long ng = p->ng;
  const REAL Bmin = Bt[Ncells - 1];
  const int tmax = p->nmin / 2;
  long long nh = 0;
  for (long long l = np - 1; l > 0; l--) {
    if (pden[l] < Bmin) break;
    const int *c = px + 3 * l;
    size_t cidx = ((size_t)c[0] * nd[1] + c[1]) * nd[2] + c[2];
    if (flag[cidx] != ng) continue;

    double den = (double)pden[l], den_tmp = (double)pden[l];
    int count = 0, count_tmp = 1, grows = 0, grows_tmp = 0;
    while (den_tmp >= Bmin) {
      if (count < count_tmp) grows = grows_tmp;
      grows_tmp++;
      den = den_tmp;
      count = count_tmp;
      den_tmp = den * ((double)count);
      int t = (int)floor(sqrt((double)grows_tmp));
      if (t > tmax) t = tmax;
      FOR_CUBE(t, di, dj, dk) {
        if (di * di + dj * dj + dk * dk != grows_tmp) continue;
        size_t id = cell_of(c, di, dj, dk, nd);
        if (flag[id] != ng)
          den_tmp += (double)-p->Mtot;
        else
          den_tmp += (double)delta[id];
        count_tmp++;
      }
      den_tmp = den_tmp / ((double)count_tmp);
    }
    for (;;) {
      if (count >= nB) return -2; /* barrier table too short */
      if (!(den < Bt[count] && count > 0)) break;
      den_tmp = den;
      count_tmp = count;
      den = den * ((double)count);
      int t = (int)floor(sqrt((double)grows));
      if (t > tmax) t = tmax;
      FOR_CUBE(t, di, dj, dk) {
        if (di * di + dj * dj + dk * dk != grows) continue;
        den -= (double)delta[cell_of(c, di, dj, dk, nd)];
        count--;
      }
      if (count > 0) den = den / ((double)count);
      if (count < count_tmp) grows_tmp = grows;
      grows--;
    }
    if (count == 0) continue;

    int t = (int)floor(sqrt((double)grows_tmp));
    if (2 * t + 1 > p->nmin) return -3; /* sphere wider than the box */
    FOR_CUBE(t, di, dj, dk) {
      if (di * di + dj * dj + dk * dk >= grows_tmp) continue;
      flag[cell_of(c, di, dj, dk, nd)] = nh;
    }
    flag[cidx] = -nh;
    if (count >= Nmin) {
      hx[3 * nh] = c[0];
      hx[3 * nh + 1] = c[1];
      hx[3 * nh + 2] = c[2];
      hcount[nh] = count;
      hr2[nh] = grows_tmp;
      nh++;
    } else {
      FOR_CUBE(t, di, dj, dk) {
        if (di * di + dj * dj + dk * dk >= grows_tmp) continue;
        flag[cell_of(c, di, dj, dk, nd)] = ng + 1;
      }
    }
  }
  return nh;
}

#ifdef ORACLE_SHARED
/* Cumulative number of lattice points with r^2 <= s over the representable
 * r^2 values (the reference ships it as a constant table,
 * cells_in_spheres.c:4; it is regenerated here by brute force, as the
 * reference's own Compute_Spheres does, find_halos.c:382-421). */
int ora_cells_in_spheres(int *out, int cap, int smax) {
  /* histogram of i^2 + j^2 + k^2 over the cube |i|,|j|,|k| <= sqrt(smax) */
  int t = (int)floor(sqrt((double)smax)) + 1;
  long long *hist = (long long *)calloc((size_t)smax, sizeof(long long));
  for (int i = -t; i <= t; i++)
    for (int j = -t; j <= t; j++)
      for (int k = -t; k <= t; k++) {
        long long s = (long long)i * i + (long long)j * j + (long long)k * k;
        if (s < smax) hist[s]++;
      }
  int m = 0;
  long long cum = 0;
  for (int s = 0; s < smax; s++) {
    if (hist[s] == 0) continue;
    cum += hist[s];
    if (m < cap) out[m] = (int)cum;
    m++;
  }
  free(hist);
  return m;
}
#endif

/* Combined Tausworthe generator of L'Ecuyer (1996, Math. Comp. 65, 203),
 * seeded through the LCG x -> 69069 x mod 2^32 and warmed up by six draws,
 * as GSL documents for gsl_rng_taus; uniform = draw / 2^32.  GSL itself is
 * absent here; the stream is pinned by GSL's own test-suite value (seed 1:
 * the 10000th draw is 2733957125). */
typedef struct {
  uint32_t s1, s2, s3;
} ora_taus;

static uint32_t taus_next(ora_taus *r) {
  uint32_t b;
  b = ((r->s1 << 13) ^ r->s1) >> 19;
  r->s1 = ((r->s1 & 4294967294u) << 12) ^ b;
  b = ((r->s2 << 2) ^ r->s2) >> 25;
  r->s2 = ((r->s2 & 4294967288u) << 4) ^ b;
  b = ((r->s3 << 3) ^ r->s3) >> 11;
  r->s3 = ((r->s3 & 4294967280u) << 17) ^ b;
  return r->s1 ^ r->s2 ^ r->s3;
}

static void taus_seed(ora_taus *r, unsigned long s) {
  if (s == 0) s = 1;
  r->s1 = (uint32_t)((69069ul * (uint32_t)s) & 0xffffffffUL);
  r->s2 = (uint32_t)((69069ul * r->s1) & 0xffffffffUL);
  r->s3 = (uint32_t)((69069ul * r->s2) & 0xffffffffUL);
  for (int i = 0; i < 6; i++) taus_next(r);
}

#ifdef ORACLE_SHARED
void ora_taus_uniforms(unsigned long seed, double *out, long long n) {
  ora_taus r;
  taus_seed(&r, seed);
  for (long long i = 0; i < n; i++) out[i] = taus_next(&r) / 4294967296.0;
}
#endif

/* Halo masses: find_halos.c:442-476.  Each halo's mass is drawn between
 * the mass of its sphere and of the next representable sphere, uniformly in
 * the cumulative Tinker mass function, using the i-th uniform for halo i.
 * `sphere`: cumulative counts (ora_cells_in_spheres).  Returns 0, or 1 if a
 * spline was evaluated outside its knots (an abort in the reference). */
int FN(ora_halo_masses)(const FN(ora_params) * p, const double *Mknots, const double *Int,
                        int Nr, const int *sphere, int nsphere, int Ncells,
                        const int *hcount, long long nh, const double *u, REAL *Mh) {
  ora_spline sI, sInv;
  spline_init(&sI, Mknots, Int, Nr);
  spline_init(&sInv, Int, Mknots, Nr);
  int err = 0;
  for (long long h = 0; h < nh; h++) {
    int idx = -1;
    for (int i = 0; i < Ncells && i < nsphere; i++)
      if (sphere[i] == hcount[h]) {
        idx = i + 1;
        break;
      }
    if (idx < 0 || idx >= nsphere) {
      err = 2;
      Mh[h] = 0;
      continue;
    }
    REAL m0 = hcount[h] * p->Mcell, m1 = sphere[idx] * p->Mcell;
    double I0 = spline_eval(&sI, (double)m0, &err);
    REAL v = I0 + (spline_eval(&sI, (double)m1, &err) - I0) * u[h];
    Mh[h] = spline_eval(&sInv, (double)v, &err);
  }
  spline_free(&sI);
  spline_free(&sInv);
  return err;
}

/* ------------------------------------------------------------------------ */
/* LPT: lpt.c:4-255                                                          */
/* ------------------------------------------------------------------------ */

/* Divide the half-complex grid by -k^2 (zero at k = 0 and k > k_smooth),
 * times `scale` (1 for 1LPT, lpt.c:12-44; Normk for 2LPT, lpt.c:141-173). */
void FN(ora_poisson_k)(const FN(ora_params) * p, REAL *ck, REAL k_smooth, int use_scale,
                       REAL scale) {
  const int n0 = p->nd[0], n1 = p->nd[1], n2 = p->nd[2], nz2 = p->nz2;
  for (int i = 0; i < n0; i++) {
    REAL kx = (2 * i < n0) ? i * p->kl[0] : (i - n0) * p->kl[0];
    for (int j = 0; j < n1; j++) {
      REAL ky = (2 * j < n1) ? j * p->kl[1] : (j - n1) * p->kl[1];
      for (int k = 0; k < nz2; k++) {
        REAL kz = k * p->kl[2];
        if (k == n2 / 2) kz = -(REAL)n2 / 2.0 * p->kl[2];
        REAL kmod = sqrt(pow(kx, 2.0) + pow(ky, 2.0) + pow(kz, 2.0));
        REAL *v = ck + 2 * (((size_t)i * n1 + j) * nz2 + k);
        if ((kx != 0.0 || ky != 0.0 || kz != 0.0) && (kmod <= k_smooth)) {
          REAL fact = -pow(kmod, 2.0);
          if (use_scale) {
            v[0] = v[0] / fact * scale;
            v[1] = v[1] / fact * scale;
          } else {
            v[0] = v[0] / fact;
            v[1] = v[1] / fact;
          }
        } else {
          v[0] = 0.0;
          v[1] = 0.0;
        }
      }
    }
  }
}

/* 6th-order central difference of f along axis `ax` at cell (i, j, k), in
 * double as in the reference (the 1.0*f products promote).  The y stencil's
 * -1 neighbour indexes rows with nd[0] (lpt.c:66); for cubic grids this is
 * the same row, and only cubic grids are accepted by the HIP library. */
static inline double fd6(const REAL *f, int stride3, int comp, const int *nd, int i,
                         int j, int k, int ax) {
  size_t o[6];
  const int n0 = nd[0], n1 = nd[1], n2 = nd[2];
  static const int sh[6] = {3, 2, 1, -1, -2, -3};
  for (int s = 0; s < 6; s++) {
    int a = i, b = j, c = k;
    if (ax == 0) a = wrap1(i, sh[s], n0);
    if (ax == 1) b = wrap1(j, sh[s], n1);
    if (ax == 2) c = wrap1(k, sh[s], n2);
    size_t row = (ax == 1 && s == 3) ? ((size_t)i * n0 + b) : ((size_t)a * n1 + b);
    o[s] = (row * n2 + c) * (size_t)stride3 + comp;
  }
  return 1.0 * f[o[0]] - 9.0 * f[o[1]] + 45.0 * f[o[2]] - 45.0 * f[o[3]] + 9.0 * f[o[4]] -
         1.0 * f[o[5]];
}

/* First-order displacement from phi (lpt.c:56-93):
 *   S_a = -FD_a(phi) * Normx / (60 Lc).
 * Halo bookkeeping (posh/velh sums per halo in cell order) lives in
 * ora_halo_disp below; this routine writes the full field. */
void FN(ora_lpt1_disp)(const FN(ora_params) * p, const REAL *phi, REAL *S) {
  const int *nd = p->nd;
  for (int i = 0; i < nd[0]; i++)
    for (int j = 0; j < nd[1]; j++)
      for (int k = 0; k < nd[2]; k++) {
        size_t c = ((size_t)i * nd[1] + j) * nd[2] + k;
        for (int a = 0; a < 3; a++) {
          REAL v = -(fd6(phi, 1, 0, nd, i, j, k, a)) * p->Normx / (60.0 * p->Lc);
          S[3 * c + a] = v;
        }
      }
}

/* Second-order source (lpt.c:116-134):
 *   phi2 = Sxx Syy + Sxx Szz + Syy Szz - Sxy^2 - Sxz^2 - Syz^2,
 * with S_ab = FD_b(S_a) / (60 Lc). */
void FN(ora_lpt2_source)(const FN(ora_params) * p, const REAL *S, REAL *phi2) {
  const int *nd = p->nd;
  for (int i = 0; i < nd[0]; i++)
    for (int j = 0; j < nd[1]; j++)
      for (int k = 0; k < nd[2]; k++) {
        size_t c = ((size_t)i * nd[1] + j) * nd[2] + k;
        REAL sxx = fd6(S, 3, 0, nd, i, j, k, 0) / (60.0 * p->Lc);
        REAL sxy = fd6(S, 3, 0, nd, i, j, k, 1) / (60.0 * p->Lc);
        REAL sxz = fd6(S, 3, 0, nd, i, j, k, 2) / (60.0 * p->Lc);
        REAL syy = fd6(S, 3, 1, nd, i, j, k, 1) / (60.0 * p->Lc);
        REAL syz = fd6(S, 3, 1, nd, i, j, k, 2) / (60.0 * p->Lc);
        REAL szz = fd6(S, 3, 2, nd, i, j, k, 2) / (60.0 * p->Lc);
        phi2[c] = (sxx * syy + sxx * szz + syy * szz - pow(sxy, 2.0) - pow(sxz, 2.0) -
                   pow(syz, 2.0));
      }
}

/* Second-order displacement added to S (lpt.c:182-223):
 *   S2_a = -3/7 Omz^(-1/143) FD_a(phi2) Normx / (60 Lc). */
void FN(ora_lpt2_disp)(const FN(ora_params) * p, const REAL *phi2, REAL *S2) {
  const int *nd = p->nd;
  for (int i = 0; i < nd[0]; i++)
    for (int j = 0; j < nd[1]; j++)
      for (int k = 0; k < nd[2]; k++) {
        size_t c = ((size_t)i * nd[1] + j) * nd[2] + k;
        for (int a = 0; a < 3; a++) {
          REAL v = (fd6(phi2, 1, 0, nd, i, j, k, a)) * p->Normx / (60.0 * p->Lc);
          v = -3.0 / 7.0 * pow(p->Omz, -1.0 / 143) * v;
          S2[3 * c + a] = v;
        }
      }
}

/* Per-halo displacement sums in cell order (lpt.c:70-74, 200-203) and the
 * final halo position (lpt.c:245-255): posh = wrap(x Lc + Lc/2 + sum/count).
 * S1 and S2 (S2 may be NULL) are full displacement fields.  Velocities
 * follow with the growth-rate factors of lpt.c:76-78 and 206-208. */
/* Compute_Posh (lpt.c:245-255): posh = wrap(x Lc + Lc/2 + posh / count),
 * velh /= count. */
static void halo_means(const FN(ora_params) * p, long long nh, const int *hx, const int *hcount,
                       REAL *posh, REAL *velh) {
  for (long long h = 0; h < nh; h++)
    for (int a = 0; a < 3; a++) {
      REAL base = hx[3 * h + a] * p->Lc + p->Lc / 2.0;
      REAL r = base + posh[3 * h + a] / hcount[h];
      if (r >= p->L[a]) r -= p->L[a];
      if (r < 0) r += p->L[a];
      posh[3 * h + a] = r;
      if (velh) velh[3 * h + a] = velh[3 * h + a] / hcount[h];
    }
}

void FN(ora_halo_disp)(const FN(ora_params) * p, const long long *flag, const REAL *S1,
                       const REAL *S2, long long nh, const int *hx, const int *hcount,
                       REAL *posh, REAL *velh) {
  const long long ng = p->ng;
  for (long long h = 0; h < 3 * nh; h++) {
    posh[h] = 0.0;
    if (velh) velh[h] = 0.0;
  }
  for (int pass = 0; pass < 2; pass++) {
    const REAL *S = pass == 0 ? S1 : S2;
    if (!S) continue;
    for (long long c = 0; c < ng; c++) {
      long long h = flag[c] < 0 ? -flag[c] : flag[c];
      if (h >= ng) continue;
      for (int a = 0; a < 3; a++) {
        posh[3 * h + a] += S[3 * c + a];
        if (!velh) continue;
        if (pass == 0)
          velh[3 * h + a] += pow(p->Omz, 0.5454) * p->Hz / (1.0 + p->redshift) * S[3 * c + a];
        else
          velh[3 * h + a] +=
              2.0 * pow(p->Omz, 6.0 / 11.0) * p->Hz / (1.0 + p->redshift) * S[3 * c + a];
      }
    }
  }
  halo_means(p, nh, hx, hcount, posh, velh);
}

/* Halo sums of given displacements and velocities (the IN_disp branch of
 * Generate_Halos_Box_from_Grid, box.c:96-110): posh += S and velh += V over
 * the cells with |flag| < ng in cell order, then Compute_Posh. */
void FN(ora_halo_disp_given)(const FN(ora_params) * p, const long long *flag, const REAL *S,
                             const REAL *V, long long nh, const int *hx, const int *hcount,
                             REAL *posh, REAL *velh) {
  const long long ng = p->ng;
  for (long long h = 0; h < 3 * nh; h++) {
    posh[h] = 0.0;
    if (velh) velh[h] = 0.0;
  }
  for (long long c = 0; c < ng; c++) {
    long long h = flag[c] < 0 ? -flag[c] : flag[c];
    if (h >= ng) continue;
    for (int a = 0; a < 3; a++) {
      posh[3 * h + a] += S[3 * c + a];
      if (velh) velh[3 * h + a] += V[3 * c + a];
    }
  }
  halo_means(p, nh, hx, hcount, posh, velh);
}

/* Particle positions q + S with one periodic wrap (lpt.c:229-242). */
void FN(ora_lpt_positions)(const FN(ora_params) * p, REAL *S) {
  const int *nd = p->nd;
  for (int i = 0; i < nd[0]; i++)
    for (int j = 0; j < nd[1]; j++)
      for (int k = 0; k < nd[2]; k++) {
        size_t c = ((size_t)i * nd[1] + j) * nd[2] + k;
        int q[3] = {i, j, k};
        for (int a = 0; a < 3; a++) {
          REAL base = q[a] * p->Lc + p->Lc / 2.0;
          REAL r = base + S[3 * c + a];
          if (r >= p->L[a]) r -= p->L[a];
          if (r < 0) r += p->L[a];
          S[3 * c + a] = r;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Mass assignment: gridmodule.c:9-53, 146-270; spectrum_h.c:4-80            */
/* ------------------------------------------------------------------------ */
static REAL wrap_while(REAL x, REAL y, REAL L) { /* spectrum_h.c:18-28 */
  REAL r = x + y;
  while (r >= L) r -= L;
  while (r < 0.0) r += L;
  return r;
}

static int wrap_int(int x, int y, int nd) { /* spectrum_h.c:4-15 */
  int r = x + y;
  while (r >= nd) r -= nd;
  while (r < 0) r += nd;
  return r;
}

static void cell_index(const REAL *x, int *xt, REAL Ld, int nd) { /* spectrum_h.c:73-80 */
  for (int a = 0; a < 3; a++) {
    xt[a] = floor(x[a] / Ld);
    if (xt[a] == nd) xt[a] -= 1;
  }
}

static void deposit_ngp(REAL *g, const REAL *x, int nd, REAL L, REAL m) {
  REAL Ld = L / nd;
  int xt[3];
  cell_index(x, xt, Ld, nd);
  ORA_ATOMIC g[xt[0] * nd * nd + xt[1] * nd + xt[2]] += m;
}

static void deposit_cic(REAL *g, const REAL *x, int nd, REAL L, REAL m) {
  REAL Ld = L / nd, dx[3], t[3];
  int xt[3], sg[3];
  cell_index(x, xt, Ld, nd);
  for (int a = 0; a < 3; a++) {
    dx[a] = x[a] / Ld - (xt[a] + 1.0 / 2.0);
    if (dx[a] >= 0.0) {
      sg[a] = 1;
      t[a] = 1.0 - dx[a];
    } else {
      sg[a] = -1;
      dx[a] = -dx[a];
      t[a] = 1.0 - dx[a];
    }
  }
  int x0 = xt[0], x1 = wrap_int(xt[0], sg[0], nd);
  int y0 = xt[1], y1 = wrap_int(xt[1], sg[1], nd);
  int z0 = xt[2], z1 = wrap_int(xt[2], sg[2], nd);
  ORA_ATOMIC g[x0 * nd * nd + y0 * nd + z0] += m * t[0] * t[1] * t[2];
  ORA_ATOMIC g[x1 * nd * nd + y0 * nd + z0] += m * dx[0] * t[1] * t[2];
  ORA_ATOMIC g[x0 * nd * nd + y1 * nd + z0] += m * t[0] * dx[1] * t[2];
  ORA_ATOMIC g[x0 * nd * nd + y0 * nd + z1] += m * t[0] * t[1] * dx[2];
  ORA_ATOMIC g[x1 * nd * nd + y1 * nd + z0] += m * dx[0] * dx[1] * t[2];
  ORA_ATOMIC g[x1 * nd * nd + y0 * nd + z1] += m * dx[0] * t[1] * dx[2];
  ORA_ATOMIC g[x0 * nd * nd + y1 * nd + z1] += m * t[0] * dx[1] * dx[2];
  ORA_ATOMIC g[x1 * nd * nd + y1 * nd + z1] += m * dx[0] * dx[1] * dx[2];
}

/* Triangular-shaped cloud (not in the reference; textbook definition,
 * Hockney & Eastwood 1988): per axis, u = x/Ld - 1/2, c = floor(u + 1/2),
 * d = u - c, weights 1/2 (1/2 - d)^2, 3/4 - d^2, 1/2 (1/2 + d)^2 on cells
 * c-1, c, c+1. */
static void deposit_tsc(REAL *g, const REAL *x, int nd, REAL L, REAL m) {
  REAL Ld = L / nd;
  int c[3];
  REAL w[3][3];
  for (int a = 0; a < 3; a++) {
    double u = (double)x[a] / Ld - 0.5;
    double cf = floor(u + 0.5);
    double d = u - cf;
    c[a] = (int)cf;
    w[a][0] = 0.5 * (0.5 - d) * (0.5 - d);
    w[a][1] = 0.75 - d * d;
    w[a][2] = 0.5 * (0.5 + d) * (0.5 + d);
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      for (int e = 0; e < 3; e++) {
        int ix = wrap_int(c[0], a - 1, nd), iy = wrap_int(c[1], b - 1, nd),
            iz = wrap_int(c[2], e - 1, nd);
        ORA_ATOMIC g[(size_t)ix * nd * nd + (size_t)iy * nd + iz] += m * w[0][a] * w[1][b] * w[2][e];
      }
}

/* Spherical (window 3) and exponential (window 4) assignment
 * (gridmodule.c:57-119): every cell of the cube of half-side
 * times = floor(1/2 + R_eff/Ld) around the particle's cell whose centre lies
 * within R_eff (minimum image per axis as the reference writes it, L - dx)
 * gets m, or m exp(-r^2 / 2R^2) with R_eff = R R_times.  The offsets are
 * evaluated in double and stored in REAL, r^2 in REAL, as in the reference.
 * Returns the deposited weight. */
static long double deposit_sphere(REAL *g, const REAL *x, int nd, REAL L, REAL R, REAL R_times,
                                  REAL m, int expw) {
  REAL Ld = L / nd;
  int xt[3];
  cell_index(x, xt, Ld, nd);
  const int times = expw ? (int)floor(0.5 + R * R_times / Ld) : (int)floor(0.5 + R / Ld);
  long double M = 0.0;
  for (int i = -times; i <= times; i++)
    for (int j = -times; j <= times; j++)
      for (int k = -times; k <= times; k++) {
        const int c[3] = {wrap_int(xt[0], i, nd), wrap_int(xt[1], j, nd),
                          wrap_int(xt[2], k, nd)};
        REAL dx[3];
        for (int a = 0; a < 3; a++) {
          dx[a] = x[a] - (c[a] + 1.0 / 2.0) * Ld;
          if (dx[a] > L / 2.0) dx[a] = L - dx[a];
          if (dx[a] < -L / 2.0) dx[a] = L + dx[a];
        }
        REAL r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        size_t at = (size_t)c[0] * nd * nd + (size_t)c[1] * nd + c[2];
        if (!expw) {
          if (r2 <= R * R) {
            ORA_ATOMIC g[at] += m;
            M += (long double)m;
          }
        } else if (r2 <= R * R * R_times * R_times) {
          REAL w = exp(-0.5 * r2 / R / R);
          ORA_ATOMIC g[at] += m * w;
          M += (long double)m * w;
        }
      }
  return M;
}

static long double deposit_one(REAL *g, int nd, REAL L, const REAL *x, REAL m, int window,
                               REAL R, REAL R_times) {
  if (window == 1)
    deposit_ngp(g, x, nd, L, m);
  else if (window == 2)
    deposit_cic(g, x, nd, L, m);
  else if (window == 3 || window == 4)
    return deposit_sphere(g, x, nd, L, R, R_times, m, window == 4);
  else
    deposit_tsc(g, x, nd, L, m);
  return (long double)m;
}

/* Tracer_Grid (gridmodule.c:146-375) with nthreads = 1 semantics (the
 * reference's multi-threaded scatter races, gridmodule.c:181-187).
 * window: 1 NGP, 2 CIC, 3 sphere, 4 exponential (R, R_times), 5 TSC.
 * grid: [ntype][1+interlacing][nd^3], zeroed by the caller.  pos is
 * modified in place by the redshift-space shift, as in the reference.
 * Returns 0, or -1 for an unsupported window. */
int FN(ora_tracer_grid)(REAL *grid, int nd, REAL L, int direction, REAL *pos,
                        const REAL *vel, long long np, const REAL *mass, const int *type,
                        int ntype, int window, REAL R, REAL R_times, int interlacing, REAL Om0,
                        REAL z, int folds) {
  if (window < 1 || window > 5) return -1;
  L = L / folds;
  REAL Lb = L / nd;
  size_t ng = (size_t)nd * nd * nd;
  long double M1[64], M2[64];
  if (ntype > 64) return -2;
  for (int t = 0; t < ntype; t++) M1[t] = M2[t] = 0.0;
  if (direction != -1) {
    REAL Hz = 100.0 * sqrt(Om0 * pow(1.0 + z, 3) + (1.0 - Om0));
    ORA_PAR_FOR
    for (long long i = 0; i < np; i++)
      pos[3 * i + direction] = wrap_while(
          pos[3 * i + direction], vel[3 * i + direction] * pow(1.0 + z, 3.0) / Hz, L);
  }
  const int nsub = 1 + (interlacing ? 1 : 0);
#ifdef ORACLE_OMP
#pragma omp parallel for schedule(static) reduction(+ : M1[:64], M2[:64])
#endif
  for (long long j = 0; j < np; j++) {
    int t = type ? type[j] : 0;
    REAL m = mass ? mass[j] : 1.0;
    REAL post[3];
    for (int a = 0; a < 3; a++) post[a] = wrap_while(pos[3 * j + a], 0.0, L);
    M1[t] += deposit_one(grid + ((size_t)t * nsub) * ng, nd, L, post, m, window, R, R_times);
    if (interlacing) {
      for (int a = 0; a < 3; a++) post[a] = wrap_while(pos[3 * j + a], -Lb / 2.0, L);
      M2[t] += deposit_one(grid + ((size_t)t * nsub + 1) * ng, nd, L, post, m, window, R,
                           R_times);
    }
  }
  if (!mass) {
    for (int t = 0; t < ntype; t++)
      for (int s = 0; s < nsub; s++) {
        REAL *g = grid + ((size_t)t * nsub + s) * ng;
        long double M = s == 0 ? M1[t] : M2[t];
        ORA_PAR_FOR
        for (size_t c = 0; c < ng; c++) g[c] = g[c] / (M / ng) - 1.0;
      }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Power spectrum: spectrum_h.c:30-167 (window, interlacing combine,         */
/* deconvolution) and powermodule.c:4-108 (mode loop and binning).            */
/* ------------------------------------------------------------------------ */
static REAL sincr(REAL x) {
  REAL r;
  if (x != 0.0)
    r = sin(x) / x;
  else
    r = 1.0;
  return r;
}

/* W (spectrum_h.c:51-70); windows 3 and 4 with the smoothing length R. */
static REAL window_k(REAL k1, REAL k2, REAL k3, REAL Lb, REAL R, int window) {
  REAL r = 1.0, kmod;
  if (window == 0) r = 1.0;
  if (window == 1) r = sincr(k1 * Lb / 2.0) * sincr(k2 * Lb / 2.0) * sincr(k3 * Lb / 2.0);
  if (window == 2)
    r = pow(sincr(k1 * Lb / 2.0) * sincr(k2 * Lb / 2.0) * sincr(k3 * Lb / 2.0), 2.0);
  if (window == 3) {
    kmod = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
    r = 3.0 / kmod / kmod / R / R * (sincr(kmod * R) - cos(kmod * R));
  }
  if (window == 4) {
    kmod = k1 * k1 + k2 * k2 + k3 * k3;
    r = exp(-kmod * R * R / 2.0);
  }
  if (window == 5)
    r = pow(sincr(k1 * Lb / 2.0) * sincr(k2 * Lb / 2.0) * sincr(k3 * Lb / 2.0), 3.0);
  return r;
}

/* Legendre P_l(mu) by the Bonnet recurrence. */
static double legendre(int l, double x) {
  if (l == 0) return 1.0;
  double p0 = 1.0, p1 = x;
  for (int n = 1; n < l; n++) {
    double p2 = ((2 * n + 1) * x * p1 - n * p0) / (n + 1);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

/* Compute_gridk for one tracer (spectrum_h.c:83-167): f1 (and f2 for the
 * interlaced copy, may be NULL) are the raw r2c outputs [nd][nd][nd/2+1][2];
 * out receives the combined, deconvolved, normalised modes. */
void FN(ora_gridk)(int nd, REAL L, int window, REAL R, const REAL *f1, const REAL *f2,
                   REAL *out) {
  REAL Normk = pow(L / (nd * nd), 3.0 / 2.0);
  REAL kn = 2 * M_PI / L;
  REAL Lb = L / nd;
  const int nz = nd / 2 + 1;
  ORA_PAR_FOR
  for (int i = 0; i < nd; i++) {
    REAL kx = (2 * i < nd) ? i * kn : (i - nd) * kn;
    for (int j = 0; j < nd; j++) {
      REAL ky = (2 * j < nd) ? j * kn : (j - nd) * kn;
      for (int k = 0; k < nz; k++) {
        REAL kz = (2 * k < nd) ? k * kn : (k - nd) * kn;
        size_t t = (size_t)nd * nz * i + (size_t)nz * j + k;
        REAL W = window_k(kx, ky, kz, Lb, R, window);
        if (f2) {
          out[2 * t] = (f1[2 * t] + f2[2 * t] * cos((kx + ky + kz) * Lb / 2.0) +
                        f2[2 * t + 1] * sin((kx + ky + kz) * Lb / 2.0)) /
                       2.0 * Normk / W;
          out[2 * t + 1] = (f1[2 * t + 1] + f2[2 * t + 1] * cos((kx + ky + kz) * Lb / 2.0) -
                            f2[2 * t] * sin((kx + ky + kz) * Lb / 2.0)) /
                           2.0 * Normk / W;
        } else {
          out[2 * t] = (f1[2 * t]) * Normk / W;
          out[2 * t + 1] = (f1[2 * t + 1]) * Normk / W;
        }
      }
    }
  }
}

/* Mode loop (powermodule.c:23-107).  gk: ntype grids of combined modes.
 * P: [NPs][ls][Nk] long double accumulators, Kmean[Nk], count[Nk] zeroed by
 * the caller.  The bin index is floor((|k| - kmin)/dk) evaluated in REAL,
 * as the reference does in fft_real (spectrum_h.c:39-48). */
void FN(ora_power_bins)(const REAL *const *gk, int ntype, int nd, REAL L, int Nk,
                        REAL k_min, REAL k_max, int l_max, int direction, double *P_out,
                        double *K_out, long long *count) {
  REAL kn = 2 * M_PI / L;
  REAL dk = (k_max - k_min) / Nk;
  int NPs = (ntype * (ntype + 1)) / 2;
  int ls = (int)floor(l_max / 2) + 1;
  const int nz = nd / 2 + 1;
  long double *P = (long double *)calloc((size_t)NPs * ls * Nk, sizeof(long double));
  long double *Km = (long double *)calloc((size_t)Nk, sizeof(long double));
#ifdef ORACLE_OMP
  const size_t nP = (size_t)NPs * ls * Nk;
#pragma omp parallel for schedule(static) reduction(+ : P[:nP], Km[:Nk], count[:Nk])
#endif
  for (int i = 0; i < nd; i++) {
    REAL kx = (2 * i < nd) ? i * kn : (i - nd) * kn;
    for (int j = 0; j < nd; j++) {
      REAL ky = (2 * j < nd) ? j * kn : (j - nd) * kn;
      for (int k = 0; k < nz; k++) {
        REAL kz = (2 * k < nd) ? k * kn : (k - nd) * kn;
        if ((k == 0 || k == nd / 2) &&
            ((i > nd / 2 && j > nd / 2) || (i == 0 || i == nd / 2) ||
             (j == 0 || j == nd / 2) || (i > nd / 2 && j < nd / 2)))
          continue;
        REAL kmod = sqrt(kx * kx + ky * ky + kz * kz);
        REAL tmp = (kmod - k_min) / dk;
        int b = floor(tmp);
        REAL mu = 0.0;
        if (direction == 2) mu = kz / kmod;
        if (direction == 1) mu = ky / kmod;
        if (direction == 0) mu = kx / kmod;
        if (b < Nk && b >= 0) {
          size_t t = (size_t)i * nd * nz + (size_t)j * nz + k;
          int pair = 0;
          for (int l = 0; l < ntype; l++)
            for (int m = 0; m <= l; m++) {
              for (int n = 0; n < ls; n++)
                P[((size_t)pair * ls + n) * Nk + b] +=
                    (long double)(gk[l][2 * t] * gk[m][2 * t] +
                                  gk[l][2 * t + 1] * gk[m][2 * t + 1]) *
                    legendre(2 * n, mu);
              pair++;
            }
          Km[b] += (long double)kmod;
          count[b] += 1;
        }
      }
    }
  }
  for (int b = 0; b < Nk; b++) {
    for (int q = 0; q < NPs; q++)
      for (int n = 0; n < ls; n++) {
        size_t o = ((size_t)q * ls + n) * Nk + b;
        P_out[o] = count[b] > 0 ? (double)((4.0 * n + 1.0) * P[o] / count[b]) : 0.0;
      }
    K_out[b] = count[b] > 0 ? (double)(Km[b] / count[b]) : 0.0;
  }
  free(P);
  free(Km);
}

/* ------------------------------------------------------------------------ */
/* Halo abundance: abundance_compute (spectrum.c:535-600) and                 */
/* Measure_Abundance (abundance.c:4-36).                                      */
/* ------------------------------------------------------------------------ */
/* Mmin/Mmax < 0: 0.9999 x the smallest and 1.0001 x the largest mass.  The
 * bin width is a float in both builds (abundance.c:7); a mass lands in bin
 * floor((ln M - ln Mmin) / dlnM), counted only for 0 < bin < Nm (bin 0
 * stays empty, abundance.c:23).  Mmean sums in REAL in catalogue order. */
void FN(ora_abundance)(const REAL *Mh, long long nh, REAL Mmin, REAL Mmax, int Nm, REAL Lc,
                       int ndx, int ndy, int ndz, REAL *Mmean, REAL *dn, REAL *dn_err) {
  REAL Lx = Lc * ndx, Ly = Lc * ndy, Lz = Lc * ndz;
  if (Mmin < 0.0) {
    Mmin = 1e+20;
    for (long long i = 0; i < nh; i++)
      if (Mh[i] < Mmin) Mmin = Mh[i];
    Mmin = Mmin * 0.9999;
  }
  if (Mmax < 0.0) {
    for (long long i = 0; i < nh; i++)
      if (Mh[i] > Mmax) Mmax = Mh[i];
    Mmax = Mmax * 1.0001;
  }
  float dlnM = (log(Mmax) - log(Mmin)) / Nm;
  int *N = (int *)calloc(Nm, sizeof(int));
  for (int i = 0; i < Nm; i++) Mmean[i] = 0.0;
  for (long long i = 0; i < nh; i++) {
    int ind = floor((log(Mh[i]) - log(Mmin)) / dlnM);
    if (ind > 0 && ind < Nm) {
      N[ind] += 1;
      Mmean[ind] += Mh[i];
    }
  }
  for (int i = 0; i < Nm; i++) {
    if (N[i] > 0) Mmean[i] = Mmean[i] / N[i];
    dn[i] = N[i] / (Lx * Ly * Lz) / dlnM;
    dn_err[i] = sqrt(N[i]) / (Lx * Ly * Lz) / dlnM;
  }
  free(N);
}
