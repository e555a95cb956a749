// Gaussian field and LPT kernels (gfx950).
//
//  field_mode      one half-complex mode of the density field from unit
//                  white noise (density_grid.c:94-198), Hermitian copies
//                  in the kz = 0, nd/2 planes computed from the partner's
//                  noise so that every mode is independent
//  k_field         delta_k * scale for the c2r that gives delta(x)
//  k_poisson_*     phi_k = delta_k / (-k^2) (lpt.c:12-44, 141-173)
//  fd_grad1        S1 = -D6(phi) / (60 Lc) (lpt.c:56-93)
//  lpt2_source     sum_{a<b} (S_aa S_bb - S_ab^2) (lpt.c:116-134)
//  fd_grad2        S2 = -3/7 Omz^(-1/143) D6(phi2) / (60 Lc), S = S1 + S2,
//                  velocities (lpt.c:182-223)
//  to_positions    q + S with one periodic wrap (lpt.c:229-242)
//
// All kernels are HBM-bound streaming kernels: one thread per mode/cell.
#include "common.hpp"
#include "kernels.hpp"

namespace exh {

template <typename T>
struct FieldArgs {
  int n0, n1, n2, nz2;
  T kl0, kl1, kl2;
  T k_smooth;
  const double* sx;  // spline knots, values, second derivatives (device)
  const double* sy;
  const double* sm;
  int nk;
  const T* noise;
  int fixed;
  T phase;
};

__device__ inline double spline_dev(const double* x, const double* y, const double* m, int n,
                                    double xv) {
  if (xv < x[0]) xv = x[0];
  if (xv > x[n - 1]) xv = x[n - 1];
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
  double b = (y[hi] - y[lo]) / h - h * (2.0 * m[lo] + m[hi]) / 6.0;
  double c = m[lo] / 2.0;
  double d = (m[hi] - m[lo]) / (6.0 * h);
  return y[lo] + t * (b + t * (c + t * d));
}

template <typename T>
__device__ inline T sqrt_t(T v);
template <>
__device__ inline float sqrt_t<float>(float v) {
  return sqrtf(v);
}
template <>
__device__ inline double sqrt_t<double>(double v) {
  return sqrt(v);
}

// Generated (non-copied) mode at (i, j, k): A(|k|) * noise, with the
// reference's float/double placement (density_grid.c:152-198).
template <typename T>
__device__ inline void generated_mode(const FieldArgs<T>& a, int i, int j, int k, T& re, T& im) {
  T kx = (i < a.n0 / 2) ? (T)i * a.kl0 : (T)(i - a.n0) * a.kl0;
  T ky = (j < a.n1 / 2) ? (T)j * a.kl1 : (T)(j - a.n1) * a.kl1;
  T kz = (T)k * a.kl2;
  if (k == a.n2 / 2) kz = (T)(-a.n2 / 2.0 * (double)a.kl2);
  T kmod = sqrt_t<T>(kx * kx + ky * ky + kz * kz);
  T A = 0;
  if (kmod <= a.k_smooth && kmod > (T)0)
    A = (T)sqrt(spline_dev(a.sx, a.sy, a.sm, a.nk, (double)kmod) / 2.0);
  size_t id = ((size_t)i * a.n1 + j) * (size_t)a.nz2 + k;
  T g0 = a.noise[2 * id], g1 = a.noise[2 * id + 1];
  bool selfconj = (i == 0 || i == a.n0 / 2) && (j == 0 || j == a.n1 / 2) &&
                  (k == 0 || k == a.n2 / 2);
  if (selfconj) {
    re = a.fixed ? (T)(1.4142135623730951 * (double)A) : (T)(1.4142135623730951 * (double)A * (double)g0);
    im = 0;
  } else if (a.fixed) {
    T theta = (T)atan2((double)g1, (double)g0);
    re = (T)((double)A * cos((double)theta));
    im = (T)((double)A * sin((double)theta));
  } else {
    re = A * g0;
    im = A * g1;
  }
  if (a.phase != (T)0) {
    re = (T)((double)re * cos((double)a.phase) - (double)im * sin((double)a.phase));
    im = (T)((double)re * sin((double)a.phase) + (double)im * cos((double)a.phase));
  }
}

// Mode (i, j, k) of the field including the Hermitian copy rule of the
// kz = 0 and kz = nd/2 planes (density_grid.c:109-150).
template <typename T>
__device__ inline void field_mode(const FieldArgs<T>& a, int i, int j, int k, T& re, T& im) {
  int pi = i, pj = j;
  bool cj = false;
  if ((k == 0 || k == a.n2 / 2) && (i > a.n0 / 2 || j > a.n1 / 2)) {
    if (i > a.n0 / 2 && j > a.n1 / 2) {
      pi = a.n0 - i;
      pj = a.n1 - j;
      cj = true;
    } else if (i == 0 || i == a.n0 / 2) {
      pj = a.n1 - j;
      cj = true;
    } else if (j == 0 || j == a.n1 / 2) {
      pi = a.n0 - i;
      cj = true;
    } else if (i > a.n0 / 2 && j < a.n1 / 2) {
      pi = a.n0 - i;
      pj = a.n1 - j;
      cj = true;
    }
  }
  generated_mode(a, pi, pj, k, re, im);
  if (cj) im = -im;
}

template <typename T>
__global__ void k_field(FieldArgs<T> a, T scale, typename cplx<T>::type* out,
                        typename cplx<T>::type* raw) {
  size_t nh = (size_t)a.n0 * a.n1 * a.nz2;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(id % a.nz2);
    size_t r = id / a.nz2;
    int j = (int)(r % a.n1);
    int i = (int)(r / a.n1);
    T re, im;
    field_mode(a, i, j, k, re, im);
    if (raw) raw[id] = {re, im};
    out[id] = {re * scale, im * scale};
  }
}

// -k^2 as the LPT code evaluates it (lpt.c:14-31): pow() in double, the
// modulus rounded to fft_real, then squared in double.
template <typename T>
__device__ inline bool lpt_fact(int n0, int n1, int n2, T kl0, T kl1, T kl2, T k_smooth, int i,
                                int j, int k, T& fact) {
  T kx = (2 * i < n0) ? (T)i * kl0 : (T)(i - n0) * kl0;
  T ky = (2 * j < n1) ? (T)j * kl1 : (T)(j - n1) * kl1;
  T kz = (T)k * kl2;
  if (k == n2 / 2) kz = (T)(-(T)n2 / 2.0 * (double)kl2);
  T kmod = (T)sqrt((double)kx * kx + (double)ky * ky + (double)kz * kz);
  if ((kx != (T)0 || ky != (T)0 || kz != (T)0) && kmod <= k_smooth) {
    fact = (T)(-((double)kmod * (double)kmod));
    return true;
  }
  return false;
}

// phi_k = delta_k / (-k^2) * scale with delta_k taken from the white-noise
// modes (no stored copy of delta_k is needed).
template <typename T>
__global__ void k_poisson_modes(FieldArgs<T> a, T k_smooth, T scale,
                                typename cplx<T>::type* out) {
  size_t nh = (size_t)a.n0 * a.n1 * a.nz2;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(id % a.nz2);
    size_t r = id / a.nz2;
    int j = (int)(r % a.n1);
    int i = (int)(r / a.n1);
    T fact;
    if (lpt_fact<T>(a.n0, a.n1, a.n2, a.kl0, a.kl1, a.kl2, k_smooth, i, j, k, fact)) {
      T re, im;
      field_mode(a, i, j, k, re, im);
      out[id] = {re / fact * scale, im / fact * scale};
    } else {
      out[id] = {(T)0, (T)0};
    }
  }
}

// In-place phi_k = c_k * pre / (-k^2) * scale for a given half-complex grid.
template <typename T>
__global__ void k_poisson_grid(int n0, int n1, int n2, T kl0, T kl1, T kl2, T k_smooth, T pre,
                               T scale, typename cplx<T>::type* ck) {
  int nz2 = n2 / 2 + 1;
  size_t nh = (size_t)n0 * n1 * nz2;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(id % nz2);
    size_t r = id / nz2;
    int j = (int)(r % n1);
    int i = (int)(r / n1);
    T fact;
    if (lpt_fact<T>(n0, n1, n2, kl0, kl1, kl2, k_smooth, i, j, k, fact)) {
      auto v = ck[id];
      ck[id] = {v.x * pre / fact * scale, v.y * pre / fact * scale};
    } else {
      ck[id] = {(T)0, (T)0};
    }
  }
}

// Sixth-order central difference D6 f = (f+3 - f-3) - 9 (f+2 - f-2)
// + 45 (f+1 - f-1), periodic, at stride `st` along an axis of length n.
template <typename T>
__device__ inline T d6(const T* f, size_t base, int c, int n, size_t st, int comp, int ncomp) {
  auto at = [&](int o) { return f[(base + (size_t)wrap1(c, o, n) * st) * ncomp + comp]; };
  return (at(3) - at(-3)) - (T)9 * (at(2) - at(-2)) + (T)45 * (at(1) - at(-1));
}

template <typename T>
__global__ void k_fd_grad1(int n0, int n1, int n2, const T* phi, T inv60Lc, T* S, T* V, T f1) {
  size_t ng = (size_t)n0 * n1 * n2;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ng;
       c += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(c % n2);
    size_t r = c / n2;
    int j = (int)(r % n1);
    int i = (int)(r / n1);
    T gx = d6(phi, c - (size_t)i * n1 * n2, i, n0, (size_t)n1 * n2, 0, 1);
    T gy = d6(phi, c - (size_t)j * n2, j, n1, (size_t)n2, 0, 1);
    T gz = d6(phi, c - (size_t)k, k, n2, 1, 0, 1);
    T sx = -gx * inv60Lc, sy = -gy * inv60Lc, sz = -gz * inv60Lc;
    S[3 * c] = sx;
    S[3 * c + 1] = sy;
    S[3 * c + 2] = sz;
    if (V) {
      V[3 * c] = f1 * sx;
      V[3 * c + 1] = f1 * sy;
      V[3 * c + 2] = f1 * sz;
    }
  }
}

template <typename T>
__global__ void k_lpt2_source(int n0, int n1, int n2, const T* S, T inv60Lc, T* src) {
  size_t ng = (size_t)n0 * n1 * n2;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ng;
       c += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(c % n2);
    size_t r = c / n2;
    int j = (int)(r % n1);
    int i = (int)(r / n1);
    size_t bx = c - (size_t)i * n1 * n2, by = c - (size_t)j * n2, bz = c - (size_t)k;
    size_t sxs = (size_t)n1 * n2, sys = (size_t)n2;
    T sxx = d6(S, bx, i, n0, sxs, 0, 3) * inv60Lc;
    T sxy = d6(S, by, j, n1, sys, 0, 3) * inv60Lc;
    T sxz = d6(S, bz, k, n2, 1, 0, 3) * inv60Lc;
    T syy = d6(S, by, j, n1, sys, 1, 3) * inv60Lc;
    T syz = d6(S, bz, k, n2, 1, 1, 3) * inv60Lc;
    T szz = d6(S, bz, k, n2, 1, 2, 3) * inv60Lc;
    src[c] = sxx * syy + sxx * szz + syy * szz - sxy * sxy - sxz * sxz - syz * syz;
  }
}

template <typename T>
__global__ void k_fd_grad2(int n0, int n1, int n2, const T* phi2, T coef, T* S, T* V, T f2) {
  size_t ng = (size_t)n0 * n1 * n2;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ng;
       c += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(c % n2);
    size_t r = c / n2;
    int j = (int)(r % n1);
    int i = (int)(r / n1);
    T gx = d6(phi2, c - (size_t)i * n1 * n2, i, n0, (size_t)n1 * n2, 0, 1) * coef;
    T gy = d6(phi2, c - (size_t)j * n2, j, n1, (size_t)n2, 0, 1) * coef;
    T gz = d6(phi2, c - (size_t)k, k, n2, 1, 0, 1) * coef;
    S[3 * c] += gx;
    S[3 * c + 1] += gy;
    S[3 * c + 2] += gz;
    if (V) {
      V[3 * c] += f2 * gx;
      V[3 * c + 1] += f2 * gy;
      V[3 * c + 2] += f2 * gz;
    }
  }
}

template <typename T>
__global__ void k_to_positions(int n0, int n1, int n2, T Lc, T L0, T L1, T L2, T* S) {
  size_t ng = (size_t)n0 * n1 * n2;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ng;
       c += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(c % n2);
    size_t r = c / n2;
    int j = (int)(r % n1);
    int i = (int)(r / n1);
    int q[3] = {i, j, k};
    T L[3] = {L0, L1, L2};
    for (int a = 0; a < 3; a++) {
      T base = (T)((double)((T)q[a] * Lc) + (double)Lc / 2.0);
      T v = base + S[3 * c + a];
      if (v >= L[a]) v -= L[a];
      if (v < (T)0) v += L[a];
      S[3 * c + a] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static unsigned stream_grid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <typename T>
static FieldArgs<T> make_field_args(exh_ctx* c, const exh_box* b, const exh_field_src* s,
                                    double k_smooth) {
  FieldArgs<T> a{};
  a.n0 = b->nd[0];
  a.n1 = b->nd[1];
  a.n2 = b->nd[2];
  a.nz2 = b->nz2;
  a.kl0 = (T)b->kl[0];
  a.kl1 = (T)b->kl[1];
  a.kl2 = (T)b->kl[2];
  a.k_smooth = (T)k_smooth;
  a.noise = (const T*)s->data;
  a.fixed = s->fixed;
  a.phase = (T)s->phase;
  // spline of P(k) on knots rounded as the wrappers pass them (fft_real)
  int nk = s->Nk;
  std::vector<double> tab(3 * (size_t)nk);
  for (int i = 0; i < nk; i++) {
    tab[i] = (double)(T)s->k[i];
    tab[nk + i] = (double)(T)s->P[i];
  }
  spline_second_derivs(tab.data(), tab.data() + nk, nk, tab.data() + 2 * nk);
  double* d = (double*)ctx_buf(c, BUF_TAB, tab.size() * sizeof(double));
  EXH_CHECK_HIP(hipMemcpyAsync(d, tab.data(), tab.size() * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));  // tab is a host temporary
  a.sx = d;
  a.sy = d + nk;
  a.sm = d + 2 * nk;
  a.nk = nk;
  // the reference aborts (GSL EDOM) when |k| of the grid leaves the table
  double kmax = 0;
  for (int ax = 0; ax < 3; ax++) {
    double kk = (double)(T)b->kl[ax] * (b->nd[ax] / 2);
    kmax += kk * kk;
  }
  kmax = std::sqrt(kmax);
  double kmin = (double)std::min({(T)b->kl[0], (T)b->kl[1], (T)b->kl[2]});
  if (kmin > k_smooth) kmin = 0;
  EXH_REQUIRE(tab[0] <= kmin * (1 + 1e-6) && std::min(kmax, k_smooth) <= tab[nk - 1] * (1 - 1e-7) + 1e-300,
              EXH_ERR_RANGE,
              "the box's |k| range [%g, %g] lies outside the tabulated P(k) [%g, %g]", kmin,
              std::min(kmax, k_smooth), tab[0], tab[nk - 1]);
  return a;
}

template <typename T>
void launch_field(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                  void* delta, void* deltak_raw) {
  using C2 = typename cplx<T>::type;
  auto a = make_field_args<T>(c, b, s, k_smooth);
  size_t nh = (size_t)b->nd[0] * b->nd[1] * b->nz2;
  C2* ck = (C2*)ctx_buf(c, BUF_CK0, nh * sizeof(C2));
  {
    TimedScope ts(c, "field_modes");
    hipLaunchKernelGGL(k_field<T>, dim3(stream_grid(nh)), dim3(256), 0, c->stream, a,
                       (T)b->Normx, ck, (C2*)deltak_raw);
    EXH_LAUNCH_CHECK();
  }
  fft_c2r(c, b->nd[0], b->nd[1], b->nd[2], b->precision, 1, ck, delta);
}

// Displacements from a field source.  S: [ng][3] (displacements, or
// positions when out_pos).  V optional.
template <typename T>
void launch_lpt(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                int do_2lpt, int out_pos, T* S, T* V) {
  using C2 = typename cplx<T>::type;
  const int n0 = b->nd[0], n1 = b->nd[1], n2 = b->nd[2];
  size_t nh = (size_t)n0 * n1 * b->nz2, ng = (size_t)b->ng;
  C2* ck = (C2*)ctx_buf(c, BUF_CK0, nh * sizeof(C2));
  T* phi = (T*)ctx_buf(c, BUF_R0, ng * sizeof(T));
  T kl0 = (T)b->kl[0], kl1 = (T)b->kl[1], kl2 = (T)b->kl[2];
  T Normx = (T)b->Normx, Normk = (T)b->Normk;
  {
    TimedScope ts(c, "lpt1_kspace");
    if (s->kind == EXH_SRC_MODES) {
      auto a = make_field_args<T>(c, b, s, 1e5);
      hipLaunchKernelGGL(k_poisson_modes<T>, dim3(stream_grid(nh)), dim3(256), 0, c->stream,
                         a, (T)k_smooth, Normx, ck);
    } else if (s->kind == EXH_SRC_DELTAK) {
      EXH_CHECK_HIP(hipMemcpyAsync(ck, s->data, nh * sizeof(C2), hipMemcpyDeviceToDevice,
                                   c->stream));
      hipLaunchKernelGGL(k_poisson_grid<T>, dim3(stream_grid(nh)), dim3(256), 0, c->stream, n0,
                         n1, n2, kl0, kl1, kl2, (T)k_smooth, (T)1, Normx, ck);
    } else {
      // Compute_Denk (density_grid.c:220-235): r2c then * Normk
      fft_r2c(c, n0, n1, n2, b->precision, 1, s->data, ck);
      hipLaunchKernelGGL(k_poisson_grid<T>, dim3(stream_grid(nh)), dim3(256), 0, c->stream, n0,
                         n1, n2, kl0, kl1, kl2, (T)k_smooth, Normk, Normx, ck);
    }
    EXH_LAUNCH_CHECK();
  }
  fft_c2r(c, n0, n1, n2, b->precision, 1, ck, phi);
  T inv60Lc = (T)(1.0 / (60.0 * b->Lc));
  T f1 = (T)(std::pow((T)b->Omz, 0.5454) * (T)b->Hz / (1.0 + (T)b->redshift));
  {
    TimedScope ts(c, "lpt1_grad");
    hipLaunchKernelGGL(k_fd_grad1<T>, dim3(stream_grid(ng)), dim3(256), 0, c->stream, n0, n1,
                       n2, phi, inv60Lc, S, V, f1);
    EXH_LAUNCH_CHECK();
  }
  if (do_2lpt) {
    T* src = (T*)ctx_buf(c, BUF_R1, ng * sizeof(T));
    {
      TimedScope ts(c, "lpt2_source");
      hipLaunchKernelGGL(k_lpt2_source<T>, dim3(stream_grid(ng)), dim3(256), 0, c->stream, n0,
                         n1, n2, S, inv60Lc, src);
      EXH_LAUNCH_CHECK();
    }
    fft_r2c(c, n0, n1, n2, b->precision, 1, src, ck);
    {
      TimedScope ts(c, "lpt2_kspace");
      hipLaunchKernelGGL(k_poisson_grid<T>, dim3(stream_grid(nh)), dim3(256), 0, c->stream, n0,
                         n1, n2, kl0, kl1, kl2, (T)k_smooth, Normk, Normx, ck);
      EXH_LAUNCH_CHECK();
    }
    fft_c2r(c, n0, n1, n2, b->precision, 1, ck, src);
    T coef = (T)(-3.0 / 7.0 * std::pow((T)b->Omz, -1.0 / 143) / (60.0 * b->Lc));
    T f2 = (T)(2.0 * std::pow((T)b->Omz, 6.0 / 11.0) * (T)b->Hz / (1.0 + (T)b->redshift));
    {
      TimedScope ts(c, "lpt2_grad");
      hipLaunchKernelGGL(k_fd_grad2<T>, dim3(stream_grid(ng)), dim3(256), 0, c->stream, n0, n1,
                         n2, src, coef, S, V, f2);
      EXH_LAUNCH_CHECK();
    }
  }
  if (out_pos) launch_positions<T>(c, b, S);
}

template <typename T>
void launch_positions(exh_ctx* c, const exh_box* b, T* S) {
  TimedScope ts(c, "positions");
  hipLaunchKernelGGL(k_to_positions<T>, dim3(stream_grid((size_t)b->ng)), dim3(256), 0,
                     c->stream, b->nd[0], b->nd[1], b->nd[2], (T)b->Lc, (T)b->L[0], (T)b->L[1],
                     (T)b->L[2], S);
  EXH_LAUNCH_CHECK();
}

template void launch_field<float>(exh_ctx*, const exh_box*, const exh_field_src*, double, void*, void*);
template void launch_field<double>(exh_ctx*, const exh_box*, const exh_field_src*, double, void*, void*);
template void launch_lpt<float>(exh_ctx*, const exh_box*, const exh_field_src*, double, int, int, float*, float*);
template void launch_lpt<double>(exh_ctx*, const exh_box*, const exh_field_src*, double, int, int, double*, double*);
template void launch_positions<float>(exh_ctx*, const exh_box*, float*);
template void launch_positions<double>(exh_ctx*, const exh_box*, double*);

}  // namespace exh
