// Gaussian field and LPT kernels (gfx950).
//
//  field_mode      one half-complex mode of the density field from unit
//                  white noise (density_grid.c:94-198), Hermitian copies
//                  in the kz = 0, nd/2 planes computed from the partner's
//                  noise so that every mode is independent
//  k_field         delta_k * scale for the c2r that gives delta(x)
//  k_poisson_*     phi_k = delta_k / (-k^2) (lpt.c:12-44, 141-173)
//  k_grad_tiles    S1 = -D6(phi) / (60 Lc) (lpt.c:56-93); S2 = -3/7
//                  Omz^(-1/143) D6(phi2) / (60 Lc) added to S, velocities
//                  (lpt.c:182-223); the last pass also writes the
//                  positions q + S with one periodic wrap (lpt.c:229-242)
//  k_source_tiles  sum_{a<b} (S_aa S_bb - S_ab^2) (lpt.c:116-134)
//  k_source_full   the same for float fields whose stencil tiles are all
//                  full, the thread's own cells staged from registers
//
// Mode kernels stream over the half-complex grid; the stencil kernels
// march planes through LDS.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"
#include "kspace.hpp"

namespace exh {

template <typename T>
struct FieldArgs {
  int n0, n1, n2, nz2;
  T kl0, kl1, kl2;
  T k_smooth;
  const double* sx;  // spline knots, values and per-interval coefficients (device)
  const double* sy;
  const double* sb;
  const double* sc;
  const double* sd;
  int nk;
  int stage;  // copy the spline into LDS (5 nk doubles)
  const T* noise;
  int fixed;
  T phase;
  // truncated P(k) (R_max < 1e5): the DC mode takes the amplitude at
  // |k| = (kl0 kl1 kl2)^(1/3) / 4 (density_grid.c:158-161)
  int dc;
  T dc_kmod;
};

// Natural cubic spline evaluated from per-interval coefficients
// b, c, d computed on the host with the same expressions as the oracle's
// spline_eval (y + t (b + t (c + t d))), so the values are identical.
struct SplineView {
  const double *x, *y, *b, *c, *d;
  int n;
};

__device__ inline double spline_eval_v(const SplineView& s, double xv) {
  if (xv < s.x[0]) xv = s.x[0];
  if (xv > s.x[s.n - 1]) xv = s.x[s.n - 1];
  int lo = 0, hi = s.n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s.x[mid] > xv)
      hi = mid;
    else
      lo = mid;
  }
  const double t = xv - s.x[lo];
  return s.y[lo] + t * (s.b[lo] + t * (s.c[lo] + t * s.d[lo]));
}

// The spline tables in LDS (STAGE) or in global memory.  STAGE is a
// template argument so that the staged kernel's loads are known to be LDS
// loads (a pointer that may be either becomes a slower flat access).
template <bool STAGE, typename T>
__device__ inline SplineView stage_spline(const FieldArgs<T>& a, double* lds) {
  SplineView v{a.sx, a.sy, a.sb, a.sc, a.sd, a.nk};
  if (!STAGE) return v;
  const double* src[5] = {a.sx, a.sy, a.sb, a.sc, a.sd};
  for (int q = 0; q < 5; q++)
    for (int e = threadIdx.x; e < a.nk; e += blockDim.x) lds[q * a.nk + e] = src[q][e];
  __syncthreads();
  return SplineView{lds, lds + a.nk, lds + 2 * a.nk, lds + 3 * a.nk, lds + 4 * a.nk, a.nk};
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

// Amplitude A = sqrt(P(|k|) / 2) of mode (i, j, k) (density_grid.c:94-166).
template <typename T>
__device__ inline T mode_amp(const FieldArgs<T>& a, const SplineView& sp, int i, int j, int k) {
  T kx = (i < a.n0 / 2) ? (T)i * a.kl0 : (T)(i - a.n0) * a.kl0;
  T ky = (j < a.n1 / 2) ? (T)j * a.kl1 : (T)(j - a.n1) * a.kl1;
  T kz = (T)k * a.kl2;
  if (k == a.n2 / 2) kz = (T)(-a.n2 / 2.0 * (double)a.kl2);
  T kmod = sqrt_t<T>(kx * kx + ky * ky + kz * kz);
  T A = 0;
  if (kmod <= a.k_smooth) {
    if (kmod > (T)0)
      A = (T)sqrt(spline_eval_v(sp, (double)kmod) / 2.0);
    else if (a.dc)
      A = (T)sqrt(spline_eval_v(sp, (double)a.dc_kmod) / 2.0);
  }
  return A;
}

// Generated (non-copied) mode at (i, j, k): A(|k|) * noise, with the
// reference's float/double placement (density_grid.c:152-198).
template <typename T>
__device__ inline void generated_mode_g(const FieldArgs<T>& a, const SplineView& sp, int i, int j,
                                        int k, T g0, T g1, T& re, T& im) {
  const T A = mode_amp(a, sp, i, j, k);
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

// The same with the mode's noise read here.
template <typename T>
__device__ inline void generated_mode(const FieldArgs<T>& a, const SplineView& sp, int i, int j,
                                      int k, T& re, T& im) {
  const size_t id = ((size_t)i * a.n1 + j) * (size_t)a.nz2 + k;
  generated_mode_g(a, sp, i, j, k, a.noise[2 * id], a.noise[2 * id + 1], re, im);
}

// True where the kz = 0, nd/2 planes copy the conjugate of the mirror mode
// (-i, -j) instead of drawing one (density_grid.c:109-150).
template <typename T>
__device__ inline bool copied_mode(const FieldArgs<T>& a, int i, int j, int k) {
  if (!((k == 0 || k == a.n2 / 2) && (i > a.n0 / 2 || j > a.n1 / 2))) return false;
  return (i > a.n0 / 2 && j > a.n1 / 2) || (i == 0 || i == a.n0 / 2) ||
         (j == 0 || j == a.n1 / 2) || (i > a.n0 / 2 && j < a.n1 / 2);
}

// The Hermitian copy rule of the kz = 0 and kz = nd/2 planes
// (density_grid.c:109-150): mode (i, j, k) is generated mode (pi, pj, k),
// conjugated where cj.
template <typename T>
__device__ inline void mode_source(const FieldArgs<T>& a, int i, int j, int k, int& pi, int& pj,
                                   bool& cj) {
  pi = i;
  pj = j;
  cj = false;
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
}

// Mode (i, j, k) of the field including that copy rule.
template <typename T>
__device__ inline void field_mode(const FieldArgs<T>& a, const SplineView& sp, int i, int j, int k,
                                  T& re, T& im) {
  int pi, pj;
  bool cj;
  mode_source(a, i, j, k, pi, pj, cj);
  generated_mode(a, sp, pi, pj, k, re, im);
  if (cj) im = -im;
}

// Mode (i, j, k) of flat half-complex index id (multiply-high divisions
// below 2^31 modes).
__device__ inline void mode_coords(size_t id, size_t nh, int n1, int nz2, const FastDiv& dz,
                                   const FastDiv& dy, int& i, int& j, int& k) {
  if (nh < (1ull << 31)) {
    const unsigned r = fdiv((unsigned)id, dz);
    k = (int)((unsigned)id - r * (unsigned)nz2);
    i = (int)fdiv(r, dy);
    j = (int)(r - (unsigned)i * (unsigned)n1);
  } else {
    k = (int)(id % nz2);
    const size_t r = id / nz2;
    j = (int)(r % n1);
    i = (int)(r / n1);
  }
}

// Mode kernels: a flat grid-stride walk over the half-complex modes, so
// consecutive lanes take consecutive modes (a row of nd/2+1 modes does not
// fit a workgroup evenly).
template <typename T, bool STAGE>
__global__ __launch_bounds__(256) void k_field(FieldArgs<T> a, T scale, FastDiv dz, FastDiv dy,
                                               typename cplx<T>::type* out,
                                               typename cplx<T>::type* raw) {
  extern __shared__ double spl[];
  const SplineView sp = stage_spline<STAGE>(a, spl);
  const size_t nh = (size_t)a.n0 * a.n1 * a.nz2;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  // two modes per iteration, both modes' noise loaded before either's stores
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh; id += 2 * st) {
    const size_t idb = id + st;
    const bool two = idb < nh;
    int i[2], j[2], k[2], pi[2], pj[2];
    bool cj[2];
    T g0[2], g1[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (u == 1 && !two) break;
      mode_coords(u ? idb : id, nh, a.n1, a.nz2, dz, dy, i[u], j[u], k[u]);
      mode_source(a, i[u], j[u], k[u], pi[u], pj[u], cj[u]);
      const size_t src = ((size_t)pi[u] * a.n1 + pj[u]) * (size_t)a.nz2 + k[u];
      g0[u] = a.noise[2 * src];
      g1[u] = a.noise[2 * src + 1];
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      if (u == 1 && !two) break;
      T re, im;
      generated_mode_g(a, sp, pi[u], pj[u], k[u], g0[u], g1[u], re, im);
      if (cj[u]) im = -im;
      const size_t o = u ? idb : id;
      if (raw) raw[o] = {re, im};
      out[o] = {re * scale, im * scale};
    }
  }
}

// phi_k = delta_k / (-k^2) * scale with delta_k taken from the white-noise
// modes (no stored copy of delta_k is needed).
template <typename T, bool STAGE>
__global__ __launch_bounds__(256) void k_poisson_modes(FieldArgs<T> a, T k_smooth, T scale,
                                                       FastDiv dz, FastDiv dy,
                                                       typename cplx<T>::type* out) {
  extern __shared__ double spl[];
  const SplineView sp = stage_spline<STAGE>(a, spl);
  const size_t nh = (size_t)a.n0 * a.n1 * a.nz2;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    mode_coords(id, nh, a.n1, a.nz2, dz, dy, i, j, k);
    T fact;
    if (lpt_fact<T>(a.n0, a.n1, a.n2, a.kl0, a.kl1, a.kl2, k_smooth, i, j, k, fact)) {
      T re, im;
      field_mode(a, sp, i, j, k, re, im);
      out[id] = {re / fact * scale, im / fact * scale};
    } else {
      out[id] = {(T)0, (T)0};
    }
  }
}

// Adjoint of noise -> delta_k (field_mode) -> phi_k = delta_k / (-k^2) *
// scale -> phi = c2r(phi_k): from gk = r2c(dL/dphi), dL/dphi_k = w gk
// (w = 2 for 0 < kz < nd/2, 1 in the kz = 0, nd/2 planes: the transpose of
// the c2r), dL/ddelta_k = dL/dphi_k scale / (-k^2), and dL/dnoise through
// the amplitudes and the copy rule: a drawn mode collects its own gradient
// and the conjugate of its copy's (the mirror mode in those planes); copied
// and zero-amplitude modes get zero.  Unit noise without `fixed` or a
// phase.
template <typename T, bool STAGE>
__global__ __launch_bounds__(256) void k_field_adjoint(FieldArgs<T> a, T k_smooth, T scale,
                                                       FastDiv dz, FastDiv dy,
                                                       const typename cplx<T>::type* gk,
                                                       T* gnoise) {
  extern __shared__ double spl[];
  const SplineView sp = stage_spline<STAGE>(a, spl);
  const size_t nh = (size_t)a.n0 * a.n1 * a.nz2;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    mode_coords(id, nh, a.n1, a.nz2, dz, dy, i, j, k);
    double g0 = 0.0, g1 = 0.0;
    T fact;
    if (!copied_mode(a, i, j, k) &&
        lpt_fact<T>(a.n0, a.n1, a.n2, a.kl0, a.kl1, a.kl2, k_smooth, i, j, k, fact)) {
      const double A = (double)mode_amp(a, sp, i, j, k);
      const double w = (k == 0 || 2 * k == a.n2) ? 1.0 : 2.0;
      const double f = A * w * (double)scale / (double)fact;
      const bool selfconj = (i == 0 || i == a.n0 / 2) && (j == 0 || j == a.n1 / 2) &&
                            (k == 0 || k == a.n2 / 2);
      const auto v = gk[id];
      if (selfconj) {
        g0 = 1.4142135623730951 * f * (double)v.x;
      } else {
        g0 = f * (double)v.x;
        g1 = f * (double)v.y;
        if (k == 0 || k == a.n2 / 2) {
          const int mi = i == 0 ? 0 : a.n0 - i, mj = j == 0 ? 0 : a.n1 - j;
          if (copied_mode(a, mi, mj, k)) {
            const auto u = gk[((size_t)mi * a.n1 + mj) * (size_t)a.nz2 + k];
            g0 += f * (double)u.x;
            g1 -= f * (double)u.y;
          }
        }
      }
    }
    gnoise[2 * id] = (T)g0;
    gnoise[2 * id + 1] = (T)g1;
  }
}

// In-place phi_k = c_k * pre / (-k^2) * scale for a given half-complex
// grid (flat index over modes; 32-bit index arithmetic below 2^31 modes).
template <typename T>
__global__ __launch_bounds__(256) void k_poisson_grid(int n0, int n1, int n2, T kl0, T kl1, T kl2,
                                                      T k_smooth, T pre, T scale, FastDiv dz,
                                                      FastDiv dy, typename cplx<T>::type* ck) {
  const int nz2 = n2 / 2 + 1;
  const size_t nh = (size_t)n0 * n1 * nz2;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    if (nh < (1ull << 31)) {
      const unsigned r = fdiv((unsigned)id, dz);
      k = (int)((unsigned)id - r * (unsigned)nz2);
      i = (int)fdiv(r, dy);
      j = (int)(r - (unsigned)i * (unsigned)n1);
    } else {
      k = (int)(id % nz2);
      const size_t r = id / nz2;
      j = (int)(r % n1);
      i = (int)(r / n1);
    }
    T fact;
    if (lpt_fact<T>(n0, n1, n2, kl0, kl1, kl2, k_smooth, i, j, k, fact)) {
      auto v = ck[id];
      ck[id] = {v.x * pre / fact * scale, v.y * pre / fact * scale};
    } else {
      ck[id] = {(T)0, (T)0};
    }
  }
}

// ---------------------------------------------------------------------------
// Stencil tiles.  A workgroup owns a ST_TJ x ST_TK tile of (j, k) and walks
// i over a slab of g.ti planes (ST_TI, halved down to 2 while the grid has
// fewer than 512 workgroups, so small boxes still fill the GPU; each slab
// reads 6 planes beyond its own).  Plane i, with halos of 3 cells in j and
// k, is staged in LDS for the y and z differences; the x difference uses a
// register window of the thread's points over i-3..i+3, so each value is
// read from HBM about once.  256 threads: 64 along k, 4 along j, each
// thread owning the 4 rows j = ty + 4 q of its column.  Differences are
// evaluated in double in the reference's order,
//   1 f[+3] - 9 f[+2] + 45 f[+1] - 45 f[-1] + 9 f[-2] - 1 f[-3]
// (lpt.c:64-68), and scaled by 1 / (60 Lc) in double.
// ---------------------------------------------------------------------------
static constexpr int ST_TJ = 16, ST_TK = 64, ST_TI = 32, ST_H = 3;
static constexpr int ST_PJ = ST_TJ + 2 * ST_H, ST_PK = ST_TK + 2 * ST_H;
static constexpr int ST_Q = ST_TJ / 4;

__device__ inline double d6w(double m3, double m2, double m1, double p1, double p2, double p3) {
  return 1.0 * p3 - 9.0 * p2 + 45.0 * p1 - 45.0 * m1 + 9.0 * m2 - 1.0 * m3;
}

struct TileGeom {
  int n0, n1, n2, ntk, ntj, nti, ti;
};

__device__ inline void tile_origin(const TileGeom& g, int& i0, int& j0, int& k0) {
  const int b = blockIdx.x;
  const int bk = b % g.ntk, bj = (b / g.ntk) % g.ntj, bi = b / (g.ntk * g.ntj);
  i0 = bi * g.ti;
  j0 = bj * ST_TJ;
  k0 = bk * ST_TK;
}

// Staging of one plane (ST_PJ x ST_PK cells with periodic halos, NC
// interleaved components per cell) into LDS in the field's own order:
// thread t stages the plane's values t, t + 256, ..., so the LDS index is
// the loop index and the offset of each value within a plane of the field
// is the same for every plane of the slab.  The offsets are computed once
// per workgroup (cells outside a partial tile are staged too, harmlessly);
// a plane then costs one load and one LDS store per value.
template <int NC>
struct PlaneStage {
  static constexpr int TOT = ST_PJ * ST_PK * NC;
  static constexpr int NE = (TOT + 255) / 256;
  int off[NE];
  __device__ void init(const TileGeom& g, int j0, int k0) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      const int v = threadIdx.x + 256 * e;
      const int cellv = v / NC, comp = v - NC * cellv;
      const int pj = cellv / ST_PK, pk = cellv - ST_PK * pj;
      int j = j0 + pj - ST_H, k = k0 + pk - ST_H;
      j = j < 0 ? j + g.n1 : (j >= g.n1 ? j - g.n1 : j);
      k = k < 0 ? k + g.n2 : (k >= g.n2 ? k - g.n2 : k);
      j = min(j, g.n1 - 1);  // partial tiles smaller than the halo
      k = min(k, g.n2 - 1);
      off[e] = v < TOT ? (j * g.n2 + k) * NC + comp : 0;
    }
  }
  // plane i of f ([ng][NC]) into pl[TOT]
  template <typename T>
  __device__ void stage(const TileGeom& g, const T* f, int i, T* pl) const {
    const T* base = f + (size_t)i * g.n1 * g.n2 * NC;
#pragma unroll
    for (int e = 0; e < NE; e++) {
      const int v = threadIdx.x + 256 * e;
      if (e < NE - 1 || v < TOT) pl[v] = base[off[e]];
    }
  }
  // the same in two halves, so the loads can be issued before the barrier
  // that frees the LDS plane (an unused slot reads the plane's first value)
  template <typename T>
  __device__ void load(const T* base, T (&r)[NE]) const {
#pragma unroll
    for (int e = 0; e < NE; e++) r[e] = base[off[e]];
  }
  template <typename T>
  __device__ void store(const T (&r)[NE], T* pl) const {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      const int v = threadIdx.x + 256 * e;
      if (e < NE - 1 || v < TOT) pl[v] = r[e];
    }
  }
};

// Staging of a full tile's halo only (n1 a multiple of ST_TJ, n2 of
// ST_TK): the tile's own cells are written into the plane from the
// registers of the thread that owns them.  Halo cells of a plane: rows
// 0 .. ST_H - 1 and ST_H + ST_TJ .. ST_PJ - 1 whole, the rows between them
// their ST_H cells at each end; halo value h is component h % NC of halo
// cell h / NC.
template <int NC>
struct HaloStage {
  static constexpr int CELLS = ST_PJ * ST_PK - ST_TJ * ST_TK;
  static constexpr int TOT = CELLS * NC;
  static constexpr int NE = (TOT + 255) / 256;
  int off[NE], lds[NE];  // offset within a plane of the field, LDS index (-1: unused)
  __device__ void init(const TileGeom& g, int j0, int k0) {
#pragma unroll
    for (int e = 0; e < NE; e++) {
      const int h = threadIdx.x + 256 * e;
      const int hc = h / NC, comp = h - NC * hc;
      constexpr int TOP = ST_H * ST_PK, MID = ST_TJ * 2 * ST_H;
      int pj, pk;
      if (hc < TOP) {
        pj = hc / ST_PK;
        pk = hc - pj * ST_PK;
      } else if (hc < TOP + MID) {
        const int t = hc - TOP, r = t % (2 * ST_H);
        pj = ST_H + t / (2 * ST_H);
        pk = r < ST_H ? r : ST_TK + r;
      } else {
        const int t = hc - TOP - MID;
        pj = ST_H + ST_TJ + t / ST_PK;
        pk = t - (t / ST_PK) * ST_PK;
      }
      int j = j0 + pj - ST_H, k = k0 + pk - ST_H;
      j = j < 0 ? j + g.n1 : (j >= g.n1 ? j - g.n1 : j);
      k = k < 0 ? k + g.n2 : (k >= g.n2 ? k - g.n2 : k);
      const bool ok = h < TOT;
      off[e] = ok ? (j * g.n2 + k) * NC + comp : 0;
      lds[e] = ok ? (pj * ST_PK + pk) * NC + comp : -1;
    }
  }
  template <typename T>
  __device__ void load(const T* base, T (&r)[NE]) const {
#pragma unroll
    for (int e = 0; e < NE; e++) r[e] = base[off[e]];
  }
  template <typename T>
  __device__ void store(const T (&r)[NE], T* pl) const {
#pragma unroll
    for (int e = 0; e < NE; e++)
      if (e < NE - 1 || lds[e] >= 0) pl[lds[e]] = r[e];
  }
};

__device__ inline int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Gradient of a scalar field phi: MODE 1 writes S = -D phi / (60 Lc) and
// V = f1 S (lpt.c:56-93); MODE 2 adds -3/7 Omz^(-1/143) D phi2 / (60 Lc) to
// S and f2 times it to V (lpt.c:182-223; coef carries the prefactor).
// Positions written by the last gradient pass: q + S wrapped into the box
// (Compute_Pos, lpt.c:229-255): the lattice point (q + 1/2) Lc in the
// field's precision, plus S, wrapped once.
template <typename T>
struct PosOut {
  T* P;  // [ng][3], may alias S; null: none
  T Lc, L0, L1, L2;
};

template <typename T>
__device__ inline T lattice_pos(int q, T Lc, T s, T L) {
  const T base = (T)((double)((T)q * Lc) + (double)Lc / 2.0);
  T v = base + s;
  if (v >= L) v -= L;
  if (v < (T)0) v += L;
  return v;
}

// FULL (n1 a multiple of ST_TJ, n2 of ST_TK): the plane's halo is loaded
// and the thread's own cells come from the x window, which holds them.
// Either way the plane's loads are issued with the window's, before the
// barrier that frees the LDS plane.  The second-order pass, which also
// reads S and V back, runs faster with at most 4 waves per SIMD (measured
// at 512^3: 0.84 ms against 0.90 ms at the 5 its registers allow).
template <typename T, int MODE, bool FULL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, MODE == 2 ? 4 : 8))) void
k_grad_tiles(TileGeom g, const T* phi, double inv60,
                                                    double coef, T* S, T* V, T fv,
                                                    PosOut<T> po) {
  __shared__ T pl[ST_PJ * ST_PK];
  int i0, j0, k0;
  tile_origin(g, i0, j0, k0);
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int k = k0 + tx;
  const bool kin = k < g.n2;
  const int kc = kin ? k : g.n2 - 1;
  typename std::conditional<FULL, HaloStage<1>, PlaneStage<1>>::type ps;
  ps.init(g, j0, k0);
  const size_t plane = (size_t)g.n1 * g.n2;
  int wo[ST_Q];  // the thread's cells within a plane
  for (int q = 0; q < ST_Q; q++) wo[q] = min(j0 + ty + 4 * q, g.n1 - 1) * g.n2 + kc;
  T win[ST_Q][7];
  for (int d = 0; d < 6; d++) {
    const T* pp = phi + (size_t)wrapi(i0 - 3 + d, g.n0) * plane;
    for (int q = 0; q < ST_Q; q++) win[q][d] = pp[wo[q]];
  }
  const int i1 = min(i0 + g.ti, g.n0);
  for (int i = i0; i < i1; i++) {
    T hv[decltype(ps)::NE];
    ps.load(phi + (size_t)i * plane, hv);
    const T* pp = phi + (size_t)wrapi(i + 3, g.n0) * plane;
    for (int q = 0; q < ST_Q; q++) win[q][6] = pp[wo[q]];
    // the second-order pass adds to S (and V): for float fields on full
    // tiles the plane's values of the thread's cells are loaded here, before
    // any store of the plane (which could alias them for the compiler); the
    // other variants keep their registers (and occupancy) and load them
    // below
    using V3 = typename vec3<T>::type;
    constexpr bool PRE = MODE == 2 && sizeof(T) == 4 && FULL;
    V3 s_old[PRE ? ST_Q : 1], v_old[PRE ? ST_Q : 1];
    if (PRE)
      for (int q = 0; q < ST_Q; q++) {
        const int j = j0 + ty + 4 * q;
        if (kin && j < g.n1) {
          const size_t c = ((size_t)i * g.n1 + j) * g.n2 + k;
          s_old[q] = *reinterpret_cast<const V3*>(S + 3 * c);
          if (V) v_old[q] = *reinterpret_cast<const V3*>(V + 3 * c);
        }
      }
    __syncthreads();  // previous plane's readers are done
    ps.store(hv, pl);
    if (FULL)
      for (int q = 0; q < ST_Q; q++) pl[(ty + 4 * q + ST_H) * ST_PK + tx + ST_H] = win[q][3];
    __syncthreads();
    for (int q = 0; q < ST_Q; q++) {
      const int jl = ty + 4 * q, j = j0 + jl;
      const T* w = win[q];
      const double gx = d6w(w[0], w[1], w[2], w[4], w[5], w[6]);
      const T* rowj = pl + (jl + ST_H) * ST_PK + tx + ST_H;
      const double gy = d6w(rowj[-3 * ST_PK], rowj[-2 * ST_PK], rowj[-ST_PK], rowj[ST_PK],
                            rowj[2 * ST_PK], rowj[3 * ST_PK]);
      const double gz = d6w(rowj[-3], rowj[-2], rowj[-1], rowj[1], rowj[2], rowj[3]);
      if (kin && j < g.n1) {
        // the cell's three components move as one 3-vector access
        const size_t c = ((size_t)i * g.n1 + j) * g.n2 + k;
        V3* Sp = reinterpret_cast<V3*>(S + 3 * c);
        V3* Vp = V ? reinterpret_cast<V3*>(V + 3 * c) : nullptr;
        V3 snew;
        if (MODE == 1) {
          const T sx = (T)(-gx * inv60), sy = (T)(-gy * inv60), sz = (T)(-gz * inv60);
          snew = V3{sx, sy, sz};
          *Sp = snew;
          if (Vp) *Vp = V3{fv * sx, fv * sy, fv * sz};
        } else {
          const T sx = (T)(coef * gx * inv60), sy = (T)(coef * gy * inv60),
                  sz = (T)(coef * gz * inv60);
          V3 o = PRE ? s_old[PRE ? q : 0] : *Sp;
          o.x += sx;
          o.y += sy;
          o.z += sz;
          *Sp = o;
          snew = o;
          if (Vp) {
            V3 u = PRE ? v_old[PRE ? q : 0] : *Vp;
            u.x += fv * sx;
            u.y += fv * sy;
            u.z += fv * sz;
            *Vp = u;
          }
        }
        if (po.P)  // after S: P may alias it
          *reinterpret_cast<V3*>(po.P + 3 * c) =
              V3{lattice_pos<T>(i, po.Lc, snew.x, po.L0), lattice_pos<T>(j, po.Lc, snew.y, po.L1),
                 lattice_pos<T>(k, po.Lc, snew.z, po.L2)};
      }
    }
    for (int q = 0; q < ST_Q; q++)
      for (int d = 0; d < 6; d++) win[q][d] = win[q][d + 1];
  }
}

// Second-order source phi2 = Sxx Syy + Sxx Szz + Syy Szz - Sxy^2 - Sxz^2 -
// Syz^2 with S_ab = D_b(S_a) / (60 Lc) (lpt.c:116-134): the three
// products in the field's precision, the squares in double, as the
// reference's expression (pow(x, 2.0) promotes) evaluates them.
template <typename T>
__global__ __launch_bounds__(256) void k_source_tiles(TileGeom g, const T* S, double inv60,
                                                      T* src) {
  __shared__ T pl[3 * ST_PJ * ST_PK];  // [pj][pk][component]
  int i0, j0, k0;
  tile_origin(g, i0, j0, k0);
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int k = k0 + tx;
  const bool kin = k < g.n2;
  const int kc = kin ? k : g.n2 - 1;
  PlaneStage<3> ps;
  ps.init(g, j0, k0);
  const size_t plane = (size_t)g.n1 * g.n2 * 3;
  int wo[ST_Q];  // Sx of the thread's cells within a plane
  for (int q = 0; q < ST_Q; q++) wo[q] = (min(j0 + ty + 4 * q, g.n1 - 1) * g.n2 + kc) * 3;
  T win[ST_Q][7];  // Sx over i-3..i+3
  for (int d = 0; d < 6; d++) {
    const T* pp = S + (size_t)wrapi(i0 - 3 + d, g.n0) * plane;
    for (int q = 0; q < ST_Q; q++) win[q][d] = pp[wo[q]];
  }
  constexpr int DY = 3 * ST_PK, DZ = 3;
  const int i1 = min(i0 + g.ti, g.n0);
  for (int i = i0; i < i1; i++) {
    const T* pp = S + (size_t)wrapi(i + 3, g.n0) * plane;
    for (int q = 0; q < ST_Q; q++) win[q][6] = pp[wo[q]];
    __syncthreads();
    ps.stage(g, S, i, pl);
    __syncthreads();
    for (int q = 0; q < ST_Q; q++) {
      const int jl = ty + 4 * q, j = j0 + jl;
      const T* w = win[q];
      const int o = ((jl + ST_H) * ST_PK + tx + ST_H) * 3;
      const T* x = pl + o;
      const T* y = pl + o + 1;
      const T* z = pl + o + 2;
      const T sxx = (T)(d6w(w[0], w[1], w[2], w[4], w[5], w[6]) * inv60);
      const T sxy =
          (T)(d6w(x[-3 * DY], x[-2 * DY], x[-DY], x[DY], x[2 * DY], x[3 * DY]) * inv60);
      const T sxz =
          (T)(d6w(x[-3 * DZ], x[-2 * DZ], x[-DZ], x[DZ], x[2 * DZ], x[3 * DZ]) * inv60);
      const T syy =
          (T)(d6w(y[-3 * DY], y[-2 * DY], y[-DY], y[DY], y[2 * DY], y[3 * DY]) * inv60);
      const T syz =
          (T)(d6w(y[-3 * DZ], y[-2 * DZ], y[-DZ], y[DZ], y[2 * DZ], y[3 * DZ]) * inv60);
      const T szz =
          (T)(d6w(z[-3 * DZ], z[-2 * DZ], z[-DZ], z[DZ], z[2 * DZ], z[3 * DZ]) * inv60);
      if (kin && j < g.n1) {
        const T t3 = sxx * syy + sxx * szz + syy * szz;
        const double v = (double)t3 - (double)sxy * (double)sxy - (double)sxz * (double)sxz -
                         (double)syz * (double)syz;
        src[((size_t)i * g.n1 + j) * g.n2 + k] = (T)v;
      }
    }
    for (int q = 0; q < ST_Q; q++)
      for (int d = 0; d < 6; d++) win[q][d] = win[q][d + 1];
  }
}

// The same source for float fields whose tiles are all full (n1 a
// multiple of 16, n2 of 64): each thread keeps its own four cells of planes
// i .. i + 3 in registers (one 12-byte load per cell and plane, made three
// planes ahead for the x difference), writes them into the staged plane
// from there, and loads only the plane's halo cells (516 of the 22 x 70).
// The separate strided load of S_x at plane i + 3, whose lines were fetched
// again when the plane was staged, goes away.
__global__ __launch_bounds__(256) void k_source_full(TileGeom g, const float* S, double inv60,
                                                     float* src) {
  __shared__ float pl[3 * ST_PJ * ST_PK];  // [pj][pk][component]
  int i0, j0, k0;
  tile_origin(g, i0, j0, k0);
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  HaloStage<3> hs;
  hs.init(g, j0, k0);
  const size_t plane = (size_t)g.n1 * g.n2 * 3;
  int co[ST_Q];  // the thread's cells within a plane of S
  for (int q = 0; q < ST_Q; q++) co[q] = ((j0 + ty + 4 * q) * g.n2 + k0 + tx) * 3;
  using V3 = float3;
  V3 cur[ST_Q][4];     // S at planes i .. i + 3, the thread's cells
  float xo[ST_Q][3];   // S_x at planes i - 3 .. i - 1
  for (int d = 0; d < 3; d++) {
    const float* pp = S + (size_t)wrapi(i0 - 3 + d, g.n0) * plane;
    for (int q = 0; q < ST_Q; q++) xo[q][d] = pp[co[q]];
  }
  for (int d = 0; d < 3; d++) {
    const float* pp = S + (size_t)wrapi(i0 + d, g.n0) * plane;
    for (int q = 0; q < ST_Q; q++) cur[q][d] = *reinterpret_cast<const V3*>(pp + co[q]);
  }
  constexpr int DY = 3 * ST_PK, DZ = 3;
  const int i1 = min(i0 + g.ti, g.n0);
  for (int i = i0; i < i1; i++) {
    // every load of the plane is in flight before the first wait: the halo
    // values (an unused slot reads the plane's first value), then plane i + 3
    float hv[HaloStage<3>::NE];
    {
      hs.load(S + (size_t)i * plane, hv);
      const float* pp = S + (size_t)wrapi(i + 3, g.n0) * plane;
      for (int q = 0; q < ST_Q; q++) cur[q][3] = *reinterpret_cast<const V3*>(pp + co[q]);
    }
    __syncthreads();  // previous plane's readers are done
    {
      hs.store(hv, pl);
      for (int q = 0; q < ST_Q; q++) {
        float* o = pl + ((ty + 4 * q + ST_H) * ST_PK + tx + ST_H) * 3;
        o[0] = cur[q][0].x;
        o[1] = cur[q][0].y;
        o[2] = cur[q][0].z;
      }
    }
    __syncthreads();
    for (int q = 0; q < ST_Q; q++) {
      const int jl = ty + 4 * q, j = j0 + jl;
      const float* x = pl + ((jl + ST_H) * ST_PK + tx + ST_H) * 3;
      const float* y = x + 1;
      const float* z = x + 2;
      const float sxx = (float)(d6w(xo[q][0], xo[q][1], xo[q][2], cur[q][1].x, cur[q][2].x,
                                    cur[q][3].x) * inv60);
      const float sxy =
          (float)(d6w(x[-3 * DY], x[-2 * DY], x[-DY], x[DY], x[2 * DY], x[3 * DY]) * inv60);
      const float sxz =
          (float)(d6w(x[-3 * DZ], x[-2 * DZ], x[-DZ], x[DZ], x[2 * DZ], x[3 * DZ]) * inv60);
      const float syy =
          (float)(d6w(y[-3 * DY], y[-2 * DY], y[-DY], y[DY], y[2 * DY], y[3 * DY]) * inv60);
      const float syz =
          (float)(d6w(y[-3 * DZ], y[-2 * DZ], y[-DZ], y[DZ], y[2 * DZ], y[3 * DZ]) * inv60);
      const float szz =
          (float)(d6w(z[-3 * DZ], z[-2 * DZ], z[-DZ], z[DZ], z[2 * DZ], z[3 * DZ]) * inv60);
      const float t3 = sxx * syy + sxx * szz + syy * szz;
      const double v = (double)t3 - (double)sxy * (double)sxy - (double)sxz * (double)sxz -
                       (double)syz * (double)syz;
      src[((size_t)i * g.n1 + j) * g.n2 + k0 + tx] = (float)v;
    }
    for (int q = 0; q < ST_Q; q++) {
      xo[q][0] = xo[q][1];
      xo[q][1] = xo[q][2];
      xo[q][2] = cur[q][0].x;
      cur[q][0] = cur[q][1];
      cur[q][1] = cur[q][2];
      cur[q][2] = cur[q][3];
    }
  }
}

static TileGeom tile_geom(int n0, int n1, int n2) {
  TileGeom g;
  g.n0 = n0;
  g.n1 = n1;
  g.n2 = n2;
  g.ntk = (n2 + ST_TK - 1) / ST_TK;
  g.ntj = (n1 + ST_TJ - 1) / ST_TJ;
  g.ti = ST_TI;
  g.nti = (n0 + g.ti - 1) / g.ti;
  while (g.ti > 2 && (long long)g.ntk * g.ntj * g.nti < 512) {
    g.ti /= 2;
    g.nti = (n0 + g.ti - 1) / g.ti;
  }
  return g;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// Mode kernels: grid (each workgroup stages the spline once) and LDS.
static unsigned mode_grid(long long nh) {
  return (unsigned)std::max(1LL, std::min((nh + 255) / 256, 4096LL));
}

template <typename T>
static size_t spline_lds(const FieldArgs<T>& a) {
  return a.stage ? 5 * (size_t)a.nk * sizeof(double) : 0;
}

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
  // spline of P(k) on knots rounded as the wrappers pass them (fft_real),
  // built on the host and kept on the device while k and P are unchanged
  int nk = s->Nk;
  EXH_REQUIRE(nk >= 2, EXH_ERR_ARG, "P(k) needs at least two knots");
  const size_t cap0 = c->cap[BUF_TAB];
  double* d = (double*)ctx_buf(c, BUF_TAB, 5 * (size_t)nk * sizeof(double));
  if (c->cap[BUF_TAB] != cap0) c->spline_key.clear();  // reallocated
  std::vector<unsigned char> key;
  const int kp[2] = {(int)sizeof(T), nk};
  key_append(key, kp, 2);
  // Inter_Power truncates xi(r) when R_max (a T) is below 1e5
  const double rmax = (double)(T)s->R_max;
  const bool trunc = s->R_max > 0 && rmax < 100000.0;
  key_append(key, &rmax, trunc ? 1 : 0);
  key_append(key, s->k, nk);
  key_append(key, s->P, nk);
  if (key != c->spline_key) {
    // [x | y | b | c | d]: interval coefficients of y + t (b + t (c + t d)),
    // with the expressions of the oracle's spline_eval
    std::vector<double> tab(5 * (size_t)nk, 0.0), m(nk);
    double *x = tab.data(), *y = x + nk, *cb = y + nk, *cc = cb + nk, *cd = cc + nk;
    for (int i = 0; i < nk; i++) {
      x[i] = (double)(T)s->k[i];
      y[i] = (double)(T)s->P[i];
    }
    if (trunc) {
      // on the double copies of the T knots, splined without rounding
      std::vector<double> xin(x, x + nk), yin(y, y + nk);
      for (int i = 1; i < nk; i++)
        EXH_REQUIRE(xin[i] > xin[i - 1] && xin[0] > 0, EXH_ERR_ARG,
                    "the FFTLog truncation needs increasing positive k");
      pk_truncate(xin.data(), yin.data(), nk, rmax, x, y);
    }
    spline_second_derivs(x, y, nk, m.data());
    for (int lo = 0; lo + 1 < nk; lo++) {
      const int hi = lo + 1;
      const double h = x[hi] - x[lo];
      cb[lo] = (y[hi] - y[lo]) / h - h * (2.0 * m[lo] + m[hi]) / 6.0;
      cc[lo] = m[lo] / 2.0;
      cd[lo] = (m[hi] - m[lo]) / (6.0 * h);
    }
    EXH_CHECK_HIP(hipMemcpyAsync(d, tab.data(), tab.size() * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));  // tab is a host temporary
    c->spline_key = key;
    c->spline_x0 = x[0];
    c->spline_x1 = x[nk - 1];
  }
  a.sx = d;
  a.sy = d + nk;
  a.sb = d + 2 * nk;
  a.sc = d + 3 * nk;
  a.sd = d + 4 * nk;
  a.nk = nk;
  a.stage = 5 * (size_t)nk * sizeof(double) <= 48 * 1024 ? 1 : 0;
  if (trunc) {
    // the reference's roundings: the product of the three T frequencies in
    // T, pow in double rounded to T, the division by 4.0 in double, stored
    // as T (density_grid.c:159)
    const T prod = (T)((T)((T)b->kl[0] * (T)b->kl[1]) * (T)b->kl[2]);
    a.dc = 1;
    a.dc_kmod = (T)((double)(T)std::pow((double)prod, 1.0 / 3.0) / 4.0);
  }
  // the reference aborts (GSL EDOM) when |k| of the grid leaves the table
  double kmax = 0;
  for (int ax = 0; ax < 3; ax++) {
    double kk = (double)(T)b->kl[ax] * (b->nd[ax] / 2);
    kmax += kk * kk;
  }
  kmax = std::sqrt(kmax);
  double kmin = (double)std::min({(T)b->kl[0], (T)b->kl[1], (T)b->kl[2]});
  if (a.dc) kmin = std::min(kmin, (double)a.dc_kmod);
  if (kmin > k_smooth) kmin = 0;
  const double x0 = c->spline_x0, x1 = c->spline_x1;
  EXH_REQUIRE(x0 <= kmin * (1 + 1e-6) && std::min(kmax, k_smooth) <= x1 * (1 - 1e-7) + 1e-300,
              EXH_ERR_RANGE,
              "the box's |k| range [%g, %g] lies outside the tabulated P(k) [%g, %g]", kmin,
              std::min(kmax, k_smooth), x0, x1);
  return a;
}

// after_modes, when given, is called once the mode kernel is launched and
// before the inverse FFT (work on other streams that needs only deltak_raw
// can start there; the FFT works in its own buffer).
template <typename T>
void launch_field(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                  void* delta, void* deltak_raw, const std::function<void()>& after_modes) {
  using C2 = typename cplx<T>::type;
  auto a = make_field_args<T>(c, b, s, k_smooth);
  size_t nh = (size_t)b->nd[0] * b->nd[1] * b->nz2;
  C2* ck = (C2*)ctx_buf(c, BUF_CKF, nh * sizeof(C2));
  {
    TimedScope ts(c, "field_modes");
    auto kf = a.stage ? k_field<T, true> : k_field<T, false>;
    hipLaunchKernelGGL(kf, dim3(mode_grid((long long)nh)), dim3(256), spline_lds(a), c->stream, a,
                       (T)b->Normx, fast_div(b->nz2), fast_div(b->nd[1]), ck, (C2*)deltak_raw);
    EXH_LAUNCH_CHECK();
  }
  if (after_modes) after_modes();
  fft_c2r(c, b->nd[0], b->nd[1], b->nd[2], b->precision, 1, ck, delta);
}

// Displacements from a field source.  S: [ng][3] (displacements, or
// positions when out_pos).  V optional.
template <typename T>
void launch_lpt(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                int do_2lpt, int out_pos, T* S, T* V, T* P) {
  // the last gradient pass also writes the positions: into P, or over S
  // when out_pos
  if (out_pos && !P) P = S;
  const PosOut<T> none{nullptr, 0, 0, 0, 0};
  const PosOut<T> pos{P, (T)b->Lc, (T)b->L[0], (T)b->L[1], (T)b->L[2]};
  using C2 = typename cplx<T>::type;
  const int n0 = b->nd[0], n1 = b->nd[1], n2 = b->nd[2];
  size_t nh = (size_t)n0 * n1 * b->nz2, ng = (size_t)b->ng;
  C2* ck = (C2*)ctx_buf(c, BUF_CK0, nh * sizeof(C2));
  T* phi = (T*)ctx_buf(c, BUF_R0, ng * sizeof(T));
  T kl0 = (T)b->kl[0], kl1 = (T)b->kl[1], kl2 = (T)b->kl[2];
  T Normx = (T)b->Normx, Normk = (T)b->Normk;
  // phi_k = c_k pre / (-k^2) Normx, then phi(x): the division is fused into
  // the first inverse FFT pass where the strided passes cover the grid
  PoissonParams pp{n0, n1, n2, (double)kl0, (double)kl1, (double)kl2, (double)(T)k_smooth,
                   1.0, (double)Normx};
  auto poisson_c2r = [&](const void* src, T* out) {
    if (fft_c2r_poisson(c, b->precision, src, ck, out, pp)) return;
    {
      TimedScope ts(c, "lpt_kspace");
      if (src != ck)
        EXH_CHECK_HIP(hipMemcpyAsync(ck, src, nh * sizeof(C2), hipMemcpyDeviceToDevice,
                                     c->stream));
      hipLaunchKernelGGL(k_poisson_grid<T>, dim3(stream_grid(nh)), dim3(256), 0, c->stream, n0,
                         n1, n2, kl0, kl1, kl2, (T)k_smooth, (T)pp.pre, Normx, fast_div(b->nz2),
                         fast_div(n1), ck);
      EXH_LAUNCH_CHECK();
    }
    fft_c2r(c, n0, n1, n2, b->precision, 1, ck, out);
  };
  if (s->kind == EXH_SRC_MODES) {
    {
      TimedScope ts(c, "lpt1_kspace");
      auto a = make_field_args<T>(c, b, s, 1e5);
      auto kp = a.stage ? k_poisson_modes<T, true> : k_poisson_modes<T, false>;
      hipLaunchKernelGGL(kp, dim3(mode_grid((long long)nh)), dim3(256), spline_lds(a), c->stream,
                         a, (T)k_smooth, Normx, fast_div(b->nz2), fast_div(n1), ck);
      EXH_LAUNCH_CHECK();
    }
    fft_c2r(c, n0, n1, n2, b->precision, 1, ck, phi);
  } else if (s->kind == EXH_SRC_DELTAK) {
    poisson_c2r(s->data, phi);  // c_k = delta_k as given (pre = 1)
  } else {
    // Compute_Denk (density_grid.c:220-235): r2c then * Normk
    fft_r2c(c, n0, n1, n2, b->precision, 1, s->data, ck);
    pp.pre = (double)Normk;
    poisson_c2r(ck, phi);
  }
  const double inv60 = 1.0 / (60.0 * (double)(T)b->Lc);
  const TileGeom tg = tile_geom(n0, n1, n2);
  const unsigned tblocks = (unsigned)(tg.ntk * tg.ntj * tg.nti);
  const bool full = n1 % ST_TJ == 0 && n2 % ST_TK == 0;  // every tile full
  T f1 = (T)(std::pow((T)b->Omz, 0.5454) * (T)b->Hz / (1.0 + (T)b->redshift));
  {
    TimedScope ts(c, "lpt1_grad");
    if (full)
      hipLaunchKernelGGL((k_grad_tiles<T, 1, true>), dim3(tblocks), dim3(256), 0, c->stream, tg,
                         phi, inv60, 1.0, S, V, f1, do_2lpt ? none : pos);
    else
      hipLaunchKernelGGL((k_grad_tiles<T, 1, false>), dim3(tblocks), dim3(256), 0, c->stream, tg,
                         phi, inv60, 1.0, S, V, f1, do_2lpt ? none : pos);
    EXH_LAUNCH_CHECK();
  }
  if (do_2lpt) {
    T* src = (T*)ctx_buf(c, BUF_R1, ng * sizeof(T));
    {
      TimedScope ts(c, "lpt2_source");
      if (sizeof(T) == 4 && full)
        hipLaunchKernelGGL(k_source_full, dim3(tblocks), dim3(256), 0, c->stream, tg,
                           (const float*)S, inv60, (float*)src);
      else
        hipLaunchKernelGGL(k_source_tiles<T>, dim3(tblocks), dim3(256), 0, c->stream, tg, S,
                           inv60, src);
      EXH_LAUNCH_CHECK();
    }
    fft_r2c(c, n0, n1, n2, b->precision, 1, src, ck);
    pp.pre = (double)Normk;
    poisson_c2r(ck, src);
    const double coef = -3.0 / 7.0 * std::pow((double)(T)b->Omz, -1.0 / 143);
    T f2 = (T)(2.0 * std::pow((T)b->Omz, 6.0 / 11.0) * (T)b->Hz / (1.0 + (T)b->redshift));
    {
      TimedScope ts(c, "lpt2_grad");
      if (full)
        hipLaunchKernelGGL((k_grad_tiles<T, 2, true>), dim3(tblocks), dim3(256), 0, c->stream,
                           tg, src, inv60, coef, S, V, f2, pos);
      else
        hipLaunchKernelGGL((k_grad_tiles<T, 2, false>), dim3(tblocks), dim3(256), 0, c->stream,
                           tg, src, inv60, coef, S, V, f2, pos);
      EXH_LAUNCH_CHECK();
    }
  }
}


// dL/dnoise of the white-noise modes from gk = r2c(dL/dphi1), phi1 the
// first-order potential the LPT computes from those modes (k_poisson_modes).
template <typename T>
void launch_field_adjoint(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                          const void* gk, T* gnoise) {
  EXH_REQUIRE(!s->fixed && s->phase == 0.0, EXH_ERR_UNSUPPORTED,
              "the adjoint takes unit white noise without fixed amplitudes or a phase");
  auto a = make_field_args<T>(c, b, s, 1e5);
  const size_t nh = (size_t)b->nd[0] * b->nd[1] * b->nz2;
  TimedScope ts(c, "field_adjoint");
  auto kf = a.stage ? k_field_adjoint<T, true> : k_field_adjoint<T, false>;
  hipLaunchKernelGGL(kf, dim3(mode_grid((long long)nh)), dim3(256), spline_lds(a), c->stream, a,
                     (T)k_smooth, (T)b->Normx, fast_div(b->nz2), fast_div(b->nd[1]),
                     (const typename cplx<T>::type*)gk, gnoise);
  EXH_LAUNCH_CHECK();
}

template void launch_field_adjoint<float>(exh_ctx*, const exh_box*, const exh_field_src*, double,
                                          const void*, float*);
template void launch_field_adjoint<double>(exh_ctx*, const exh_box*, const exh_field_src*,
                                           double, const void*, double*);
template void launch_field<float>(exh_ctx*, const exh_box*, const exh_field_src*, double, void*,
                                  void*, const std::function<void()>&);
template void launch_field<double>(exh_ctx*, const exh_box*, const exh_field_src*, double, void*,
                                   void*, const std::function<void()>&);
template void launch_lpt<float>(exh_ctx*, const exh_box*, const exh_field_src*, double, int, int,
                                float*, float*, float*);
template void launch_lpt<double>(exh_ctx*, const exh_box*, const exh_field_src*, double, int, int,
                                 double*, double*, double*);

}  // namespace exh
