// Mass assignment and power-spectrum estimator (gfx950).
//
//  k_deposit     NGP / CIC / TSC scatter of particles (and of the copy
//                shifted by -Lb/2 for interlacing) with device-scope atomics
//                (gridmodule.c:9-53, 146-270); float atomics in FP32, double
//                atomics in FP64
//  k_deposit_tiles  the same weights for particles in Lagrangian lattice
//                order, accumulated per lattice tile in LDS (double) and
//                flushed with one coalesced atomic per touched cell
//  k_normalise   rho / rho_bar - 1 for unweighted grids (gridmodule.c:197-202)
//  k_power_bins  one pass over the half-complex modes of all tracer grids:
//                interlacing combine, window deconvolution and Normk
//                (spectrum_h.c:112-157), the reference's mode mask and bin
//                index (powermodule.c:47-55), and the pair products times
//                P_2l(mu) (powermodule.c:66-72).  Each wavefront takes one
//                (i, j) row at a time; along a row the bin index never
//                decreases, so bins are summed by a segmented scan across
//                the 64 lanes and added by the segment's last lane into a
//                wave-private LDS accumulator.  Waves and blocks are then
//                reduced in a fixed order: the double sums are bit-for-bit
//                reproducible.
#include "common.hpp"
#include "kernels.hpp"
#include "util.hpp"

#include <array>
#include <cmath>
#include <cstdlib>
#include <mutex>

namespace exh {

// ---------------------------------------------------------------------------
// deposit
// ---------------------------------------------------------------------------
// Periodic wrap into [0, L).  Values within one box length of the range
// take the reference's subtract/add loop (spectrum_h.c:18-28); farther
// values are reduced with floor first, so that the loop is bounded (the
// reference's loop does not terminate for huge or infinite values).
// Non-finite input returns -1, which the callers skip.
template <typename T>
__device__ inline T wrap_while(T x, T L) {
  if (!isfinite(x)) return (T)-1;
  if (x >= (T)2 * L || x < -L) x = x - floor(x / L) * L;
  while (x >= L) x -= L;
  while (x < (T)0) x += L;
  return x;
}

// Hardware float/double atomic add (global_atomic_add_f32/f64).  HIP's
// atomicAdd(float*) defaults to a compare-and-swap loop, which serialises
// under the contention of neighbouring particles; the grids live in
// coarse-grained hipMalloc memory, where the hardware instruction is exact.
template <typename T>
__device__ inline void atomic_add_t(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

template <typename T>
__device__ inline void dep_ngp(T* g, const T* x, int nd, T Ld, T m) {
  int xt[3];
  for (int a = 0; a < 3; a++) {
    xt[a] = (int)floor(x[a] / Ld);
    if (xt[a] == nd) xt[a] -= 1;
  }
  atomic_add_t(g + ((size_t)xt[0] * nd + xt[1]) * nd + xt[2], m);
}

template <typename T>
__device__ inline void dep_cic(T* g, const T* x, int nd, T Ld, T m) {
  int xt[3], o[3];
  T dx[3], t[3];
  for (int a = 0; a < 3; a++) {
    T q = x[a] / Ld;
    xt[a] = (int)floor(q);
    if (xt[a] == nd) xt[a] -= 1;
    dx[a] = (T)((double)q - ((double)xt[a] + 0.5));
    if (dx[a] >= (T)0) {
      o[a] = xt[a] + 1 == nd ? 0 : xt[a] + 1;
    } else {
      o[a] = xt[a] == 0 ? nd - 1 : xt[a] - 1;
      dx[a] = -dx[a];
    }
    t[a] = (T)(1.0 - (double)dx[a]);
  }
  const size_t n2 = (size_t)nd * nd;
  const size_t X0 = xt[0] * n2, X1 = o[0] * n2, Y0 = (size_t)xt[1] * nd, Y1 = (size_t)o[1] * nd;
  const size_t Z0 = xt[2], Z1 = o[2];
  atomic_add_t(g + X0 + Y0 + Z0, m * t[0] * t[1] * t[2]);
  atomic_add_t(g + X1 + Y0 + Z0, m * dx[0] * t[1] * t[2]);
  atomic_add_t(g + X0 + Y1 + Z0, m * t[0] * dx[1] * t[2]);
  atomic_add_t(g + X0 + Y0 + Z1, m * t[0] * t[1] * dx[2]);
  atomic_add_t(g + X1 + Y1 + Z0, m * dx[0] * dx[1] * t[2]);
  atomic_add_t(g + X1 + Y0 + Z1, m * dx[0] * t[1] * dx[2]);
  atomic_add_t(g + X0 + Y1 + Z1, m * t[0] * dx[1] * dx[2]);
  atomic_add_t(g + X1 + Y1 + Z1, m * dx[0] * dx[1] * dx[2]);
}

template <typename T>
__device__ inline void dep_tsc(T* g, const T* x, int nd, T Ld, T m) {
  int c[3];
  T w[3][3];
  for (int a = 0; a < 3; a++) {
    double u = (double)x[a] / (double)Ld - 0.5;
    double cf = floor(u + 0.5);
    double d = u - cf;
    c[a] = (int)cf;
    w[a][0] = (T)(0.5 * (0.5 - d) * (0.5 - d));
    w[a][1] = (T)(0.75 - d * d);
    w[a][2] = (T)(0.5 * (0.5 + d) * (0.5 + d));
  }
  for (int a = 0; a < 3; a++) {
    int ix = c[0] + a - 1;
    ix = ix < 0 ? ix + nd : (ix >= nd ? ix - nd : ix);
    for (int b = 0; b < 3; b++) {
      int iy = c[1] + b - 1;
      iy = iy < 0 ? iy + nd : (iy >= nd ? iy - nd : iy);
      T wab = m * w[0][a] * w[1][b];
      for (int e = 0; e < 3; e++) {
        int iz = c[2] + e - 1;
        iz = iz < 0 ? iz + nd : (iz >= nd ? iz - nd : iz);
        atomic_add_t(g + ((size_t)ix * nd + iy) * nd + iz, wab * w[2][e]);
      }
    }
  }
}

template <typename T, int WIN>
__global__ void k_deposit(const T* pos, const T* vel, const T* mass, const int* type,
                          long long np, int nd, T L, T Lb, int direction, T rsd, int nsub,
                          T* grid, unsigned long long* counts) {
  const size_t ng = (size_t)nd * nd * nd;
  const T Ld = L / (T)nd;
  // particle counts per type (the normalisation's total mass): registers
  // for up to 4 types, an LDS histogram beyond; one atomic per block and
  // type at the end instead of one per particle on a single address
  __shared__ unsigned long long sh_cnt[64];
  unsigned long long reg_cnt[4] = {0, 0, 0, 0};
  const bool count = !mass && counts;
  const int ntype_reg = 4;
  if (threadIdx.x < 64) sh_cnt[threadIdx.x] = 0;
  __syncthreads();
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < np;
       j += (long long)gridDim.x * blockDim.x) {
    T x[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    if (direction >= 0) x[direction] = wrap_while<T>(x[direction] + vel[3 * j + direction] * rsd, L);
    int t = type ? type[j] : 0;
    T m = mass ? mass[j] : (T)1;
    T p[3];
    for (int a = 0; a < 3; a++) p[a] = wrap_while<T>(x[a], L);
    if (p[0] < (T)0 || p[1] < (T)0 || p[2] < (T)0) continue;  // non-finite position
    T* g = grid + (size_t)t * nsub * ng;
    if (WIN == 1) dep_ngp(g, p, nd, Ld, m);
    if (WIN == 2) dep_cic(g, p, nd, Ld, m);
    if (WIN == 5) dep_tsc(g, p, nd, Ld, m);
    if (nsub == 2) {
      T sh = (T)(-(double)Lb / 2.0);
      for (int a = 0; a < 3; a++) p[a] = wrap_while<T>(x[a] + sh, L);
      if (WIN == 1) dep_ngp(g + ng, p, nd, Ld, m);
      if (WIN == 2) dep_cic(g + ng, p, nd, Ld, m);
      if (WIN == 5) dep_tsc(g + ng, p, nd, Ld, m);
    }
    if (count) {
      if (t < ntype_reg) {
        for (int q = 0; q < ntype_reg; q++) reg_cnt[q] += (q == t);
      } else {
        atomicAdd(&sh_cnt[t], 1ull);
      }
    }
  }
  if (!count) return;
  for (int q = 0; q < ntype_reg; q++)
    if (reg_cnt[q]) atomicAdd(&sh_cnt[q], reg_cnt[q]);
  __syncthreads();
  if (threadIdx.x < 64 && sh_cnt[threadIdx.x]) atomicAdd(&counts[threadIdx.x], sh_cnt[threadIdx.x]);
}

// Spherical (window 3) and exponential (window 4) assignment
// (gridmodule.c:57-119): every cell of the cube of half-side
// times = floor(1/2 + R_eff / Ld) around the particle's cell whose centre lies
// within R_eff gets m (sphere, R_eff = R) or m exp(-r^2 / 2R^2) (R_eff =
// R R_times).  As in the reference, each offset is evaluated in double and
// stored in T, the periodic image is L - dx, and r^2 is summed in T.
// Returns the deposited weight.
__device__ inline int wrap_any(int x, int n) {
  int r = x % n;
  return r < 0 ? r + n : r;
}

template <typename T, bool EXPW>
__device__ inline double dep_sphere(T* g, const T* x, int nd, T L, T Ld, T R, T Rt, T m) {
  int xt[3];
  for (int a = 0; a < 3; a++) {
    xt[a] = (int)floor(x[a] / Ld);
    if (xt[a] == nd) xt[a] -= 1;
  }
  const int times = EXPW ? (int)floor(0.5 + (double)(R * Rt / Ld)) : (int)floor(0.5 + (double)(R / Ld));
  const T R2 = R * R, Re2 = R * R * Rt * Rt;
  const double halfL = (double)L / 2.0;
  auto offset = [&](int a, int c) {
    T d = (T)((double)x[a] - ((double)c + 0.5) * (double)Ld);
    if ((double)d > halfL) d = L - d;
    if ((double)d < -halfL) d = L + d;
    return d;
  };
  double M = 0.0;
  for (int i = -times; i <= times; i++) {
    const int ci = wrap_any(xt[0] + i, nd);
    const T d0 = offset(0, ci);
    for (int j = -times; j <= times; j++) {
      const int cj = wrap_any(xt[1] + j, nd);
      const T d1 = offset(1, cj);
      T* row = g + ((size_t)ci * nd + cj) * nd;
      for (int k = -times; k <= times; k++) {
        const int ck = wrap_any(xt[2] + k, nd);
        const T d2 = offset(2, ck);
        const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (!EXPW) {
          if (r2 <= R2) {
            atomic_add_t(row + ck, m);
            M += (double)m;
          }
        } else if (r2 <= Re2) {
          const T w = (T)exp(-0.5 * (double)r2 / (double)R / (double)R);
          atomic_add_t(row + ck, m * w);
          M += (double)m * (double)w;
        }
      }
    }
  }
  return M;
}

// One thread per particle; wsum[t * nsub + s] receives the deposited weight
// of each grid (the normalisation's total, in place of the particle count).
template <typename T, bool EXPW>
__global__ void k_deposit_sphere(const T* pos, const T* vel, const T* mass, const int* type,
                                 long long np, int nd, T L, T Lb, T R, T Rt, int direction, T rsd,
                                 int nsub, int ngrids, T* grid, double* wsum) {
  __shared__ double sh[128];
  const size_t ng = (size_t)nd * nd * nd;
  const T Ld = L / (T)nd;
  if (threadIdx.x < 128) sh[threadIdx.x] = 0.0;
  __syncthreads();
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < np;
       j += (long long)gridDim.x * blockDim.x) {
    T x[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    if (direction >= 0) x[direction] = wrap_while<T>(x[direction] + vel[3 * j + direction] * rsd, L);
    const int t = type ? type[j] : 0;
    const T m = mass ? mass[j] : (T)1;
    T p[3];
    for (int a = 0; a < 3; a++) p[a] = wrap_while<T>(x[a], L);
    if (p[0] < (T)0 || p[1] < (T)0 || p[2] < (T)0) continue;  // non-finite position
    T* g = grid + (size_t)t * nsub * ng;
    atomicAdd(&sh[t * nsub], dep_sphere<T, EXPW>(g, p, nd, L, Ld, R, Rt, m));
    if (nsub == 2) {
      const T sft = (T)(-(double)Lb / 2.0);
      for (int a = 0; a < 3; a++) p[a] = wrap_while<T>(x[a] + sft, L);
      atomicAdd(&sh[t * nsub + 1], dep_sphere<T, EXPW>(g + ng, p, nd, L, Ld, R, Rt, m));
    }
  }
  __syncthreads();
  if (threadIdx.x < ngrids && sh[threadIdx.x] != 0.0) atomicAdd(&wsum[threadIdx.x], sh[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// LDS-privatised deposit for particles stored in Lagrangian lattice order
// (particle (i, j, k) of an nL^3 lattice at index (i nL + j) nL + k, as the
// LPT positions are).  One workgroup takes a TL^3 tile of the lattice: its
// particles, displaced by a smooth field, land in a compact Eulerian box.
// The block finds that box (cells relative to the tile centre, unwrapped
// across the periodic boundary), accumulates the window weights there with
// LDS atomics, and adds each touched cell to the grid with one coalesced
// global atomic (about 3 per particle for CIC at nd = nL instead of 8
// conflicting ones).  Tiles whose box exceeds the LDS deposit directly.
// The boxes accumulate in double for either grid precision: LDS float
// atomic adds (ds_add_f32) run at a small fraction of the rate of
// ds_add_f64 on gfx950, and the double sums round once when flushed.
// Weights are computed exactly as in dep_ngp/dep_cic/dep_tsc.
// ---------------------------------------------------------------------------
__device__ inline int rel_cell(int c, int qc, int nd) {
  int d = c - qc;
  if (d >= nd / 2) d -= nd;
  if (d < -nd / 2) d += nd;
  return d;
}

// Cells handled here lie within (-nd/2, 3nd/2) (a stencil neighbour, or a
// box cell a relative offset of less than nd/2 from a tile centre), so one
// conditional step wraps them exactly.
__device__ inline int wrap_cell(int c, int nd) { return c < 0 ? c + nd : (c >= nd ? c - nd : c); }

// Scaled coordinate of a wrapped position, the one division of a window
// evaluation: x / Ld for NGP and CIC, x / Ld - 0.5 in double for TSC.
template <typename T, int WIN>
__device__ inline auto scale_q(T x, T Ld) {
  if constexpr (WIN == 5)
    return (double)x / (double)Ld - 0.5;
  else
    return x / Ld;
}

// Cells and weights along one axis from the scaled coordinate, with the
// expressions of dep_ngp/dep_cic/dep_tsc.
template <typename T, int WIN, typename Q>
__device__ inline int stencil_q(Q q, int nd, int* cell, T* w) {
  if constexpr (WIN == 1) {
    int xt = (int)floor(q);
    if (xt == nd) xt -= 1;
    cell[0] = xt;
    w[0] = (T)1;
    return 1;
  } else if constexpr (WIN == 2) {
    int xt = (int)floor(q);
    if (xt == nd) xt -= 1;
    // dx = q - (xt + 1/2) and 1 - dx in T: both differences are exact in
    // double and round once, so T's correctly rounded subtraction gives the
    // reference's double-then-round values bit for bit
    T dx = q - ((T)xt + (T)0.5);
    int o;
    if (dx >= (T)0) {
      o = xt + 1 == nd ? 0 : xt + 1;
    } else {
      o = xt == 0 ? nd - 1 : xt - 1;
      dx = -dx;
    }
    cell[0] = xt;
    w[0] = (T)1 - dx;
    cell[1] = o;
    w[1] = dx;
    return 2;
  } else {
    const double u = q;
    const double cf = floor(u + 0.5);
    const double d = u - cf;
    const int c = (int)cf;
    cell[0] = wrap_cell(c - 1, nd);
    cell[1] = wrap_cell(c, nd);
    cell[2] = wrap_cell(c + 1, nd);
    w[0] = (T)(0.5 * (0.5 - d) * (0.5 - d));
    w[1] = (T)(0.75 - d * d);
    w[2] = (T)(0.5 * (0.5 + d) * (0.5 + d));
    return 3;
  }
}

// Cells and weights of one particle along one axis for window WIN.
template <typename T, int WIN>
__device__ inline int axis_stencil(T x, T Ld, int nd, int* cell, T* w) {
  return stencil_q<T, WIN>(scale_q<T, WIN>(x, Ld), nd, cell, w);
}

template <int N>
__device__ inline void block_minmax(int* v_min, int* v_max, int* sh) {
  // sh: 2 * N * (blockDim/64) ints; reduces N axes at once
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int a = 0; a < N; a++) {
    int mn = v_min[a], mx = v_max[a];
    for (int o = 32; o > 0; o >>= 1) {
      mn = min(mn, __shfl_xor(mn, o, 64));
      mx = max(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) {
      sh[(2 * a) * nw + wave] = mn;
      sh[(2 * a + 1) * nw + wave] = mx;
    }
  }
  __syncthreads();
  for (int a = 0; a < N; a++) {
    int mn = sh[(2 * a) * nw], mx = sh[(2 * a + 1) * nw];
    for (int w = 1; w < nw; w++) {
      mn = min(mn, sh[(2 * a) * nw + w]);
      mx = max(mx, sh[(2 * a + 1) * nw + w]);
    }
    v_min[a] = mn;
    v_max[a] = mx;
  }
  __syncthreads();
}

static constexpr int TILE_L = 16;        // lattice tile side
static constexpr int TILE_THREADS = 1024;
static constexpr int TILE_PPT = TILE_L * TILE_L * TILE_L / TILE_THREADS;  // particles per thread

template <typename T, int WIN>
__global__ __launch_bounds__(TILE_THREADS) void k_deposit_tiles(
    const T* pos, const T* vel, const T* mass, int nL, int nd, int W, T L, T Lb, int direction,
    T rsd, int nsub, int cap, T* grid, unsigned long long* counts) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* box = (double*)lds_raw;
  __shared__ int red[2 * 3 * (TILE_THREADS / 64)];
  __shared__ unsigned long long nvalid;
  constexpr int NW = WIN == 1 ? 1 : (WIN == 2 ? 2 : 3);
  const size_t ng = (size_t)nd * nd * nd;
  const T Ld = L / (T)nd;
  const int nt = nL / TILE_L;
  const int tz = blockIdx.x % nt, ty = (blockIdx.x / nt) % nt, tx = blockIdx.x / (nt * nt);
  // tile centre in grid cells: the middle of the grid tile of the same index
  const int qc[3] = {tx * W + W / 2, ty * W + W / 2, tz * W + W / 2};
  if (threadIdx.x == 0) nvalid = 0;
  // particles of this thread: lanes along k (coalesced rows)
  T xs[TILE_PPT][3];
  T ms[TILE_PPT];
  bool ok[TILE_PPT];
  unsigned long long myvalid = 0;
  for (int p = 0; p < TILE_PPT; p++) {
    int lin = threadIdx.x + p * TILE_THREADS;
    int c = lin % TILE_L, b = (lin / TILE_L) % TILE_L, a = lin / (TILE_L * TILE_L);
    size_t j = ((size_t)(tx * TILE_L + a) * nL + (ty * TILE_L + b)) * nL + (tz * TILE_L + c);
    T x[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    if (direction >= 0) x[direction] = wrap_while<T>(x[direction] + vel[3 * j + direction] * rsd, L);
    ok[p] = isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]);
    for (int q = 0; q < 3; q++) xs[p][q] = x[q];
    ms[p] = mass ? mass[j] : (T)1;
    myvalid += ok[p];
  }
  for (int s = 0; s < nsub; s++) {
    T* g = grid + (size_t)s * ng;
    const T sh = s == 0 ? (T)0 : (T)(-(double)Lb / 2.0);
    // box extent of this sub-grid's stencils
    int vmin[3] = {1 << 30, 1 << 30, 1 << 30}, vmax[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int p = 0; p < TILE_PPT; p++) {
      if (!ok[p]) continue;
      for (int q = 0; q < 3; q++) {
        T pa = wrap_while<T>(xs[p][q] + sh, L);
        int cell[3];
        T w[3];
        int n = axis_stencil<T, WIN>(pa, Ld, nd, cell, w);
        for (int u = 0; u < n; u++) {
          int r = rel_cell(cell[u], qc[q], nd);
          vmin[q] = min(vmin[q], r);
          vmax[q] = max(vmax[q], r);
        }
      }
    }
    block_minmax<3>(vmin, vmax, red);
    if (vmax[0] < vmin[0]) continue;  // no valid particle
    const int e0 = vmax[0] - vmin[0] + 1, e1 = vmax[1] - vmin[1] + 1, e2 = vmax[2] - vmin[2] + 1;
    const bool fits = (long long)e0 * e1 * e2 <= cap;
    if (fits) {
      const int vol = e0 * e1 * e2;
      for (int v = threadIdx.x; v < vol; v += TILE_THREADS) box[v] = 0.0;
      __syncthreads();
    }
    for (int p = 0; p < TILE_PPT; p++) {
      if (!ok[p]) continue;
      int cell[3][3];
      T w[3][3];
      for (int q = 0; q < 3; q++) {
        T pa = wrap_while<T>(xs[p][q] + sh, L);
        axis_stencil<T, WIN>(pa, Ld, nd, cell[q], w[q]);
      }
      for (int u = 0; u < NW; u++) {
        for (int b = 0; b < NW; b++) {
          T wab = ms[p] * w[0][u] * w[1][b];
          for (int c = 0; c < NW; c++) {
            T val = wab * w[2][c];
            const int lx = rel_cell(cell[0][u], qc[0], nd) - vmin[0];
            const int ly = rel_cell(cell[1][b], qc[1], nd) - vmin[1];
            const int lz = rel_cell(cell[2][c], qc[2], nd) - vmin[2];
            // the box bounds were taken over these same stencils; the
            // range test only guards the LDS index
            if (fits && (unsigned)lx < (unsigned)e0 && (unsigned)ly < (unsigned)e1 &&
                (unsigned)lz < (unsigned)e2) {
              atomicAdd(&box[(lx * e1 + ly) * e2 + lz], (double)val);
            } else {
              atomic_add_t(g + ((size_t)cell[0][u] * nd + cell[1][b]) * nd + cell[2][c], val);
            }
          }
        }
      }
    }
    if (fits) {
      __syncthreads();
      const int vol = e0 * e1 * e2;
      for (int v = threadIdx.x; v < vol; v += TILE_THREADS) {
        const double val = box[v];
        if (val == 0.0) continue;
        const int lz = v % e2, ly = (v / e2) % e1, lx = v / (e1 * e2);
        const int cx = wrap_cell(qc[0] + vmin[0] + lx, nd), cy = wrap_cell(qc[1] + vmin[1] + ly, nd),
                  cz = wrap_cell(qc[2] + vmin[2] + lz, nd);
        atomic_add_t(g + ((size_t)cx * nd + cy) * nd + cz, (T)val);
      }
      __syncthreads();
    }
  }
  if (!mass && counts) {
    atomicAdd(&nvalid, myvalid);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&counts[0], nvalid);
  }
}

// Window deposit: the tile's particles go into a fixed LDS window of
// E = W + 2 h cells per axis (W = grid cells per tile side, margin h), placed at
// the tile's mean displacement.  The mean is taken over one sampled
// particle per wave, so no pass over all particles precedes the deposit;
// each particle's wrap and division are evaluated once per sub-grid.  The
// rare stencil that leaves the window goes straight to the grid.  The
// window is flushed row by row, lanes along z, one global atomic per
// touched cell.
//
// A tile is PX x PY x PZ particles of the lattice and the window holds
// ACC (float: ds_add_f32, or double: ds_add_f64) sums.  The flush's global
// atomics scale with the window's cells per grid cell, (W + 2h)^3 / W^3 for
// cubic tiles, so larger tiles in float windows issue fewer of them.
static constexpr int WIN_THREADS = 1024;
static constexpr int WIN_HALO = 4;  // default margin h

// Window sums: float or double (LDS float/double atomic adds), or unsigned
// 32-bit fixed point with 22 fraction bits (ds_add_u32; integer sums are
// exact, so the window does not depend on the order of the adds).  Fixed
// point serves unweighted deposits only, whose weights lie in [0, 1]: a
// cell holds 1023 particles' weight before the sum would wrap.  Every add
// checks the value it replaced, so the add that wraps is seen; its tile
// then deposits straight to the grid instead of through the window.
template <typename ACC>
struct WinAcc {
  template <typename T>
  __device__ static ACC to(T v) {
    return (ACC)v;
  }
  __device__ static double from(ACC a) { return (double)a; }
  // add v to *p; true if the sum wrapped
  __device__ static bool add(ACC* p, ACC v) {
    atomicAdd(p, v);
    return false;
  }
};
template <>
struct WinAcc<unsigned> {
  template <typename T>
  __device__ static unsigned to(T v) {
    return (unsigned)((float)v * 4194304.0f + 0.5f);
  }
  __device__ static double from(unsigned a) { return (double)a * (1.0 / 4194304.0); }
  __device__ static bool add(unsigned* p, unsigned v) {
    const unsigned old = atomicAdd(p, v);
    return old + v < old;
  }
};

template <typename T, int WIN, typename ACC, int PX, int PY, int PZ>
__device__ __forceinline__ void deposit_window_tile(
    int tile, const T* pos, const T* vel, const T* mass, int nL, int nd, int Wx, int Wy, int Wz,
    int wh, T L, T Lb, int direction, T rsd, int nsub, T* grid, unsigned long long* counts) {
  constexpr int NW = WIN == 1 ? 1 : (WIN == 2 ? 2 : 3);
  constexpr int NWAVE = WIN_THREADS / 64;
  constexpr int PPT = PX * PY * PZ / WIN_THREADS;
  static_assert(PPT * WIN_THREADS == PX * PY * PZ, "tile must be a multiple of the block");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  ACC* box = (ACC*)lds_raw;
  __shared__ float samp[3][NWAVE];
  __shared__ int wo_s[3];
  __shared__ unsigned long long nvalid;
  const size_t ng = (size_t)nd * nd * nd;
  const T Ld = L / (T)nd;
  const int nty = nL / PY, ntz = nL / PZ;
  const int W[3] = {Wx, Wy, Wz};
  const int Ex = Wx + 2 * wh, Ey = Wy + 2 * wh, Ez = Wz + 2 * wh, sx = Ey * Ez;
  const int tz = tile % ntz, ty = (tile / ntz) % nty, tx = tile / (ntz * nty);
  // tile centre in grid cells: the middle of the grid tile of the same index
  const int qc[3] = {tx * Wx + Wx / 2, ty * Wy + Wy / 2, tz * Wz + Wz / 2};
  if (threadIdx.x == 0) nvalid = 0;
  T xs[PPT][3];
  T ms[PPT];
  bool ok[PPT];
  unsigned long long myvalid = 0;
#pragma unroll
  for (int p = 0; p < PPT; p++) {
    const int lin = threadIdx.x + p * WIN_THREADS;
    const int c = lin % PZ, b = (lin / PZ) % PY, a = lin / (PZ * PY);
    const size_t j = ((size_t)(tx * PX + a) * nL + (ty * PY + b)) * nL + (tz * PZ + c);
    T x[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
    if (direction >= 0) x[direction] = wrap_while<T>(x[direction] + vel[3 * j + direction] * rsd, L);
    ok[p] = isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]);
#pragma unroll
    for (int q = 0; q < 3; q++) xs[p][q] = x[q];
    ms[p] = mass ? mass[j] : (T)1;
    myvalid += ok[p];
  }
  // displacement (in cells) of each wave's first particle: its cell relative
  // to the tile centre minus its lattice point's offset from the centre
  if ((threadIdx.x & 63) == 0) {
    const int lin = threadIdx.x;
    const int lat[3] = {lin / (PZ * PY), (lin / PZ) % PY, lin % PZ};
#pragma unroll
    for (int q = 0; q < 3; q++) {
      float r = 1e30f;
      if (ok[0]) {
        int cc = (int)floor(wrap_while<T>(xs[0][q], L) / Ld);
        cc = cc < nd ? cc : nd - 1;
        r = (float)rel_cell(cc, qc[q], nd) -
            ((lat[q] + 0.5f) * (float)nd / (float)nL - 0.5f * (float)W[q]);
      }
      samp[q][threadIdx.x >> 6] = r;
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    float sum = 0.f;
    int n = 0;
    for (int w = 0; w < NWAVE; w++) {
      const float v = samp[threadIdx.x][w];
      if (v < 1e29f) {
        sum += v;
        n++;
      }
    }
    const float mean = n ? sum / (float)n : 0.f;
    wo_s[threadIdx.x] = (int)floorf(mean + 0.5f) - W[threadIdx.x] / 2 - wh;
  }
  __syncthreads();
  const int wo[3] = {wo_s[0], wo_s[1], wo_s[2]};
  const int E[3] = {Ex, Ey, Ez};
  const int vol = Ex * sx;
  const T shift = (T)(-(double)Lb / 2.0);
  // CIC scales by the reciprocal of the cell size (its weights are
  // continuous in the position, so the rounding of x / Ld against x * (1 /
  // Ld) moves them by an ulp); NGP and TSC divide as the reference does
  const T invLd = (T)1 / Ld;
  for (int s = 0; s < nsub; s++) {
    T* g = grid + (size_t)s * ng;
    for (int v = threadIdx.x; v < vol; v += WIN_THREADS) box[v] = (ACC)0;
    __syncthreads();
    bool wrapped = false;
#pragma unroll
    for (int p = 0; p < PPT; p++) {
      if (!ok[p]) continue;
      int cell[3][3], off[3][3];
      T w[3][3];
      bool in = true;
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const T pa = wrap_while<T>(s == 0 ? xs[p][q] : xs[p][q] + shift, L);
        if constexpr (WIN == 2)
          stencil_q<T, WIN>(pa * invLd, nd, cell[q], w[q]);
        else
          stencil_q<T, WIN>(scale_q<T, WIN>(pa, Ld), nd, cell[q], w[q]);
        const int st = q == 0 ? sx : (q == 1 ? Ez : 1);
#pragma unroll
        for (int u = 0; u < NW; u++) {
          const int l = rel_cell(cell[q][u], qc[q], nd) - wo[q];
          in = in && (unsigned)l < (unsigned)E[q];
          off[q][u] = l * st;
        }
      }
      if (in) {
#pragma unroll
        for (int u = 0; u < NW; u++) {
#pragma unroll
          for (int b = 0; b < NW; b++) {
            const T wab = ms[p] * w[0][u] * w[1][b];
            const int o = off[0][u] + off[1][b];
#pragma unroll
            for (int c = 0; c < NW; c++)
              wrapped |= WinAcc<ACC>::add(&box[o + off[2][c]], WinAcc<ACC>::to(wab * w[2][c]));
          }
        }
      } else {
        for (int u = 0; u < NW; u++)
          for (int b = 0; b < NW; b++) {
            const T wab = ms[p] * w[0][u] * w[1][b];
            for (int c = 0; c < NW; c++)
              atomic_add_t(g + ((size_t)cell[0][u] * nd + cell[1][b]) * nd + cell[2][c],
                           wab * w[2][c]);
          }
      }
    }
    if (__syncthreads_or(wrapped)) {
      // a fixed-point window cell wrapped: the window is void; every
      // in-window stencil of the tile goes to the grid directly instead
      // (the out-of-window ones already did)
#pragma unroll
      for (int p = 0; p < PPT; p++) {
        if (!ok[p]) continue;
        int cell[3][3];
        T w[3][3];
        bool in = true;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const T pa = wrap_while<T>(s == 0 ? xs[p][q] : xs[p][q] + shift, L);
          // the stencil of the window pass, so that `in` matches it
          if constexpr (WIN == 2)
            stencil_q<T, WIN>(pa * invLd, nd, cell[q], w[q]);
          else
            stencil_q<T, WIN>(scale_q<T, WIN>(pa, Ld), nd, cell[q], w[q]);
          for (int u = 0; u < NW; u++) {
            const int l = rel_cell(cell[q][u], qc[q], nd) - wo[q];
            in = in && (unsigned)l < (unsigned)E[q];
          }
        }
        if (!in) continue;
        for (int u = 0; u < NW; u++)
          for (int b = 0; b < NW; b++) {
            const T wab = ms[p] * w[0][u] * w[1][b];
            for (int c = 0; c < NW; c++)
              atomic_add_t(g + ((size_t)cell[0][u] * nd + cell[1][b]) * nd + cell[2][c],
                           wab * w[2][c]);
          }
      }
      continue;
    }
    // flush: (y, z) rows over the threads, z fastest in 32-lane groups
    const int lz0 = threadIdx.x & 31;
    for (int ly = threadIdx.x >> 5; ly < Ey; ly += WIN_THREADS / 32) {
      const int cy = wrap_cell(qc[1] + wo[1] + ly, nd);
      for (int lz = lz0; lz < Ez; lz += 32) {
        const int cz = wrap_cell(qc[2] + wo[2] + lz, nd);
        const ACC* bp = box + ly * Ez + lz;
        int cx = wrap_cell(qc[0] + wo[0], nd);
        for (int lx = 0; lx < Ex; lx++) {
          const ACC val = bp[lx * sx];
          if (val != (ACC)0)
            atomic_add_t(g + ((size_t)cx * nd + cy) * nd + cz, (T)WinAcc<ACC>::from(val));
          cx = cx + 1 == nd ? 0 : cx + 1;
        }
      }
    }
    __syncthreads();
  }
  if (!mass && counts) {
    atomicAdd(&nvalid, myvalid);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&counts[0], nvalid);
  }
}

// One workgroup per tile, or (EXH_DEPOSIT_WG) fewer workgroups that walk the
// tiles, leaving CUs to the kernels of other streams.
template <typename T, int WIN, typename ACC, int PX, int PY, int PZ>
__global__ __launch_bounds__(WIN_THREADS) void k_deposit_window(
    const T* pos, const T* vel, const T* mass, int nL, int nd, int Wx, int Wy, int Wz, int wh,
    T L, T Lb, int direction, T rsd, int nsub, T* grid, unsigned long long* counts, int ntiles) {
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    deposit_window_tile<T, WIN, ACC, PX, PY, PZ>(tile, pos, vel, mass, nL, nd, Wx, Wy, Wz, wh, L,
                                                 Lb, direction, rsd, nsub, grid, counts);
    __syncthreads();
  }
}

template <typename T>
__global__ void k_zero(T* g, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    g[i] = (T)0;
}

// The same for a 16-byte aligned array, with 16-byte stores (the values
// past the last whole 16 bytes one by one).
template <typename T>
__global__ void k_zero16(T* g, size_t n) {
  const size_t nv = n * sizeof(T) / 16;
  const size_t t0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  uint4* v = reinterpret_cast<uint4*>(g);
  for (size_t i = t0; i < nv; i += st) v[i] = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = nv * 16 / sizeof(T) + t0; i < n; i += st) g[i] = (T)0;
}

// grid (blockIdx.y) of ngrids: delta = rho / (M / ng) - 1
template <typename T>
__global__ void k_normalise(T* grid, size_t ng, int nsub, const unsigned long long* counts) {
  const double M = (double)counts[blockIdx.y / nsub];
  const double mean = M / (double)ng;
  T* g = grid + (size_t)blockIdx.y * ng;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ng;
       i += (size_t)gridDim.x * blockDim.x)
    g[i] = (T)((double)g[i] / mean - 1.0);
}

// The same with 16-byte loads and stores, for 16-byte aligned grids whose
// size is a multiple of 16 bytes; each value's expression is unchanged.
template <typename T>
__global__ void k_normalise16(T* grid, size_t ng, int nsub, const unsigned long long* counts) {
  constexpr int W = 16 / (int)sizeof(T);
  const double M = (double)counts[blockIdx.y / nsub];
  const double mean = M / (double)ng;
  uint4* g = reinterpret_cast<uint4*>(grid + (size_t)blockIdx.y * ng);
  const size_t nv = ng / W;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nv;
       i += (size_t)gridDim.x * blockDim.x) {
    uint4 u = g[i];
    T t[W];
    __builtin_memcpy(t, &u, 16);
#pragma unroll
    for (int e = 0; e < W; e++) t[e] = (T)((double)t[e] / mean - 1.0);
    __builtin_memcpy(&u, t, 16);
    g[i] = u;
  }
}

// grid (blockIdx.y): delta = rho / (W / ng) - 1, W the deposited weight
template <typename T>
__global__ void k_normalise_w(T* grid, size_t ng, const double* wsum) {
  const double mean = wsum[blockIdx.y] / (double)ng;
  T* g = grid + (size_t)blockIdx.y * ng;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ng;
       i += (size_t)gridDim.x * blockDim.x)
    g[i] = (T)((double)g[i] / mean - 1.0);
}

// Side of the particle lattice when np is a perfect cube, else 0.
static int lattice_side(long long np) {
  if (np <= 0) return 0;
  long long n = (long long)std::llround(std::cbrt((double)np));
  for (long long m = n > 1 ? n - 1 : 1; m <= n + 1; m++)
    if (m * m * m == np) return (int)m;
  return 0;
}

// LDS bytes of the tiled deposit's Eulerian box: EXH_DEPOSIT_LDS_KB (0
// disables the tiled path), default 144 KB of the 160 KB per workgroup.
static int deposit_lds_bytes() {
  static int v = [] {
    const char* e = std::getenv("EXH_DEPOSIT_LDS_KB");
    int kb = e ? std::atoi(e) : 144;
    if (kb < 0) kb = 0;
    if (kb > 156) kb = 156;
    return kb * 1024;
  }();
  return v;
}

// The window deposit for lattice-ordered particles (EXH_DEPOSIT_WINDOW=0
// keeps the extent-sized boxes of k_deposit_tiles).
static bool deposit_window_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("EXH_DEPOSIT_WINDOW");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

// Workgroups of the window deposit (EXH_DEPOSIT_WG; 0 = one per tile).
static int deposit_workgroups() {
  static const int v = [] {
    const char* e = std::getenv("EXH_DEPOSIT_WG");
    const int n = e ? std::atoi(e) : 0;
    return n < 0 ? 0 : n;
  }();
  return v;
}

// Window margin h in cells (EXH_DEPOSIT_MARGIN, 1..8, default WIN_HALO).
static int deposit_window_margin() {
  static const int v = [] {
    const char* e = std::getenv("EXH_DEPOSIT_MARGIN");
    const int h = e ? std::atoi(e) : WIN_HALO;
    return h < 1 ? 1 : (h > 8 ? 8 : h);
  }();
  return v;
}

// Scratch of the allocation-free deposit (exh_grid_compute_noalloc): the
// per-type particle counts (64 words) or the per-grid deposited weights
// (128 doubles) that the normalisation reads.  Each device holds its own
// copy; host code takes its address once per device from
// hipGetSymbolAddress and never uses the array's name.  Calls that
// overlap on one device share it and race.
static __device__ __align__(256) unsigned long long scratch[NOALLOC_GRID_WORDS];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

void* noalloc_grid_scratch(int device) {
  EXH_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, EXH_ERR_ARG,
              "device %d out of range for the allocation-free entry point", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    EXH_CHECK_HIP(hipGetDevice(&cur));
    EXH_REQUIRE(cur == device, EXH_ERR_INTERNAL, "device %d is not current", device);
    void* p = nullptr;
    EXH_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EXH_REQUIRE(p != nullptr, EXH_ERR_HIP, "no address for the deposit scratch on device %d",
                device);
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

// Tile shape of the window deposit (EXH_DEPOSIT_TILE: 16 = 16^3 particles,
// the default; 3216 = 32 x 16 x 16; 32 = 32 x 32 x 16) and its
// accumulator (EXH_DEPOSIT_ACC: 1 fixed point, the default for float grids
// without masses; 8 double; 4 float).  At nd = nL = 512 (CIC, interlaced),
// alone: 3.18 ms for 32 x 16 x 16 in fixed point, 3.06 ms for 32 x 32 x 16,
// 3.63 ms for 16^3, 3.89 ms for 16^3 in double, 12.6 ms for 16^3 in float
// (ds_add_f32 runs far slower than ds_add_u32 or ds_add_f64 here).  Beside
// the halo finder's sort the 32 x 32 x 16 workgroups (153 KB of LDS each)
// hold CUs long enough to slow the sort by 2.6 ms; the 16^3 workgroups
// (55 KB, two per CU) leave it most room: the fused nd = 512 step runs
// 26.85-26.99 ms with them against 27.21-27.27 ms with 32 x 16 x 16.
struct WinShape {
  int px, py, pz;
};
static WinShape tile_shape(int t) {
  if (t == 32) return WinShape{32, 32, 16};
  if (t == 3216) return WinShape{32, 16, 16};
  return WinShape{16, 16, 16};
}
static WinShape deposit_window_shape() {
  static const WinShape v = [] {
    const char* e = std::getenv("EXH_DEPOSIT_TILE");
    return tile_shape(e ? std::atoi(e) : 16);
  }();
  return v;
}
static int deposit_window_acc() {
  static const int v = [] {
    const char* e = std::getenv("EXH_DEPOSIT_ACC");
    const int a = e ? std::atoi(e) : 1;
    return (a == 4 || a == 1) ? a : 8;
  }();
  return v;
}
// LDS of the window (the kernel's static shared memory takes < 1 KB of the
// 160 KB)
static constexpr int WIN_LDS_MAX = 159 * 1024;

template <typename T>
struct DepositWinArgs {
  const T *pos, *vel, *mass;
  int nL, nd, Wx, Wy, Wz, wh;
  T L, Lb;
  int direction;
  T rsd;
  int nsub;
  T* grid;
  unsigned long long* counts;
  int ntiles;
};

template <typename T, int WN, typename ACC, int PX, int PY, int PZ>
static void launch_window(exh_ctx* c, unsigned nwg, int lds, const DepositWinArgs<T>& a) {
  auto k = k_deposit_window<T, WN, ACC, PX, PY, PZ>;
  EXH_CHECK_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    lds));
  hipLaunchKernelGGL(k, dim3(nwg), dim3(WIN_THREADS), lds, c->stream, a.pos, a.vel, a.mass, a.nL,
                     a.nd, a.Wx, a.Wy, a.Wz, a.wh, a.L, a.Lb, a.direction, a.rsd, a.nsub, a.grid,
                     a.counts, a.ntiles);
}

template <typename T, int WN>
static void launch_window_shape(exh_ctx* c, int acc, const WinShape& ws, unsigned nwg, int lds,
                                const DepositWinArgs<T>& a) {
  if (ws.px == 32 && ws.py == 32) {
    if (acc == 1) launch_window<T, WN, unsigned, 32, 32, 16>(c, nwg, lds, a);
    else if (acc == 4) launch_window<T, WN, float, 32, 32, 16>(c, nwg, lds, a);
    else launch_window<T, WN, double, 32, 32, 16>(c, nwg, lds, a);
  } else if (ws.px == 32) {
    if (acc == 1) launch_window<T, WN, unsigned, 32, 16, 16>(c, nwg, lds, a);
    else if (acc == 4) launch_window<T, WN, float, 32, 16, 16>(c, nwg, lds, a);
    else launch_window<T, WN, double, 32, 16, 16>(c, nwg, lds, a);
  } else {
    if (acc == 1) launch_window<T, WN, unsigned, 16, 16, 16>(c, nwg, lds, a);
    else if (acc == 4) launch_window<T, WN, float, 16, 16, 16>(c, nwg, lds, a);
    else launch_window<T, WN, double, 16, 16, 16>(c, nwg, lds, a);
  }
}

template <typename T>
static void launch_window_any(exh_ctx* c, int window, int acc, const WinShape& ws, unsigned nwg,
                              int lds, const DepositWinArgs<T>& a) {
  if (window == 1) launch_window_shape<T, 1>(c, acc, ws, nwg, lds, a);
  if (window == 2) launch_window_shape<T, 2>(c, acc, ws, nwg, lds, a);
  if (window == 5) launch_window_shape<T, 5>(c, acc, ws, nwg, lds, a);
}

template <typename T>
void run_grid(exh_ctx* c, const T* pos, const T* vel, const T* mass, const int* type,
              long long np, int ntype, int nd, double L, double Om0, double z, int direction,
              int window, double R, double R_times, int interlacing, int folds, T* grid,
              void* scratch_dev) {
  EXH_REQUIRE(window >= EXH_WINDOW_NGP && window <= EXH_WINDOW_TSC, EXH_ERR_UNSUPPORTED,
              "window %d: NGP (1), CIC (2), sphere (3), exponential (4) or TSC (5)", window);
  EXH_REQUIRE(window < 3 || window > 4 || (R > 0 && (window == 3 || R_times > 0)), EXH_ERR_ARG,
              "the sphere and exponential windows need R > 0 (and R_times > 0)");
  EXH_REQUIRE(nd > 0 && ntype >= 1 && folds >= 1, EXH_ERR_ARG, "bad grid arguments");
  EXH_REQUIRE(direction < 0 || vel != nullptr, EXH_ERR_ARG, "redshift space needs velocities");
  const int nsub = interlacing ? 2 : 1;
  const size_t ng = (size_t)nd * nd * nd;
  T Lf = (T)((T)L / (T)folds);
  T Lb = Lf / (T)nd;
  // reference: vel * (1+z)^3 / H(Om0, z) (gridmodule.c:167-169)
  T Hz = (T)(100.0 * std::sqrt((T)Om0 * std::pow(1.0 + (T)z, 3) + (1.0 - (T)Om0)));
  T rsd = (T)(std::pow(1.0 + (T)z, 3.0) / (double)Hz);
  unsigned long long* counts = scratch_dev ? (unsigned long long*)scratch_dev
                                           : (unsigned long long*)ctx_buf(c, BUF_GCNT, 64 * 8);
  EXH_REQUIRE(ntype <= 64, EXH_ERR_ARG, "at most 64 tracer types");
  {
    TimedScope ts(c, "grid_zero");
    const size_t nz = ng * ntype * nsub;
    if (reinterpret_cast<uintptr_t>(grid) % 16 == 0)
      hipLaunchKernelGGL(k_zero16<T>, dim3(grid_for(nz * sizeof(T) / 16 + 1, 256 * 4)),
                         dim3(256), 0, c->stream, grid, nz);
    else
      hipLaunchKernelGGL(k_zero<T>, dim3(grid_for(nz, 256 * 8)), dim3(256), 0, c->stream, grid,
                         nz);
    hipLaunchKernelGGL(k_zero<unsigned long long>, dim3(1), dim3(64), 0, c->stream, counts,
                       (size_t)64);
    EXH_LAUNCH_CHECK();
  }
  unsigned gsz = grid_for(np, 256);
  if (gsz > 65536 * 4) gsz = 65536 * 4;
  // particles on a lattice whose side is a multiple of the tile side, with
  // a grid side divisible by the number of tiles per axis, are deposited
  // tile by tile through LDS (correct for any order; fast when the array is
  // in lattice order, as LPT particles are)
  if (window == 3 || window == 4) {
    // spherical / exponential windows: normalised by the deposited weight
    double* wsum = scratch_dev ? (double*)scratch_dev : (double*)ctx_buf(c, BUF_GCNT, 128 * 8);
    hipLaunchKernelGGL(k_zero<double>, dim3(1), dim3(128), 0, c->stream, wsum, (size_t)128);
    {
      TimedScope ts(c, window == 3 ? "sphere_deposit" : "exp_deposit");
      const unsigned gs = std::min(grid_for(np, 256), 65536u);
#define EXH_DEPS(E)                                                                              \
  hipLaunchKernelGGL((k_deposit_sphere<T, E>), dim3(gs), dim3(256), 0, c->stream, pos, vel, mass, \
                     type, np, nd, Lf, Lb, (T)R, (T)R_times, direction, rsd, nsub, ntype * nsub,  \
                     grid, wsum)
      if (window == 3) EXH_DEPS(false);
      if (window == 4) EXH_DEPS(true);
#undef EXH_DEPS
      EXH_LAUNCH_CHECK();
    }
    if (!mass) {
      TimedScope ts(c, "grid_normalise");
      hipLaunchKernelGGL(k_normalise_w<T>, dim3(grid_for(ng, 256 * 8), ntype * nsub), dim3(256),
                         0, c->stream, grid, ng, wsum);
      EXH_LAUNCH_CHECK();
    }
    return;
  }
  const int nL = lattice_side(np);
  const int nt = nL / TILE_L;
  const bool tiles = type == nullptr && nL >= TILE_L && nL % TILE_L == 0 && nd % nt == 0 &&
                     deposit_lds_bytes() > 0;
  {
    TimedScope ts(c, window == 2 ? "cic_deposit" : (window == 1 ? "ngp_deposit" : "tsc_deposit"));
    const int W = tiles ? nd / nt : 0;
    const int wh = deposit_window_margin();
    // the window's tile shape and accumulator (EXH_DEPOSIT_TILE,
    // EXH_DEPOSIT_ACC), where the lattice and the grid divide into it and
    // the window fits the LDS; nd >= E keeps every window cell within
    // (-nd, 2nd), where the flush's single-step wrap is exact
    // (float windows for float grids only; the default shape when the
    // chosen one does not fit)
    WinShape ws = c->tune_deposit_tile ? tile_shape(c->tune_deposit_tile) : deposit_window_shape();
    int acc = sizeof(T) == 8 ? 8 : (c->tune_deposit_acc ? c->tune_deposit_acc : deposit_window_acc());
    if (acc == 1 && mass) acc = 8;  // fixed point holds weights in [0, 1] only
    int Wq[3] = {0, 0, 0}, win_lds = 0;
    auto fits = [&](const WinShape& w, int a) {
      if (!(tiles && deposit_window_enabled() && nL % w.px == 0 && nL % w.py == 0 &&
            nL % w.pz == 0))
        return false;
      const int pq[3] = {w.px, w.py, w.pz};
      long long cells = 1;
      for (int q = 0; q < 3; q++) {
        const int ntq = nL / pq[q];
        if (nd % ntq != 0) return false;
        Wq[q] = nd / ntq;
        if (nd < Wq[q] + 2 * wh) return false;
        cells *= Wq[q] + 2 * wh;
      }
      win_lds = (int)std::min<long long>(cells * (a == 1 ? 4 : a), 1 << 30);
      return win_lds <= WIN_LDS_MAX;
    };
    bool win = fits(ws, acc);
    if (!win && (ws.px != 16 || ws.py != 16 || ws.pz != 16)) {
      ws = WinShape{16, 16, 16};
      win = fits(ws, acc);
    }
    if (!win && acc != 8) {
      acc = 8;
      win = fits(ws, acc);
    }
    if (win) {
      const unsigned nblk = (unsigned)((nL / ws.px) * (nL / ws.py) * (nL / ws.pz));
      const unsigned nwg = deposit_workgroups() > 0 ? std::min(nblk, (unsigned)deposit_workgroups())
                                                    : nblk;
      const DepositWinArgs<T> da{pos, vel, mass, nL, nd, Wq[0], Wq[1], Wq[2], wh, Lf, Lb,
                                 direction, rsd, nsub, grid, counts, (int)nblk};
      launch_window_any<T>(c, window, acc, ws, nwg, win_lds, da);
    } else if (tiles) {
      const int lds = deposit_lds_bytes();
      const int cap = lds / (int)sizeof(double);
      const unsigned nblk = (unsigned)nt * nt * nt;
#define EXH_DEPT(WN)                                                                           \
  do {                                                                                         \
    EXH_CHECK_HIP(hipFuncSetAttribute((const void*)k_deposit_tiles<T, WN>,                     \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, lds));       \
    hipLaunchKernelGGL((k_deposit_tiles<T, WN>), dim3(nblk), dim3(TILE_THREADS), lds,          \
                       c->stream, pos, vel, mass, nL, nd, W, Lf, Lb, direction, rsd, nsub, cap, \
                       grid, counts);                                                          \
  } while (0)
      if (window == 1) EXH_DEPT(1);
      if (window == 2) EXH_DEPT(2);
      if (window == 5) EXH_DEPT(5);
#undef EXH_DEPT
    } else {
#define EXH_DEP(W)                                                                         \
  hipLaunchKernelGGL((k_deposit<T, W>), dim3(gsz), dim3(256), 0, c->stream, pos, vel, mass, \
                     type, np, nd, Lf, Lb, direction, rsd, nsub, grid, counts)
      if (window == 1) EXH_DEP(1);
      if (window == 2) EXH_DEP(2);
      if (window == 5) EXH_DEP(5);
#undef EXH_DEP
    }
    EXH_LAUNCH_CHECK();
  }
  if (!mass) {
    TimedScope ts(c, "grid_normalise");
    if (reinterpret_cast<uintptr_t>(grid) % 16 == 0 && (ng * sizeof(T)) % 16 == 0)
      hipLaunchKernelGGL(k_normalise16<T>, dim3(grid_for(ng * sizeof(T) / 16, 256 * 2),
                                                ntype * nsub),
                         dim3(256), 0, c->stream, grid, ng, nsub, counts);
    else
      hipLaunchKernelGGL(k_normalise<T>, dim3(grid_for(ng, 256 * 8), ntype * nsub), dim3(256),
                         0, c->stream, grid, ng, nsub, counts);
    EXH_LAUNCH_CHECK();
  }
}

// ---------------------------------------------------------------------------
// power spectrum
// ---------------------------------------------------------------------------
struct PowerArgs {
  int nd, nz, ntype, nsub, Nk, ls, NPs, NV, direction, kbin;
  const double2* ex;  // per-axis phase e^{-i k Lb/2}: x/y [nd], z [nz]
  const double2* ez;
  const double* wx;   // per-axis window^-1 factors
  const double* wz;
  double Normk;
  int rwin;         // 3 (sphere) or 4 (exponential): window per mode, else 0
  double knT, RT;   // kn and R in the grid's precision
  // bin index arithmetic
  float knf, kminf, dkf;
  double kn, kmin, dk;
  FastDiv dz, dnd;  // divisions by nz and nd (flat mode index -> i, j, k)
};

__device__ inline double legendre_d(int l, double x) {
  if (l == 0) return 1.0;
  double p0 = 1.0, p1 = x;
  for (int n = 1; n < l; n++) {
    double p2 = ((2 * n + 1) * x * p1 - n * p0) / (n + 1);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

// One DPP move of a double (its two halves) across the wave; lanes without
// a source read 0.  Controls: row_shr:n = 0x110 + n (within rows of 16
// lanes), row_bcast:15 = 0x142 (lane 15 of each row to the next row),
// row_bcast:31 = 0x143 (lane 31 to rows 2 and 3).
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(
      (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo));
}

// Inclusive segmented sum over a full wave: lane l gets x summed over lanes
// sstart .. l (sstart <= l, the start of l's segment), in a fixed order:
// shifts by 1, 2, 4 and 8 within rows of 16 lanes, then the last lane of
// row 0 into row 1 and of row 2 into row 3, then lane 31 into rows 2 and 3,
// each added only where the source lies in the lane's segment.  DPP moves
// run on the vector ALU; the cross-lane shuffles they replace went through
// the LDS crossbar.
__device__ inline double seg_scan_d(double x, int lane, int sstart) {
  const int rl = lane & 15, rs = lane & ~15;
  double y;
  y = dpp_d<0x111>(x);
  if (rl >= 1 && lane - 1 >= sstart) x += y;
  y = dpp_d<0x112>(x);
  if (rl >= 2 && lane - 2 >= sstart) x += y;
  y = dpp_d<0x114>(x);
  if (rl >= 4 && lane - 4 >= sstart) x += y;
  y = dpp_d<0x118>(x);
  if (rl >= 8 && lane - 8 >= sstart) x += y;
  y = dpp_d<0x142>(x);
  if ((rs == 16 || rs == 48) && sstart < rs) x += y;
  y = dpp_d<0x143>(x);
  if (rs >= 32 && sstart < 32) x += y;
  return x;
}

static constexpr int PW_MAXV = 64;  // max values per mode (pairs x multipoles + 1)

// Bin of half-complex mode (i, j, k), or -1: the reference's mode mask
// (powermodule.c:47-55) and bin index floor((|k| - k_min) / dk) in float
// (kbin 4, spectrum_h.c:39-48) or double (kbin 8); |k| and mu along the
// line of sight as the binning evaluates them.
__device__ inline int mode_bin(const PowerArgs& a, int i, int j, int k, double& kmod, double& mu) {
  const int nd = a.nd;
  const bool masked = (k == 0 || k == nd / 2) &&
                      ((i > nd / 2 && j > nd / 2) || (i == 0 || i == nd / 2) ||
                       (j == 0 || j == nd / 2) || (i > nd / 2 && j < nd / 2));
  if (masked) return -1;
  const int fi = (2 * i < nd) ? i : i - nd, fj = (2 * j < nd) ? j : j - nd;
  const int fk = (2 * k < nd) ? k : k - nd;
  int b;
  if (a.kbin == 4) {
    float kx = (float)fi * a.knf, ky = (float)fj * a.knf, kz = (float)fk * a.knf;
    float km = sqrtf(kx * kx + ky * ky + kz * kz);
    float tmp = (km - a.kminf) / a.dkf;
    b = (int)floorf(tmp);
    kmod = km;
    if (a.direction == 0) mu = (double)(kx / km);
    if (a.direction == 1) mu = (double)(ky / km);
    if (a.direction == 2) mu = (double)(kz / km);
  } else {
    double kx = fi * a.kn, ky = fj * a.kn, kz = fk * a.kn;
    double km = sqrt(kx * kx + ky * ky + kz * kz);
    b = (int)floor((km - a.kmin) / a.dk);
    kmod = km;
    if (a.direction == 0) mu = kx / km;
    if (a.direction == 1) mu = ky / km;
    if (a.direction == 2) mu = kz / km;
  }
  return (b < 0 || b >= a.Nk) ? -1 : b;
}

// W of the sphere (3) and exponential (4) windows at mode (fi, fj, fk), in
// the grid's precision with the reference's double intermediates
// (spectrum_h.c:60-67).
template <typename T>
__device__ inline T window_r(const PowerArgs& a, int fi, int fj, int fk) {
  const T kn = (T)a.knT, R = (T)a.RT;
  const T k1 = (T)fi * kn, k2 = (T)fj * kn, k3 = (T)fk * kn;
  if (a.rwin == 3) {
    const T kmod = (T)sqrt((double)(k1 * k1 + k2 * k2 + k3 * k3));
    const T kR = kmod * R;
    const T sc = kR != (T)0 ? (T)(sin((double)kR) / (double)kR) : (T)1;
    return (T)(3.0 / kmod / kmod / R / R * ((double)sc - cos((double)kR)));
  }
  const T kmod = k1 * k1 + k2 * k2 + k3 * k3;
  return (T)exp((double)(-kmod * R * R) / 2.0);
}

// ONE: one tracer and the monopole only (two values per mode: P, |k|), with
// the counts fixed at compile time so the per-tracer arrays stay in
// registers; the arithmetic is the general path's.
template <typename T, bool ONE>
__global__ void k_power_bins(PowerArgs a, const typename cplx<T>::type* gk, double* part,
                             unsigned* pcnt) {
  extern __shared__ double lds[];
  using C2 = typename cplx<T>::type;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int NV = ONE ? 2 : a.NV, Nk = a.Nk, LS = ONE ? 1 : a.ls;
  double* acc = lds + (size_t)wave * Nk * NV;
  unsigned* cacc = (unsigned*)(lds + (size_t)nw * Nk * NV) + (size_t)wave * Nk;
  for (int q = lane; q < Nk * NV; q += 64) acc[q] = 0.0;
  for (int q = lane; q < Nk; q += 64) cacc[q] = 0;
  __syncthreads();
  const int nd = a.nd, nz = a.nz;
  const size_t nh = (size_t)nd * nd * nz;
  // waves walk 64-mode chunks of the flat mode index (a row of nd/2+1
  // modes would leave a chunk with one active lane); bins change at row
  // boundaries, which the segmented scan below handles like any change
  const long long nchunks = (long long)((nh + 63) / 64);
  const long long gw = (long long)blockIdx.x * nw + wave, tw = (long long)gridDim.x * nw;
  for (long long ch = gw; ch < nchunks; ch += tw) {
    {
      const size_t id = (size_t)ch * 64 + lane;
      int i = 0, j = 0, k = nz;  // k = nz: no mode
      long long row = (long long)((nh - 1) / nz);
      if (id < nh) {
        if (nh < (1ull << 31)) {
          const unsigned r = fdiv((unsigned)id, a.dz);
          k = (int)((unsigned)id - r * (unsigned)nz);
          i = (int)fdiv(r, a.dnd);
          j = (int)(r - (unsigned)i * (unsigned)nd);
          row = r;
        } else {
          k = (int)(id % nz);
          const size_t r = id / nz;
          j = (int)(r % nd);
          i = (int)(r / nd);
          row = (long long)r;
        }
      }
      // the chunk's rows: within one row bins never decrease, so a bin has
      // at most one segment per row; tails update the accumulators one row
      // at a time (deterministically, without two lanes on one address)
      const long long row0 = __shfl(row, 0, 64);
      const int part = (int)(row - row0);
      const int nparts = (int)(__shfl(row, 63, 64) - row0) + 1;
      const int fi = (2 * i < nd) ? i : i - nd, fj = (2 * j < nd) ? j : j - nd;
      double kmod = 0, mu = 0;
      const int b = k < nz ? mode_bin(a, i, j, k, kmod, mu) : -1;
      const bool valid = b >= 0;
      // segment bookkeeping
      const int bprev = __shfl_up(b, 1, 64);
      const bool head = valid && (lane == 0 || bprev != b);
      const unsigned long long heads = __ballot(head);
      const unsigned long long upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
      const int sstart = upto ? 63 - __clzll((long long)upto) : 0;
      const int bnext = __shfl_down(b, 1, 64);
      const bool tail = valid && (lane == 63 || bnext != b);
      // combined modes of every tracer
      double dr[8], di[8];
      const int nt = ONE ? 1 : a.ntype;
      if constexpr (ONE && sizeof(T) == 4) {
        // one float tracer: the combine and deconvolution in float (the
        // grid's precision), |delta|^2 summed in double below
        if (valid) {
          const double2 e1d = a.ex[i], e2d = a.ex[j], e3d = a.ez[k];
          const float2 e1 = make_float2((float)e1d.x, (float)e1d.y);
          const float2 e2 = make_float2((float)e2d.x, (float)e2d.y);
          const float2 e3 = make_float2((float)e3d.x, (float)e3d.y);
          const float wk = a.rwin ? (float)(a.Normk / (double)window_r<T>(a, fi, fj,
                                                                           (2 * k < nd) ? k : k - nd))
                                  : (float)a.wx[i] * (float)a.wx[j] * (float)a.wz[k] * (float)a.Normk;
          const float pr = e1.x * e2.x - e1.y * e2.y, pi = e1.x * e2.y + e1.y * e2.x;
          const float er = pr * e3.x - pi * e3.y, ei = pr * e3.y + pi * e3.x;
          const C2 f1 = gk[id];
          float re = (float)f1.x, im = (float)f1.y;
          if (a.nsub == 2) {
            const C2 f2 = gk[nh + id];
            re = (re + ((float)f2.x * er - (float)f2.y * ei)) * 0.5f;
            im = (im + ((float)f2.x * ei + (float)f2.y * er)) * 0.5f;
          }
          dr[0] = (double)(re * wk);
          di[0] = (double)(im * wk);
        }
      } else if (valid) {
        const double2 e1 = a.ex[i], e2 = a.ex[j], e3 = a.ez[k];
        const double wk = a.rwin ? a.Normk / (double)window_r<T>(a, fi, fj, (2 * k < nd) ? k : k - nd)
                                 : a.wx[i] * a.wx[j] * a.wz[k] * a.Normk;
        // e^{-i theta} = e1 e2 e3
        double pr = e1.x * e2.x - e1.y * e2.y, pi = e1.x * e2.y + e1.y * e2.x;
        double er = pr * e3.x - pi * e3.y, ei = pr * e3.y + pi * e3.x;
        for (int t = 0; t < nt; t++) {
          const C2 f1 = gk[(size_t)(t * a.nsub) * nh + id];
          double re = (double)f1.x, im = (double)f1.y;
          if (a.nsub == 2) {
            const C2 f2 = gk[(size_t)(t * a.nsub + 1) * nh + id];
            re = (re + ((double)f2.x * er - (double)f2.y * ei)) / 2.0;
            im = (im + ((double)f2.x * ei + (double)f2.y * er)) / 2.0;
          }
          dr[t] = re * wk;
          di[t] = im * wk;
        }
      }
      // values: pair products x P_2l(mu), then |k|
      for (int v = 0; v < NV; v++) {
        double x = 0.0;
        if (valid) {
          if (v == NV - 1) {
            x = kmod;
          } else {
            int pair = v / LS, l2 = 2 * (v % LS);
            // pair index -> (l, m <= l) in the reference's order
            int l = 0;
            while ((l + 1) * (l + 2) / 2 <= pair) l++;
            int m = pair - l * (l + 1) / 2;
            x = (dr[l] * dr[m] + di[l] * di[m]) * legendre_d(l2, mu);
          }
        }
        x = seg_scan_d(x, lane, sstart);
        for (int p = 0; p < nparts; p++)
          if (tail && part == p) acc[(size_t)b * NV + v] += x;
      }
      for (int p = 0; p < nparts; p++)
        if (tail && part == p) cacc[b] += (unsigned)(lane - sstart + 1);
    }
  }
  __syncthreads();
  // fixed-order reduction over waves, one partial per block
  double* out = part + (size_t)blockIdx.x * Nk * NV;
  unsigned* cout = pcnt + (size_t)blockIdx.x * Nk;
  for (int q = threadIdx.x; q < Nk * NV; q += blockDim.x) {
    double s = 0.0;
    for (int w = 0; w < nw; w++) s += lds[(size_t)w * Nk * NV + q];
    out[q] = s;
  }
  for (int q = threadIdx.x; q < Nk; q += blockDim.x) {
    unsigned s = 0;
    unsigned* cbase = (unsigned*)(lds + (size_t)nw * Nk * NV);
    for (int w = 0; w < nw; w++) s += cbase[(size_t)w * Nk + q];
    cout[q] = s;
  }
}

// One block per output value: 256 threads stride over the block partials,
// then a fixed-shape tree; the result does not depend on timing.
__global__ void k_power_reduce(const double* part, const unsigned* pcnt, int nblocks, int Nk,
                               int NV, double* sum, unsigned long long* cnt) {
  __shared__ double sd[256];
  __shared__ unsigned long long sc[256];
  const int q = blockIdx.x;
  const bool is_cnt = q >= Nk * NV;
  double s = 0.0;
  unsigned long long u = 0;
  for (int bl = threadIdx.x; bl < nblocks; bl += 256) {
    if (is_cnt)
      u += pcnt[(size_t)bl * Nk + (q - Nk * NV)];
    else
      s += part[(size_t)bl * Nk * NV + q];
  }
  sd[threadIdx.x] = s;
  sc[threadIdx.x] = u;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sd[threadIdx.x] += sd[threadIdx.x + w];
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (is_cnt)
      cnt[q - Nk * NV] = sc[0];
    else
      sum[q] = sd[0];
  }
}

template <typename T>
static double sinc_t(T x) {
  T r;
  if (x != (T)0)
    r = (T)(std::sin((double)x) / (double)x);
  else
    r = (T)1;
  return r;
}

// Binning arguments of the P(k) passes: per-axis tables with the
// reference's fft_real rounding of kn and Lb (uploaded to the context's
// table slot), window, Normk and the bin-edge arithmetic.
template <typename T>
static PowerArgs make_power_args(exh_ctx* c, int ntype, int nd, double L, int window, double R,
                                 int interlacing, int Nk, double k_min, double k_max, int l_max,
                                 int direction, int kbin_precision) {
  const int nsub = interlacing ? 2 : 1;
  const int nz = nd / 2 + 1;
  const int NPs = ntype * (ntype + 1) / 2, ls = l_max / 2 + 1;
  const int NV = NPs * ls + 1;

  // per-axis tables with the reference's fft_real rounding of kn and Lb
  T Lr = (T)L;
  T kn = (T)(2 * M_PI / (double)Lr);
  T Lb = Lr / (T)nd;
  T Normk = (T)std::pow((double)(Lr / (T)(nd * nd)), 3.0 / 2.0);
  // NGP, CIC, TSC: per-axis sinc powers; the sphere and exponential windows
  // are functions of |k|, evaluated per mode in the kernel
  int pw = (window == 0 || window == 3 || window == 4) ? 0 : (window == 1 ? 1 : (window == 2 ? 2 : 3));
  std::vector<double2> ex(nd), ez(nz);
  std::vector<double> wx(nd), wz(nz);
  for (int i = 0; i < nd; i++) {
    T kx = (2 * i < nd) ? (T)i * kn : (T)(i - nd) * kn;
    double th = (double)kx * (double)Lb / 2.0;
    ex[i] = make_double2(std::cos(th), -std::sin(th));
    wx[i] = 1.0 / std::pow(sinc_t<T>((T)((double)kx * (double)Lb / 2.0)), pw);
  }
  for (int k = 0; k < nz; k++) {
    T kz = (2 * k < nd) ? (T)k * kn : (T)(k - nd) * kn;
    double th = (double)kz * (double)Lb / 2.0;
    ez[k] = make_double2(std::cos(th), -std::sin(th));
    wz[k] = 1.0 / std::pow(sinc_t<T>((T)((double)kz * (double)Lb / 2.0)), pw);
  }
  size_t tb = (2 * (size_t)nd + 2 * nz) * sizeof(double2);
  char* tab = (char*)ctx_buf(c, BUF_PTAB, tb);
  double2* dex = (double2*)tab;
  double2* dez = dex + nd;
  double* dwx = (double*)(dez + nz);
  double* dwz = dwx + nd;
  EXH_CHECK_HIP(hipMemcpyAsync(dex, ex.data(), nd * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(dez, ez.data(), nz * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(dwx, wx.data(), nd * sizeof(double), hipMemcpyHostToDevice, c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(dwz, wz.data(), nz * sizeof(double), hipMemcpyHostToDevice, c->stream));

  PowerArgs a{};
  a.nd = nd;
  a.nz = nz;
  a.ntype = ntype;
  a.nsub = nsub;
  a.Nk = Nk;
  a.ls = ls;
  a.NPs = NPs;
  a.NV = NV;
  a.direction = direction;
  a.kbin = kbin_precision;
  a.dz = fast_div((unsigned)nz);
  a.dnd = fast_div((unsigned)nd);
  a.ex = dex;
  a.ez = dez;
  a.wx = dwx;
  a.wz = dwz;
  a.Normk = (double)Normk;
  a.rwin = (window == 3 || window == 4) ? window : 0;
  a.knT = (double)kn;
  a.RT = (double)(T)R;
  // bin edges: fft_real arithmetic for kbin 4 (spectrum_h.c:39-48, powermodule.c:13)
  a.knf = (float)(2 * M_PI / (double)(float)L);
  a.kminf = (float)k_min;
  a.dkf = ((float)k_max - (float)k_min) / (float)Nk;
  a.kn = 2 * M_PI / L;
  a.kmin = k_min;
  a.dk = (k_max - k_min) / Nk;
  if (sizeof(T) == 4) {
    // the FP32 build rounds kn to float before use
    a.kn = (double)(float)a.kn;
  }
  return a;
}

template <typename T>
void run_power(exh_ctx* c, const T* grid, int ntype, int nd, double L, int window, double R,
               int interlacing, int Nk, double k_min, double k_max, int l_max, int direction,
               int kbin_precision, double* k_out, double* P_out, long long* count_out,
               bool deferred) {
  using C2 = typename cplx<T>::type;
  EXH_REQUIRE(nd >= 2 && nd % 2 == 0, EXH_ERR_ARG, "nd must be even");
  EXH_REQUIRE(ntype >= 1 && ntype <= 8, EXH_ERR_UNSUPPORTED, "1 to 8 tracer types supported");
  EXH_REQUIRE(Nk >= 1 && l_max >= 0, EXH_ERR_ARG, "bad binning");
  EXH_REQUIRE(window >= 0 && window <= 5, EXH_ERR_UNSUPPORTED, "window %d cannot be deconvolved",
              window);
  const int nsub = interlacing ? 2 : 1;
  const int nz = nd / 2 + 1;
  const size_t nh = (size_t)nd * nd * nz;
  const int NPs = ntype * (ntype + 1) / 2, ls = l_max / 2 + 1;
  const int NV = NPs * ls + 1;
  EXH_REQUIRE(NV <= PW_MAXV, EXH_ERR_UNSUPPORTED, "too many spectra x multipoles");
  C2* gk = (C2*)ctx_buf(c, BUF_GRIDK, nh * ntype * nsub * sizeof(C2));
  fft_r2c(c, nd, nd, nd, sizeof(T) == 4 ? 4 : 8, ntype * nsub, grid, gk);
  PowerArgs a = make_power_args<T>(c, ntype, nd, L, window, R, interlacing, Nk, k_min, k_max,
                                   l_max, direction, kbin_precision);

  // LDS: per wave Nk*NV doubles + Nk counters
  size_t per_wave = (size_t)Nk * NV * 8 + (size_t)Nk * 4;
  int waves = 4;
  while (waves > 1 && per_wave * waves > 60 * 1024) waves--;
  EXH_REQUIRE(per_wave <= 60 * 1024, EXH_ERR_UNSUPPORTED,
              "Nk x spectra x multipoles too large for the LDS accumulator (%zu bytes)", per_wave);
  long long nrows = (long long)nd * nd;
  int nblocks = (int)std::min<long long>((nrows + waves - 1) / waves, 2048);
  size_t poff = 0;
  auto ptake = [&](size_t bytes) {
    size_t r = poff;
    poff += (bytes + 255) & ~(size_t)255;
    return r;
  };
  size_t o_part = ptake((size_t)nblocks * Nk * NV * 8), o_pcnt = ptake((size_t)nblocks * Nk * 4),
         o_sum = ptake((size_t)Nk * NV * 8), o_cnt = ptake((size_t)Nk * 8);
  char* pbuf = (char*)ctx_buf(c, BUF_PART, poff);
  double* part = (double*)(pbuf + o_part);
  unsigned* pcnt = (unsigned*)(pbuf + o_pcnt);
  double* sum = (double*)(pbuf + o_sum);
  unsigned long long* cnt = (unsigned long long*)(pbuf + o_cnt);
  {
    TimedScope ts(c, "power_bins");
    if (NV == 2)
      hipLaunchKernelGGL((k_power_bins<T, true>), dim3(nblocks), dim3(64 * waves),
                         per_wave * waves, c->stream, a, gk, part, pcnt);
    else
      hipLaunchKernelGGL((k_power_bins<T, false>), dim3(nblocks), dim3(64 * waves),
                         per_wave * waves, c->stream, a, gk, part, pcnt);
    EXH_LAUNCH_CHECK();
  }
  {
    TimedScope ts(c, "power_reduce");
    hipLaunchKernelGGL(k_power_reduce, dim3((unsigned)(Nk * NV + Nk)), dim3(256), 0,
                       c->stream, part, pcnt, nblocks, Nk, NV, sum, cnt);
    EXH_LAUNCH_CHECK();
  }
  c->power_pending.o_sum = o_sum;
  c->power_pending.o_cnt = o_cnt;
  c->power_pending.Nk = Nk;
  c->power_pending.NV = NV;
  c->power_pending.NPs = NPs;
  c->power_pending.ls = ls;
  if (!deferred) power_copy_out(c, k_out, P_out, count_out);
}

// Copies the sums of the last run_power to the host (waits for the stream)
// and forms k, P and the mode counts per bin.
void power_copy_out(exh_ctx* c, double* k_out, double* P_out, long long* count_out) {
  const auto& pp = c->power_pending;
  const int Nk = pp.Nk, NV = pp.NV, NPs = pp.NPs, ls = pp.ls;
  char* pbuf = (char*)c->buf[BUF_PART];
  const double* sum = (const double*)(pbuf + pp.o_sum);
  const unsigned long long* cnt = (const unsigned long long*)(pbuf + pp.o_cnt);
  size_t hbytes = (size_t)Nk * NV * 8 + (size_t)Nk * 8;
  char* h = (char*)ctx_pinned(c, hbytes);
  EXH_CHECK_HIP(hipMemcpyAsync(h, sum, (size_t)Nk * NV * 8, hipMemcpyDeviceToHost, c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(h + (size_t)Nk * NV * 8, cnt, (size_t)Nk * 8, hipMemcpyDeviceToHost,
                               c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  const double* hs = (const double*)h;
  const unsigned long long* hc = (const unsigned long long*)(h + (size_t)Nk * NV * 8);
  for (int b = 0; b < Nk; b++) {
    long long n = (long long)hc[b];
    if (count_out) count_out[b] = n;
    if (k_out) k_out[b] = n > 0 ? hs[(size_t)b * NV + NV - 1] / n : 0.0;
    for (int p = 0; p < NPs; p++)
      for (int l = 0; l < ls; l++) {
        double s = hs[(size_t)b * NV + p * ls + l];
        if (P_out) P_out[((size_t)p * ls + l) * Nk + b] = n > 0 ? (4.0 * l + 1.0) * s / n : 0.0;
      }
  }
}

// ---------------------------------------------------------------------------
// adjoints of the monopole P(k) and of the mass assignment
// ---------------------------------------------------------------------------
// In place over the modes F = r2c(delta) of one grid (no interlacing):
// G = (dL/dF) / w with dL/dF = gPn[b] wk^2 F for every mode the forward
// pass bins (same mask and bin index), 0 elsewhere.  With
// P_b = sum_b |F wk|^2 / N_b, gPn[b] = 2 dL/dP_b / N_b.  w is the weight
// of a stored mode in the c2r (2 for 0 < kz < nd/2, 1 in the kz = 0, nd/2
// planes), so that c2r(G) is the gradient with respect to the real grid:
// the transpose of the r2c.
template <typename T>
__global__ void k_power_adjoint(PowerArgs a, const double* gPn, typename cplx<T>::type* gk) {
  using C2 = typename cplx<T>::type;
  const int nd = a.nd, nz = a.nz;
  const size_t nh = (size_t)nd * nd * nz;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    if (nh < (1ull << 31)) {
      const unsigned r = fdiv((unsigned)id, a.dz);
      k = (int)((unsigned)id - r * (unsigned)nz);
      i = (int)fdiv(r, a.dnd);
      j = (int)(r - (unsigned)i * (unsigned)nd);
    } else {
      k = (int)(id % nz);
      const size_t r = id / nz;
      j = (int)(r % nd);
      i = (int)(r / nd);
    }
    double kmod, mu;
    const int b = mode_bin(a, i, j, k, kmod, mu);
    C2 out{(T)0, (T)0};
    if (b >= 0) {
      const int fi = (2 * i < nd) ? i : i - nd, fj = (2 * j < nd) ? j : j - nd;
      const double wk = a.rwin ? a.Normk / (double)window_r<T>(a, fi, fj, (2 * k < nd) ? k : k - nd)
                               : a.wx[i] * a.wx[j] * a.wz[k] * a.Normk;
      const double w = (k == 0 || 2 * k == nd) ? 1.0 : 2.0;
      const double f = gPn[b] * wk * wk / w;
      const C2 F = gk[id];
      out = C2{(T)(f * (double)F.x), (T)(f * (double)F.y)};
    }
    gk[id] = out;
  }
}

// CIC weights of one axis and their derivatives with respect to the
// coordinate, for the cell xt and its neighbour o toward the particle
// (dep_cic's arithmetic).
template <typename T>
__device__ inline void cic_axis(T x, T Ld, int nd, int& xt, int& o, T& t, T& d, double& dt) {
  T q = x / Ld;
  xt = (int)floor(q);
  if (xt == nd) xt -= 1;
  T dx = (T)((double)q - ((double)xt + 0.5));
  double sgn = 1.0;
  if (dx >= (T)0) {
    o = xt + 1 == nd ? 0 : xt + 1;
  } else {
    o = xt == 0 ? nd - 1 : xt - 1;
    dx = -dx;
    sgn = -1.0;
  }
  t = (T)(1.0 - (double)dx);
  d = dx;
  dt = -sgn / (double)Ld;  // d t / dx; d d / dx = -dt
}

// dL/dS of every particle of the nL^3 lattice (the LPT particles, lattice
// order) from g = dL/drho on the grid: the transpose of the deposit, the
// weights' derivatives at the particle times g at the cells it feeds, times
// `scale` (1 / rho_bar for normalised grids).  The positions are recomputed
// from the displacements, q + S with q = (i + 1/2) Lc, in double: for a
// float field this places each particle relative to the weights' kinks (the
// cell centres) to the precision of S rather than to that of a float
// coordinate up to L, so the gradient is that of the exact positions.
template <typename T, int WIN>
__global__ void k_deposit_adjoint(const T* S, int nL, double Lc, int nd, double L, const T* g,
                                  double scale, T* gpos) {
  using D = double;
  const D Ld = L / (D)nd;
  const size_t n2 = (size_t)nd * nd;
  const long long np = (long long)nL * nL * nL;
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < np;
       p += (long long)gridDim.x * blockDim.x) {
    const long long r = p / nL;
    const int q[3] = {(int)(r / nL), (int)(r % nL), (int)(p % nL)};
    D x[3];
    bool ok = true;
    for (int a = 0; a < 3; a++) {
      D v = ((D)q[a] * Lc + Lc / 2.0) + (D)S[3 * p + a];
      x[a] = wrap_while<D>(v, L);
      ok = ok && x[a] >= 0.0;
    }
    double gx[3] = {0.0, 0.0, 0.0};
    if (ok && WIN == 2) {
      int xt[3], o[3];
      D t[3], d[3];
      double dt[3];
      for (int a = 0; a < 3; a++) cic_axis<D>(x[a], Ld, nd, xt[a], o[a], t[a], d[a], dt[a]);
      // the eight cells: c[ax] selects xt (0) or o (1) per axis
      for (int c = 0; c < 8; c++) {
        const int s0 = c & 1, s1 = (c >> 1) & 1, s2 = (c >> 2) & 1;
        const size_t id = (size_t)(s0 ? o[0] : xt[0]) * n2 + (size_t)(s1 ? o[1] : xt[1]) * nd +
                          (size_t)(s2 ? o[2] : xt[2]);
        const double gv = (double)g[id];
        const double w0 = s0 ? d[0] : t[0], w1 = s1 ? d[1] : t[1], w2 = s2 ? d[2] : t[2];
        const double e0 = s0 ? -dt[0] : dt[0], e1 = s1 ? -dt[1] : dt[1], e2 = s2 ? -dt[2] : dt[2];
        gx[0] += gv * e0 * w1 * w2;
        gx[1] += gv * w0 * e1 * w2;
        gx[2] += gv * w0 * w1 * e2;
      }
    } else if (ok && WIN == 5) {
      int cc[3];
      double w[3][3], dw[3][3];
      for (int a = 0; a < 3; a++) {
        const double u = x[a] / Ld - 0.5;
        const double cf = floor(u + 0.5);
        const double dd = u - cf;
        cc[a] = (int)cf;
        w[a][0] = 0.5 * (0.5 - dd) * (0.5 - dd);
        w[a][1] = 0.75 - dd * dd;
        w[a][2] = 0.5 * (0.5 + dd) * (0.5 + dd);
        const double du = 1.0 / (double)Ld;
        dw[a][0] = -(0.5 - dd) * du;
        dw[a][1] = -2.0 * dd * du;
        dw[a][2] = (0.5 + dd) * du;
      }
      for (int ia = 0; ia < 3; ia++) {
        int ix = cc[0] + ia - 1;
        ix = ix < 0 ? ix + nd : (ix >= nd ? ix - nd : ix);
        for (int ib = 0; ib < 3; ib++) {
          int iy = cc[1] + ib - 1;
          iy = iy < 0 ? iy + nd : (iy >= nd ? iy - nd : iy);
          for (int ie = 0; ie < 3; ie++) {
            int iz = cc[2] + ie - 1;
            iz = iz < 0 ? iz + nd : (iz >= nd ? iz - nd : iz);
            const double gv = (double)g[(size_t)ix * n2 + (size_t)iy * nd + iz];
            gx[0] += gv * dw[0][ia] * w[1][ib] * w[2][ie];
            gx[1] += gv * w[0][ia] * dw[1][ib] * w[2][ie];
            gx[2] += gv * w[0][ia] * w[1][ib] * dw[2][ie];
          }
        }
      }
    }
    for (int a = 0; a < 3; a++) gpos[3 * p + a] = (T)(gx[a] * scale);
  }
}

// Gradient of the monopole P(k) of `grid` with respect to the grid values:
// g = c2r(G) with G from k_power_adjoint; gP: host [Nk] dL/dP_b.  Runs
// the forward binning first (the counts N_b, and P_out / count_out when
// given).  The modes are left in the context's BUF_GRIDK and overwritten.
template <typename T>
void run_power_adjoint(exh_ctx* c, const T* grid, int nd, double L, int window, double R,
                       int Nk, double k_min, double k_max, int kbin_precision, const double* gP,
                       T* g, double* P_out, long long* count_out) {
  using C2 = typename cplx<T>::type;
  std::vector<long long> cnt(Nk);
  std::vector<double> Pk(Nk);
  run_power<T>(c, grid, 1, nd, L, window, R, 0, Nk, k_min, k_max, 0, -1, kbin_precision, nullptr,
               Pk.data(), cnt.data(), false);
  if (P_out) std::copy(Pk.begin(), Pk.end(), P_out);
  if (count_out) std::copy(cnt.begin(), cnt.end(), count_out);
  std::vector<double> gPn(Nk);
  for (int b = 0; b < Nk; b++) gPn[b] = cnt[b] > 0 ? 2.0 * gP[b] / (double)cnt[b] : 0.0;
  PowerArgs a = make_power_args<T>(c, 1, nd, L, window, R, 0, Nk, k_min, k_max, 0, -1,
                                   kbin_precision);
  double* dg = (double*)ctx_pinned(c, (size_t)Nk * sizeof(double));
  std::copy(gPn.begin(), gPn.end(), dg);
  double* dgd = (double*)ctx_buf(c, BUF_ADJ_TAB, (size_t)Nk * sizeof(double));
  EXH_CHECK_HIP(hipMemcpyAsync(dgd, dg, (size_t)Nk * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
  const size_t nh = (size_t)nd * nd * (nd / 2 + 1);
  C2* gk = (C2*)c->buf[BUF_GRIDK];
  {
    TimedScope ts(c, "power_adjoint");
    hipLaunchKernelGGL(k_power_adjoint<T>, dim3(std::min<size_t>((nh + 255) / 256, 65536)),
                       dim3(256), 0, c->stream, a, dgd, gk);
    EXH_LAUNCH_CHECK();
  }
  fft_c2r(c, nd, nd, nd, sizeof(T) == 4 ? 4 : 8, 1, gk, g);
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));  // dg is pinned host staging
}

// dL/dS (gS [nL^3][3]) of the LPT particles from dL/ddelta (g) of their
// unweighted, normalised grid: delta = rho ng / np - 1.  Lc: the lattice
// spacing (the box's cell), L the box side.
template <typename T>
void run_deposit_adjoint(exh_ctx* c, const T* S, int nL, double Lc, int nd, double L, int window,
                         const T* g, T* gS) {
  EXH_REQUIRE(window == EXH_WINDOW_CIC || window == EXH_WINDOW_TSC, EXH_ERR_UNSUPPORTED,
              "the adjoint supports the CIC (2) and TSC (5) windows");
  const long long np = (long long)nL * nL * nL;
  const double scale = (double)nd * nd * nd / (double)np;
  const unsigned gs = (unsigned)std::min<long long>((np + 255) / 256, 65536);
  TimedScope ts(c, "deposit_adjoint");
  if (window == EXH_WINDOW_CIC)
    hipLaunchKernelGGL((k_deposit_adjoint<T, 2>), dim3(gs), dim3(256), 0, c->stream, S, nL, Lc,
                       nd, L, g, scale, gS);
  else
    hipLaunchKernelGGL((k_deposit_adjoint<T, 5>), dim3(gs), dim3(256), 0, c->stream, S, nL, Lc,
                       nd, L, g, scale, gS);
  EXH_LAUNCH_CHECK();
}

template void run_power_adjoint<float>(exh_ctx*, const float*, int, double, int, double, int,
                                       double, double, int, const double*, float*, double*,
                                       long long*);
template void run_power_adjoint<double>(exh_ctx*, const double*, int, double, int, double, int,
                                        double, double, int, const double*, double*, double*,
                                        long long*);
template void run_deposit_adjoint<float>(exh_ctx*, const float*, int, double, int, double, int,
                                         const float*, float*);
template void run_deposit_adjoint<double>(exh_ctx*, const double*, int, double, int, double, int,
                                          const double*, double*);

// ---------------------------------------------------------------------------
// halo abundance: abundance_compute (spectrum.c:535-600) and
// Measure_Abundance (abundance.c:4-36)
// ---------------------------------------------------------------------------
// Per-block smallest and largest mass (at most 1024 blocks, reduced on the
// host).
template <typename T>
__global__ void k_mass_range(const T* Mh, long long n, T* bmin, T* bmax) {
  __shared__ T smin[256], smax[256];
  T lo = (T)1e30, hi = (T)-1e30;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const T m = Mh[i];
    lo = m < lo ? m : lo;
    hi = m > hi ? m : hi;
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smin[threadIdx.x] = smin[threadIdx.x + w] < smin[threadIdx.x] ? smin[threadIdx.x + w]
                                                                    : smin[threadIdx.x];
      smax[threadIdx.x] = smax[threadIdx.x + w] > smax[threadIdx.x] ? smax[threadIdx.x + w]
                                                                    : smax[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bmin[blockIdx.x] = smin[0];
    bmax[blockIdx.x] = smax[0];
  }
}

// Counts and mass sums per bin floor((ln M - ln Mmin) / dlnM), kept for
// 0 < bin < Nm as the reference does; LDS histograms, one atomic per block
// and bin.
template <typename T>
__global__ void k_abundance(const T* Mh, long long n, double lnMmin, float dlnM, int Nm,
                            unsigned long long* cnt, double* msum) {
  extern __shared__ double ab_lds[];
  double* ssum = ab_lds;
  unsigned* scnt = (unsigned*)(ab_lds + Nm);
  for (int b = threadIdx.x; b < Nm; b += blockDim.x) {
    ssum[b] = 0.0;
    scnt[b] = 0;
  }
  __syncthreads();
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double f = floor((log((double)Mh[i]) - lnMmin) / (double)dlnM);
    if (f > 0.0 && f < (double)Nm) {
      const int b = (int)f;
      atomicAdd(&scnt[b], 1u);
      atomicAdd(&ssum[b], (double)Mh[i]);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < Nm; b += blockDim.x)
    if (scnt[b]) {
      atomicAdd(&cnt[b], (unsigned long long)scnt[b]);
      atomicAdd(&msum[b], ssum[b]);
    }
}

template <typename T>
void run_abundance(exh_ctx* c, const T* Mh, long long nh, double Mmin_in, double Mmax_in, int Nm,
                   double Lc_in, int ndx, int ndy, int ndz, double* Mmean_out, double* dn_out,
                   double* dn_err_out) {
  EXH_REQUIRE(Nm >= 1 && Nm <= 4096, EXH_ERR_ARG, "Nm must lie in [1, 4096]");
  EXH_REQUIRE(nh >= 0 && (nh == 0 || Mh != nullptr), EXH_ERR_ARG, "null masses");
  T Mmin = (T)Mmin_in, Mmax = (T)Mmax_in;
  const T Lc = (T)Lc_in;
  const T Lx = Lc * (T)ndx, Ly = Lc * (T)ndy, Lz = Lc * (T)ndz;
  const unsigned nb = std::max(1u, std::min(grid_for(nh, 256), 1024u));
  char* buf = (char*)ctx_buf(c, BUF_MISC, 256 + 2 * 1024 * sizeof(T) + (size_t)Nm * 16);
  if (Mmin < (T)0 || Mmax < (T)0) {
    // abundance_compute's defaults: 0.9999 x the smallest, 1.0001 x the largest
    T* bmin = (T*)(buf + 256);
    T* bmax = bmin + 1024;
    hipLaunchKernelGGL(k_mass_range<T>, dim3(nb), dim3(256), 0, c->stream, Mh, nh, bmin, bmax);
    EXH_LAUNCH_CHECK();
    std::vector<T> hmin(nb), hmax(nb);
    EXH_CHECK_HIP(hipMemcpyAsync(hmin.data(), bmin, nb * sizeof(T), hipMemcpyDeviceToHost,
                                 c->stream));
    EXH_CHECK_HIP(hipMemcpyAsync(hmax.data(), bmax, nb * sizeof(T), hipMemcpyDeviceToHost,
                                 c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (Mmin < (T)0) {
      T lo = (T)1e+20;
      for (long long b = 0; b < (long long)nb && nh > 0; b++) lo = hmin[b] < lo ? hmin[b] : lo;
      Mmin = (T)((double)lo * 0.9999);
    }
    if (Mmax < (T)0) {
      for (long long b = 0; b < (long long)nb && nh > 0; b++) Mmax = hmax[b] > Mmax ? hmax[b] : Mmax;
      Mmax = (T)((double)Mmax * 1.0001);
    }
  }
  const float dlnM = (float)((std::log((double)Mmax) - std::log((double)Mmin)) / Nm);
  unsigned long long* cnt = (unsigned long long*)(buf + 256 + 2 * 1024 * sizeof(T));
  double* msum = (double*)(cnt + Nm);
  hipLaunchKernelGGL(k_zero<unsigned long long>, dim3(grid_for(Nm, 256)), dim3(256), 0, c->stream,
                     cnt, (size_t)Nm);
  hipLaunchKernelGGL(k_zero<double>, dim3(grid_for(Nm, 256)), dim3(256), 0, c->stream, msum,
                     (size_t)Nm);
  if (nh > 0)
    hipLaunchKernelGGL(k_abundance<T>, dim3(nb), dim3(256), (size_t)Nm * 12, c->stream, Mh, nh,
                       std::log((double)Mmin), dlnM, Nm, cnt, msum);
  EXH_LAUNCH_CHECK();
  std::vector<unsigned long long> hc(Nm);
  std::vector<double> hs(Nm);
  EXH_CHECK_HIP(hipMemcpyAsync(hc.data(), cnt, Nm * 8, hipMemcpyDeviceToHost, c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(hs.data(), msum, Nm * 8, hipMemcpyDeviceToHost, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  // the reference's arithmetic: T sums and quotients, float dlnM
  const T V = Lx * Ly * Lz;
  for (int b = 0; b < Nm; b++) {
    const int N = (int)hc[b];
    const T mean = N > 0 ? (T)((T)hs[b] / (T)N) : (T)0;
    if (Mmean_out) Mmean_out[b] = (double)mean;
    if (dn_out) dn_out[b] = (double)(T)((T)N / V / (T)dlnM);
    if (dn_err_out) dn_err_out[b] = (double)(T)(std::sqrt((double)N) / (double)V / (double)dlnM);
  }
}

template void run_abundance<float>(exh_ctx*, const float*, long long, double, double, int,
                                   double, int, int, int, double*, double*, double*);
template void run_abundance<double>(exh_ctx*, const double*, long long, double, double, int,
                                    double, int, int, int, double*, double*, double*);

// ---------------------------------------------------------------------------
// bispectrum: bi_compute (spectrum.c:291-406) / Bi_Spectrum (bimodule.c)
// ---------------------------------------------------------------------------
// Combined, deconvolved modes of every tracer (Compute_gridk,
// spectrum_h.c:83-167, in the grid's precision as the reference evaluates
// it) and the shell of every mode, Indice(|k|, k_min, dk) (bimodule.c:72-75);
// per-shell sums of |k| and mode counts.
struct ShellArgs {
  int nd, nz, ntype, nsub, Nk, window;
  double kn, Lb, Normk, R, kmin, dk;  // kn, Lb, Normk, R, kmin, dk hold values of T
  FastDiv dz, dnd;
};

template <typename T>
__device__ inline T sinc_dev(T x) {
  return x != (T)0 ? (T)(sin((double)x) / (double)x) : (T)1;
}

template <typename T>
__global__ void k_shell_modes(ShellArgs a, const typename cplx<T>::type* raw,
                              typename cplx<T>::type* gk, short* shell, double* ksum,
                              unsigned long long* kcnt) {
  using C2 = typename cplx<T>::type;
  extern __shared__ double sm_lds[];
  double* sks = sm_lds;
  unsigned* scs = (unsigned*)(sm_lds + a.Nk);
  for (int q = threadIdx.x; q < a.Nk; q += blockDim.x) {
    sks[q] = 0.0;
    scs[q] = 0;
  }
  __syncthreads();
  const size_t nh = (size_t)a.nd * a.nd * a.nz;
  const T kn = (T)a.kn, Lb = (T)a.Lb, Normk = (T)a.Normk, kmin = (T)a.kmin, dk = (T)a.dk;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    const unsigned r = fdiv((unsigned)id, a.dz);
    const int k = (int)((unsigned)id - r * (unsigned)a.nz);
    const int i = (int)fdiv(r, a.dnd);
    const int j = (int)(r - (unsigned)i * (unsigned)a.nd);
    const T kx = (T)((2 * i < a.nd) ? i : i - a.nd) * kn;
    const T ky = (T)((2 * j < a.nd) ? j : j - a.nd) * kn;
    const T kz = (T)((2 * k < a.nd) ? k : k - a.nd) * kn;
    // W (spectrum_h.c:51-70)
    T W = (T)1;
    if (a.window == 1 || a.window == 2 || a.window == 5) {
      const T s3 = sinc_dev<T>((T)((double)kx * (double)Lb / 2.0)) *
                   sinc_dev<T>((T)((double)ky * (double)Lb / 2.0)) *
                   sinc_dev<T>((T)((double)kz * (double)Lb / 2.0));
      W = a.window == 1 ? s3 : (T)pow((double)s3, a.window == 2 ? 2.0 : 3.0);
    } else if (a.window == 3) {
      const T km = (T)sqrt((double)(kx * kx + ky * ky + kz * kz));
      const T kR = km * (T)a.R;
      W = (T)(3.0 / km / km / (T)a.R / (T)a.R * ((double)sinc_dev<T>(kR) - cos((double)kR)));
    } else if (a.window == 4) {
      const T km = kx * kx + ky * ky + kz * kz;
      W = (T)exp((double)(-km * (T)a.R * (T)a.R) / 2.0);
    }
    for (int t = 0; t < a.ntype; t++) {
      const C2 f1 = raw[(size_t)(t * a.nsub) * nh + id];
      C2 o;
      if (a.nsub == 2) {
        const C2 f2 = raw[(size_t)(t * a.nsub + 1) * nh + id];
        const double th = (double)((kx + ky + kz) * Lb) / 2.0;
        const double c = cos(th), sn = sin(th);
        o.x = (T)(((double)f1.x + (double)f2.x * c + (double)f2.y * sn) / 2.0 * (double)Normk /
                  (double)W);
        o.y = (T)(((double)f1.y + (double)f2.y * c - (double)f2.x * sn) / 2.0 * (double)Normk /
                  (double)W);
      } else {
        o.x = f1.x * Normk / W;
        o.y = f1.y * Normk / W;
      }
      gk[(size_t)t * nh + id] = o;
    }
    const T km = (T)sqrt((double)(kx * kx + ky * ky + kz * kz));
    const double f = floor((double)((km - kmin) / dk));
    const int b = (f >= 0.0 && f < (double)a.Nk) ? (int)f : -1;
    shell[id] = (short)b;
    if (b >= 0) {
      atomicAdd(&sks[b], (double)km);
      atomicAdd(&scs[b], 1u);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < a.Nk; q += blockDim.x)
    if (scs[q]) {
      atomicAdd(&ksum[q], sks[q]);
      atomicAdd(&kcnt[q], (unsigned long long)scs[q]);
    }
}

// The modes of shell s of every tracer and the shell's indicator (1 + 0i),
// zero elsewhere: ntype + 1 half-complex grids for one batched c2r.
template <typename T>
__global__ void k_shell_fill(const typename cplx<T>::type* gk, const short* shell, size_t nh,
                             int ntype, int s, typename cplx<T>::type* out) {
  using C2 = typename cplx<T>::type;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < nh;
       id += (size_t)gridDim.x * blockDim.x) {
    const bool in = shell[id] == s;
    for (int t = 0; t < ntype; t++) {
      C2 v = gk[(size_t)t * nh + id];
      if (!in) v.x = v.y = (T)0;
      out[(size_t)t * nh + id] = v;
    }
    C2 one;
    one.x = in ? (T)1 : (T)0;
    one.y = (T)0;
    out[(size_t)ntype * nh + id] = one;
  }
}

// Sums over the cells of products of two or three shell fields, one double
// per product ("term": fields fa, fb and fc, fc < 0 for a pair).  A block
// stages BS_CH cells of every field in LDS, then each thread runs its
// terms over those cells (all lanes read the same cell: LDS broadcast).
static constexpr int BS_TPT = 4;  // terms per thread

template <typename T>
__global__ __launch_bounds__(256) void k_shell_products(const T* fields, int nfield, size_t ng,
                                                        int ch, const int4* terms, int nterm,
                                                        double* sums) {
  extern __shared__ __align__(16) unsigned char bs_raw[];
  T* lf = (T*)bs_raw;  // [nfield][ch]
  int4 tm[BS_TPT];
  double acc[BS_TPT];
  for (int q = 0; q < BS_TPT; q++) {
    const int t = blockIdx.y * 256 * BS_TPT + q * 256 + threadIdx.x;
    tm[q] = t < nterm ? terms[t] : make_int4(-1, 0, 0, 0);
    acc[q] = 0.0;
  }
  const size_t nch = (ng + ch - 1) / ch;
  for (size_t chunk = blockIdx.x; chunk < nch; chunk += gridDim.x) {
    const size_t x0 = chunk * ch;
    const int nx = (int)min((size_t)ch, ng - x0);
    __syncthreads();
    for (int f = 0; f < nfield; f++)
      for (int c = threadIdx.x; c < ch; c += 256)
        lf[f * ch + c] = c < nx ? fields[(size_t)f * ng + x0 + c] : (T)0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BS_TPT; q++) {
      if (tm[q].x < 0) continue;
      const T* A = lf + tm[q].x * ch;
      const T* B = lf + tm[q].y * ch;
      double s = 0.0;
      if (tm[q].z >= 0) {
        const T* Cf = lf + tm[q].z * ch;
        for (int c = 0; c < nx; c++) s += (double)A[c] * (double)B[c] * (double)Cf[c];
      } else {
        for (int c = 0; c < nx; c++) s += (double)A[c] * (double)B[c];
      }
      acc[q] += s;
    }
  }
  for (int q = 0; q < BS_TPT; q++)
    if (tm[q].x >= 0) atomicAdd(&sums[blockIdx.y * 256 * BS_TPT + q * 256 + threadIdx.x], acc[q]);
}

// Configuration order of bimodule.c:49-317.
void bispectrum_configs(int Nk, std::vector<std::array<int, 3>>& tri) {
  tri.clear();
  for (int a = 0; a < Nk; a++) {
    for (int b = 0; b < a; b++) {
      if (b < a / 2) {
        tri.push_back({a, a, b});
        continue;
      }
      for (int c = std::max(a - b - 1, 0); c < b; c++) tri.push_back({a, b, c});
      tri.push_back({a, b, b});
      tri.push_back({a, a, b});
    }
    tri.push_back({a, a, a});
  }
}

template <typename T>
int run_bispectrum(exh_ctx* c, const T* grid, int ntype, int nd, double L, int window, double R,
                   int interlacing, int Nk, double k_min, double k_max, long long cap,
                   double* kP, double* Pk, double* Nk_out, double* kB, double* Bk,
                   double* Ntri_out) {
  using C2 = typename cplx<T>::type;
  EXH_REQUIRE(nd >= 2 && nd % 2 == 0, EXH_ERR_ARG, "nd must be even");
  EXH_REQUIRE(ntype >= 1 && ntype <= 4, EXH_ERR_UNSUPPORTED, "1 to 4 tracer types supported");
  EXH_REQUIRE(Nk >= 1 && Nk <= 4096, EXH_ERR_ARG, "Nk must lie in [1, 4096]");
  EXH_REQUIRE(window >= 0 && window <= 5, EXH_ERR_UNSUPPORTED, "window %d", window);
  EXH_REQUIRE(k_max > k_min, EXH_ERR_ARG, "k_max must exceed k_min");
  std::vector<std::array<int, 3>> tri;
  bispectrum_configs(Nk, tri);
  const long long ntri = (long long)tri.size();
  EXH_REQUIRE(cap >= ntri, EXH_ERR_ARG, "output capacity %lld below %lld configurations", cap,
              ntri);
  const int nsub = interlacing ? 2 : 1;
  const int nz = nd / 2 + 1;
  const size_t nh = (size_t)nd * nd * nz, ng = (size_t)nd * nd * nd;
  const int nf = ntype + 1;  // fields per shell: tracers and the indicator
  const size_t shells_bytes = (size_t)Nk * nf * ng * sizeof(T);
  EXH_REQUIRE(shells_bytes <= (size_t)96 << 30, EXH_ERR_UNSUPPORTED,
              "bispectrum: %zu bytes of shell fields exceed the 96 GiB budget", shells_bytes);
  // 1. r2c of every grid, combined modes and shells
  C2* raw = (C2*)ctx_buf(c, BUF_GRIDK, nh * ntype * nsub * sizeof(C2));
  fft_r2c(c, nd, nd, nd, sizeof(T) == 4 ? 4 : 8, ntype * nsub, grid, raw);
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t r = off;
    off += (b + 255) & ~(size_t)255;
    return r;
  };
  const size_t o_gk = take(nh * ntype * sizeof(C2)), o_sh = take(nh * sizeof(short)),
               o_w = take(nh * nf * sizeof(C2)), o_ks = take((size_t)Nk * 8),
               o_kc = take((size_t)Nk * 8);
  char* kb = (char*)ctx_buf(c, BUF_SHELLK, off);
  C2* gk = (C2*)(kb + o_gk);
  short* shell = (short*)(kb + o_sh);
  C2* work = (C2*)(kb + o_w);
  double* ksum = (double*)(kb + o_ks);
  unsigned long long* kcnt = (unsigned long long*)(kb + o_kc);
  hipLaunchKernelGGL(k_zero<double>, dim3(grid_for(Nk, 256)), dim3(256), 0, c->stream, ksum,
                     (size_t)Nk);
  hipLaunchKernelGGL(k_zero<unsigned long long>, dim3(grid_for(Nk, 256)), dim3(256), 0,
                     c->stream, kcnt, (size_t)Nk);
  ShellArgs sa{};
  const T Lr = (T)L;
  sa.nd = nd;
  sa.nz = nz;
  sa.ntype = ntype;
  sa.nsub = nsub;
  sa.Nk = Nk;
  sa.window = window;
  sa.kn = (double)(T)(2 * M_PI / (double)Lr);
  sa.Lb = (double)(T)(Lr / (T)nd);
  sa.Normk = (double)(T)std::pow((double)(Lr / (T)(nd * nd)), 3.0 / 2.0);
  sa.R = (double)(T)R;
  sa.kmin = (double)(T)k_min;
  sa.dk = (double)(T)(((T)k_max - (T)k_min) / (T)Nk);
  sa.dz = fast_div((unsigned)nz);
  sa.dnd = fast_div((unsigned)nd);
  EXH_REQUIRE(nh < (1ull << 31), EXH_ERR_UNSUPPORTED, "grid too large for the shell kernels");
  {
    TimedScope ts(c, "bispec_shells");
    hipLaunchKernelGGL(k_shell_modes<T>, dim3(std::min(grid_for(nh, 256), 4096u)), dim3(256),
                       (size_t)Nk * 12, c->stream, sa, raw, gk, shell, ksum, kcnt);
    EXH_LAUNCH_CHECK();
    // 2. the real field of every shell and tracer, and of its indicator
    T* fields = (T*)ctx_buf(c, BUF_SHELLS, shells_bytes);
    for (int s = 0; s < Nk; s++) {
      hipLaunchKernelGGL(k_shell_fill<T>, dim3(std::min(grid_for(nh, 256), 8192u)), dim3(256), 0,
                         c->stream, gk, shell, nh, ntype, s, work);
      EXH_LAUNCH_CHECK();
      fft_c2r(c, nd, nd, nd, sizeof(T) == 4 ? 4 : 8, nf, work, fields + (size_t)s * nf * ng);
    }
  }
  // 3. the products: per configuration ntype^3 triples and the indicator's,
  // per shell the ntype (ntype+1)/2 pairs and the indicator's square
  const int NBs = ntype * ntype * ntype, NPs = ntype * (ntype + 1) / 2;
  std::vector<int4> terms;
  for (const auto& t : tri) {
    for (int f1 = 0; f1 < ntype; f1++)
      for (int f2 = 0; f2 < ntype; f2++)
        for (int f3 = 0; f3 < ntype; f3++)
          terms.push_back(make_int4(t[0] * nf + f1, t[1] * nf + f2, t[2] * nf + f3, 0));
    terms.push_back(make_int4(t[0] * nf + ntype, t[1] * nf + ntype, t[2] * nf + ntype, 0));
  }
  const size_t tri_terms = terms.size();
  for (int s = 0; s < Nk; s++) {
    for (int f1 = 0; f1 < ntype; f1++)
      for (int f2 = 0; f2 <= f1; f2++) terms.push_back(make_int4(s * nf + f1, s * nf + f2, -1, 0));
    terms.push_back(make_int4(s * nf + ntype, s * nf + ntype, -1, 0));
  }
  const int nterm = (int)terms.size();
  size_t toff = 0;
  auto ttake = [&](size_t b) {
    size_t r = toff;
    toff += (b + 255) & ~(size_t)255;
    return r;
  };
  const size_t o_terms = ttake((size_t)nterm * sizeof(int4)), o_sums = ttake((size_t)nterm * 8);
  char* tb = (char*)ctx_buf(c, BUF_PART, toff);
  int4* dterms = (int4*)(tb + o_terms);
  double* dsums = (double*)(tb + o_sums);
  EXH_CHECK_HIP(hipMemcpyAsync(dterms, terms.data(), (size_t)nterm * sizeof(int4),
                               hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_zero<double>, dim3(grid_for(nterm, 256)), dim3(256), 0, c->stream, dsums,
                     (size_t)nterm);
  const int nfield = Nk * nf;
  int ch = 256;
  while (ch > 16 && (size_t)nfield * ch * sizeof(T) > 96 * 1024) ch >>= 1;
  EXH_REQUIRE((size_t)nfield * ch * sizeof(T) <= 96 * 1024, EXH_ERR_UNSUPPORTED,
              "bispectrum: %d shell fields exceed the LDS staging", nfield);
  {
    TimedScope ts(c, "bispec_products");
    const T* fields = (const T*)c->buf[BUF_SHELLS];
    const int lds = nfield * ch * (int)sizeof(T);
    EXH_CHECK_HIP(hipFuncSetAttribute((const void*)k_shell_products<T>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    const unsigned ty = (unsigned)((nterm + 256 * BS_TPT - 1) / (256 * BS_TPT));
    const size_t nch = (ng + ch - 1) / ch;
    const unsigned tx = (unsigned)std::max<size_t>(1, std::min<size_t>(nch, 2048 / ty + 1));
    hipLaunchKernelGGL(k_shell_products<T>, dim3(tx, ty), dim3(256), lds, c->stream, fields,
                       nfield, ng, ch, dterms, nterm, dsums);
    EXH_LAUNCH_CHECK();
  }
  std::vector<double> hs(nterm), hks(Nk);
  std::vector<unsigned long long> hkc(Nk);
  EXH_CHECK_HIP(hipMemcpyAsync(hs.data(), dsums, (size_t)nterm * 8, hipMemcpyDeviceToHost,
                               c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(hks.data(), ksum, (size_t)Nk * 8, hipMemcpyDeviceToHost,
                               c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(hkc.data(), kcnt, (size_t)Nk * 8, hipMemcpyDeviceToHost,
                               c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  // 4. normalisation (bimodule.c:356-370) and outputs in T (spectrum.c:368-383)
  std::vector<double> km(Nk);
  for (int s = 0; s < Nk; s++) km[s] = (double)(T)(hkc[s] ? hks[s] / (double)hkc[s] : NAN);
  const double L15 = std::pow((double)Lr, 1.5);
  for (long long n = 0; n < ntri; n++) {
    const double* v = hs.data() + (size_t)n * (NBs + 1);
    const double IB = v[NBs];
    for (int j = 0; j < NBs; j++) Bk[(size_t)j * ntri + n] = (double)(T)(v[j] / IB * L15);
    Ntri_out[n] = (double)(T)(IB / (double)ng);
    for (int q = 0; q < 3; q++) kB[(size_t)q * ntri + n] = km[tri[n][q]];
  }
  for (int s = 0; s < Nk; s++) {
    const double* v = hs.data() + tri_terms + (size_t)s * (NPs + 1);
    const double IP = v[NPs];
    for (int j = 0; j < NPs; j++) Pk[(size_t)j * Nk + s] = (double)(T)(v[j] / IP);
    Nk_out[s] = (double)(T)(IP / (double)ng);
    kP[s] = km[s];
  }
  return (int)ntri;
}

template int run_bispectrum<float>(exh_ctx*, const float*, int, int, double, int, double, int,
                                   int, double, double, long long, double*, double*, double*,
                                   double*, double*, double*);
template int run_bispectrum<double>(exh_ctx*, const double*, int, int, double, int, double, int,
                                    int, double, double, long long, double*, double*, double*,
                                    double*, double*, double*);

template void run_grid<float>(exh_ctx*, const float*, const float*, const float*, const int*,
                              long long, int, int, double, double, double, int, int, double,
                              double, int, int, float*, void*);
template void run_grid<double>(exh_ctx*, const double*, const double*, const double*, const int*,
                               long long, int, int, double, double, double, int, int, double,
                               double, int, int, double*, void*);
template void run_power<float>(exh_ctx*, const float*, int, int, double, int, double, int, int,
                               double, double, int, int, int, double*, double*, long long*, bool);
template void run_power<double>(exh_ctx*, const double*, int, int, double, int, double, int, int,
                                double, double, int, int, int, double*, double*, long long*,
                                bool);

}  // namespace exh
