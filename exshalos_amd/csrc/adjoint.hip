// Adjoint of white-noise modes -> 1LPT/2LPT -> mass assignment -> P(k)
// (gfx950).
//
// The forward path is the library's own: k_poisson_modes + c2r (phi1),
// k_grad_tiles (S1 = -D phi1 / (60 Lc), lpt.c:56-93), k_source_tiles (the
// 2LPT source, lpt.c:116-134), r2c, the Poisson factor and c2r (phi2),
// k_grad_tiles (S2 = c D phi2 / (60 Lc), lpt.c:182-223), positions
// (lpt.c:229-242), the deposit (gridmodule.c:21-53, 146-270) and the
// monopole binning (powermodule.c:4-108).  The reference has no adjoint;
// this is the transpose of each step, in reverse:
//
//   dL/dP_b -> dL/dF (k_power_adjoint) -> dL/ddelta = c2r(dL/dF / w)
//   -> dL/dpos (k_deposit_adjoint, positions recomputed from the lattice
//      and S in double, not stored)
//   -> dL/dphi2 = -c D^T... = -c/(60 Lc) sum_a D_a(dL/dS_a)   (D^T = -D)
//   -> dL/dsrc = c2r(K2 r2c(dL/dphi2))   (K2 real and even: self-adjoint)
//   -> dL/dS_ab (k_src_adjoint_a) -> dL/dS1 += -1/(60 Lc) D_b(dL/dS_ab)
//   -> dL/dphi1 = 1/(60 Lc) sum_a D_a(dL/dS1_a)
//   -> dL/dnoise (k_field_adjoint on r2c(dL/dphi1))
//
// D is the reference's sixth-order difference, antisymmetric on the
// periodic grid, so its transpose is -D.  Stencils are evaluated in double
// from global loads (the adjoint's grids are small; the forward's
// LDS-tiled passes are reused as they are).
#include "common.hpp"
#include "kernels.hpp"
#include "kspace.hpp"

#include <cmath>

namespace exh {

__device__ inline double d6(double m3, double m2, double m1, double p1, double p2, double p3) {
  return 1.0 * p3 - 9.0 * p2 + 45.0 * p1 - 45.0 * m1 + 9.0 * m2 - 1.0 * m3;
}

__device__ inline int wr(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// D along `axis` of component `comp` (of `nc` interleaved) of f at (i, j, k)
// on the cubic grid of side n.
template <typename T>
__device__ inline double d6_at(const T* f, int nc, int comp, int n, int i, int j, int k,
                               int axis) {
  double v[6];
  const int sh[6] = {-3, -2, -1, 1, 2, 3};
  for (int s = 0; s < 6; s++) {
    int a = i, b = j, c = k;
    if (axis == 0) a = wr(i + sh[s], n);
    if (axis == 1) b = wr(j + sh[s], n);
    if (axis == 2) c = wr(k + sh[s], n);
    v[s] = (double)f[(((size_t)a * n + b) * n + c) * nc + comp];
  }
  return d6(v[0], v[1], v[2], v[3], v[4], v[5]);
}

__device__ inline void cell_ijk(size_t c, int n, int& i, int& j, int& k) {
  k = (int)(c % n);
  const size_t r = c / n;
  j = (int)(r % n);
  i = (int)(r / n);
}

// out = scale * sum_a D_a(v_a), v an interleaved [ng][3] field.
template <typename T>
__global__ __launch_bounds__(256) void k_fd_div3(int n, const T* v, double scale, T* out) {
  const size_t ng = (size_t)n * n * n;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ng;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(c, n, i, j, k);
    const double d = d6_at(v, 3, 0, n, i, j, k, 0) + d6_at(v, 3, 1, n, i, j, k, 1) +
                     d6_at(v, 3, 2, n, i, j, k, 2);
    out[c] = (T)(scale * d);
  }
}

// dL/dS_ab of the 2LPT source phi2 = Sxx Syy + Sxx Szz + Syy Szz - Sxy^2 -
// Sxz^2 - Syz^2, S_ab = D_b(S1_a) / (60 Lc) rounded as k_source_tiles
// rounds them: g6 = [xx, yy, zz, xy, xz, yz] planes of dL/dS_ab.
template <typename T>
__global__ __launch_bounds__(256) void k_src_adjoint_a(int n, const T* S1, const T* gsrc,
                                                       double inv60, T* g6) {
  const size_t ng = (size_t)n * n * n;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ng;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(c, n, i, j, k);
    const double sxx = (double)(T)(d6_at(S1, 3, 0, n, i, j, k, 0) * inv60);
    const double sxy = (double)(T)(d6_at(S1, 3, 0, n, i, j, k, 1) * inv60);
    const double sxz = (double)(T)(d6_at(S1, 3, 0, n, i, j, k, 2) * inv60);
    const double syy = (double)(T)(d6_at(S1, 3, 1, n, i, j, k, 1) * inv60);
    const double syz = (double)(T)(d6_at(S1, 3, 1, n, i, j, k, 2) * inv60);
    const double szz = (double)(T)(d6_at(S1, 3, 2, n, i, j, k, 2) * inv60);
    const double g = (double)gsrc[c];
    g6[0 * ng + c] = (T)(g * (syy + szz));
    g6[1 * ng + c] = (T)(g * (sxx + szz));
    g6[2 * ng + c] = (T)(g * (sxx + syy));
    g6[3 * ng + c] = (T)(-2.0 * g * sxy);
    g6[4 * ng + c] = (T)(-2.0 * g * sxz);
    g6[5 * ng + c] = (T)(-2.0 * g * syz);
  }
}

// gS (= dL/dpos on entry) += dL/dS1 through the source:
// -1/(60 Lc) (D_x g_xx + D_y g_xy + D_z g_xz, D_y g_yy + D_z g_yz, D_z g_zz).
template <typename T>
__global__ __launch_bounds__(256) void k_src_adjoint_b(int n, const T* g6, double inv60, T* gS) {
  const size_t ng = (size_t)n * n * n;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < ng;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_ijk(c, n, i, j, k);
    const T* gxx = g6;
    const T* gyy = g6 + ng;
    const T* gzz = g6 + 2 * ng;
    const T* gxy = g6 + 3 * ng;
    const T* gxz = g6 + 4 * ng;
    const T* gyz = g6 + 5 * ng;
    const double ax = d6_at(gxx, 1, 0, n, i, j, k, 0) + d6_at(gxy, 1, 0, n, i, j, k, 1) +
                      d6_at(gxz, 1, 0, n, i, j, k, 2);
    const double ay = d6_at(gyy, 1, 0, n, i, j, k, 1) + d6_at(gyz, 1, 0, n, i, j, k, 2);
    const double az = d6_at(gzz, 1, 0, n, i, j, k, 2);
    gS[3 * c + 0] = (T)((double)gS[3 * c + 0] - inv60 * ax);
    gS[3 * c + 1] = (T)((double)gS[3 * c + 1] - inv60 * ay);
    gS[3 * c + 2] = (T)((double)gS[3 * c + 2] - inv60 * az);
  }
}

static unsigned cells_grid(size_t ng) {
  return (unsigned)std::min<size_t>((ng + 255) / 256, 65536);
}

template <typename T>
void run_pk_adjoint(exh_ctx* c, const exh_box* b, const exh_field_src* s, double k_smooth,
                    int do_2lpt, int window, int Nk, double k_min, double k_max,
                    int kbin_precision, const double* gP, T* gnoise, double* P_out,
                    long long* count_out) {
  const int n = b->nd[0];
  EXH_REQUIRE(b->nd[1] == n && b->nd[2] == n && n % 2 == 0, EXH_ERR_UNSUPPORTED,
              "the adjoint needs a cubic grid of even side");
  EXH_REQUIRE(s->kind == EXH_SRC_MODES, EXH_ERR_ARG, "the adjoint takes white-noise modes");
  const size_t ng = (size_t)b->ng;
  const int prec = sizeof(T) == 4 ? 4 : 8;
  const double L = (double)(T)b->L[0];
  const double inv60 = 1.0 / (60.0 * (double)(T)b->Lc);
  const double coef = -3.0 / 7.0 * std::pow((double)(T)b->Omz, -1.0 / 143);

  // forward, recomputed: S1 (for the source), positions, grid
  T* S1 = do_2lpt ? (T*)ctx_buf(c, BUF_ADJ_S1, 3 * ng * sizeof(T)) : nullptr;
  T* S = (T*)ctx_buf(c, BUF_S, 3 * ng * sizeof(T));
  T* pos = (T*)ctx_buf(c, BUF_ADJ_POS, 3 * ng * sizeof(T));
  if (do_2lpt) launch_lpt<T>(c, b, s, k_smooth, 0, 0, S1, nullptr, nullptr);
  launch_lpt<T>(c, b, s, k_smooth, do_2lpt, 0, S, nullptr, pos);
  T* grid = (T*)ctx_buf(c, BUF_ADJ_GRID, ng * sizeof(T));
  run_grid<T>(c, pos, nullptr, nullptr, nullptr, (long long)ng, 1, n, L, (double)(T)b->Om0,
              (double)(T)b->redshift, -1, window, 0.0, 0.0, 0, 1, grid);

  // P(k) -> grid -> positions
  T* g = (T*)ctx_buf(c, BUF_ADJ_G, ng * sizeof(T));
  run_power_adjoint<T>(c, grid, n, L, window, 0.0, Nk, k_min, k_max, kbin_precision, gP, g, P_out,
                       count_out);
  T* gS = (T*)ctx_buf(c, BUF_ADJ_GS, 3 * ng * sizeof(T));
  run_deposit_adjoint<T>(c, S, n, (double)(T)b->Lc, n, L, window, g, gS);

  using C2 = typename cplx<T>::type;
  const size_t nh = (size_t)n * n * (n / 2 + 1);
  C2* ck = (C2*)ctx_buf(c, BUF_CK0, nh * sizeof(C2));
  if (do_2lpt) {
    TimedScope ts(c, "lpt2_adjoint");
    // dL/dphi2, then dL/dsrc by the same Poisson operator as the forward
    hipLaunchKernelGGL(k_fd_div3<T>, dim3(cells_grid(ng)), dim3(256), 0, c->stream, n, gS,
                       -coef * inv60, g);
    EXH_LAUNCH_CHECK();
    fft_r2c(c, n, n, n, prec, 1, g, ck);
    const PoissonParams pp{n, n, n, (double)(T)b->kl[0], (double)(T)b->kl[1],
                           (double)(T)b->kl[2], (double)(T)k_smooth, (double)(T)b->Normk,
                           (double)(T)b->Normx};
    if (!fft_c2r_poisson(c, prec, ck, ck, g, pp))
      EXH_REQUIRE(false, EXH_ERR_UNSUPPORTED,
                  "the adjoint needs a grid side the FFT passes cover (powers of two)");
    T* g6 = (T*)ctx_buf(c, BUF_ADJ_G6, 6 * ng * sizeof(T));
    hipLaunchKernelGGL(k_src_adjoint_a<T>, dim3(cells_grid(ng)), dim3(256), 0, c->stream, n, S1,
                       g, inv60, g6);
    hipLaunchKernelGGL(k_src_adjoint_b<T>, dim3(cells_grid(ng)), dim3(256), 0, c->stream, n, g6,
                       inv60, gS);
    EXH_LAUNCH_CHECK();
  }
  {
    TimedScope ts(c, "lpt1_adjoint");
    hipLaunchKernelGGL(k_fd_div3<T>, dim3(cells_grid(ng)), dim3(256), 0, c->stream, n, gS, inv60,
                       g);
    EXH_LAUNCH_CHECK();
  }
  fft_r2c(c, n, n, n, prec, 1, g, ck);
  launch_field_adjoint<T>(c, b, s, k_smooth, ck, gnoise);
}

template void run_pk_adjoint<float>(exh_ctx*, const exh_box*, const exh_field_src*, double, int,
                                    int, int, double, double, int, const double*, float*, double*,
                                    long long*);
template void run_pk_adjoint<double>(exh_ctx*, const exh_box*, const exh_field_src*, double, int,
                                     int, int, double, double, int, const double*, double*,
                                     double*, long long*);

}  // namespace exh
