// Device helpers of k-space operations shared by the LPT kernels and the
// FFT passes that fuse them.
#pragma once

#include "common.hpp"

namespace exh {

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

// Host-side parameters of phi_k = c_k * pre / (-k^2) * scale, with phi_k = 0
// at k = 0 and for |k| > k_smooth (lpt.c:12-44).  The real values are
// already rounded to the grid's precision.
struct PoissonParams {
  int n0, n1, n2;
  double kl0, kl1, kl2, k_smooth, pre, scale;
};

// fft.hip: the axis-0 inverse pass with the Poisson factor applied as the
// points are loaded from `src`, written to `dst` (may equal src).
void fft_axis0_inverse_poisson(exh_ctx* c, int prec, const void* src, void* dst,
                               const PoissonParams& p);
// runtime.cpp: phi(x) from c_k through that pass, the axis-1 pass and the
// contiguous-axis real transform; `work` is a half-complex grid (may equal
// src).  False, with nothing launched, where the strided passes do not
// cover the grid; the caller then runs the Poisson kernel and fft_c2r.
bool fft_c2r_poisson(exh_ctx* c, int prec, const void* src, void* work, void* out,
                     const PoissonParams& p);

}  // namespace exh
