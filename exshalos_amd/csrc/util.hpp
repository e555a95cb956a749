// Device-wide exclusive scan and block-level helpers (64-wide wavefronts).
#pragma once

#include "common.hpp"

namespace exh {

// out[0..n] = exclusive prefix sums of in[0..n-1]; out[n] = total.
// `in` and `out` may not alias.  Uses BUF_SCAN scratch (or `scratch`).
template <typename T>
void exclusive_scan(exh_ctx* c, const T* in, T* out, long long n);

// Exclusive prefix of `v` over the 256 threads of a block; returns this
// thread's prefix and writes the block total to *total.  `sh` must hold
// 4 entries of T in shared memory.
template <typename T>
__device__ inline T block_exclusive_256(T v, T* sh, T* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T x = v;
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wave] = x;
  __syncthreads();
  T base = 0, tot = 0;
  for (int w = 0; w < 4; w++) {
    T s = sh[w];
    if (w < wave) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

__device__ inline unsigned long long lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace exh
