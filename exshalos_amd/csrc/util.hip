// Device-wide exclusive scan: one workgroup for up to ONE_ITEMS * 256
// entries; beyond, per-block sums, recursive scan of the sums, then a
// block-local rescan that adds the block offset.
#include "util.hpp"

namespace exh {

static constexpr int SCAN_ITEMS = 8;
static constexpr int SCAN_TILE = 256 * SCAN_ITEMS;

template <typename T>
__global__ void k_scan_reduce(const T* in, long long n, T* sums) {
  __shared__ T sh[4];
  long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  T s = 0;
  for (int q = 0; q < SCAN_ITEMS; q++)
    if (base + q < n) s += in[base + q];
  T tot;
  block_exclusive_256<T>(s, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <typename T>
__global__ void k_scan_apply(const T* in, long long n, const T* offs, T* out) {
  __shared__ T sh[4];
  long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T s = 0;
  for (int q = 0; q < SCAN_ITEMS; q++) {
    v[q] = (base + q < n) ? in[base + q] : (T)0;
    s += v[q];
  }
  T tot;
  T off = offs ? offs[blockIdx.x] : (T)0;
  T pre = block_exclusive_256<T>(s, sh, &tot) + off;
  for (int q = 0; q < SCAN_ITEMS; q++) {
    if (base + q < n) out[base + q] = pre;
    pre += v[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) out[n] = off + tot;
}

// Whole scan in one small workgroup (256 threads, ONE_ITEMS consecutive
// entries each, independent loads): it fits beside the large workgroups
// of kernels running on another stream.
static constexpr int ONE_ITEMS = 32;
static constexpr int ONE_THREADS = 256;

template <typename T>
__global__ __launch_bounds__(ONE_THREADS) void k_scan_one(const T* in, long long n, T* out) {
  __shared__ T sh[4];
  const int base = threadIdx.x * ONE_ITEMS, nn = (int)n;
  T v[ONE_ITEMS];
#pragma unroll
  for (int q = 0; q < ONE_ITEMS; q++) v[q] = base + q < nn ? in[base + q] : (T)0;
  T s = 0;
#pragma unroll
  for (int q = 0; q < ONE_ITEMS; q++) s += v[q];
  T tot;
  T pre = block_exclusive_256<T>(s, sh, &tot);
#pragma unroll
  for (int q = 0; q < ONE_ITEMS; q++) {
    if (base + q < nn) out[base + q] = pre;
    pre += v[q];
  }
  if (threadIdx.x == 0) out[n] = tot;
}

template <typename T>
static void scan_rec(exh_ctx* c, const T* in, T* out, long long n, T* scratch) {
  long long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb < 1) nb = 1;
  T* sums = scratch;
  T* offs = scratch + nb;
  T* rest = scratch + 2 * nb + 1;
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(256), 0, c->stream, in, n, sums);
  EXH_LAUNCH_CHECK();
  if (nb > 1) scan_rec<T>(c, sums, offs, nb, rest);
  hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(256), 0, c->stream, in, n,
                     nb > 1 ? offs : (const T*)nullptr, out);
  EXH_LAUNCH_CHECK();
}

template <typename T>
void exclusive_scan(exh_ctx* c, const T* in, T* out, long long n) {
  if (n <= 0) {
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(256), 0, c->stream, in, 0LL,
                       (const T*)nullptr, out);
    EXH_LAUNCH_CHECK();
    return;
  }
  if (n <= (long long)ONE_ITEMS * ONE_THREADS) {
    hipLaunchKernelGGL(k_scan_one<T>, dim3(1), dim3(ONE_THREADS), 0, c->stream, in, n, out);
    EXH_LAUNCH_CHECK();
    return;
  }
  // scratch: geometric series of levels, < 3 * n / SCAN_TILE + 64 entries
  long long need = 64;
  for (long long m = n; m > 1; m = (m + SCAN_TILE - 1) / SCAN_TILE) need += 2 * ((m + SCAN_TILE - 1) / SCAN_TILE) + 2;
  T* scratch = (T*)ctx_buf(c, BUF_SCAN, (size_t)need * sizeof(T));
  scan_rec<T>(c, in, out, n, scratch);
}

template void exclusive_scan<int>(exh_ctx*, const int*, int*, long long);
template void exclusive_scan<long long>(exh_ctx*, const long long*, long long*, long long);

}  // namespace exh
