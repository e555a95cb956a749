// Strided complex FFT passes along the two outer axes of a half-complex grid.
//
// A 3D real transform of [n0][n1][n2] runs as three passes: a batched 1D real
// transform along the contiguous axis (the row kernels at the end of this
// file), and one complex pass along each strided axis.  A pass sees the grid
// as
//
//     lines x [N points, `stride` elements apart] x [ncol contiguous columns]
//
// (axis 1: lines = n0, stride = ncol = n2/2+1; axis 0: lines = 1,
// stride = ncol = n1 (n2/2+1)) and transforms every column in place.
// Lengths are 2^a or 3 * 2^a (a leading radix-3 stage).
//
// One workgroup owns BZ adjacent columns of one line: 128 bytes per point,
// so every global access is a full 128-byte row segment.  The transform is a
// Stockham autosort FFT (Govindaraju et al., "High performance discrete
// Fourier transforms on graphics processors", SC 2008) with radix-8 stages,
// one leading radix-2 or radix-4 stage where log2 of N's power-of-two part is
// not a multiple of three, and a radix-3 stage before them where 3 divides N
// (the mixed-radix Stockham indexing is the same for every radix).  The
// first stage reads its operands straight from global memory and
// the last stage writes straight back; only the intermediate stages go
// through LDS, whose rows are padded by one element against bank conflicts.
// Twiddles W_N^k are computed on the host in double and rounded once.
//
// Direction and scaling follow rocFFT: forward uses exp(-2 pi i k n / N),
// inverse exp(+2 pi i k n / N), neither is normalised.
#include <cmath>
#include <type_traits>

#include "common.hpp"
#include "kspace.hpp"

namespace exh {

namespace {

constexpr int FFT_THREADS = 256;
constexpr int AXIS_THREADS = 256;  // threads of a strided-axis workgroup

template <typename T>
struct FftCfg;
template <>
struct FftCfg<float> {
  static constexpr int BZ = 16;  // columns per workgroup: 16 x 8 bytes
  static constexpr int PAD = 1;
};
template <>
struct FftCfg<double> {
  static constexpr int BZ = 8;  // 8 x 16 bytes
  static constexpr int PAD = 0;  // keeps two workgroups per CU at N = 512
};

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// power-of-two factor of a length 2^a or 3 * 2^a
constexpr int pow2_part(int n) { return n % 3 == 0 ? n / 3 : n; }

// Stages: one radix-3 stage where 3 divides N, then radix-8 stages with one
// leading radix-2 or radix-4 stage where log2 of the power-of-two part is
// not a multiple of three.
template <int N>
constexpr int fft_nstages() {
  return (N % 3 == 0 ? 1 : 0) + (ilog2(pow2_part(N)) % 3 ? 1 : 0) + ilog2(pow2_part(N)) / 3;
}

template <int N>
constexpr int fft_radix(int s) {
  if (N % 3 == 0) {
    if (s == 0) return 3;
    s -= 1;
  }
  return (ilog2(pow2_part(N)) % 3 && s == 0) ? (1 << (ilog2(pow2_part(N)) % 3)) : 8;
}

template <typename C2>
__device__ __forceinline__ C2 cadd(C2 a, C2 b) {
  C2 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}

template <typename C2>
__device__ __forceinline__ C2 csub(C2 a, C2 b) {
  C2 r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}

// a * w (forward) or a * conj(w) (inverse)
template <bool INV, typename C2>
__device__ __forceinline__ C2 ctw(C2 a, C2 w) {
  C2 r;
  if (INV) {
    r.x = a.x * w.x + a.y * w.y;
    r.y = a.y * w.x - a.x * w.y;
  } else {
    r.x = a.x * w.x - a.y * w.y;
    r.y = a.x * w.y + a.y * w.x;
  }
  return r;
}

// a * W_4: -i forward, +i inverse
template <bool INV, typename C2>
__device__ __forceinline__ C2 crot(C2 a) {
  C2 r;
  if (INV) {
    r.x = -a.y;
    r.y = a.x;
  } else {
    r.x = a.y;
    r.y = -a.x;
  }
  return r;
}

// a * W_8: (1 -+ i) / sqrt 2
template <bool INV, typename C2>
__device__ __forceinline__ C2 cw8(C2 a) {
  using T = decltype(a.x);
  const T h = (T)0.70710678118654752440;
  C2 r;
  if (INV) {
    r.x = (a.x - a.y) * h;
    r.y = (a.x + a.y) * h;
  } else {
    r.x = (a.x + a.y) * h;
    r.y = (a.y - a.x) * h;
  }
  return r;
}

template <int R, bool INV, typename C2>
__device__ __forceinline__ void dft(C2* v) {
  if constexpr (R == 2) {
    const C2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const C2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const C2 t2 = cadd(v[1], v[3]), t3 = crot<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 3) {
    // X1 = a - (b + c)/2 -+ i (sqrt 3 / 2)(b - c), X2 the conjugate rotation
    using T = decltype(v[0].x);
    const T s3 = (T)0.86602540378443864676;
    const C2 a = v[0], t1 = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    C2 h, u;
    h.x = a.x - (T)0.5 * t1.x;
    h.y = a.y - (T)0.5 * t1.y;
    if (INV) {  // +i s3 d
      u.x = -s3 * d.y;
      u.y = s3 * d.x;
    } else {  // -i s3 d
      u.x = s3 * d.y;
      u.y = -s3 * d.x;
    }
    v[0] = cadd(a, t1);
    v[1] = cadd(h, u);
    v[2] = csub(h, u);
  } else {
    static_assert(R == 8, "radix 2, 3, 4 or 8");
    C2 e[4] = {v[0], v[2], v[4], v[6]};
    C2 o[4] = {v[1], v[3], v[5], v[7]};
    dft<4, INV>(e);
    dft<4, INV>(o);
    o[1] = cw8<INV>(o[1]);
    o[2] = crot<INV>(o[2]);
    o[3] = crot<INV>(cw8<INV>(o[3]));
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
}

// Operators applied to each point as the first stage loads it (point index
// n along the axis; the column is bound once per thread).
struct OpNone {
  __device__ void bind(long long) {}
  template <typename C2>
  __device__ C2 operator()(C2 v, int) const {
    return v;
  }
};

// phi_k = c_k * pre / (-k^2) * scale on the pass along axis 0, whose columns
// are the (j, k) pairs of a plane.  k^2 = kx^2 + (ky^2 + kz^2) in double,
// the column's part bound once; the factor pre / (-k^2) * scale is formed
// once per point and multiplies both parts.  The standalone Poisson kernel
// (lpt_fact) rounds |k| to T and divides each part, so the two agree to a
// few ulp of T; its |k| <= k_smooth test is repeated exactly for the points
// whose k^2 lies within 1e-6 of k_smooth^2, so both zero the same modes.
template <typename T>
struct OpPoisson {
  int n0, n1, n2, nz2;
  T kl0, kl1, kl2, k_smooth, pre, scale;
  double ks2;  // k_smooth^2
  int j, k;
  double kyz2;
  __device__ void bind(long long col) {
    j = (int)(col / nz2);
    k = (int)(col - (long long)j * nz2);
    const T ky = (2 * j < n1) ? (T)j * kl1 : (T)(j - n1) * kl1;
    T kz = (T)k * kl2;
    if (k == n2 / 2) kz = (T)(-(T)n2 / 2.0 * (double)kl2);
    kyz2 = (double)ky * ky + (double)kz * kz;
  }
  template <typename C2>
  __device__ C2 operator()(C2 v, int i) const {
    const T kx = (2 * i < n0) ? (T)i * kl0 : (T)(i - n0) * kl0;
    const double s = (double)kx * kx + kyz2;
    bool in = s > 0.0 && s <= ks2;
    if (fabs(s - ks2) <= 1e-6 * ks2) {
      T fact;
      in = lpt_fact<T>(n0, n1, n2, kl0, kl1, kl2, k_smooth, i, j, k, fact);
    }
    C2 r;
    if (in) {
      const T m = pre / (T)(-s) * scale;
      r.x = v.x * m;
      r.y = v.y * m;
    } else {
      r.x = 0;
      r.y = 0;
    }
    return r;
  }
};

// One Stockham stage over the workgroup's BZ columns: butterfly j of a
// column reads points j + r N/R, multiplies by W_{NS R}^{(j mod NS) r} and
// writes points (j / NS) NS R + (j mod NS) + r NS.  The first stage reads
// `gs` (through the load operator), the last writes `gd`.
template <typename C2, int N, int R, int NS, int BZ, int LD, int TH, bool INV, bool FROM_GLOBAL,
          bool TO_GLOBAL, typename Op>
__device__ __forceinline__ void fft_stage(const C2* gs, C2* gd, long long stride, bool ok,
                                          C2* __restrict__ lds, const C2* __restrict__ tw,
                                          int c, int grp, const Op& op) {
  constexpr int G = TH / BZ;  // threads per column
  constexpr int NBF = N / R;            // butterflies per column
  constexpr int PER = (NBF + G - 1) / G;
  C2 v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int j = grp + p * G;
    if (NBF % G == 0 || j < NBF) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int n = j + r * NBF;
        if constexpr (FROM_GLOBAL) {
          if (ok) {
            v[p][r] = op(gs[(long long)n * stride], n);
          } else {
            v[p][r].x = 0;
            v[p][r].y = 0;
          }
        } else {
          v[p][r] = lds[n * LD + c];
        }
      }
    }
  }
  if constexpr (!FROM_GLOBAL) __syncthreads();  // every read before any write
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int j = grp + p * G;
    if (NBF % G == 0 || j < NBF) {
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; r++) v[p][r] = ctw<INV>(v[p][r], tw[k * r * (N / (NS * R))]);
      }
      dft<R, INV>(v[p]);
      const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int n = base + r * NS;
        if constexpr (TO_GLOBAL) {
          if (ok) gd[(long long)n * stride] = v[p][r];
        } else {
          lds[n * LD + c] = v[p][r];
        }
      }
    }
  }
  if constexpr (!TO_GLOBAL) __syncthreads();
}

template <typename C2, int N, int BZ, int LD, int TH, bool INV, int S, int NS, typename Op>
__device__ __forceinline__ void fft_stages(const C2* gs, C2* gd, long long stride, bool ok,
                                           C2* __restrict__ lds, const C2* __restrict__ tw,
                                           int c, int grp, const Op& op) {
  constexpr int NST = fft_nstages<N>();
  constexpr int R = fft_radix<N>(S);
  fft_stage<C2, N, R, NS, BZ, LD, TH, INV, S == 0, S == NST - 1>(gs, gd, stride, ok, lds, tw, c,
                                                                 grp, op);
  if constexpr (S + 1 < NST)
    fft_stages<C2, N, BZ, LD, TH, INV, S + 1, NS * R>(gs, gd, stride, ok, lds, tw, c, grp, op);
}

template <typename T, int N>
constexpr size_t fft_lds_bytes() {
  using C2 = typename cplx<T>::type;
  return ((size_t)N * (FftCfg<T>::BZ + FftCfg<T>::PAD) + N) * sizeof(C2);
}

// Workgroup w runs on XCD w mod 8; each XCD gets a contiguous range of
// (line, column block) pairs so that neighbouring blocks, which share
// 128-byte lines when a row is not line-aligned, meet in one L2.
template <typename T, int N, bool INV, typename Op, int TH>
__global__ __launch_bounds__(TH) void k_fft_axis(const typename cplx<T>::type* src,
                                                          typename cplx<T>::type* dst,
                                                          long long stride, long long ncol,
                                                          long long line_stride, int nbx,
                                                          long long nblocks,
                                                          const typename cplx<T>::type* tw,
                                                          Op op) {
  using C2 = typename cplx<T>::type;
  constexpr int BZ = FftCfg<T>::BZ;
  constexpr int LD = BZ + FftCfg<T>::PAD;
  extern __shared__ __align__(16) unsigned char fft_lds_raw[];
  C2* lds = reinterpret_cast<C2*>(fft_lds_raw);
  C2* tws = lds + N * LD;
  const long long w = blockIdx.x;
  const long long q = nblocks / 8, rem = nblocks % 8;
  const long long xcd = w % 8;
  const long long logical = xcd * q + (xcd < rem ? xcd : rem) + w / 8;
  const long long line = logical / nbx;
  const long long xb = logical - line * nbx;
  const int t = threadIdx.x, c = t % BZ, grp = t / BZ;
  for (int i = t; i < N; i += TH) tws[i] = tw[i];
  __syncthreads();
  const long long col = xb * BZ + c;
  const bool ok = col < ncol;
  const long long off = line * line_stride + (ok ? col : 0);
  op.bind(ok ? col : 0);
  fft_stages<C2, N, BZ, LD, TH, INV, 0, 1>(src + off, dst + off, stride, ok, lds, tws, c, grp,
                                           op);
}

// W_N^k = exp(-2 pi i k / N), k < N, rounded from double; one table per
// (N, precision), kept for the life of the context.
template <typename T>
const typename cplx<T>::type* twiddles(exh_ctx* c, int N) {
  using C2 = typename cplx<T>::type;
  const int key = N * 16 + (int)sizeof(T);
  auto it = c->fft_tw.find(key);
  if (it != c->fft_tw.end()) return (const C2*)it->second;
  std::vector<C2> h(N);
  for (int k = 0; k < N; k++) {
    const double a = -2.0 * M_PI * (double)k / (double)N;
    h[k].x = (T)std::cos(a);
    h[k].y = (T)std::sin(a);
  }
  void* d = nullptr;
  EXH_CHECK_HIP(hipMalloc(&d, N * sizeof(C2)));
  c->fft_tw[key] = d;
  EXH_CHECK_HIP(hipMemcpyAsync(d, h.data(), N * sizeof(C2), hipMemcpyHostToDevice, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));  // h is a host temporary
  return (const C2*)d;
}

template <typename T, int N, bool INV, typename Op>
void launch_axis(exh_ctx* c, const void* src, void* dst, long long stride, long long ncol,
                 long long line_stride, long long nlines, const Op& op) {
  constexpr int BZ = FftCfg<T>::BZ;
  const int nbx = (int)((ncol + BZ - 1) / BZ);
  const long long nblocks = (long long)nbx * nlines;
  EXH_REQUIRE(nblocks < (1LL << 31), EXH_ERR_ARG, "FFT pass too large");
  const size_t lds = fft_lds_bytes<T, N>();
  // the pass that applies the Poisson factor (double arithmetic per point)
  // runs with twice the waves where every thread still has radix-8
  // butterflies (N / 8 >= threads per column), which hide that work: 1LPT
  // and 2LPT inverse transforms at nd = 512 0.99 -> 0.86 ms; the plain
  // passes gain nothing
  constexpr bool WIDE =
      std::is_same<Op, OpPoisson<T>>::value && N / 8 >= 2 * AXIS_THREADS / BZ;
  constexpr int TH = WIDE ? 2 * AXIS_THREADS : AXIS_THREADS;
  auto kern = k_fft_axis<T, N, INV, Op, TH>;
  EXH_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
  hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(TH), lds, c->stream,
                     (const typename cplx<T>::type*)src, (typename cplx<T>::type*)dst, stride,
                     ncol, line_stride, nbx, nblocks, twiddles<T>(c, N), op);
  EXH_LAUNCH_CHECK();
}

template <typename T, bool INV, typename Op>
void axis_dispatch(exh_ctx* c, int N, const void* src, void* dst, long long stride,
                   long long ncol, long long line_stride, long long nlines, const Op& op) {
  switch (N) {
#define EXH_AXIS_CASE(NN) \
  case NN: return launch_axis<T, NN, INV>(c, src, dst, stride, ncol, line_stride, nlines, op);
    EXH_AXIS_CASE(16)
    EXH_AXIS_CASE(32)
    EXH_AXIS_CASE(64)
    EXH_AXIS_CASE(128)
    EXH_AXIS_CASE(256)
    EXH_AXIS_CASE(512)
    EXH_AXIS_CASE(1024)
    EXH_AXIS_CASE(48)
    EXH_AXIS_CASE(96)
    EXH_AXIS_CASE(192)
    EXH_AXIS_CASE(384)
    EXH_AXIS_CASE(768)
#undef EXH_AXIS_CASE
    default: EXH_REQUIRE(false, EXH_ERR_ARG, "no strided FFT pass of length %d", N);
  }
}

template <typename T>
OpPoisson<T> poisson_op(const PoissonParams& p) {
  OpPoisson<T> o{};
  o.n0 = p.n0;
  o.n1 = p.n1;
  o.n2 = p.n2;
  o.nz2 = p.n2 / 2 + 1;
  o.kl0 = (T)p.kl0;
  o.kl1 = (T)p.kl1;
  o.kl2 = (T)p.kl2;
  o.k_smooth = (T)p.k_smooth;
  o.pre = (T)p.pre;
  o.scale = (T)p.scale;
  // |k| <= k_smooth as k^2 <= k_smooth^2 (a negative cut keeps no mode)
  o.ks2 = o.k_smooth < (T)0 ? -1.0 : (double)o.k_smooth * (double)o.k_smooth;
  return o;
}

// ---------------------------------------------------------------------------
// Real transforms along the contiguous axis (length N, M = N/2): the N reals
// of a row read as M complex points z_n = x_2n + i x_2n+1, an M-point
// Stockham FFT in LDS, and the even/odd split
//   forward:  X_k = (A + B)/2 - i W_N^k (A - B)/2,  A = Z_k, B = conj(Z_{M-k})
//   inverse:  Z_k = (A + B) + i W_N^-k (A - B),      A = X_k, B = conj(X_{M-k})
// (k = 0..M forward, 0..M-1 inverse; the inverse ignores the imaginary
// parts of X_0 and X_M as a real inverse transform does).  A workgroup owns
// RB rows, G = min(M/8, 64) threads per row, lanes along the row, so every
// global access is contiguous.  LDS rows are padded by one element in nine.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rpad(int n) { return n + (n >> 3); }

template <typename C2, int M, int R, int NS, int G, bool INV, bool FROM_GLOBAL, bool TO_GLOBAL>
__device__ __forceinline__ void row_stage(const C2* gin, C2* gout, bool ok, C2* __restrict__ lrow,
                                          const C2* __restrict__ tw, int j0, int twstep) {
  constexpr int NBF = M / R;
  constexpr int PER = (NBF + G - 1) / G;
  C2 v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int j = j0 + p * G;
    if (NBF % G == 0 || j < NBF) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int n = j + r * NBF;
        v[p][r] = FROM_GLOBAL ? gin[n] : lrow[rpad(n)];
      }
    }
  }
  if constexpr (!FROM_GLOBAL) __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int j = j0 + p * G;
    if (NBF % G == 0 || j < NBF) {
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; r++)
          v[p][r] = ctw<INV>(v[p][r], tw[k * r * (M / (NS * R)) * twstep]);
      }
      dft<R, INV>(v[p]);
      const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int n = base + r * NS;
        if constexpr (TO_GLOBAL) {
          if (ok) gout[n] = v[p][r];
        } else
          lrow[rpad(n)] = v[p][r];
      }
    }
  }
  if constexpr (!TO_GLOBAL) __syncthreads();
}

template <typename C2, int M, int G, bool INV, bool FIRST_GLOBAL, bool LAST_GLOBAL, int S, int NS>
__device__ __forceinline__ void row_stages(const C2* gin, C2* gout, bool ok, C2* __restrict__ lrow,
                                           const C2* __restrict__ tw, int j0, int twstep) {
  constexpr int NST = fft_nstages<M>();
  constexpr int R = fft_radix<M>(S);
  row_stage<C2, M, R, NS, G, INV, FIRST_GLOBAL && S == 0, LAST_GLOBAL && S == NST - 1>(
      gin, gout, ok, lrow, tw, j0, twstep);
  if constexpr (S + 1 < NST)
    row_stages<C2, M, G, INV, FIRST_GLOBAL, LAST_GLOBAL, S + 1, NS * R>(gin, gout, ok, lrow, tw,
                                                                      j0, twstep);
}

constexpr int pow2_floor(int n) { return n < 2 ? 1 : 2 * pow2_floor(n / 2); }

template <int M>
struct RowCfg {
  // threads per row: a power of two, so that the rows tile the workgroup
  static constexpr int G = pow2_floor(M / 8 < 64 ? M / 8 : 64);
  static constexpr int RB = FFT_THREADS / G;          // rows per workgroup
  static constexpr int LDR = M + M / 8;               // padded LDS row
};

// r2c: real rows (pitch 2M) -> half-complex rows (pitch M+1).
// c2r: half-complex rows (pitch M+1) -> real rows (pitch 2M), unnormalised.
template <typename T, int M, bool INV>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_rows(const void* in, void* out,
                                                          long long nrows,
                                                          const typename cplx<T>::type* tw) {
  using C2 = typename cplx<T>::type;
  using RC = RowCfg<M>;
  constexpr int N = 2 * M;
  extern __shared__ __align__(16) unsigned char fft_lds_raw[];
  C2* lds = reinterpret_cast<C2*>(fft_lds_raw);
  C2* tws = lds + RC::RB * RC::LDR;
  const int t = threadIdx.x, rl = t / RC::G, j0 = t % RC::G;
  for (int i = t; i < N; i += FFT_THREADS) tws[i] = tw[i];
  const long long row = (long long)blockIdx.x * RC::RB + rl;
  const bool ok = row < nrows;
  const long long rr = ok ? row : 0;
  C2* lrow = lds + rl * RC::LDR;
  if constexpr (!INV) {
    const C2* zin = reinterpret_cast<const C2*>(static_cast<const T*>(in)) + rr * M;
    C2* xout = static_cast<C2*>(out) + rr * (M + 1);
    __syncthreads();  // twiddles
    // M-point forward FFT of z (the twiddles W_M^k are W_N^2k)
    row_stages<C2, M, RC::G, false, true, false, 0, 1>(zin, nullptr, ok, lrow, tws, j0, 2);
    for (int k = j0; k <= M; k += RC::G) {
      const C2 A = lrow[rpad(k == M ? 0 : k)];
      C2 B = lrow[rpad(k == 0 ? 0 : M - k)];
      B.y = -B.y;
      C2 E, D, O, X;
      E.x = (A.x + B.x) * (T)0.5;
      E.y = (A.y + B.y) * (T)0.5;
      D.x = (A.x - B.x) * (T)0.5;
      D.y = (A.y - B.y) * (T)0.5;
      O.x = D.y;  // -i D
      O.y = -D.x;
      const C2 WO = ctw<false>(O, tws[k]);
      X.x = E.x + WO.x;
      X.y = E.y + WO.y;
      if (ok) xout[k] = X;
    }
  } else {
    const C2* xin = static_cast<const C2*>(in) + rr * (M + 1);
    C2* zout = reinterpret_cast<C2*>(static_cast<T*>(out)) + rr * M;
    __syncthreads();  // twiddles
    for (int k = j0; k < M; k += RC::G) {
      C2 A = ok ? xin[k] : C2{};
      C2 B = ok ? xin[M - k] : C2{};
      if (k == 0) {
        A.y = 0;
        B.y = 0;
      }
      B.y = -B.y;
      C2 S, D, Z;
      S.x = A.x + B.x;
      S.y = A.y + B.y;
      D.x = A.x - B.x;
      D.y = A.y - B.y;
      const C2 WD = ctw<true>(D, tws[k]);  // W_N^-k (A - B)
      Z.x = S.x - WD.y;                      // + i W D
      Z.y = S.y + WD.x;
      lrow[rpad(k)] = Z;
    }
    __syncthreads();
    // M-point inverse FFT, the last stage writes the real row
    row_stages<C2, M, RC::G, true, false, true, 0, 1>(nullptr, zout, ok, lrow, tws, j0, 2);
  }
}

template <typename T, int M, bool INV>
void launch_rows(exh_ctx* c, const void* in, void* out, long long nrows) {
  using C2 = typename cplx<T>::type;
  using RC = RowCfg<M>;
  const long long nblk = (nrows + RC::RB - 1) / RC::RB;
  EXH_REQUIRE(nblk < (1LL << 31), EXH_ERR_ARG, "FFT row pass too large");
  const size_t lds = ((size_t)RC::RB * RC::LDR + 2 * M) * sizeof(C2);
  auto kern = k_fft_rows<T, M, INV>;
  EXH_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
  hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(FFT_THREADS), lds, c->stream, in, out, nrows,
                     twiddles<T>(c, 2 * M));
  EXH_LAUNCH_CHECK();
}

template <typename T, bool INV>
void rows_dispatch(exh_ctx* c, int N, const void* in, void* out, long long nrows) {
  switch (N / 2) {
#define EXH_ROWS_CASE(MM) \
  case MM: return launch_rows<T, MM, INV>(c, in, out, nrows);
    EXH_ROWS_CASE(16)
    EXH_ROWS_CASE(32)
    EXH_ROWS_CASE(64)
    EXH_ROWS_CASE(128)
    EXH_ROWS_CASE(256)
    EXH_ROWS_CASE(512)
    EXH_ROWS_CASE(24)
    EXH_ROWS_CASE(48)
    EXH_ROWS_CASE(96)
    EXH_ROWS_CASE(192)
    EXH_ROWS_CASE(384)
#undef EXH_ROWS_CASE
    default: EXH_REQUIRE(false, EXH_ERR_ARG, "no real row transform of length %d", N);
  }
}

}  // namespace

// 2^a, or 3 * 2^a with a >= 4
static bool fft_length_ok(int n, int min_pow2) {
  const int p = n % 3 == 0 ? n / 3 : n;
  return n <= 1024 && p >= (n % 3 == 0 ? 16 : min_pow2) && (p & (p - 1)) == 0;
}

bool fft_axis_supported(int n) { return fft_length_ok(n, 16); }

void fft_axis_pass(exh_ctx* c, int prec, bool inverse, int N, void* data, long long stride,
                   long long ncol, long long line_stride, long long nlines) {
  const OpNone op;
  if (prec == 4) {
    if (inverse)
      axis_dispatch<float, true>(c, N, data, data, stride, ncol, line_stride, nlines, op);
    else
      axis_dispatch<float, false>(c, N, data, data, stride, ncol, line_stride, nlines, op);
  } else {
    if (inverse)
      axis_dispatch<double, true>(c, N, data, data, stride, ncol, line_stride, nlines, op);
    else
      axis_dispatch<double, false>(c, N, data, data, stride, ncol, line_stride, nlines, op);
  }
}

void fft_axis0_inverse_poisson(exh_ctx* c, int prec, const void* src, void* dst,
                               const PoissonParams& p) {
  const long long plane = (long long)p.n1 * (p.n2 / 2 + 1);
  if (prec == 4)
    axis_dispatch<float, true>(c, p.n0, src, dst, plane, plane, (long long)p.n0 * plane, 1,
                               poisson_op<float>(p));
  else
    axis_dispatch<double, true>(c, p.n0, src, dst, plane, plane, (long long)p.n0 * plane, 1,
                                poisson_op<double>(p));
}

bool fft_rows_supported(int n) { return fft_length_ok(n, 32); }

void fft_rows_pass(exh_ctx* c, int prec, bool inverse, int N, const void* in, void* out,
                   long long nrows) {
  if (prec == 4) {
    if (inverse)
      rows_dispatch<float, true>(c, N, in, out, nrows);
    else
      rows_dispatch<float, false>(c, N, in, out, nrows);
  } else {
    if (inverse)
      rows_dispatch<double, true>(c, N, in, out, nrows);
    else
      rows_dispatch<double, false>(c, N, in, out, nrows);
  }
}

void fft_axis_release(exh_ctx* c) {
  for (auto& kv : c->fft_tw) (void)hipFree(kv.second);
  c->fft_tw.clear();
}

}  // namespace exh
