// Exact emulation of the reference's peak sort on the GPU.
//
// The reference sorts the peaks with a recursive Lomuto quicksort: pivot =
// last element, elements "<= pivot" swapped to the front in scan order,
// ascending and unstable (find_halos.c:159-197).  Ties among peak densities
// are frequent in float, and their final order decides the halo numbering
// and which uniform each halo's mass draw uses, so the permutation must be
// reproduced exactly, not merely a sorted order.
//
// One Lomuto partition of a[lo..hi] (pivot p = a[hi]) is computed in
// parallel from a closed form:
//  * the "<=" elements keep their scan order and land at lo + rank;
//  * the ">" elements form a FIFO queue: appended when scanned, and rotated
//    (front moved to the back) by every "<=" element scanned after the first
//    ">" element, plus once by the final pivot swap.  With r rotations and
//    A(q) = number of ">" elements appended before rotation q, an element
//    appended at stream index s0 = t + R_t (t its rank among ">", R_t the
//    rotations before it) is re-appended at s + A(s) while s < r; its final
//    slot is lo + c + 1 + (s - r), c = number of "<=" elements.
// Every stream index below r is visited once, so a partition costs O(n).
//
// Large segments (> SMALL) are processed level by level with tile kernels
// across the whole device; segments of 2..SMALL elements are finished by one
// wavefront each in LDS with the same partition rule.
#include "common.hpp"
#include "kernels.hpp"
#include "util.hpp"

namespace exh {

static constexpr int TILE = 2048;  // elements per tile, 256 threads x 8
static constexpr int PER_THREAD = TILE / 256;

struct SegInfo {
  long long lo, hi;  // inclusive range, pivot at hi
  long long c, r, j0;
  long long first_tile;
};

// ---------------------------------------------------------------------------
// level kernels
// ---------------------------------------------------------------------------
__global__ void k_seg_ntiles(const long long* seg_lo, const long long* seg_hi, int nseg,
                             long long* ntiles, long long* j0) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  long long len = seg_hi[s] - seg_lo[s];  // elements lo..hi-1
  ntiles[s] = (len + TILE - 1) / TILE;
  j0[s] = seg_hi[s];  // "no > element" until found
}

__global__ void k_fill_int(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ inline int find_seg(const long long* tile_off, int nseg, long long t) {
  int lo = 0, hi = nseg;  // tile_off[lo] <= t < tile_off[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (tile_off[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename K>
__global__ void k_tile_count(const K* keys, const long long* seg_lo, const long long* seg_hi,
                             const long long* tile_off, int nseg, long long* tile_le,
                             unsigned long long* j0) {
  __shared__ long long sh[4];
  long long t = blockIdx.x;
  int s = find_seg(tile_off, nseg, t);
  long long lo = seg_lo[s], hi = seg_hi[s];
  K pivot = keys[hi];
  long long start = lo + (t - tile_off[s]) * TILE;
  long long cnt = 0, first_gt = hi;
  for (int q = 0; q < PER_THREAD; q++) {
    long long j = start + (long long)threadIdx.x * PER_THREAD + q;
    if (j < hi) {
      if (keys[j] <= pivot)
        cnt++;
      else if (j < first_gt)
        first_gt = j;
    }
  }
  long long tot;
  block_exclusive_256<long long>(cnt, sh, &tot);
  if (threadIdx.x == 0) tile_le[t] = tot;
  // first ">" element of the segment
  if (first_gt < hi) atomicMin(&j0[s], (unsigned long long)first_gt);
}

template <typename K>
__global__ void k_seg_final(const K* keysA, const int* valsA, K* keysB, int* valsB,
                            const long long* seg_lo, const long long* seg_hi,
                            const long long* tile_off, const long long* tile_pre,
                            const unsigned long long* j0s, int nseg, SegInfo* info,
                            long long* arot, long long small_max, long long* next_lo,
                            long long* next_hi, int* n_next, long long* small_lo,
                            long long* small_hi, int* n_small) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  long long lo = seg_lo[s], hi = seg_hi[s];
  long long ft = tile_off[s], lt = tile_off[s + 1];
  long long c = tile_pre[lt] - tile_pre[ft];
  long long m = (hi - lo) - c;
  long long j0 = (long long)j0s[s];
  long long r = m > 0 ? (c - (j0 - lo)) + 1 : 0;
  SegInfo in;
  in.lo = lo;
  in.hi = hi;
  in.c = c;
  in.r = r;
  in.j0 = j0;
  in.first_tile = ft;
  info[s] = in;
  keysB[lo + c] = keysA[hi];
  valsB[lo + c] = valsA[hi];
  if (m > 0) arot[lo + r - 1] = m;
  // children
  long long cl[2] = {lo, lo + c + 1}, ch[2] = {lo + c - 1, hi};
  for (int q = 0; q < 2; q++) {
    long long len = ch[q] - cl[q] + 1;
    if (len > small_max) {
      int p = atomicAdd(n_next, 1);
      next_lo[p] = cl[q];
      next_hi[p] = ch[q];
    } else if (len >= 2) {
      int p = atomicAdd(n_small, 1);
      small_lo[p] = cl[q];
      small_hi[p] = ch[q];
    }
  }
}

template <typename K>
__global__ void k_tile_scatter(const K* keysA, const int* valsA, K* keysB, int* valsB,
                               const long long* tile_off, const long long* tile_pre, int nseg,
                               const SegInfo* info, long long* arot, long long* s0) {
  __shared__ long long sh[4];
  long long t = blockIdx.x;
  int s = find_seg(tile_off, nseg, t);
  SegInfo g = info[s];
  K pivot = keysA[g.hi];
  long long start = g.lo + (t - g.first_tile) * TILE + (long long)threadIdx.x * PER_THREAD;
  bool le[PER_THREAD];
  long long cnt = 0;
  for (int q = 0; q < PER_THREAD; q++) {
    long long j = start + q;
    le[q] = (j < g.hi) && (keysA[j] <= pivot);
    cnt += le[q] ? 1 : 0;
  }
  long long tot;
  long long rank = block_exclusive_256<long long>(cnt, sh, &tot) + tile_pre[t] - tile_pre[g.first_tile];
  for (int q = 0; q < PER_THREAD; q++) {
    long long j = start + q;
    if (j >= g.hi) break;
    if (le[q]) {
      keysB[g.lo + rank] = keysA[j];
      valsB[g.lo + rank] = valsA[j];
      if (g.r > 0 && j > g.j0) arot[g.lo + rank - (g.j0 - g.lo)] = (j - g.lo) - rank;
      rank++;
    } else {
      long long gt_before = (j - g.lo) - rank;
      s0[j] = gt_before + (rank - (g.j0 - g.lo));
    }
  }
}

// Chase with at most `cap` dependent steps.  Elements that finish are
// placed and their s0 set to LLONG_MAX; the others keep their progress in
// s0 and flag their segment for pointer doubling (an extreme pivot makes
// the queue short and its elements rotate many times).
static constexpr int CHASE_CAP = 48;

template <typename K>
__global__ void k_tile_chase(const K* keysA, const int* valsA, K* keysB, int* valsB,
                             const long long* tile_off, int nseg, const SegInfo* info,
                             const long long* arot, long long* s0, int* seg_need,
                             int* need_any) {
  long long t = blockIdx.x;
  int sg = find_seg(tile_off, nseg, t);
  SegInfo g = info[sg];
  K pivot = keysA[g.hi];
  long long start = g.lo + (t - g.first_tile) * TILE;
  for (int q = 0; q < PER_THREAD; q++) {
    long long j = start + (long long)q * 256 + threadIdx.x;
    if (j >= g.hi) continue;
    K kv = keysA[j];
    if (kv <= pivot) continue;
    long long s = s0[j];
    int steps = 0;
    while (s < g.r && steps < CHASE_CAP) {
      s += max(arot[g.lo + s], 1LL);  // every entry is >= 1
      steps++;
    }
    if (s < g.r) {
      s0[j] = s;
      seg_need[sg] = 1;
      *need_any = 1;
      continue;
    }
    s0[j] = LLONG_MAX;
    long long dst = g.lo + g.c + 1 + (s - g.r);
    keysB[dst] = kv;
    valsB[dst] = valsA[j];
  }
}

// Pointer doubling on nxt(q) = q + A(q) over the stream indices of flagged
// segments: after k rounds one entry covers 2^k chase steps.  Values are
// stored relative to the segment start.
__global__ void k_double_init(const long long* tile_off, int nseg, const SegInfo* info,
                              const int* seg_need, const long long* arot, int* nxt) {
  long long t = blockIdx.x;
  int sg = find_seg(tile_off, nseg, t);
  if (!seg_need[sg]) return;
  SegInfo g = info[sg];
  long long q0 = (t - g.first_tile) * TILE;
  bool last = t == tile_off[sg + 1] - 1;
  long long q1 = last ? g.r : min(q0 + TILE, g.r);
  for (long long q = q0 + threadIdx.x; q < q1; q += 256) {
    nxt[g.lo + q] = (int)(q + max(arot[g.lo + q], 1LL));  // >= r: left the stream prefix
  }
}

__global__ void k_double_round(const long long* tile_off, int nseg, const SegInfo* info,
                               const int* seg_need, const int* src, int* dst, int* changed) {
  long long t = blockIdx.x;
  int sg = find_seg(tile_off, nseg, t);
  if (!seg_need[sg]) return;
  SegInfo g = info[sg];
  long long q0 = (t - g.first_tile) * TILE;
  bool last = t == tile_off[sg + 1] - 1;
  long long q1 = last ? g.r : min(q0 + TILE, g.r);
  bool open = false;
  for (long long q = q0 + threadIdx.x; q < q1; q += 256) {
    int v = src[g.lo + q];
    int w = v < g.r ? src[g.lo + v] : v;
    dst[g.lo + q] = w;
    open |= w < g.r;
  }
  if (open) *changed = 1;
}

template <typename K>
__global__ void k_tile_chase2(const K* keysA, const int* valsA, K* keysB, int* valsB,
                              const long long* tile_off, int nseg, const SegInfo* info,
                              const int* seg_need, const int* nxt, const long long* s0) {
  long long t = blockIdx.x;
  int sg = find_seg(tile_off, nseg, t);
  if (!seg_need[sg]) return;
  SegInfo g = info[sg];
  const K pivot = keysA[g.hi];
  long long start = g.lo + (t - g.first_tile) * TILE;
  for (int q = 0; q < PER_THREAD; q++) {
    long long j = start + (long long)q * 256 + threadIdx.x;
    if (j >= g.hi) continue;
    if (keysA[j] <= pivot) continue;
    long long s = s0[j];
    if (s >= g.r) continue;  // placed by the capped chase
    for (int guard = 0; s < g.r && guard < 64; guard++) s = nxt[g.lo + s];
    long long dst = g.lo + g.c + 1 + (s - g.r);
    keysB[dst] = keysA[j];
    valsB[dst] = valsA[j];
  }
}

template <typename K>
__global__ void k_tile_copyback(K* keysA, int* valsA, const K* keysB, const int* valsB,
                                const long long* tile_off, int nseg, const SegInfo* info) {
  long long t = blockIdx.x;
  int sg = find_seg(tile_off, nseg, t);
  SegInfo g = info[sg];
  long long start = g.lo + (t - g.first_tile) * TILE;
  for (int q = 0; q < PER_THREAD; q++) {
    long long j = start + (long long)q * 256 + threadIdx.x;
    if (j > g.hi) continue;
    if (j == g.hi && t != tile_off[sg + 1] - 1) continue;  // hi copied by the last tile
    keysA[j] = keysB[j];
    valsA[j] = valsB[j];
  }
  // the last tile also covers hi when the tile ends before it
  if (t == tile_off[sg + 1] - 1 && threadIdx.x == 0) {
    long long tile_end = start + TILE;
    if (tile_end <= g.hi) {
      keysA[g.hi] = keysB[g.hi];
      valsA[g.hi] = valsB[g.hi];
    }
  }
}

// ---------------------------------------------------------------------------
// small segments: one wavefront each, LDS resident
// ---------------------------------------------------------------------------
static constexpr int TINY = 64;

// Scalar Lomuto quicksort of sk/sv[lo..hi] (LDS) by one lane: the
// reference's partition verbatim.  The smaller side is sorted next and the
// larger pushed, so the stack depth stays below log2(TINY) + 1.
template <typename K>
__device__ inline void lane_quicksort(K* sk, int* sv, int lo, int hi) {
  int st_lo[8], st_hi[8];
  int top = 0;
  for (;;) {
    while (lo < hi) {
      const K pivot = sk[hi];
      int i = lo - 1;
      for (int j = lo; j < hi; j++) {
        const K kj = sk[j];
        if (kj <= pivot) {
          i++;
          const K ki = sk[i];
          sk[i] = kj;
          sk[j] = ki;
          const int vi = sv[i];
          sv[i] = sv[j];
          sv[j] = vi;
        }
      }
      {
        const K t = sk[i + 1];
        sk[i + 1] = sk[hi];
        sk[hi] = t;
        const int v = sv[i + 1];
        sv[i + 1] = sv[hi];
        sv[hi] = v;
      }
      const int p = i + 1;
      int al = lo, ah = p - 1, bl = p + 1, bh = hi;
      if (ah - al > bh - bl) {  // [al, ah] becomes the smaller side
        int x = al, y = ah;
        al = bl;
        ah = bh;
        bl = x;
        bh = y;
      }
      if (bh > bl) {
        st_lo[top] = bl;
        st_hi[top] = bh;
        top++;
      }
      lo = al;
      hi = ah;
    }
    if (top == 0) break;
    top--;
    lo = st_lo[top];
    hi = st_hi[top];
  }
}

// One wavefront per segment of 2..SMALL elements, resident in LDS.
// Sub-segments longer than TINY are partitioned by the whole wave with the
// closed-form Lomuto rule (see the file comment); shorter ones are queued
// and then sorted one per lane by lane_quicksort.
template <typename K, int SMALL>
__global__ __launch_bounds__(64) void k_small_sort(K* keys, int* vals, const long long* seg_lo,
                                                   const long long* seg_hi) {
  __shared__ K sk[SMALL], tk[SMALL];
  __shared__ int sv[SMALL], tv[SMALL], arot[SMALL], s0[SMALL];
  __shared__ int st_lo[64], st_hi[64];
  __shared__ short tl_lo[SMALL / 2 + 1], tl_hi[SMALL / 2 + 1];
  __shared__ int ntiny;
  const int lane = threadIdx.x;
  const long long base = seg_lo[blockIdx.x];
  const int n = (int)(seg_hi[blockIdx.x] - base + 1);
  for (int i = lane; i < n; i += 64) {
    sk[i] = keys[base + i];
    sv[i] = vals[base + i];
  }
  int top = 0;
  if (lane == 0) {
    ntiny = 0;
    if (n > TINY) {
      st_lo[0] = 0;
      st_hi[0] = n - 1;
    } else {
      tl_lo[0] = 0;
      tl_hi[0] = (short)(n - 1);
      ntiny = 1;
    }
  }
  top = n > TINY ? 1 : 0;
  __syncthreads();
  while (top > 0) {
    top--;
    const int lo = st_lo[top], hi = st_hi[top];
    __syncthreads();
    const K pivot = sk[hi];
    // pass 1: count "<=" and find the first ">"
    int c = 0, j0 = hi;
    for (int b = lo; b < hi; b += 64) {
      int j = b + lane;
      bool le = (j < hi) && (sk[j] <= pivot);
      bool gt = (j < hi) && !le;
      unsigned long long mle = __ballot(le), mgt = __ballot(gt);
      if (j0 == hi && mgt) j0 = b + __ffsll((long long)mgt) - 1;
      c += __popcll(mle);
    }
    const int m = (hi - lo) - c;
    const int r = m > 0 ? (c - (j0 - lo)) + 1 : 0;
    // pass 2: place "<=" elements, record rotations and queue entries
    int run = 0;
    for (int b = lo; b < hi; b += 64) {
      int j = b + lane;
      bool le = (j < hi) && (sk[j] <= pivot);
      unsigned long long mle = __ballot(le);
      int rank = run + __popcll(mle & lanemask_lt());
      if (j < hi) {
        if (le) {
          tk[lo + rank] = sk[j];
          tv[lo + rank] = sv[j];
          if (r > 0 && j > j0) arot[rank - (j0 - lo)] = (j - lo) - rank;
        } else {
          s0[j] = ((j - lo) - rank) + (rank - (j0 - lo));
        }
      }
      run += __popcll(mle);
    }
    if (lane == 0) {
      tk[lo + c] = sk[hi];
      tv[lo + c] = sv[hi];
      if (m > 0) arot[r - 1] = m;
    }
    __syncthreads();
    // pass 3: chase the queue
    for (int b = lo; b < hi; b += 64) {
      int j = b + lane;
      if (j < hi && !(sk[j] <= pivot)) {
        int s = s0[j];
        for (int guard = 0; s < r && guard <= r; guard++) s += max(arot[s], 1);
        tk[lo + c + 1 + (s - r)] = sk[j];
        tv[lo + c + 1 + (s - r)] = sv[j];
      }
    }
    __syncthreads();
    for (int i = lo + lane; i <= hi; i += 64) {
      sk[i] = tk[i];
      sv[i] = tv[i];
    }
    // children: long ones on the stack (larger first, so the smaller is
    // partitioned next), short ones to the per-lane list
    int al = lo, ah = lo + c - 1, bl = lo + c + 1, bh = hi;
    if (ah - al > bh - bl) {
      int x = al, y = ah;
      al = bl;
      ah = bh;
      bl = x;
      bh = y;
    }
    int pushed = 0;
    if (lane == 0) {
      int tt = top;
      if (bh - bl + 1 > TINY) {
        st_lo[tt] = bl;
        st_hi[tt] = bh;
        tt++;
      } else if (bh > bl) {
        tl_lo[ntiny] = (short)bl;
        tl_hi[ntiny] = (short)bh;
        ntiny++;
      }
      if (ah - al + 1 > TINY) {
        st_lo[tt] = al;
        st_hi[tt] = ah;
        tt++;
      } else if (ah > al) {
        tl_lo[ntiny] = (short)al;
        tl_hi[ntiny] = (short)ah;
        ntiny++;
      }
    }
    pushed = (bh - bl + 1 > TINY) + (ah - al + 1 > TINY);
    top += pushed;
    __syncthreads();
  }
  __syncthreads();
  for (int q = lane; q < ntiny; q += 64) lane_quicksort<K>(sk, sv, tl_lo[q], tl_hi[q]);
  __syncthreads();
  for (int i = lane; i < n; i += 64) {
    keys[base + i] = sk[i];
    vals[base + i] = sv[i];
  }
}

template <typename K>
struct SmallCfg;
template <>
struct SmallCfg<float> {
  static constexpr int N = 2048;
};
template <>
struct SmallCfg<double> {
  static constexpr int N = 1024;
};

template <typename K>
void lomuto_sort(exh_ctx* c, K* keys, int* vals, long long n) {
  TimedScope ts(c, "peak_sort");
  constexpr int SMALL = SmallCfg<K>::N;
  c->stat_sort_levels = 0;
  c->stat_sort_doubling = 0;
  if (n < 2) return;
  // scratch layout in BUF_SORT_AUX
  const long long maxseg = n / (SMALL + 1) * 2 + 4;
  const long long maxtiles = n / TILE + maxseg + 4;
  const long long maxsmall = n / 2 + 4;
  size_t bytes = 0;
  auto take = [&](size_t b) {
    size_t o = bytes;
    bytes += (b + 255) & ~(size_t)255;
    return o;
  };
  size_t o_klo = take(maxseg * 8), o_khi = take(maxseg * 8), o_nlo = take(maxseg * 8),
         o_nhi = take(maxseg * 8), o_ntiles = take((maxseg + 1) * 8),
         o_toff = take((maxseg + 2) * 8), o_j0 = take(maxseg * 8),
         o_tle = take((maxtiles + 1) * 8), o_tpre = take((maxtiles + 2) * 8),
         o_info = take(maxseg * sizeof(SegInfo)), o_slo = take(maxsmall * 8),
         o_shi = take(maxsmall * 8), o_cnt = take(64), o_arot = take((size_t)n * 8),
         o_s0 = take((size_t)n * 8), o_need = take(maxseg * 4 + 64), o_nxa = take((size_t)n * 4 + 64),
         o_nxb = take((size_t)n * 4 + 64);
  char* aux = (char*)ctx_buf(c, BUF_SORT_AUX, bytes);
  K* kB = (K*)ctx_buf(c, BUF_SORT_K, (size_t)n * sizeof(K));
  int* vB = (int*)ctx_buf(c, BUF_SORT_V, (size_t)n * sizeof(int));
  long long* cur_lo = (long long*)(aux + o_klo);
  long long* cur_hi = (long long*)(aux + o_khi);
  long long* nxt_lo = (long long*)(aux + o_nlo);
  long long* nxt_hi = (long long*)(aux + o_nhi);
  long long* ntiles = (long long*)(aux + o_ntiles);
  long long* toff = (long long*)(aux + o_toff);
  unsigned long long* j0 = (unsigned long long*)(aux + o_j0);
  long long* tle = (long long*)(aux + o_tle);
  long long* tpre = (long long*)(aux + o_tpre);
  SegInfo* info = (SegInfo*)(aux + o_info);
  long long* slo = (long long*)(aux + o_slo);
  long long* shi = (long long*)(aux + o_shi);
  int* cnt = (int*)(aux + o_cnt);  // [0] next count, [1] small count
  long long* arot = (long long*)(aux + o_arot);
  long long* s0 = (long long*)(aux + o_s0);
  int* hc = (int*)ctx_pinned(c, 64);

  int nsmall_host = 0;
  int nseg = 0;
  long long init_lo = 0, init_hi = n - 1;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, cnt, 16, 0);
  if (n > SMALL) {
    EXH_CHECK_HIP(hipMemcpyAsync(cur_lo, &init_lo, 8, hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipMemcpyAsync(cur_hi, &init_hi, 8, hipMemcpyHostToDevice, c->stream));
    nseg = 1;
  } else {
    EXH_CHECK_HIP(hipMemcpyAsync(slo, &init_lo, 8, hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipMemcpyAsync(shi, &init_hi, 8, hipMemcpyHostToDevice, c->stream));
    nsmall_host = 1;
  }
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  int* seg_need = (int*)(aux + o_need);
  int* nxa = (int*)(aux + o_nxa);
  int* nxb = (int*)(aux + o_nxb);
  int* flags = cnt + 2;  // [2] any segment needs doubling, [3] doubling not converged
  while (nseg > 0) {
    c->stat_sort_levels++;
    hipLaunchKernelGGL(k_seg_ntiles, dim3(grid_for(nseg, 256)), dim3(256), 0, c->stream, cur_lo,
                       cur_hi, nseg, ntiles, (long long*)j0);
    EXH_LAUNCH_CHECK();
    exclusive_scan<long long>(c, ntiles, toff, nseg);
    long long ntile_host = 0;
    EXH_CHECK_HIP(hipMemcpyAsync(&ntile_host, toff + nseg, 8, hipMemcpyDeviceToHost, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    hipLaunchKernelGGL(k_tile_count<K>, dim3((unsigned)ntile_host), dim3(256), 0, c->stream,
                       keys, cur_lo, cur_hi, toff, nseg, tle, j0);
    EXH_LAUNCH_CHECK();
    exclusive_scan<long long>(c, tle, tpre, ntile_host);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, cnt, 1, 0);  // next count
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, flags, 2, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(nseg, 256)), dim3(256), 0, c->stream, seg_need,
                       nseg, 0);
    hipLaunchKernelGGL(k_seg_final<K>, dim3(grid_for(nseg, 256)), dim3(256), 0, c->stream, keys,
                       vals, kB, vB, cur_lo, cur_hi, toff, tpre, j0, nseg, info, arot,
                       (long long)SMALL, nxt_lo, nxt_hi, cnt, slo, shi, cnt + 1);
    EXH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_scatter<K>, dim3((unsigned)ntile_host), dim3(256), 0, c->stream,
                       keys, vals, kB, vB, toff, tpre, nseg, info, arot, s0);
    EXH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_chase<K>, dim3((unsigned)ntile_host), dim3(256), 0, c->stream,
                       keys, vals, kB, vB, toff, nseg, info, arot, s0, seg_need, flags);
    EXH_LAUNCH_CHECK();
    EXH_CHECK_HIP(hipMemcpyAsync(hc, cnt, 12, hipMemcpyDeviceToHost, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (hc[2]) {
      // extreme pivots: finish their chases by pointer doubling
      c->stat_sort_doubling++;
      hipLaunchKernelGGL(k_double_init, dim3((unsigned)ntile_host), dim3(256), 0, c->stream, toff,
                         nseg, info, seg_need, arot, nxa);
      EXH_LAUNCH_CHECK();
      for (int round = 0; round < 40; round++) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, flags + 1, 1, 0);
        hipLaunchKernelGGL(k_double_round, dim3((unsigned)ntile_host), dim3(256), 0, c->stream,
                           toff, nseg, info, seg_need, nxa, nxb, flags + 1);
        EXH_LAUNCH_CHECK();
        std::swap(nxa, nxb);
        EXH_CHECK_HIP(hipMemcpyAsync(hc + 3, flags + 1, 4, hipMemcpyDeviceToHost, c->stream));
        EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
        if (!hc[3]) break;
        EXH_REQUIRE(round < 39, EXH_ERR_INTERNAL, "peak sort: pointer doubling did not converge");
      }
      hipLaunchKernelGGL(k_tile_chase2<K>, dim3((unsigned)ntile_host), dim3(256), 0, c->stream,
                         keys, vals, kB, vB, toff, nseg, info, seg_need, nxa, s0);
      EXH_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_tile_copyback<K>, dim3((unsigned)ntile_host), dim3(256), 0, c->stream,
                       keys, vals, kB, vB, toff, nseg, info);
    EXH_LAUNCH_CHECK();
    nseg = hc[0];
    nsmall_host = hc[1];
    EXH_REQUIRE(nseg <= maxseg && nsmall_host <= maxsmall, EXH_ERR_INTERNAL,
                "peak sort: segment lists overflow (%d, %d)", nseg, nsmall_host);
    std::swap(cur_lo, nxt_lo);
    std::swap(cur_hi, nxt_hi);
  }
  if (nsmall_host > 0) {
    hipLaunchKernelGGL((k_small_sort<K, SMALL>), dim3((unsigned)nsmall_host), dim3(64), 0,
                       c->stream, keys, vals, slo, shi);
    EXH_LAUNCH_CHECK();
  }
}

template void lomuto_sort<float>(exh_ctx*, float*, int*, long long);
template void lomuto_sort<double>(exh_ctx*, double*, int*, long long);

}  // namespace exh
