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
// A chase grows roughly geometrically (s -> s / f for a pivot at quantile
// f), so chains are short except for early queue entries of segments whose
// pivot is near the top; those finish by pointer doubling.
//
// Large segments (> SMALL) are processed level by level with tile kernels
// across the whole device; segments of 2..SMALL elements are finished by one
// wavefront each in LDS with the same partition rule.
#include "common.hpp"
#include "kernels.hpp"
#include "util.hpp"

#include <cstdlib>

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

// 64-bit wave reductions
__device__ inline long long wave_min_ll(long long v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, (long long)__shfl_xor(v, o, 64));
  return v;
}

// Per tile: number of "<=" elements and the first ">" element.  Lanes
// read consecutive elements; one atomic per block for the segment's j0.
template <typename K>
__global__ __launch_bounds__(256) void k_tile_count(const K* keys, const long long* seg_lo,
                                                    const long long* seg_hi,
                                                    const long long* tile_off, int nseg,
                                                    long long* tile_le, unsigned long long* j0) {
  __shared__ int shc[4];
  __shared__ long long shj[4];
  const long long t = blockIdx.x;
  if (t >= tile_off[nseg]) {  // the grid is an upper bound of the tile count
    if (threadIdx.x == 0) tile_le[t] = 0;
    return;
  }
  const int s = find_seg(tile_off, nseg, t);
  const long long lo = seg_lo[s], hi = seg_hi[s];
  const K pivot = keys[hi];
  const long long start = lo + (t - tile_off[s]) * TILE;
  int cnt = 0;
  long long first_gt = hi;
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = start + (long long)q * 256 + threadIdx.x;
    if (j < hi) {
      if (keys[j] <= pivot)
        cnt++;
      else
        first_gt = min(first_gt, j);
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  first_gt = wave_min_ll(first_gt);
  if (lane == 0) {
    shc[w] = cnt;
    shj[w] = first_gt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    tile_le[t] = (long long)shc[0] + shc[1] + shc[2] + shc[3];
    const long long fg = min(min(shj[0], shj[1]), min(shj[2], shj[3]));
    // most tiles come after the segment's first ">": a plain read of the
    // running minimum skips their atomic on this one address
    if (fg < hi && (unsigned long long)fg < __atomic_load_n(&j0[s], __ATOMIC_RELAXED))
      atomicMin(&j0[s], (unsigned long long)fg);
  }
}

template <typename K>
__global__ void k_seg_final(const K* keysA, const int* valsA, K* keysB, int* valsB,
                            const long long* seg_lo, const long long* seg_hi,
                            const long long* tile_off, const long long* tile_pre,
                            const unsigned long long* j0s, int nseg, SegInfo* info,
                            int* arot, long long small_max, long long* next_lo,
                            long long* next_hi, int* n_next, long long* small_lo,
                            long long* small_hi, int* n_small, unsigned long long* next_elems) {
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
  if (m > 0) arot[lo + r - 1] = (int)m;
  // children
  long long cl[2] = {lo, lo + c + 1}, ch[2] = {lo + c - 1, hi};
  for (int q = 0; q < 2; q++) {
    long long len = ch[q] - cl[q] + 1;
    if (len > small_max) {
      int p = atomicAdd(n_next, 1);
      next_lo[p] = cl[q];
      next_hi[p] = ch[q];
      atomicAdd(next_elems, (unsigned long long)len);
    } else if (len >= 2) {
      int p = atomicAdd(n_small, 1);
      small_lo[p] = cl[q];
      small_hi[p] = ch[q];
    }
  }
}

// Place the "<=" elements at lo + rank, record the rotation table A and
// each ">" element's first stream index.  Wave w of a tile takes its
// elements [w * 512, (w + 1) * 512) in eight coalesced rows of 64; ranks
// come from ballots, the wave offsets from LDS.
template <typename K>
__global__ __launch_bounds__(256) void k_tile_scatter(const K* keysA, const int* valsA, K* keysB,
                                                      int* valsB, const long long* tile_off,
                                                      const long long* tile_pre, int nseg,
                                                      const SegInfo* info, int* arot, int* s0) {
  __shared__ int shw[4];
  const long long t = blockIdx.x;
  if (t >= tile_off[nseg]) return;
  const int sg = find_seg(tile_off, nseg, t);
  const SegInfo g = info[sg];
  const K pivot = keysA[g.hi];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long base = g.lo + (t - g.first_tile) * TILE + (long long)w * (TILE / 4);
  K kv[PER_THREAD];
  unsigned long long bal[PER_THREAD];
  int wcnt = 0;
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = base + q * 64 + lane;
    const bool in = j < g.hi;
    kv[q] = in ? keysA[j] : pivot;
    bal[q] = __ballot(in && kv[q] <= pivot);
    wcnt += __popcll(bal[q]);
  }
  if (lane == 0) shw[w] = wcnt;
  __syncthreads();
  long long run = tile_pre[t] - tile_pre[g.first_tile];
  for (int v = 0; v < w; v++) run += shw[v];
  const unsigned long long below = lanemask_lt();
  const long long jlead = g.j0 - g.lo;  // "<=" elements before the first ">"
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = base + q * 64 + lane;
    if (j < g.hi) {
      const long long rank = run + __popcll(bal[q] & below);
      if ((bal[q] >> lane) & 1ull) {
        keysB[g.lo + rank] = kv[q];
        valsB[g.lo + rank] = valsA[j];
        if (g.r > 0 && j > g.j0) arot[g.lo + rank - jlead] = (int)((j - g.lo) - rank);
      } else {
        const long long gt_before = (j - g.lo) - rank;
        s0[j] = (int)(gt_before + (rank - jlead));
      }
    }
    run += __popcll(bal[q]);
  }
}

// Chase with at most `cap` dependent steps per element.  A thread advances
// its eight elements' chains in lock-step, so their loads are in flight
// together.  Elements that finish are placed and their s0 set to S0_DONE;
// the others keep their progress in s0, and the block flags its segment
// for pointer doubling (one write per block, not per element).
static constexpr int S0_DONE = 0x7fffffff;

template <typename K>
__global__ __launch_bounds__(256) void k_tile_chase(const K* keysA, const int* valsA, K* keysB,
                                                    int* valsB, const long long* tile_off,
                                                    int nseg, const SegInfo* info,
                                                    const int* arot, int* s0, int* seg_need,
                                                    int* need_any, unsigned long long* rmax,
                                                    int cap) {
  const long long t = blockIdx.x;
  if (t >= tile_off[nseg]) return;
  const int sg = find_seg(tile_off, nseg, t);
  const SegInfo g = info[sg];
  const K pivot = keysA[g.hi];
  const long long start = g.lo + (t - g.first_tile) * TILE;
  const int r = (int)g.r;
  const int* A = arot + g.lo;
  K kv[PER_THREAD];
  int s[PER_THREAD];
  bool act[PER_THREAD];
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = start + (long long)q * 256 + threadIdx.x;
    act[q] = false;
    s[q] = 0;
    if (j < g.hi) {
      kv[q] = keysA[j];
      act[q] = !(kv[q] <= pivot);
      if (act[q]) s[q] = s0[j];
    }
  }
  for (int step = 0; step < cap; step++) {
    bool any = false;
    for (int q = 0; q < PER_THREAD; q++) {
      const bool run = act[q] && s[q] < r;
      const int a = A[run ? s[q] : 0];  // r >= 1 whenever a ">" element exists
      s[q] += run ? max(a, 1) : 0;       // every entry is >= 1
      any |= run;
    }
    if (!any) break;
  }
  // steps still needed by an unfinished chain: every later step advances
  // by at least A(s), since A never decreases along the stream
  unsigned long long rem = 0;
  for (int q = 0; q < PER_THREAD; q++) {
    if (!act[q]) continue;
    const long long j = start + (long long)q * 256 + threadIdx.x;
    if (s[q] < r) {
      s0[j] = s[q];
      const int a = max(A[s[q]], 1);
      rem = max(rem, (unsigned long long)((r - s[q] + a - 1) / a));
      continue;
    }
    s0[j] = S0_DONE;
    const long long dst = g.lo + g.c + 1 + (s[q] - r);
    keysB[dst] = kv[q];
    valsB[dst] = valsA[j];
  }
  __shared__ unsigned long long sh_rem[4];
  for (int o = 32; o > 0; o >>= 1) rem = max(rem, (unsigned long long)__shfl_xor(rem, o, 64));
  if ((threadIdx.x & 63) == 0) sh_rem[threadIdx.x >> 6] = rem;
  __syncthreads();
  if (threadIdx.x == 0) {
    rem = max(max(sh_rem[0], sh_rem[1]), max(sh_rem[2], sh_rem[3]));
    if (rem > 0) {
      seg_need[sg] = 1;
      if (!__atomic_load_n(need_any, __ATOMIC_RELAXED)) atomicExch(need_any, 1);
      if (__atomic_load_n(rmax, __ATOMIC_RELAXED) < rem) atomicMax(rmax, rem);
    }
  }
}

// Chase cap: EXH_SORT_CHASE_CAP (default 16).
static int chase_cap() {
  static int v = [] {
    const char* e = std::getenv("EXH_SORT_CHASE_CAP");
    int c = e ? std::atoi(e) : 16;
    return c < 1 ? 1 : (c > 4096 ? 4096 : c);
  }();
  return v;
}

// Pointer doubling on nxt(q) = q + A(q) over the stream indices of flagged
// segments: after k rounds one entry covers at least 2^k chase steps.
// Values are stored relative to the segment start.  Each entry is either a
// stream index below r on the chain from q or the chain's first index >= r,
// so updating the array in place is exact: a round that reads an entry
// another block has already advanced only jumps further along the same
// chain, never past its exit.  A round whose predecessor left no entry
// below r (round_open[k-1] == 0) returns at once, so the host can launch a
// fixed number of rounds without waiting for the device.
__global__ void k_double_init(const long long* tile_off, int nseg, const SegInfo* info,
                              const int* seg_need, const int* need_any, const int* arot,
                              int* nxt) {
  if (!*need_any) return;
  long long t = blockIdx.x;
  if (t >= tile_off[nseg]) return;
  int sg = find_seg(tile_off, nseg, t);
  if (!seg_need[sg]) return;
  SegInfo g = info[sg];
  long long q0 = (t - g.first_tile) * TILE;
  bool last = t == tile_off[sg + 1] - 1;
  long long q1 = last ? g.r : min(q0 + TILE, g.r);
  for (long long q = q0 + threadIdx.x; q < q1; q += 256) {
    nxt[g.lo + q] = (int)q + max(arot[g.lo + q], 1);  // >= r: left the stream prefix
  }
}

__global__ void k_double_round(const long long* tile_off, int nseg, const SegInfo* info,
                               const int* seg_need, const int* need_any, int* nxt,
                               int* round_open, int k) {
  if (!*need_any || (k > 0 && !round_open[k - 1])) return;
  long long t = blockIdx.x;
  if (t >= tile_off[nseg]) return;
  int sg = find_seg(tile_off, nseg, t);
  if (!seg_need[sg]) return;
  SegInfo g = info[sg];
  long long q0 = (t - g.first_tile) * TILE;
  bool last = t == tile_off[sg + 1] - 1;
  long long q1 = last ? g.r : min(q0 + TILE, g.r);
  // all loads of a thread's entries are issued before its stores
  int* X = nxt + g.lo;
  const int r = (int)g.r;
  int v[PER_THREAD];
  for (int e = 0; e < PER_THREAD; e++) {
    const long long q = q0 + threadIdx.x + (long long)e * 256;
    v[e] = q < q1 ? X[q] : r;
  }
  int w[PER_THREAD];
  for (int e = 0; e < PER_THREAD; e++) w[e] = v[e] < r ? X[v[e]] : v[e];
  bool open = false;
  for (int e = 0; e < PER_THREAD; e++) {
    const long long q = q0 + threadIdx.x + (long long)e * 256;
    if (v[e] < r) {
      X[q] = w[e];
      open |= w[e] < r;
    }
  }
  if (__syncthreads_or(open) && threadIdx.x == 0) round_open[k] = 1;
}

// Asynchronous pointer jumping over the same entries: every thread keeps
// replacing its entries' pointers by the pointer they point to until each
// leaves the stream prefix (>= r).  Pointers only ever move forward along
// their chain and every stored value is a valid chain position, so reading
// an entry before or after another thread advanced it is equally correct;
// nothing waits for another thread.  Loads and stores are agent-scope
// atomics so that advances made on other XCDs become visible.  Entries
// still open after `guard` sweeps set *open (the host then finishes with
// the synchronous rounds of k_double_round).
__global__ __launch_bounds__(256) void k_double_async(const long long* tile_off, int nseg,
                                                      const SegInfo* info, const int* seg_need,
                                                      const int* need_any, int* nxt, int guard,
                                                      int* open) {
  if (!*need_any) return;
  const long long t = blockIdx.x;
  if (t >= tile_off[nseg]) return;
  const int sg = find_seg(tile_off, nseg, t);
  if (!seg_need[sg]) return;
  const SegInfo g = info[sg];
  const long long q0 = (t - g.first_tile) * TILE;
  const bool last = t == tile_off[sg + 1] - 1;
  const long long q1 = last ? g.r : min(q0 + TILE, g.r);
  int* X = nxt + g.lo;
  const int r = (int)g.r;
  int v[PER_THREAD];
  for (int e = 0; e < PER_THREAD; e++) {
    const long long q = q0 + threadIdx.x + (long long)e * 256;
    v[e] = q < q1 ? X[q] : r;
  }
  bool any = true;
  for (int it = 0; it < guard && any; it++) {
    int w[PER_THREAD];
#pragma unroll
    for (int e = 0; e < PER_THREAD; e++)
      w[e] = v[e] < r ? __hip_atomic_load(&X[v[e]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      : v[e];
    any = false;
#pragma unroll
    for (int e = 0; e < PER_THREAD; e++) {
      if (v[e] < r) {
        const long long q = q0 + threadIdx.x + (long long)e * 256;
        v[e] = w[e];
        __hip_atomic_store(&X[q], w[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        any |= w[e] < r;
      }
    }
  }
  if (any) atomicExch(open, 1);
}

template <typename K>
__global__ void k_tile_chase2(const K* keysA, const int* valsA, K* keysB, int* valsB,
                              const long long* tile_off, int nseg, const SegInfo* info,
                              const int* seg_need, const int* need_any, const int* nxt,
                              const int* s0, int* err) {
  if (!*need_any) return;
  long long t = blockIdx.x;
  if (t >= tile_off[nseg]) return;
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
    if (s < g.r) {  // doubling rounds too few: report, write nothing
      *err = 1;
      continue;
    }
    long long dst = g.lo + g.c + 1 + (s - g.r);
    keysB[dst] = keysA[j];
    valsB[dst] = valsA[j];
  }
}

template <typename K>
__global__ void k_tile_copyback(K* keysA, int* valsA, const K* keysB, const int* valsB,
                                const long long* tile_off, int nseg, const SegInfo* info) {
  long long t = blockIdx.x;
  if (t >= tile_off[nseg]) return;
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

// One workgroup of SMALL_WAVES wavefronts per segment of 2..SMALL
// elements, resident in LDS.  Sub-segments longer than TINY are partitioned
// by the whole workgroup with the closed-form Lomuto rule (see the file
// comment): each wave takes a contiguous chunk, ranks come from ballots and
// the chunk offsets from LDS.  Shorter sub-segments are queued and then
// sorted one per lane by lane_quicksort.  Several waves share one segment's
// LDS so that the LDS per CU, not per wave, bounds the occupancy.
static constexpr int SMALL_WAVES = 4;

template <typename K, int SMALL>
__global__ __launch_bounds__(64 * SMALL_WAVES) void k_small_sort(K* keys, int* vals,
                                                                 const long long* seg_lo,
                                                                 const long long* seg_hi) {
  static_assert(SMALL <= 32767, "LDS indices are 16-bit");
  constexpr int NT = 64 * SMALL_WAVES;
  __shared__ K sk[SMALL], tk[SMALL];
  __shared__ int sv[SMALL], tv[SMALL];
  __shared__ short arot[SMALL];
  __shared__ short tl_lo[SMALL / 2 + 1], tl_hi[SMALL / 2 + 1];
  __shared__ short st_lo[64], st_hi[64];
  __shared__ int wc[SMALL_WAVES], wj[SMALL_WAVES];
  __shared__ int ntiny, top_sh;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const unsigned long long below = lanemask_lt();
  const long long base = seg_lo[blockIdx.x];
  const int n = (int)(seg_hi[blockIdx.x] - base + 1);
  for (int i = tid; i < n; i += NT) {
    sk[i] = keys[base + i];
    sv[i] = vals[base + i];
  }
  if (tid == 0) {
    ntiny = 0;
    top_sh = 0;
    if (n > TINY) {
      st_lo[0] = 0;
      st_hi[0] = (short)(n - 1);
      top_sh = 1;
    } else {
      tl_lo[0] = 0;
      tl_hi[0] = (short)(n - 1);
      ntiny = 1;
    }
  }
  __syncthreads();
  for (;;) {
    const int top = top_sh;
    if (top == 0) break;
    const int lo = st_lo[top - 1], hi = st_hi[top - 1];
    const K pivot = sk[hi];
    const int L = hi - lo;
    const int Q = ((L + SMALL_WAVES - 1) / SMALL_WAVES + 63) & ~63;
    const int cb = lo + w * Q, ce = min(cb + Q, hi);
    // pass 1: per-chunk count of "<=" and first ">"
    int cnt = 0, j0w = hi;
    for (int b = cb; b < ce; b += 64) {
      const int j = b + lane;
      const bool in = j < ce;
      const bool le = in && sk[j] <= pivot;
      const unsigned long long mle = __ballot(le), mgt = __ballot(in && !le);
      if (j0w == hi && mgt) j0w = b + __ffsll((long long)mgt) - 1;
      cnt += __popcll(mle);
    }
    if (lane == 0) {
      wc[w] = cnt;
      wj[w] = j0w;
    }
    __syncthreads();
    int c = 0, j0 = hi, wbase = 0;
    for (int v = 0; v < SMALL_WAVES; v++) {
      if (v < w) wbase += wc[v];
      c += wc[v];
      j0 = min(j0, wj[v]);
    }
    const int m = L - c;
    const int r = m > 0 ? (c - (j0 - lo)) + 1 : 0;
    const int jlead = j0 - lo;
    // pass 2: place "<=" elements, record the rotation table
    int run = wbase;
    for (int b = cb; b < ce; b += 64) {
      const int j = b + lane;
      const bool le = j < ce && sk[j] <= pivot;
      const unsigned long long mle = __ballot(le);
      if (le) {
        const int rank = run + __popcll(mle & below);
        tk[lo + rank] = sk[j];
        tv[lo + rank] = sv[j];
        if (r > 0 && j > j0) arot[rank - jlead] = (short)((j - lo) - rank);
      }
      run += __popcll(mle);
    }
    if (tid == 0) {
      tk[lo + c] = sk[hi];
      tv[lo + c] = sv[hi];
      if (m > 0) arot[r - 1] = (short)m;
    }
    __syncthreads();
    // pass 3: chase the queue from each ">" element's first stream index
    run = wbase;
    for (int b = cb; b < ce; b += 64) {
      const int j = b + lane;
      const bool in = j < ce;
      const bool le = in && sk[j] <= pivot;
      const unsigned long long mle = __ballot(le);
      if (in && !le) {
        const int rank = run + __popcll(mle & below);
        int s = ((j - lo) - rank) + (rank - jlead);
        for (int guard = 0; s < r && guard <= r; guard++) s += max((int)arot[s], 1);
        tk[lo + c + 1 + (s - r)] = sk[j];
        tv[lo + c + 1 + (s - r)] = sv[j];
      }
      run += __popcll(mle);
    }
    __syncthreads();
    for (int i = lo + tid; i <= hi; i += NT) {
      sk[i] = tk[i];
      sv[i] = tv[i];
    }
    // children: long ones on the stack (larger first, so the smaller is
    // partitioned next), short ones to the per-lane list
    if (tid == 0) {
      int al = lo, ah = lo + c - 1, bl = lo + c + 1, bh = hi;
      if (ah - al > bh - bl) {
        int x = al, y = ah;
        al = bl;
        ah = bh;
        bl = x;
        bh = y;
      }
      int tt = top - 1;
      if (bh - bl + 1 > TINY) {
        st_lo[tt] = (short)bl;
        st_hi[tt] = (short)bh;
        tt++;
      } else if (bh > bl) {
        tl_lo[ntiny] = (short)bl;
        tl_hi[ntiny] = (short)bh;
        ntiny++;
      }
      if (ah - al + 1 > TINY) {
        st_lo[tt] = (short)al;
        st_hi[tt] = (short)ah;
        tt++;
      } else if (ah > al) {
        tl_lo[ntiny] = (short)al;
        tl_hi[ntiny] = (short)ah;
        ntiny++;
      }
      top_sh = tt;
    }
    __syncthreads();
  }
  for (int q = tid; q < ntiny; q += NT) lane_quicksort<K>(sk, sv, tl_lo[q], tl_hi[q]);
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
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
  c->stat_sort_async_open = 0;
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
         o_shi = take(maxsmall * 8), o_cnt = take(64 * 4), o_arot = take((size_t)n * 4),
         o_s0 = take((size_t)n * 4), o_need = take(maxseg * 4 + 64), o_nxa = take((size_t)n * 4 + 64);
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
  int* arot = (int*)(aux + o_arot);
  int* s0 = (int*)(aux + o_s0);
  // level state read back once per level: [0] large children, [1] small
  // segments so far, [2] some segment needs doubling, [4..5] most steps an
  // unfinished chain still needs, [6..7] elements in large children; [8] doubling error (kept
  // over all levels); [9] chains left open by the asynchronous doubling;
  // [16..63] per-round flags of the doubling
  int* hc = (int*)ctx_pinned(c, 64);
  int* seg_need = (int*)(aux + o_need);
  int* nxt = (int*)(aux + o_nxa);
  int* round_open = cnt + 16;
  unsigned long long* rmax = (unsigned long long*)(cnt + 4);
  unsigned long long* next_elems = (unsigned long long*)(cnt + 6);

  int nsmall_host = 0;
  int nseg = 0;
  long long level_elems = 0;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, cnt, 64, 0);
  {
    long long init[2] = {0, n - 1};
    long long* dlo = n > SMALL ? cur_lo : slo;
    long long* dhi = n > SMALL ? cur_hi : shi;
    EXH_CHECK_HIP(hipMemcpyAsync(dlo, &init[0], 8, hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipMemcpyAsync(dhi, &init[1], 8, hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (n > SMALL) {
      nseg = 1;
      level_elems = n;
    } else {
      nsmall_host = 1;
    }
  }
  bool pending_err = false;
  while (nseg > 0) {
    c->stat_sort_levels++;
    // tiles of this level: at most ceil(elements / TILE) + segments; the
    // tile kernels' blocks beyond the exact count return at once
    const long long ntile_ub = (level_elems + TILE - 1) / TILE + nseg;
    EXH_REQUIRE(ntile_ub <= maxtiles, EXH_ERR_INTERNAL, "peak sort: tile bound exceeded");
    hipLaunchKernelGGL(k_seg_ntiles, dim3(grid_for(nseg, 256)), dim3(256), 0, c->stream, cur_lo,
                       cur_hi, nseg, ntiles, (long long*)j0);
    EXH_LAUNCH_CHECK();
    exclusive_scan<long long>(c, ntiles, toff, nseg);
    hipLaunchKernelGGL(k_tile_count<K>, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream, keys,
                       cur_lo, cur_hi, toff, nseg, tle, j0);
    EXH_LAUNCH_CHECK();
    exclusive_scan<long long>(c, tle, tpre, ntile_ub);
    // reset the per-level counters (not cnt[1]: the small segments collect
    // over all levels) and the round flags
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, cnt, 1, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, cnt + 2, 6, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, round_open, 48, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(nseg, 256)), dim3(256), 0, c->stream, seg_need,
                       nseg, 0);
    hipLaunchKernelGGL(k_seg_final<K>, dim3(grid_for(nseg, 256)), dim3(256), 0, c->stream, keys,
                       vals, kB, vB, cur_lo, cur_hi, toff, tpre, j0, nseg, info, arot,
                       (long long)SMALL, nxt_lo, nxt_hi, cnt, slo, shi, cnt + 1, next_elems);
    EXH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_scatter<K>, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream,
                       keys, vals, kB, vB, toff, tpre, nseg, info, arot, s0);
    EXH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_chase<K>, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream, keys,
                       vals, kB, vB, toff, nseg, info, arot, s0, seg_need, cnt + 2, rmax, chase_cap());
    EXH_LAUNCH_CHECK();
    EXH_CHECK_HIP(hipMemcpyAsync(hc, cnt, 32, hipMemcpyDeviceToHost, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (hc[2]) {
      // long chains: finish them by pointer doubling; after k rounds an
      // entry covers 2^k steps, and no chain needs more than rm more steps
      c->stat_sort_doubling++;
      const unsigned long long rm = *(unsigned long long*)(hc + 4);
      int rounds = 1;
      while ((1ull << (rounds - 1)) < rm && rounds < 62) rounds++;
      EXH_REQUIRE(rounds <= 40, EXH_ERR_INTERNAL, "peak sort: segment too long for doubling");
      hipLaunchKernelGGL(k_double_init, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream, toff,
                         nseg, info, seg_need, cnt + 2, arot, nxt);
      hipLaunchKernelGGL(k_double_async, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream, toff,
                         nseg, info, seg_need, cnt + 2, nxt, 64, cnt + 9);
      EXH_LAUNCH_CHECK();
      EXH_CHECK_HIP(hipMemcpyAsync(hc + 9, cnt + 9, 4, hipMemcpyDeviceToHost, c->stream));
      EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
      if (hc[9]) {
        // some chain is still open: the synchronous rounds finish every
        // chain within the bound
        c->stat_sort_async_open++;
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, cnt + 9, 1, 0);
        for (int k = 0; k < rounds; k++) {
          hipLaunchKernelGGL(k_double_round, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream,
                             toff, nseg, info, seg_need, cnt + 2, nxt, round_open, k);
          EXH_LAUNCH_CHECK();
        }
      }
      hipLaunchKernelGGL(k_tile_chase2<K>, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream,
                         keys, vals, kB, vB, toff, nseg, info, seg_need, cnt + 2, nxt, s0, cnt + 8);
      EXH_LAUNCH_CHECK();
      pending_err = true;
    }
    hipLaunchKernelGGL(k_tile_copyback<K>, dim3((unsigned)ntile_ub), dim3(256), 0, c->stream,
                       keys, vals, kB, vB, toff, nseg, info);
    EXH_LAUNCH_CHECK();
    nseg = hc[0];
    nsmall_host = hc[1];
    level_elems = *(long long*)(hc + 6);
    EXH_REQUIRE(nseg <= maxseg && nsmall_host <= maxsmall, EXH_ERR_INTERNAL,
                "peak sort: segment lists overflow (%d, %d)", nseg, nsmall_host);
    std::swap(cur_lo, nxt_lo);
    std::swap(cur_hi, nxt_hi);
  }
  if (pending_err) {
    EXH_CHECK_HIP(hipMemcpyAsync(hc + 8, cnt + 8, 4, hipMemcpyDeviceToHost, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    EXH_REQUIRE(!hc[8], EXH_ERR_INTERNAL, "peak sort: pointer doubling did not converge");
  }
  if (nsmall_host > 0) {
    hipLaunchKernelGGL((k_small_sort<K, SMALL>), dim3((unsigned)nsmall_host), dim3(64 * SMALL_WAVES), 0,
                       c->stream, keys, vals, slo, shi);
    EXH_LAUNCH_CHECK();
  }
}

template void lomuto_sort<float>(exh_ctx*, float*, int*, long long);
template void lomuto_sort<double>(exh_ctx*, double*, int*, long long);

}  // namespace exh
