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
//
// Keys and values move between two work buffers level by level (src ->
// dst).  Elements whose place is final (pivots, children of one element)
// are also written to the output array; segments of 2..SMALL elements are
// queued with the buffer that holds them and finished by k_small_sort.
// Children of more than SMALL elements become the next level's segments:
// each reserves its list slot and its range of tiles with one packed 64-bit
// atomic (slot count in the high half, tile count in the low half), so the
// tile offsets grow with the slot index and no scan is needed.
// ---------------------------------------------------------------------------
__global__ void k_fill_int(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

struct LevelState {
  unsigned long long pk;          // next level: (segments << 32) | tiles
  unsigned long long next_elems;  // elements in the next level's segments
  unsigned long long rmax;        // most steps an unfinished chain still needs
  unsigned long long nsmall;      // small segments queued so far (all levels)
  unsigned long long need_any;    // some segment needs pointer doubling
  unsigned long long async_open;  // chains left open by k_double_walk
  unsigned long long err;         // doubling did not converge (all levels)
  unsigned long long pad;
};

struct LevelBufs {
  const long long* lo;  // current segments
  const long long* hi;
  const long long* toff;  // first tile of each segment (ascending)
  unsigned long long* j0;  // first ">" element of each segment
  int nseg;
  long long ntiles;
  long long* nlo;  // next level's segments
  long long* nhi;
  long long* ntoff;
  unsigned long long* nj0;
  long long* slo;  // small segments and the buffer holding each
  long long* shi;
  int* sbuf;
  int dst_id;  // 1 or 2: buffer id of dst for the small list
  long long small_max;
  SegInfo* info;
  int* tseg;  // segment of each tile of the level (written by k_tile_count)
  int* seg_need;
  LevelState* st;
  // keys below `floor` need no order: a left child of a pivot below it
  // (all its keys <= pivot) is left unsorted
  double floor;
};

__device__ inline long long seg_end_tile(const LevelBufs& b, int sg) {
  return sg + 1 < b.nseg ? b.toff[sg + 1] : b.ntiles;
}

__global__ void k_sort_init(long long* lo, long long* hi, long long* toff, unsigned long long* j0,
                            LevelState* st, long long n) {
  if (threadIdx.x == 0) {
    lo[0] = 0;
    hi[0] = n - 1;
    toff[0] = 0;
    j0[0] = (unsigned long long)(n - 1);
    *st = LevelState{};
  }
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
// Block 0 also clears the per-level counters (read by the host after the
// previous level).
template <typename K>
__global__ __launch_bounds__(256) void k_tile_count(const K* keys, LevelBufs b, long long* tile_le) {
  __shared__ int shc[4];
  __shared__ long long shj[4];
  const long long t = blockIdx.x;
  if (t == 0 && threadIdx.x == 0) {
    LevelState* st = b.st;
    st->pk = 0;
    st->next_elems = 0;
    st->rmax = 0;
    st->need_any = 0;
    st->async_open = 0;
  }
  const int s = find_seg(b.toff, b.nseg, t);
  if (threadIdx.x == 0) b.tseg[t] = s;
  const long long lo = b.lo[s], hi = b.hi[s];
  const K pivot = keys[hi];
  const long long start = lo + (t - b.toff[s]) * TILE;
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
    if (fg < hi && (unsigned long long)fg < __atomic_load_n(&b.j0[s], __ATOMIC_RELAXED))
      atomicMin(&b.j0[s], (unsigned long long)fg);
  }
}

__device__ inline SegInfo seg_info(const LevelBufs& b, const long long* tile_pre, int s) {
  SegInfo g;
  g.lo = b.lo[s];
  g.hi = b.hi[s];
  g.first_tile = b.toff[s];
  const long long lt = seg_end_tile(b, s);
  g.c = tile_pre[lt] - tile_pre[g.first_tile];
  const long long m = (g.hi - g.lo) - g.c;
  g.j0 = (long long)b.j0[s];
  g.r = m > 0 ? (g.c - (g.j0 - g.lo)) + 1 : 0;
  return g;
}

// Queue child [cl, ch] (held in buffer dst_id) as a next-level or small
// segment.
__device__ inline void push_child(const LevelBufs& b, long long cl, long long ch) {
  const long long len = ch - cl + 1;
  if (len > b.small_max) {
    const unsigned long long nt = (unsigned long long)((len - 1 + TILE - 1) / TILE);
    const unsigned long long old = atomicAdd(&b.st->pk, (1ull << 32) | nt);
    const int p = (int)(old >> 32);
    b.nlo[p] = cl;
    b.nhi[p] = ch;
    b.ntoff[p] = (long long)(old & 0xffffffffull);
    b.nj0[p] = (unsigned long long)ch;  // "no > element" until found
    atomicAdd(&b.st->next_elems, (unsigned long long)len);
  } else if (len >= 2) {
    const long long p = (long long)atomicAdd(&b.st->nsmall, 1ull);
    b.slo[p] = cl;
    b.shi[p] = ch;
    b.sbuf[p] = b.dst_id;
  }
}

// Place the "<=" elements at lo + rank (dst; also the output when the left
// child is a single element), record the rotation table A and each ">"
// element's first stream index.  Wave w of a tile takes its elements
// [w * 512, (w + 1) * 512) in eight coalesced rows of 64; ranks come from
// ballots, the wave offsets from LDS.  The segment's first tile also
// publishes its SegInfo, places the pivot in the output, closes A and
// queues the two children.
template <typename K>
__global__ __launch_bounds__(256) void k_tile_scatter(const K* keysA, const int* valsA, K* keysB,
                                                      int* valsB, K* keysF, int* valsF,
                                                      LevelBufs b, const long long* tile_pre,
                                                      int* arot, int* s0) {
  __shared__ int shw[4];
  const long long t = blockIdx.x;
  const int sg = b.tseg[t];
  const SegInfo g = seg_info(b, tile_pre, sg);
  const K pivot = keysA[g.hi];
  const long long m = (g.hi - g.lo) - g.c;
  if (t == g.first_tile && threadIdx.x == 0) {
    b.info[sg] = g;
    b.seg_need[sg] = 0;
    keysF[g.lo + g.c] = pivot;
    valsF[g.lo + g.c] = valsA[g.hi];
    if (m > 0) arot[g.lo + g.r - 1] = (int)m;
    if (!((double)pivot < b.floor)) push_child(b, g.lo, g.lo + g.c - 1);
    push_child(b, g.lo + g.c + 1, g.hi);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long base = g.lo + (t - g.first_tile) * TILE + (long long)w * (TILE / 4);
  K kv[PER_THREAD];
  int vv[PER_THREAD];
  unsigned long long bal[PER_THREAD];
  int wcnt = 0;
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = base + q * 64 + lane;
    const bool in = j < g.hi;
    kv[q] = in ? keysA[j] : pivot;
    bal[q] = __ballot(in && kv[q] <= pivot);
    wcnt += __popcll(bal[q]);
  }
  // the "<=" elements' values, loaded before the first store (which could
  // alias them for the compiler)
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = base + q * 64 + lane;
    vv[q] = (bal[q] >> lane) & 1ull ? valsA[j] : 0;
  }
  if (lane == 0) shw[w] = wcnt;
  __syncthreads();
  long long run = tile_pre[t] - tile_pre[g.first_tile];
  for (int v = 0; v < w; v++) run += shw[v];
  const unsigned long long below = lanemask_lt();
  const long long jlead = g.j0 - g.lo;  // "<=" elements before the first ">"
  const bool single_left = g.c == 1;
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = base + q * 64 + lane;
    if (j < g.hi) {
      const long long rank = run + __popcll(bal[q] & below);
      if ((bal[q] >> lane) & 1ull) {
        keysB[g.lo + rank] = kv[q];
        valsB[g.lo + rank] = vv[q];
        if (single_left) {
          keysF[g.lo] = kv[q];
          valsF[g.lo] = vv[q];
        }
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
// for pointer doubling (one write per block, not per element).  The block
// also initialises the doubling pointers jmp[q] = q + A(q) of its tile's
// stream indices.
static constexpr int S0_DONE = 0x7fffffff;

template <typename K>
__global__ __launch_bounds__(256) void k_tile_chase(const K* keysA, const int* valsA, K* keysB,
                                                    int* valsB, K* keysF, int* valsF,
                                                    LevelBufs b, const int* arot, int* s0,
                                                    int* jmp, int cap) {
  const long long t = blockIdx.x;
  const int sg = b.tseg[t];
  const SegInfo g = b.info[sg];
  const K pivot = keysA[g.hi];
  const long long start = g.lo + (t - g.first_tile) * TILE;
  const int r = (int)g.r;
  const int* A = arot + g.lo;
  const bool single_right = (g.hi - g.lo) - g.c == 1;
  // doubling pointers of this tile's stream indices [q0, q1)
  {
    const long long q0 = (t - g.first_tile) * TILE;
    const bool last = t == seg_end_tile(b, sg) - 1;
    const long long q1 = last ? g.r : min(q0 + TILE, g.r);
    for (long long q = q0 + threadIdx.x; q < q1; q += 256)
      jmp[g.lo + q] = (int)q + max(A[q], 1);  // >= r: left the stream prefix
  }
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
  // what the stores below need, loaded before the first of them (which
  // could alias it for the compiler): a finished element's value, an open
  // chain's next rotation entry
  int ld[PER_THREAD];
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = start + (long long)q * 256 + threadIdx.x;
    ld[q] = !act[q] ? 0 : (s[q] < r ? A[s[q]] : valsA[j]);
  }
  for (int q = 0; q < PER_THREAD; q++) {
    if (!act[q]) continue;
    const long long j = start + (long long)q * 256 + threadIdx.x;
    if (s[q] < r) {
      s0[j] = s[q];
      const int a = max(ld[q], 1);
      rem = max(rem, (unsigned long long)((r - s[q] + a - 1) / a));
      continue;
    }
    s0[j] = S0_DONE;
    const long long dst = g.lo + g.c + 1 + (s[q] - r);
    keysB[dst] = kv[q];
    valsB[dst] = ld[q];
    if (single_right) {
      keysF[dst] = kv[q];
      valsF[dst] = ld[q];
    }
  }
  __shared__ unsigned long long sh_rem[4];
  for (int o = 32; o > 0; o >>= 1) rem = max(rem, (unsigned long long)__shfl_xor(rem, o, 64));
  if ((threadIdx.x & 63) == 0) sh_rem[threadIdx.x >> 6] = rem;
  __syncthreads();
  if (threadIdx.x == 0) {
    rem = max(max(sh_rem[0], sh_rem[1]), max(sh_rem[2], sh_rem[3]));
    if (rem > 0) {
      b.seg_need[sg] = 1;
      if (!__atomic_load_n(&b.st->need_any, __ATOMIC_RELAXED)) atomicExch(&b.st->need_any, 1ull);
      if (__atomic_load_n(&b.st->rmax, __ATOMIC_RELAXED) < rem) atomicMax(&b.st->rmax, rem);
    }
  }
}

// Chase cap: EXH_SORT_CHASE_CAP (default 4).  Longer chases finish more
// elements before the pointer jumping but hold every tile of the level as
// long as its longest chain: at 512^3 the fused step took 25.0 ms with caps
// of 1 to 4, 25.2 ms with 8, 25.4 with 16 and 25.9 with 32.
static int chase_cap() {
  static int v = [] {
    const char* e = std::getenv("EXH_SORT_CHASE_CAP");
    int c = e ? std::atoi(e) : 4;
    return c < 1 ? 1 : (c > 4096 ? 4096 : c);
  }();
  return v;
}

// Asynchronous pointer jumping over the stream indices of flagged
// segments: every thread keeps replacing its entries' pointers jmp[q] by
// the pointer they point to until each leaves the stream prefix (>= r).
// Each entry is a stream index below r on the chain from q or the chain's
// first index >= r; pointers only ever move forward along their chain, so
// reading an entry before or after another thread advanced it is equally
// correct, and nothing waits for another thread.  Loads and stores are
// agent-scope atomics so that advances made on other XCDs become visible.
// Entries still open after `guard` sweeps set async_open; the host then
// finishes with the synchronous rounds of k_double_round.
// The same kernel then walks the tile's own open elements to their slots,
// as k_tile_chase2 (strict = 0) does: a walk that reaches the stream's end
// has its exact slot whether or not the jumping of other tiles has
// converged, since every pointer leads forward along its own chain.
template <typename K>
__global__ __launch_bounds__(256) void k_double_walk(const K* keysA, const int* valsA, K* keysB,
                                                     int* valsB, K* keysF, int* valsF, LevelBufs b,
                                                     int* jmp, const int* s0, int guard) {
  const long long t = blockIdx.x;
  const int sg = b.tseg[t];
  if (!b.seg_need[sg]) return;
  const SegInfo g = b.info[sg];
  const long long q0 = (t - g.first_tile) * TILE;
  const bool last = t == seg_end_tile(b, sg) - 1;
  const long long q1 = last ? g.r : min(q0 + TILE, g.r);
  int* X = jmp + g.lo;
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
  bool open = any;
  const K pivot = keysA[g.hi];
  const bool single_right = (g.hi - g.lo) - g.c == 1;
  const long long start = g.lo + q0;
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = start + (long long)q * 256 + threadIdx.x;
    if (j >= g.hi) continue;
    if (keysA[j] <= pivot) continue;
    long long s = s0[j];
    if (s >= g.r) continue;  // placed by the capped chase
    for (int gd = 0; s < g.r && gd < 64; gd++)
      s = __hip_atomic_load(&X[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (s < g.r) {  // write nothing: the synchronous rounds finish it
      open = true;
      continue;
    }
    const long long dst = g.lo + g.c + 1 + (s - g.r);
    keysB[dst] = keysA[j];
    valsB[dst] = valsA[j];
    if (single_right) {
      keysF[dst] = keysA[j];
      valsF[dst] = valsA[j];
    }
  }
  if (open) atomicExch(&b.st->async_open, 1ull);
}

// Synchronous doubling round k (in place, same invariant as above); a
// round whose predecessor left no entry open returns at once.
__global__ __launch_bounds__(256) void k_double_round(LevelBufs b, int* jmp, int* round_open, int k) {
  if (k > 0 && !round_open[k - 1]) return;
  const long long t = blockIdx.x;
  const int sg = b.tseg[t];
  if (!b.seg_need[sg]) return;
  const SegInfo g = b.info[sg];
  const long long q0 = (t - g.first_tile) * TILE;
  const bool last = t == seg_end_tile(b, sg) - 1;
  const long long q1 = last ? g.r : min(q0 + TILE, g.r);
  int* X = jmp + g.lo;
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

// Elements whose capped chase did not finish: follow the pointers to their
// slot.  Every pointer leads forward along its own chain, so an element
// whose walk reaches the stream's end within the guard has its exact slot,
// converged or not.  Walks left short set async_open (strict = 0: the host
// then runs the synchronous rounds and this kernel again) or err (strict).
template <typename K>
__global__ void k_tile_chase2(const K* keysA, const int* valsA, K* keysB, int* valsB, K* keysF,
                              int* valsF, LevelBufs b, const int* jmp, const int* s0,
                              int strict) {
  const long long t = blockIdx.x;
  const int sg = b.tseg[t];
  if (!b.seg_need[sg]) return;
  const SegInfo g = b.info[sg];
  const K pivot = keysA[g.hi];
  const bool single_right = (g.hi - g.lo) - g.c == 1;
  const long long start = g.lo + (t - g.first_tile) * TILE;
  for (int q = 0; q < PER_THREAD; q++) {
    const long long j = start + (long long)q * 256 + threadIdx.x;
    if (j >= g.hi) continue;
    if (keysA[j] <= pivot) continue;
    long long s = s0[j];
    if (s >= g.r) continue;  // placed by the capped chase
    for (int guard = 0; s < g.r && guard < 64; guard++) s = jmp[g.lo + s];
    if (s < g.r) {  // write nothing: later rounds finish it, or report
      atomicExch(strict ? &b.st->err : &b.st->async_open, 1ull);
      continue;
    }
    const long long dst = g.lo + g.c + 1 + (s - g.r);
    keysB[dst] = keysA[j];
    valsB[dst] = valsA[j];
    if (single_right) {
      keysF[dst] = keysA[j];
      valsF[dst] = valsA[j];
    }
  }
}

// ---------------------------------------------------------------------------
// small segments: one wavefront each, LDS resident
// ---------------------------------------------------------------------------
static constexpr int TINY = 32;      // sorted one per lane
static constexpr int WAVE_MAX = 256; // partitioned by one wave

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

// Ordering of LDS accesses between the lanes of one wavefront.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Lomuto partition of sk/sv[lo..hi] (LDS) by one wavefront with the
// closed-form rule of the file comment; temporaries tk/tv and the rotation
// table at arot[lo + q].  Returns c, the number of "<=" elements.
template <typename K>
__device__ inline int wave_partition(K* sk, int* sv, K* tk, int* tv, short* arot, int lo, int hi,
                                     int lane, unsigned long long below) {
  const K pivot = sk[hi];
  int c = 0, j0 = hi;
  for (int b = lo; b < hi; b += 64) {
    const int j = b + lane;
    const bool in = j < hi;
    const bool le = in && sk[j] <= pivot;
    const unsigned long long mle = __ballot(le), mgt = __ballot(in && !le);
    if (j0 == hi && mgt) j0 = b + __ffsll((long long)mgt) - 1;
    c += __popcll(mle);
  }
  const int m = (hi - lo) - c;
  const int r = m > 0 ? (c - (j0 - lo)) + 1 : 0;
  const int jlead = j0 - lo;
  int run = 0;
  for (int b = lo; b < hi; b += 64) {
    const int j = b + lane;
    const bool le = j < hi && sk[j] <= pivot;
    const unsigned long long mle = __ballot(le);
    if (le) {
      const int rank = run + __popcll(mle & below);
      tk[lo + rank] = sk[j];
      tv[lo + rank] = sv[j];
      if (r > 0 && j > j0) arot[lo + rank - jlead] = (short)((j - lo) - rank);
    }
    run += __popcll(mle);
  }
  if (lane == 0) {
    tk[lo + c] = sk[hi];
    tv[lo + c] = sv[hi];
    if (m > 0) arot[lo + r - 1] = (short)m;
  }
  wave_sync();
  run = 0;
  for (int b = lo; b < hi; b += 64) {
    const int j = b + lane;
    const bool in = j < hi;
    const bool le = in && sk[j] <= pivot;
    const unsigned long long mle = __ballot(le);
    if (in && !le) {
      const int rank = run + __popcll(mle & below);
      int s = ((j - lo) - rank) + (rank - jlead);
      for (int guard = 0; s < r && guard <= r; guard++) s += max((int)arot[lo + s], 1);
      tk[lo + c + 1 + (s - r)] = sk[j];
      tv[lo + c + 1 + (s - r)] = sv[j];
    }
    run += __popcll(mle);
  }
  wave_sync();
  for (int i = lo + lane; i <= hi; i += 64) {
    sk[i] = tk[i];
    sv[i] = tv[i];
  }
  wave_sync();
  return c;
}

// One workgroup of SMALL_WAVES wavefronts per segment of 2..SMALL
// elements, resident in LDS.  Sub-segments longer than WAVE_MAX are
// partitioned by the whole workgroup (each wave takes a contiguous chunk;
// ranks from ballots, chunk offsets from LDS); those of TINY+1..WAVE_MAX
// go to a queue from which each wave takes sub-segments and partitions
// them alone (no workgroup barriers); those of 2..TINY are sorted one per
// lane by lane_quicksort.  All use the closed-form Lomuto rule of the file
// comment, with the rotation table indexed from the sub-segment start, so
// that waves working on disjoint sub-segments never share LDS.
static constexpr int SMALL_WAVES = 4;
static constexpr int SMALL_BATCH = 1024;  // segments per small-sort launch, at least

template <typename K, int SMALL>
__global__ __launch_bounds__(64 * SMALL_WAVES) void k_small_sort(
    K* keys, int* vals, const K* keys1, const int* vals1, const K* keys2, const int* vals2,
    const long long* seg_lo, const long long* seg_hi, const int* seg_buf) {
  static_assert(SMALL <= 32767, "LDS indices are 16-bit");
  constexpr int NT = 64 * SMALL_WAVES;
  constexpr int WQ = SMALL / (TINY + 1) + 2;
  __shared__ K sk[SMALL], tk[SMALL];
  __shared__ int sv[SMALL], tv[SMALL];
  __shared__ short arot[SMALL];
  __shared__ short tl_lo[SMALL / 2 + 1], tl_hi[SMALL / 2 + 1];
  __shared__ short wq_lo[WQ], wq_hi[WQ];
  __shared__ short st_lo[64], st_hi[64];
  __shared__ short ws_lo[SMALL_WAVES][16], ws_hi[SMALL_WAVES][16];
  __shared__ int wc[SMALL_WAVES], wj[SMALL_WAVES];
  __shared__ int ntiny, nwq, wq_next, top_sh;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const unsigned long long below = lanemask_lt();
  const long long base = seg_lo[blockIdx.x];
  const int n = (int)(seg_hi[blockIdx.x] - base + 1);
  const int buf = seg_buf[blockIdx.x];
  const K* srck = buf == 0 ? keys : (buf == 1 ? keys1 : keys2);
  const int* srcv = buf == 0 ? vals : (buf == 1 ? vals1 : vals2);
  for (int i = tid; i < n; i += NT) {
    sk[i] = srck[base + i];
    sv[i] = srcv[base + i];
  }
  if (tid == 0) {
    ntiny = 0;
    nwq = 0;
    wq_next = 0;
    top_sh = 0;
    if (n > WAVE_MAX) {
      st_lo[0] = 0;
      st_hi[0] = (short)(n - 1);
      top_sh = 1;
    } else if (n > TINY) {
      wq_lo[0] = 0;
      wq_hi[0] = (short)(n - 1);
      nwq = 1;
    } else {
      tl_lo[0] = 0;
      tl_hi[0] = (short)(n - 1);
      ntiny = 1;
    }
  }
  __syncthreads();
  // ---- workgroup phase: sub-segments longer than WAVE_MAX
  for (;;) {
    const int top = top_sh;
    if (top == 0) break;
    const int lo = st_lo[top - 1], hi = st_hi[top - 1];
    const K pivot = sk[hi];
    const int L = hi - lo;
    const int Q = ((L + SMALL_WAVES - 1) / SMALL_WAVES + 63) & ~63;
    const int cb = lo + w * Q, ce = min(cb + Q, hi);
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
    int run = wbase;
    for (int b = cb; b < ce; b += 64) {
      const int j = b + lane;
      const bool le = j < ce && sk[j] <= pivot;
      const unsigned long long mle = __ballot(le);
      if (le) {
        const int rank = run + __popcll(mle & below);
        tk[lo + rank] = sk[j];
        tv[lo + rank] = sv[j];
        if (r > 0 && j > j0) arot[lo + rank - jlead] = (short)((j - lo) - rank);
      }
      run += __popcll(mle);
    }
    if (tid == 0) {
      tk[lo + c] = sk[hi];
      tv[lo + c] = sv[hi];
      if (m > 0) arot[lo + r - 1] = (short)m;
    }
    __syncthreads();
    run = wbase;
    for (int b = cb; b < ce; b += 64) {
      const int j = b + lane;
      const bool in = j < ce;
      const bool le = in && sk[j] <= pivot;
      const unsigned long long mle = __ballot(le);
      if (in && !le) {
        const int rank = run + __popcll(mle & below);
        int s = ((j - lo) - rank) + (rank - jlead);
        for (int guard = 0; s < r && guard <= r; guard++) s += max((int)arot[lo + s], 1);
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
    // partitioned next), medium ones to the wave queue, short ones to the
    // per-lane list
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
      const int cl[2] = {bl, al}, ch[2] = {bh, ah};
      for (int q = 0; q < 2; q++) {
        const int len = ch[q] - cl[q] + 1;
        if (len > WAVE_MAX) {
          st_lo[tt] = (short)cl[q];
          st_hi[tt] = (short)ch[q];
          tt++;
        } else if (len > TINY) {
          wq_lo[nwq] = (short)cl[q];
          wq_hi[nwq] = (short)ch[q];
          nwq++;
        } else if (len >= 2) {
          tl_lo[ntiny] = (short)cl[q];
          tl_hi[ntiny] = (short)ch[q];
          ntiny++;
        }
      }
      top_sh = tt;
    }
    __syncthreads();
  }
  // ---- wave phase: sub-segments of TINY+1..WAVE_MAX, one wave each
  for (;;) {
    int idx = 0;
    if (lane == 0) idx = atomicAdd(&wq_next, 1);
    idx = __shfl(idx, 0, 64);
    if (idx >= nwq) break;
    int top = 1;
    if (lane == 0) {
      ws_lo[w][0] = wq_lo[idx];
      ws_hi[w][0] = wq_hi[idx];
    }
    wave_sync();
    while (top > 0) {
      top--;
      const int lo = ws_lo[w][top], hi = ws_hi[w][top];
      wave_sync();
      const int c = wave_partition<K>(sk, sv, tk, tv, arot, lo, hi, lane, below);
      int al = lo, ah = lo + c - 1, bl = lo + c + 1, bh = hi;
      if (ah - al > bh - bl) {
        int x = al, y = ah;
        al = bl;
        ah = bh;
        bl = x;
        bh = y;
      }
      const int cl[2] = {bl, al}, ch[2] = {bh, ah};
      for (int q = 0; q < 2; q++) {
        const int len = ch[q] - cl[q] + 1;
        if (len > TINY) {
          if (lane == 0) {
            ws_lo[w][top] = (short)cl[q];
            ws_hi[w][top] = (short)ch[q];
          }
          top++;
        } else if (len >= 2 && lane == 0) {
          const int p = atomicAdd(&ntiny, 1);
          tl_lo[p] = (short)cl[q];
          tl_hi[p] = (short)ch[q];
        }
      }
      wave_sync();
    }
  }
  __syncthreads();
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
void lomuto_sort(exh_ctx* c, K* keys, int* vals, long long n, double floor) {
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
  size_t o_lo[2], o_hi[2], o_toff[2], o_j0[2];
  for (int q = 0; q < 2; q++) {
    o_lo[q] = take(maxseg * 8);
    o_hi[q] = take(maxseg * 8);
    o_toff[q] = take((maxseg + 1) * 8);
    o_j0[q] = take(maxseg * 8);
  }
  const size_t o_tle = take((maxtiles + 1) * 8), o_tpre = take((maxtiles + 2) * 8),
               o_info = take(maxseg * sizeof(SegInfo)), o_slo = take(maxsmall * 8),
               o_shi = take(maxsmall * 8), o_sbuf = take(maxsmall * 4),
               o_st = take(sizeof(LevelState)), o_round = take(64 * 4),
               o_arot = take((size_t)n * 4), o_s0 = take((size_t)n * 4),
               o_need = take(maxseg * 4 + 64), o_jmp = take((size_t)n * 4 + 64),
               o_tseg = take((maxtiles + 1) * 4);
  char* aux = (char*)ctx_buf(c, BUF_SORT_AUX, bytes);
  // work buffers: 1 = B, 2 = C; the output (0) is `keys`/`vals`
  K* kB = (K*)ctx_buf(c, BUF_SORT_K, (size_t)n * sizeof(K));
  int* vB = (int*)ctx_buf(c, BUF_SORT_V, (size_t)n * sizeof(int));
  K* kC = (K*)ctx_buf(c, BUF_SORT_K2, (size_t)n * sizeof(K));
  int* vC = (int*)ctx_buf(c, BUF_SORT_V2, (size_t)n * sizeof(int));
  long long* tle = (long long*)(aux + o_tle);
  long long* tpre = (long long*)(aux + o_tpre);
  LevelState* st = (LevelState*)(aux + o_st);
  int* round_open = (int*)(aux + o_round);
  int* arot = (int*)(aux + o_arot);
  int* s0 = (int*)(aux + o_s0);
  int* jmp = (int*)(aux + o_jmp);
  LevelState* hs = (LevelState*)ctx_pinned(c, sizeof(LevelState));

  LevelBufs b{};
  b.slo = (long long*)(aux + o_slo);
  b.shi = (long long*)(aux + o_shi);
  b.sbuf = (int*)(aux + o_sbuf);
  b.small_max = SMALL;
  b.info = (SegInfo*)(aux + o_info);
  b.seg_need = (int*)(aux + o_need);
  b.tseg = (int*)(aux + o_tseg);
  b.st = st;
  b.floor = floor;
  int cur = 0;
  long long nsmall_host = 0;
  // Small segments are final once their level has placed them: each level's
  // new ones are sorted on the context's side stream (forked after the
  // level, joined when the sort ends), beside the next levels.
  if (!c->side) {
    EXH_CHECK_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    EXH_CHECK_HIP(hipEventCreateWithFlags(&c->ev_side_fork, hipEventDisableTiming));
    EXH_CHECK_HIP(hipEventCreateWithFlags(&c->ev_side_join, hipEventDisableTiming));
  }
  struct SideJoin {
    exh_ctx* c;
    bool used = false;
    ~SideJoin() {
      if (!used) return;
      (void)hipEventRecord(c->ev_side_join, c->side);
      (void)hipStreamWaitEvent(c->stream, c->ev_side_join, 0);
    }
  } join{c};
  long long nsmall_done = 0;
  // The queued segments are launched once at least SMALL_BATCH of them are
  // pending (and the rest after the last level).  A launch lasts as long as
  // its slowest segment (one workgroup partitioning up to SMALL elements
  // level by level), so one launch per level -- 27 at 512^3 (13 M peaks),
  // most of them queued back to back after the levels had ended -- kept the
  // side stream busy for 6.7 ms; in batches of 1024, 7 launches took
  // 2.9 ms and the fused step ran 26.1 -> 25.4 ms.
  auto sort_small = [&](bool last) {
    if (nsmall_host <= nsmall_done) return;
    if (!last && nsmall_host - nsmall_done < SMALL_BATCH) return;
    EXH_CHECK_HIP(hipEventRecord(c->ev_side_fork, c->stream));
    EXH_CHECK_HIP(hipStreamWaitEvent(c->side, c->ev_side_fork, 0));
    join.used = true;
    hipLaunchKernelGGL((k_small_sort<K, SMALL>), dim3((unsigned)(nsmall_host - nsmall_done)),
                       dim3(64 * SMALL_WAVES), 0, c->side, keys, vals, kB, vB, kC, vC,
                       b.slo + nsmall_done, b.shi + nsmall_done, b.sbuf + nsmall_done);
    EXH_LAUNCH_CHECK();
    nsmall_done = nsmall_host;
  };
  if (n <= SMALL) {
    long long init[2] = {0, n - 1};
    int zero = 0;
    EXH_CHECK_HIP(hipMemcpyAsync(b.slo, &init[0], 8, hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipMemcpyAsync(b.shi, &init[1], 8, hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipMemcpyAsync(b.sbuf, &zero, 4, hipMemcpyHostToDevice, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    nsmall_host = 1;
  } else {
    // level 0 reads a copy, so pivots can be written to the output at once
    EXH_CHECK_HIP(hipMemcpyAsync(kC, keys, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice,
                                 c->stream));
    EXH_CHECK_HIP(hipMemcpyAsync(vC, vals, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice,
                                 c->stream));
    hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(64), 0, c->stream,
                       (long long*)(aux + o_lo[0]), (long long*)(aux + o_hi[0]),
                       (long long*)(aux + o_toff[0]), (unsigned long long*)(aux + o_j0[0]), st, n);
    EXH_LAUNCH_CHECK();
    b.nseg = 1;
    b.ntiles = (n - 1 + TILE - 1) / TILE;
  }
  // src/dst work buffers of the current level (ids 2 and 1 at level 0)
  int src_id = 2;
  while (b.nseg > 0) {
    c->stat_sort_levels++;
    const int dst_id = 3 - src_id;
    K* ks = src_id == 1 ? kB : kC;
    int* vs = src_id == 1 ? vB : vC;
    K* kd = dst_id == 1 ? kB : kC;
    int* vd = dst_id == 1 ? vB : vC;
    b.lo = (const long long*)(aux + o_lo[cur]);
    b.hi = (const long long*)(aux + o_hi[cur]);
    b.toff = (const long long*)(aux + o_toff[cur]);
    b.j0 = (unsigned long long*)(aux + o_j0[cur]);
    b.nlo = (long long*)(aux + o_lo[1 - cur]);
    b.nhi = (long long*)(aux + o_hi[1 - cur]);
    b.ntoff = (long long*)(aux + o_toff[1 - cur]);
    b.nj0 = (unsigned long long*)(aux + o_j0[1 - cur]);
    b.dst_id = dst_id;
    EXH_REQUIRE(b.ntiles <= maxtiles && b.ntiles < (1ll << 31), EXH_ERR_INTERNAL,
                "peak sort: tile count exceeded");
    const unsigned nt = (unsigned)b.ntiles;
    hipLaunchKernelGGL(k_tile_count<K>, dim3(nt), dim3(256), 0, c->stream, ks, b, tle);
    EXH_LAUNCH_CHECK();
    exclusive_scan<long long>(c, tle, tpre, b.ntiles);
    hipLaunchKernelGGL(k_tile_scatter<K>, dim3(nt), dim3(256), 0, c->stream, ks, vs, kd, vd,
                       keys, vals, b, tpre, arot, s0);
    EXH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_chase<K>, dim3(nt), dim3(256), 0, c->stream, ks, vs, kd, vd, keys,
                       vals, b, arot, s0, jmp, chase_cap());
    EXH_LAUNCH_CHECK();
    // long chains (segments the chase flagged; the others return at once):
    // asynchronous pointer jumping, then the walk to the slots, before the
    // level's one host read
    hipLaunchKernelGGL(k_double_walk<K>, dim3(nt), dim3(256), 0, c->stream, ks, vs, kd, vd, keys,
                       vals, b, jmp, s0, 64);
    EXH_LAUNCH_CHECK();
    EXH_CHECK_HIP(hipMemcpyAsync(hs, st, sizeof(LevelState), hipMemcpyDeviceToHost, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (hs->need_any) c->stat_sort_doubling++;
    if (hs->async_open) {
      // chains still open: synchronous doubling rounds, bounded by the
      // steps left, then the walk again (strict)
      c->stat_sort_async_open++;
      const unsigned long long rm = hs->rmax;
      int rounds = 1;
      while ((1ull << (rounds - 1)) < rm && rounds < 62) rounds++;
      EXH_REQUIRE(rounds <= 40, EXH_ERR_INTERNAL, "peak sort: segment too long for doubling");
      hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, round_open, 64, 0);
      for (int k = 0; k < rounds; k++)
        hipLaunchKernelGGL(k_double_round, dim3(nt), dim3(256), 0, c->stream, b, jmp,
                           round_open, k);
      EXH_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_tile_chase2<K>, dim3(nt), dim3(256), 0, c->stream, ks, vs, kd, vd,
                         keys, vals, b, jmp, s0, 1);
      EXH_LAUNCH_CHECK();
    }
    const int nseg_next = (int)(hs->pk >> 32);
    const long long ntiles_next = (long long)(hs->pk & 0xffffffffull);
    nsmall_host = (long long)hs->nsmall;
    EXH_REQUIRE(nseg_next <= maxseg && nsmall_host <= maxsmall, EXH_ERR_INTERNAL,
                "peak sort: segment lists overflow (%d, %lld)", nseg_next, nsmall_host);
    sort_small(false);
    b.nseg = nseg_next;
    b.ntiles = ntiles_next;
    cur = 1 - cur;
    src_id = dst_id;
  }
  sort_small(true);
  if (c->stat_sort_async_open > 0) {
    EXH_CHECK_HIP(hipMemcpyAsync(&hs->err, &st->err, 8, hipMemcpyDeviceToHost, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    EXH_REQUIRE(!hs->err, EXH_ERR_INTERNAL, "peak sort: pointer doubling did not converge");
  }
}

template void lomuto_sort<float>(exh_ctx*, float*, int*, long long, double);
template void lomuto_sort<double>(exh_ctx*, double*, int*, long long, double);

}  // namespace exh
