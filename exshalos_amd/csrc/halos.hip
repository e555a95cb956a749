// Excursion-set halo finder on the GPU (find_halos.c:96-476).
//
// 1. peaks: strict 6-neighbour periodic maxima, compacted in lexicographic
//    cell order (find_halos.c:96-156);
// 2. Lomuto quicksort emulation (sort.hip), exact tie order;
// 3. greedy spherical growth in deterministic rounds.  The reference visits
//    peaks from the densest down and each growth sees the flags of all
//    denser halos (find_halos.c:226-379).  Every round, each pending
//    candidate grows speculatively against the flags committed so far
//    (thread per candidate, the reference's exact double-precision sum
//    order) and claims the cells it could write (r^2 < G, G = first failing
//    shell; more flags can only stop the growth earlier, so the final halo
//    lies inside that ball).  A candidate whose read ball (r^2 <= G) holds
//    no claim of a denser pending candidate has seen its final flags: it
//    commits.  The densest pending candidate always commits, so the rounds
//    terminate; the result equals the sequential loop.
// 4. masses: taus uniforms by GF(2) jump-ahead, Tinker-integral splines
//    (find_halos.c:442-476);
// 5. halo positions from the displacement field (lpt.c:70-74, 245-255).
#include "common.hpp"
#include "kernels.hpp"
#include "util.hpp"

namespace exh {

static constexpr int FLAG_FREE = -1;
static constexpr int FLAG_REJECTED = -2;
static constexpr int S_TAB_MAX = 1024;  // largest r^2 a halo may reach

__device__ inline size_t cell_off(int ci, int cj, int ck, unsigned p, int n0, int n1, int n2) {
  int di = (int)(p & 255u) - 128, dj = (int)((p >> 8) & 255u) - 128,
      dk = (int)((p >> 16) & 255u) - 128;
  return ((size_t)wrap1(ci, di, n0) * n1 + wrap1(cj, dj, n1)) * n2 + wrap1(ck, dk, n2);
}

// ---------------------------------------------------------------------------
// peaks
// ---------------------------------------------------------------------------
// Cells are below 2^30 here, so indices and offsets are 32-bit.  A thread
// tests the cells c0 + q * 256 (q < PK_ITEMS) of its block's tile: the
// coordinates of the first come from two multiply-high divisions, the
// others by stepping, and the six neighbours from offsets with the
// periodic wrap as a select.  Every neighbour is loaded (no short-circuit),
// so the loads of a thread's cells are all in flight together.
static constexpr int PK_ITEMS = 8;
static constexpr int PK_TILE = 256 * PK_ITEMS;

struct PeakGeom {
  int n0, n1, n2;
  unsigned sx, sy;  // plane and row strides
  FastDiv dz, dy;
};

template <typename T>
__device__ inline bool is_peak_at(const T* d, unsigned c, T v, int i, int j, int k,
                                  const PeakGeom& g) {
  const unsigned xp = i == g.n0 - 1 ? c - (unsigned)(g.n0 - 1) * g.sx : c + g.sx;
  const unsigned xm = i == 0 ? c + (unsigned)(g.n0 - 1) * g.sx : c - g.sx;
  const unsigned yp = j == g.n1 - 1 ? c - (unsigned)(g.n1 - 1) * g.sy : c + g.sy;
  const unsigned ym = j == 0 ? c + (unsigned)(g.n1 - 1) * g.sy : c - g.sy;
  const unsigned zp = k == g.n2 - 1 ? c - (unsigned)(g.n2 - 1) : c + 1;
  const unsigned zm = k == 0 ? c + (unsigned)(g.n2 - 1) : c - 1;
  const T a = d[xp], b = d[xm], e = d[yp], f = d[ym], h = d[zp], l = d[zm];
  return (v > a) & (v > b) & (v > e) & (v > f) & (v > h) & (v > l);
}

// Peak flags of the thread's PK_ITEMS cells (cells past the grid: false).
// The pass also copies those cells of delta to eff, the growth rounds'
// effective density.
template <typename T>
__device__ inline unsigned peak_bits(const T* d, unsigned c0, unsigned ng, const PeakGeom& g,
                                     T* eff) {
  unsigned r = fdiv(c0, g.dz);
  int k = (int)(c0 - r * (unsigned)g.n2);
  int i = (int)fdiv(r, g.dy);
  int j = (int)(r - (unsigned)i * (unsigned)g.n1);
  unsigned bits = 0;
  T cv[PK_ITEMS];
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++) {
    const unsigned c = c0 + (unsigned)q * 256u;
    cv[q] = 0;
    if (c < ng) {
      cv[q] = d[c];
      if (is_peak_at(d, c, cv[q], i, j, k, g)) bits |= 1u << q;
    }
    k += 256;
    while (k >= g.n2) {
      k -= g.n2;
      if (++j == g.n1) {
        j = 0;
        ++i;
      }
    }
  }
  // stores after every load, so that no load waits behind a store
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++) {
    const unsigned c = c0 + (unsigned)q * 256u;
    if (c < ng) eff[c] = cv[q];
  }
  return bits;
}

// Peaks are found over tiles of PK_TILE consecutive cells; thread x of a
// block tests cells tile + q * 256 + x (q < PK_ITEMS), so every load row is
// coalesced.  The write pass ranks peaks in cell order from per-row wave
// ballots, which keeps the reference's lexicographic peak order.
// The count pass keeps each thread's PK_ITEMS peak flags as one byte
// (ng / 8 bytes), so the write pass reads them instead of testing the
// neighbours again.
template <typename T>
__global__ __launch_bounds__(256) void k_peak_count(const T* d, PeakGeom g, long long* bcnt,
                                                    T* eff, unsigned char* pbits) {
  static_assert(PK_ITEMS == 8, "one byte of peak flags per thread");
  __shared__ int sh[4];
  const unsigned ng = (unsigned)g.n0 * g.sx;
  const unsigned base = blockIdx.x * (unsigned)PK_TILE + threadIdx.x;
  const unsigned bits = peak_bits(d, base, ng, g, eff);
  pbits[blockIdx.x * 256u + threadIdx.x] = (unsigned char)bits;
  int cnt = __popc(bits);
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) bcnt[blockIdx.x] = (long long)sh[0] + sh[1] + sh[2] + sh[3];
}

template <typename T>
__global__ __launch_bounds__(256) void k_peak_write(const T* d, PeakGeom g, const long long* boff,
                                                    const unsigned char* pbits, T* key, int* cell,
                                                    int* flag) {
  __shared__ int wcnt[PK_ITEMS][4];
  const unsigned ng = (unsigned)g.n0 * g.sx;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned base = blockIdx.x * (unsigned)PK_TILE + threadIdx.x;
  const unsigned bits = pbits[blockIdx.x * 256u + threadIdx.x];
  // the peaks' densities, loaded before the first store (which could alias
  // them for the compiler)
  T dv[PK_ITEMS];
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++)
    dv[q] = (bits >> q) & 1u ? d[base + (unsigned)q * 256u] : (T)0;
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++) {
    const unsigned c = base + (unsigned)q * 256u;
    if (c < ng) flag[c] = FLAG_FREE;
  }
  unsigned long long bal[PK_ITEMS];
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++) {
    bal[q] = __ballot((bits >> q) & 1u);
    if (lane == 0) wcnt[q][w] = __popcll(bal[q]);
  }
  __syncthreads();
  long long o = boff[blockIdx.x];
  const unsigned long long below = lanemask_lt();
  for (int q = 0; q < PK_ITEMS; q++) {
    long long ow = o;
    for (int u = 0; u < w; u++) ow += wcnt[q][u];
    if ((bal[q] >> lane) & 1ull) {
      const long long at = ow + __popcll(bal[q] & below);
      const unsigned c = base + (unsigned)q * 256u;
      key[at] = dv[q];
      cell[at] = (int)c;
    }
    o += wcnt[q][0] + wcnt[q][1] + wcnt[q][2] + wcnt[q][3];
  }
}

// Candidates by rank (0 = densest): rank r <-> sorted position np-1-r,
// kept for sorted positions >= lb (density >= the minimum barrier, and
// never position 0, which the reference does not visit).  The sort carried
// each peak's cell along with its density.
template <typename T>
__global__ void k_cand_setup(const T* skey, const int* scell, long long np, long long lb,
                             int* ccell, T* cden) {
  for (long long l = lb + blockIdx.x * (long long)blockDim.x + threadIdx.x; l < np;
       l += (long long)gridDim.x * blockDim.x) {
    const long long r = np - 1 - l;
    ccell[r] = scell[l];
    cden[r] = skey[l];
  }
}

// Block counts of the non-negative entries of v (PK_TILE entries a block),
// for the in-order compaction of the retry list.
__global__ void k_nonneg_count(const int* v, const long long* dn, long long* bcnt) {
  __shared__ long long sh[4];
  const long long n = *dn;
  long long base = (long long)blockIdx.x * PK_TILE + (long long)threadIdx.x * PK_ITEMS;
  long long cnt = 0;
  for (int q = 0; q < PK_ITEMS; q++)
    if (base + q < n && v[base + q] >= 0) cnt++;
  long long tot;
  block_exclusive_256<long long>(cnt, sh, &tot);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

// Order of the first round: columns of BR x BR cells in (i, j) over all k
// ("bricks"), one after the other, and within a brick by blocks of ORD_RB
// consecutive ranks (similar densities, hence similar growth work per
// wave).  A counting sort: per rank block, a histogram over bricks; one
// exclusive scan over (brick, block); then each block scatters its ranks.
static constexpr int ORD_RB = 4096;
static constexpr int RT_RING = 16;   // device ring of per-round pending counts
static constexpr int RT_BATCH = 4;   // rounds per host read after the first
static constexpr int ORD_MAXB = 8192;  // bricks (LDS ints)

__device__ inline int brick_of(int cell, int n1, int n2, int bsh, int nby) {
  const int r = cell / n2;
  const int j = r % n1, i = r / n1;
  return (i >> bsh) * nby + (j >> bsh);
}

__global__ void k_brick_hist(const int* ccell, long long nc, int n1, int n2, int bsh, int nby,
                             int nbr, int nblk, int* counts) {
  __shared__ int hist[ORD_MAXB];
  for (int b = threadIdx.x; b < nbr; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const long long r0 = (long long)blockIdx.x * ORD_RB;
  const long long r1 = r0 + ORD_RB < nc ? r0 + ORD_RB : nc;
  for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x)
    atomicAdd(&hist[brick_of(ccell[r], n1, n2, bsh, nby)], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < nbr; b += blockDim.x)
    counts[(long long)b * nblk + blockIdx.x] = hist[b];
}

template <typename T>
__global__ void k_brick_scatter(const int* ccell, const T* cden, long long nc, int n1, int n2,
                                int bsh, int nby, int nbr, int nblk, const int* offs, int* prank,
                                int* pcell, T* pden) {
  __shared__ int cur[ORD_MAXB];
  for (int b = threadIdx.x; b < nbr; b += blockDim.x)
    cur[b] = offs[(long long)b * nblk + blockIdx.x];
  __syncthreads();
  const long long r0 = (long long)blockIdx.x * ORD_RB;
  const long long r1 = r0 + ORD_RB < nc ? r0 + ORD_RB : nc;
  // the block's cells and densities first, all loads in flight together
  // (the stores below could alias them for the compiler); then the ranks
  // in the same order as a strided loop would take them
  constexpr int PER = ORD_RB / 256;
  int cells[PER];
  T dens[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const long long r = r0 + threadIdx.x + 256 * q;
    if (r < r1) {
      cells[q] = ccell[r];
      dens[q] = cden[r];
    }
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const long long r = r0 + threadIdx.x + 256 * q;
    if (r < r1) {
      const int pos = atomicAdd(&cur[brick_of(cells[q], n1, n2, bsh, nby)], 1);
      prank[pos] = (int)r;
      pcell[pos] = cells[q];
      pden[pos] = dens[q];
    }
  }
}

// In-order compaction of the candidates to retry (retry[t] >= 0), carrying
// their cell and density along.
template <typename T>
__global__ void k_retry_write(const int* retry, const long long* dn, const long long* boff,
                              long long nbk, const int* cell_in, const T* den_in, int* rank_out,
                              int* cell_out, T* den_out, long long* dnext) {
  __shared__ long long sh[4];
  const long long n = *dn;
  if (blockIdx.x == 0 && threadIdx.x == 0) *dnext = boff[nbk];  // next round's count
  long long base = (long long)blockIdx.x * PK_TILE + (long long)threadIdx.x * PK_ITEMS;
  // the thread's entries, cells and densities are all loaded before the
  // first store (which could alias them for the compiler)
  int rv[PK_ITEMS], cv[PK_ITEMS];
  T dv[PK_ITEMS];
  long long cnt = 0;
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++) {
    rv[q] = base + q < n ? retry[base + q] : -1;
    cnt += rv[q] >= 0;
  }
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++)
    if (rv[q] >= 0) {
      cv[q] = cell_in[base + q];
      dv[q] = den_in[base + q];
    }
  long long tot;
  long long o = block_exclusive_256<long long>(cnt, sh, &tot) + boff[blockIdx.x];
#pragma unroll
  for (int q = 0; q < PK_ITEMS; q++)
    if (rv[q] >= 0) {
      rank_out[o] = rv[q];
      cell_out[o] = cv[q];
      den_out[o] = dv[q];
      o++;
    }
}

// number of keys below Bmin (the first candidate's sorted position); *out
// starts at zero
template <typename T>
__global__ void k_count_below(const T* key, long long n, T Bmin, unsigned long long* out) {
  __shared__ unsigned sh[4];
  unsigned cnt = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    cnt += key[i] < Bmin;
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicAdd(out, (unsigned long long)(sh[0] + sh[1] + sh[2] + sh[3]));
}

__global__ void k_fill_i32(int* p, long long n, int v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, long long n, unsigned long long v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_iota(int* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = (int)i;
}

// ---------------------------------------------------------------------------
// growth rounds
// ---------------------------------------------------------------------------
template <typename T>
struct GrowArgs {
  int n0, n1, n2;
  const T* delta;
  T* eff;          // delta, or -Mtot where a halo (or a rejected sphere) has flagged the cell
  int* flag;
  unsigned long long* claim;
  ShellTab sh;
  const T* Bt;
  int nB;
  T Bmin;
  T negMtot;
  int Nmin;
  // the round's candidates by list position: rank (0 = densest), centre
  // cell and density, and the growth results handed from k_grow to k_commit
  const int* prank;
  const int* pcell;
  const T* pden;
  int* pG;         // first failing shell (read radius^2); -1 = skipped
  int* pcnt;       // count
  int* pr2;        // grows_tmp
  // final results by rank
  int* cnt;
  int* r2;
  unsigned char* kept;
  int* err;
};

template <typename T>
__device__ inline void decode_cell(int cell, int n1, int n2, int& i, int& j, int& k) {
  k = cell % n2;
  int r = cell / n2;
  j = r % n1;
  i = r / n1;
}

// Workgroup w runs on XCD w mod 8: give each XCD a contiguous range of
// blocks, so that a round's consecutive candidates (neighbours in space)
// share one L2.
__device__ inline int xcd_block(unsigned nb) {
  const unsigned w = blockIdx.x;
  const unsigned q = nb / 8, rem = nb % 8, x = w % 8;
  return (int)(x * q + (x < rem ? x : rem) + w / 8);
}

// The round's count is read on the device (*dn): the grid may be larger
// (sized by an earlier round's count) and its extra blocks return.
// Claims are 64-bit (round tag << 32 | ~rank) or, with fewer than 2^24
// candidates, 32-bit (round tag << 24 | (2^24 - 1 - rank)): atomicMax keeps
// the densest claim of the current round, and any claim of the round beats
// those of earlier rounds.
template <typename T, bool C32>
__global__ void k_grow(GrowArgs<T> a, const long long* dn, unsigned long long tag_hi) {
  const int npend = (int)*dn;
  const unsigned nba = (unsigned)((npend + 255) / 256);
  if (blockIdx.x >= nba) return;
  int t = xcd_block(nba) * blockDim.x + threadIdx.x;
  if (t >= npend) return;
  const int rank = a.prank[t];
  const int cell = a.pcell[t];
  if (a.flag[cell] != FLAG_FREE) {  // already inside a denser halo
    a.pG[t] = -1;
    a.kept[rank] = 0;
    return;
  }
  int ci, cj, ck;
  decode_cell<T>(cell, a.n1, a.n2, ci, cj, ck);
  const double Bmin = (double)a.Bmin;
  double den = (double)a.pden[t], den_tmp = den;
  int count = 0, count_tmp = 1, grows = 0, grows_tmp = 0;
  while (den_tmp >= Bmin) {
    if (count < count_tmp) grows = grows_tmp;
    grows_tmp++;
    if (grows_tmp >= a.sh.stab) {
      atomicOr(a.err, 1);
      a.pG[t] = -1;
      return;
    }
    den = den_tmp;
    count = count_tmp;
    den_tmp = den * ((double)count);
    // the shell's cells in cube order; loads issued in batches, summed in
    // order (flagged cells hold -Mtot in eff, as the reference adds)
    const int o0 = a.sh.start[grows_tmp], o1 = a.sh.start[grows_tmp + 1];
    for (int ob = o0; ob < o1; ob += 8) {
      T v[8];
      const int nb = min(8, o1 - ob);
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (e < nb) v[e] = a.eff[cell_off(ci, cj, ck, a.sh.offs[ob + e], a.n0, a.n1, a.n2)];
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (e < nb) den_tmp += (double)v[e];
      count_tmp += nb;
    }
    den_tmp = den_tmp / ((double)count_tmp);
  }
  const int G = grows_tmp;
  for (;;) {
    if (count >= a.nB) {
      atomicOr(a.err, 2);
      a.pG[t] = -1;
      return;
    }
    if (!(den < (double)a.Bt[count] && count > 0)) break;
    if (grows < 0) {  // cannot happen while count > 0 on an unclamped table
      atomicOr(a.err, 32);
      a.pG[t] = -1;
      return;
    }
    den_tmp = den;
    count_tmp = count;
    den = den * ((double)count);
    const int o0 = a.sh.start[grows], o1 = a.sh.start[grows + 1];
    for (int ob = o0; ob < o1; ob += 8) {
      T v[8];
      const int nb = min(8, o1 - ob);
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (e < nb) v[e] = a.delta[cell_off(ci, cj, ck, a.sh.offs[ob + e], a.n0, a.n1, a.n2)];
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (e < nb) den -= (double)v[e];
      count -= nb;
    }
    if (count > 0) den = den / ((double)count);
    if (count < count_tmp) grows_tmp = grows;
    grows--;
  }
  a.pG[t] = G;
  a.pcnt[t] = count;
  a.pr2[t] = grows_tmp;
  // claim every cell the final halo could occupy (r^2 < G); the offsets
  // of 8 cells are loaded before their atomics (which could alias them for
  // the compiler)
  const int oend = a.sh.start[G];
  for (int ob = 0; ob < oend; ob += 8) {
    const int nb = min(8, oend - ob);
    unsigned off[8];
#pragma unroll
    for (int e = 0; e < 8; e++)
      if (e < nb) off[e] = a.sh.offs[ob + e];
    if constexpr (C32) {
      unsigned* cl = reinterpret_cast<unsigned*>(a.claim);
      const unsigned tag = (unsigned)tag_hi | (0xffffffu - (unsigned)rank);
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (e < nb) atomicMax(&cl[cell_off(ci, cj, ck, off[e], a.n0, a.n1, a.n2)], tag);
    } else {
      const unsigned long long tag = tag_hi | (unsigned long long)(0xffffffffu - (unsigned)rank);
#pragma unroll
      for (int e = 0; e < 8; e++)
        if (e < nb) atomicMax(&a.claim[cell_off(ci, cj, ck, off[e], a.n0, a.n1, a.n2)], tag);
    }
  }
}

// retry[t]: the rank again when candidate t must retry next round, else -1
// (compacted in order, so the next round keeps this round's order).
template <typename T, bool C32>
__global__ void k_commit(GrowArgs<T> a, const long long* dn, unsigned long long tag_hi,
                         int* retry, int tmax) {
  const int npend = (int)*dn;
  const unsigned nba = (unsigned)((npend + 255) / 256);
  if (blockIdx.x >= nba) return;
  int t = xcd_block(nba) * blockDim.x + threadIdx.x;
  if (t >= npend) return;
  const int rank = a.prank[t];
  const int G = a.pG[t];
  if (G < 0) {  // skipped or failed
    retry[t] = -1;
    return;
  }
  const int cell = a.pcell[t];
  int ci, cj, ck;
  decode_cell<T>(cell, a.n1, a.n2, ci, cj, ck);
  // safe iff no denser pending candidate claims a cell of the read ball
  const int oend = a.sh.start[G + 1];
  bool blocked = false;
  if constexpr (C32) {
    const unsigned* cl = reinterpret_cast<const unsigned*>(a.claim);
    const unsigned lim = 0xffffffu - (unsigned)rank, th = (unsigned)tag_hi;
    for (int ob = 0; ob < oend && !blocked; ob += 8) {
      unsigned v[8];
      const int nb = min(8, oend - ob);
#pragma unroll
      for (int e = 0; e < 8; e++)
        v[e] = e < nb ? cl[cell_off(ci, cj, ck, a.sh.offs[ob + e], a.n0, a.n1, a.n2)] : 0u;
#pragma unroll
      for (int e = 0; e < 8; e++)
        blocked |= (v[e] & 0xff000000u) == th && (v[e] & 0xffffffu) > lim;
    }
  } else {
    const unsigned lim = 0xffffffffu - (unsigned)rank;
    for (int ob = 0; ob < oend && !blocked; ob += 8) {
      unsigned long long v[8];
      const int nb = min(8, oend - ob);
#pragma unroll
      for (int e = 0; e < 8; e++)
        v[e] = e < nb ? a.claim[cell_off(ci, cj, ck, a.sh.offs[ob + e], a.n0, a.n1, a.n2)] : 0ull;
#pragma unroll
      for (int e = 0; e < 8; e++)
        blocked |=
            (v[e] & 0xffffffff00000000ull) == tag_hi && (unsigned)(v[e] & 0xffffffffu) > lim;
    }
  }
  retry[t] = blocked ? rank : -1;
  if (blocked) return;
  const int count = a.pcnt[t];
  if (count == 0) {
    a.kept[rank] = 0;
    return;
  }
  const int R2 = a.pr2[t];
  if (isqrt_floor(R2 - 1) > tmax) {  // flag sphere would wrap (reference writes past its cube)
    atomicOr(a.err, 4);
    a.kept[rank] = 0;
    return;
  }
  bool keep = count >= a.Nmin;
  int v = keep ? (rank << 1) : FLAG_REJECTED;
  const int fend = a.sh.start[R2];
  for (int ob = 0; ob < fend; ob += 8) {  // offsets of 8 cells before their stores
    const int nb = min(8, fend - ob);
    unsigned off[8];
#pragma unroll
    for (int e = 0; e < 8; e++)
      if (e < nb) off[e] = a.sh.offs[ob + e];
#pragma unroll
    for (int e = 0; e < 8; e++)
      if (e < nb) {
        const size_t id = cell_off(ci, cj, ck, off[e], a.n0, a.n1, a.n2);
        a.flag[id] = v;
        a.eff[id] = a.negMtot;
      }
  }
  if (keep) a.flag[cell] = (rank << 1) | 1;
  a.kept[rank] = keep ? 1 : 0;
  a.cnt[rank] = count;
  a.r2[rank] = R2;
}

// ---------------------------------------------------------------------------
// catalogue, masses, flags, positions
// ---------------------------------------------------------------------------
__global__ void k_kept_to_int(const unsigned char* kept, long long n, long long* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = kept[i];
}

template <typename T>
__global__ void k_gather_halos(const unsigned char* kept, const long long* hidx, long long nc,
                               const int* ccell, const int* cnt, const int* r2, int n1, int n2,
                               int* hx, int* hcount, int* hr2) {
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < nc;
       r += (long long)gridDim.x * blockDim.x) {
    if (!kept[r]) continue;
    const long long h = hidx[r];
    const int cell = ccell[r], hc = cnt[r], hr = r2[r];  // all loads before the stores
    int i, j, k;
    decode_cell<T>(cell, n1, n2, i, j, k);
    hx[3 * h] = i;
    hx[3 * h + 1] = j;
    hx[3 * h + 2] = k;
    hcount[h] = hc;
    hr2[h] = hr;
  }
}

// taus uniforms: thread b produces draws [b*64, b*64+64)
__global__ void k_taus(const unsigned* mats, unsigned s1, unsigned s2, unsigned s3, long long n,
                       double* u) {
  long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long first = b * 64;
  if (first >= n) return;
  unsigned st[3] = {s1, s2, s3};
  for (int comp = 0; comp < 3; comp++) {
    const unsigned* M = mats + comp * 40 * 32;
    unsigned v = st[comp];
    for (int bit = 0; bit < 40; bit++) {
      if (!((first >> bit) & 1)) continue;
      unsigned r = 0;
      for (int col = 0; col < 32; col++)
        if ((v >> col) & 1u) r ^= M[bit * 32 + col];
      v = r;
    }
    st[comp] = v;
  }
  for (int q = 0; q < 64 && first + q < n; q++) {
    unsigned b1 = ((st[0] << 13) ^ st[0]) >> 19;
    st[0] = ((st[0] & 4294967294u) << 12) ^ b1;
    unsigned b2 = ((st[1] << 2) ^ st[1]) >> 25;
    st[1] = ((st[1] & 4294967288u) << 4) ^ b2;
    unsigned b3 = ((st[2] << 3) ^ st[2]) >> 11;
    st[2] = ((st[2] & 4294967280u) << 17) ^ b3;
    u[first + q] = (double)(st[0] ^ st[1] ^ st[2]) / 4294967296.0;
  }
}

__device__ inline double spline_eval_d(const double* x, const double* y, const double* m, int n,
                                       double xv, int* bad) {
  if (xv < x[0] || xv > x[n - 1]) {
    *bad = 1;
    xv = xv < x[0] ? x[0] : x[n - 1];
  }
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (x[mid] > xv)
      hi = mid;
    else
      lo = mid;
  }
  double h = x[hi] - x[lo];
  double t = xv - x[lo];
  double b = (y[hi] - y[lo]) / h - h * (2.0 * m[lo] + m[hi]) / 6.0;
  double c = m[lo] / 2.0;
  double d = (m[hi] - m[lo]) / (6.0 * h);
  return y[lo] + t * (b + t * (c + t * d));
}

// Mh = InvI(I(M_count) + (I(M_next) - I(M_count)) u)   (find_halos.c:455-476)
template <typename T>
__global__ void k_masses(const int* hcount, long long nh, const int* sphere, int nsph,
                         int Ncells, const double* itab, const int* ibad, const double* Mk,
                         const double* Int, const double* mInv, int Nr, const double* u, T* Mh,
                         int* err) {
  for (long long h = blockIdx.x * (long long)blockDim.x + threadIdx.x; h < nh;
       h += (long long)gridDim.x * blockDim.x) {
    int cnt = hcount[h];
    // first index with sphere[i] == cnt among the first Ncells entries
    int lo = 0, hi = (Ncells < nsph ? Ncells : nsph);
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (sphere[mid] < cnt)
        lo = mid + 1;
      else
        hi = mid;
    }
    int bad = 0;
    if (lo >= (Ncells < nsph ? Ncells : nsph) || sphere[lo] != cnt || lo + 1 >= nsph) {
      atomicOr(err, 8);
      Mh[h] = (T)0;
      continue;
    }
    // I of the masses of spheres lo and lo + 1, from the table
    const double I0 = itab[lo];
    T v = (T)(I0 + (itab[lo + 1] - I0) * u[h]);
    bad = ibad[lo] | ibad[lo + 1];
    Mh[h] = (T)spline_eval_d(Int, Mk, mInv, Nr, (double)v, &bad);
    if (bad) atomicOr(err, 16);
  }
}

// I(M) at the mass (T)sphere[q] * Mcell of every sphere size q < nt, as
// k_masses would evaluate it for a halo of that size (a table, since most
// halos share a few sizes); ibad[q] flags a mass outside the spline's range.
template <typename T>
__global__ void k_mass_table(const int* sphere, int nt, T Mcell, const double* Mk,
                             const double* Int, const double* mI, int Nr, double* itab,
                             int* ibad) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nt; q += gridDim.x * blockDim.x) {
    int bad = 0;
    const T m = (T)sphere[q] * Mcell;
    itab[q] = spline_eval_d(Mk, Int, mI, Nr, (double)m, &bad);
    ibad[q] = bad;
  }
}

__global__ void k_flag_out(const int* flag, long long ng, const long long* hidx, long long* out) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ng;
       c += (long long)gridDim.x * blockDim.x) {
    int v = flag[c];
    long long o;
    if (v == FLAG_FREE)
      o = ng;
    else if (v == FLAG_REJECTED)
      o = ng + 1;
    else {
      long long h = hidx[v >> 1];
      o = (v & 1) ? -h : h;
    }
    out[c] = o;
  }
}

// halo displacement sums over the member sphere, in double
template <typename T>
__global__ void k_halo_posh(const int* hx, const int* hcount, const int* hr2, long long nh,
                            ShellTab sh, int n0, int n1, int n2, const T* S, const T* V, T Lc,
                            T L0, T L1, T L2, int lagrangian, T* posh, T* velh) {
  for (long long h = blockIdx.x * (long long)blockDim.x + threadIdx.x; h < nh;
       h += (long long)gridDim.x * blockDim.x) {
    int ci = hx[3 * h], cj = hx[3 * h + 1], ck = hx[3 * h + 2];
    double s[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    if (!lagrangian) {
      // cells in batches of 8, each cell's three components in one load,
      // all of a batch's loads in flight before the sums, which keep the
      // cell order per component
      using V3 = typename vec3<T>::type;
      const int end = sh.start[hr2[h]];
      for (int ob = 0; ob < end; ob += 8) {
        const int nb = min(8, end - ob);
        V3 sv[8], vv[8];
#pragma unroll
        for (int e = 0; e < 8; e++)
          if (e < nb) {
            const size_t id = cell_off(ci, cj, ck, sh.offs[ob + e], n0, n1, n2);
            sv[e] = *reinterpret_cast<const V3*>(S + 3 * id);
            if (V) vv[e] = *reinterpret_cast<const V3*>(V + 3 * id);
          }
#pragma unroll
        for (int e = 0; e < 8; e++)
          if (e < nb) {
            s[0] += (double)sv[e].x;
            s[1] += (double)sv[e].y;
            s[2] += (double)sv[e].z;
            if (V) {
              v[0] += (double)vv[e].x;
              v[1] += (double)vv[e].y;
              v[2] += (double)vv[e].z;
            }
          }
      }
    }
    int q[3] = {ci, cj, ck};
    T L[3] = {L0, L1, L2};
    for (int a = 0; a < 3; a++) {
      T base = (T)((double)((T)q[a] * Lc) + (double)Lc / 2.0);
      T r = base + (T)(s[a] / hcount[h]);
      if (r >= L[a]) r -= L[a];
      if (r < (T)0) r += L[a];
      posh[3 * h + a] = r;
      if (velh) velh[3 * h + a] = (T)(v[a] / hcount[h]);
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Upload the shell table for (stab, tmax) unless the context already holds
// it; returns the number of offsets.
static long long ensure_shells(exh_ctx* c, int stab, int tmax) {
  HaloState& S = c->halo;
  if (S.sh.stab == stab && S.sh_tmax == tmax) return c->shell_noffs;
  std::vector<int> start;
  std::vector<unsigned> packed;
  shell_offsets(stab, tmax, start, packed);
  size_t bytes = (start.size() + packed.size()) * 4;
  int* d = (int*)ctx_buf(c, BUF_SHELL, bytes);
  EXH_CHECK_HIP(hipMemcpyAsync(d, start.data(), start.size() * 4, hipMemcpyHostToDevice, c->stream));
  EXH_CHECK_HIP(hipMemcpyAsync(d + start.size(), packed.data(), packed.size() * 4,
                               hipMemcpyHostToDevice, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  S.sh.start = d;
  S.sh.offs = (const unsigned*)(d + start.size());
  S.sh.stab = stab;
  S.sh_tmax = tmax;
  c->shell_noffs = (long long)packed.size();
  return c->shell_noffs;
}

// Constant tables, built once per process.
static const std::vector<int>& sphere_table() {
  static const std::vector<int> t = cells_in_spheres(10000);
  return t;
}

static const std::vector<unsigned>& taus_table() {
  static const std::vector<unsigned> t = [] {
    std::vector<unsigned> m(3 * 40 * 32);
    taus_jump_tables(m.data());
    return m;
  }();
  return t;
}

template <typename T>
// phase(p), when given, is called as the finder reaches p: 0 = start,
// 1 = after the peak sort, 2 = before the growth rounds (launches of other
// streams can be placed there).
long long run_find_halos(exh_ctx* c, const exh_box* b, const T* delta, const exh_halo_opts* o,
                         const std::function<void(int)>& phase) {
  if (phase) phase(0);
  const int n0 = b->nd[0], n1 = b->nd[1], n2 = b->nd[2];
  const long long ng = b->ng;
  EXH_REQUIRE(ng < (1LL << 30), EXH_ERR_ARG, "grid too large for int32 cell indices");
  HaloState& S = c->halo;
  S.ng = ng;
  S.have_posh = S.have_velh = false;
  const int tmax = b->nmin / 2;
  const int stab = std::min(S_TAB_MAX, 3 * tmax * tmax + 2);
  const long long noffs = ensure_shells(c, stab, tmax);

  // tables (built on the host, kept on the device while their inputs are
  // unchanged): barrier per count, mass-function knots and their splines,
  // sphere counts, taus jump matrices
  const int nBt = (int)noffs + 2;
  const std::vector<int>& sph = sphere_table();
  const std::vector<unsigned>& mats = taus_table();
  size_t toff = 0;
  auto ttake = [&](size_t bytes) {
    size_t r = toff;
    toff += (bytes + 255) & ~(size_t)255;
    return r;
  };
  const int Nr = halo_table_radii(o);  // knots of the sigma(M) and Tinker tables
  const size_t o_bt = ttake(nBt * sizeof(T)), o_mk = ttake(Nr * 8), o_int = ttake(Nr * 8),
               o_mi = ttake(Nr * 8), o_minv = ttake(Nr * 8), o_sph = ttake(sph.size() * 4),
               o_mats = ttake(mats.size() * 4);
  std::vector<unsigned char> key;
  const double oscal[3] = {o->a, o->beta, o->alpha};
  const int sizes[3] = {o->Nk, nBt, o->sigma_radii};
  key_append(key, b, 1);
  key_append(key, oscal, 3);
  key_append(key, sizes, 3);
  key_append(key, o->k, o->Nk);
  key_append(key, o->P, o->Nk);
  const size_t cap0 = c->cap[BUF_HALOTAB];
  char* tabs = (char*)ctx_buf(c, BUF_HALOTAB, toff);
  if (c->cap[BUF_HALOTAB] != cap0) c->halotab_key.clear();  // reallocated
  if (key != c->halotab_key) {
    std::vector<T> sig(std::max(1, (int)std::floor(1e16 / (T)b->Mcell)) + 1), Bt(nBt);
    std::vector<double> Mk(Nr), Int(Nr);
    int Ncells = halo_tables(b, o, sig.data(), Bt.data(), nBt, Mk.data(), Int.data());
    EXH_REQUIRE(Ncells >= 1, Ncells < 0 ? EXH_ERR_RANGE : EXH_ERR_ARG,
                "sigma(M) table: cell mass %g gives %d cells below M_max", (double)b->Mcell,
                Ncells);
    std::vector<double> mI(Nr), mInv(Nr);
    spline_second_derivs(Mk.data(), Int.data(), Nr, mI.data());
    spline_second_derivs(Int.data(), Mk.data(), Nr, mInv.data());
    auto up = [&](size_t o2, const void* src, size_t bytes) {
      EXH_CHECK_HIP(hipMemcpyAsync(tabs + o2, src, bytes, hipMemcpyHostToDevice, c->stream));
    };
    up(o_bt, Bt.data(), nBt * sizeof(T));
    up(o_mk, Mk.data(), Nr * 8);
    up(o_int, Int.data(), Nr * 8);
    up(o_mi, mI.data(), Nr * 8);
    up(o_minv, mInv.data(), Nr * 8);
    up(o_sph, sph.data(), sph.size() * 4);
    up(o_mats, mats.data(), mats.size() * 4);
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));  // the host tables are temporaries
    c->halotab_key = key;
    c->halotab_ncells = Ncells;
    c->halotab_bmin = (double)Bt[Ncells - 1];
  }
  const int Ncells = c->halotab_ncells;
  char* misc = (char*)ctx_buf(c, BUF_MISC, 256);
  const size_t o_cnt = 0;
  long long* dcnt = (long long*)(misc + o_cnt);  // [0..7] scratch counters
  int* derr = (int*)(misc + o_cnt + 128);
  long long* hbuf = (long long*)ctx_pinned(c, 256);

  // 1. peaks
  long long nblk = (ng + PK_TILE - 1) / PK_TILE;
  const PeakGeom pg{n0, n1, n2, (unsigned)((long long)n1 * n2), (unsigned)n2, fast_div(n2),
                    fast_div(n1)};
  // block counts and offsets, then the peak flag bytes
  const size_t o_pb = ((size_t)(2 * nblk + 2) * 8 + 255) & ~(size_t)255;
  char* hb = (char*)ctx_buf(c, BUF_HALO, o_pb + (size_t)nblk * 256);
  long long* bc = (long long*)hb;
  unsigned char* pbits = (unsigned char*)(hb + o_pb);
  int* flag = (int*)ctx_buf(c, BUF_FLAG, (size_t)ng * 4);
  T* eff = (T*)ctx_buf(c, BUF_EFF, (size_t)ng * sizeof(T));
  unsigned long long* claim = (unsigned long long*)ctx_buf(c, BUF_CLAIM, (size_t)ng * 8);
  {
    TimedScope ts(c, "peaks");
    hipLaunchKernelGGL(k_peak_count<T>, dim3((unsigned)nblk), dim3(256), 0, c->stream, delta, pg,
                       bc, eff, pbits);
    EXH_LAUNCH_CHECK();
    exclusive_scan<long long>(c, bc, bc + nblk + 1, nblk);
  }
  EXH_CHECK_HIP(hipMemcpyAsync(hbuf, bc + nblk + 1 + nblk, 8, hipMemcpyDeviceToHost, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  const long long np = hbuf[0];
  c->stat_peaks = np;
  T* pkey = (T*)ctx_buf(c, BUF_PEAK_KEY, (size_t)std::max(np, 1LL) * sizeof(T));
  int* pcell = (int*)ctx_buf(c, BUF_PEAK_CELL, (size_t)std::max(np, 1LL) * sizeof(int));
  {
    TimedScope ts(c, "peaks_write");
    hipLaunchKernelGGL(k_peak_write<T>, dim3((unsigned)nblk), dim3(256), 0, c->stream, delta, pg,
                       bc + nblk + 1, pbits, pkey, pcell, flag);
    EXH_LAUNCH_CHECK();
  }
  // 2. sort (densities in lexicographic peak order; each peak's cell rides
  // along as the value).  Peaks below the smallest barrier are never
  // visited: their count is the first visited sorted position, and the
  // sort leaves their order open.
  const T Bmin = (T)c->halotab_bmin;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, (unsigned long long*)dcnt, 1,
                     0ull);
  hipLaunchKernelGGL(k_count_below<T>, dim3(grid_for(np, 256 * 16)), dim3(256), 0, c->stream,
                     pkey, np, Bmin, (unsigned long long*)dcnt);
  EXH_LAUNCH_CHECK();
  lomuto_sort<T>(c, pkey, pcell, np, (double)Bmin);
  if (phase) phase(1);

  // 3. candidates (the sort stages its state in the same pinned buffer)
  EXH_CHECK_HIP(hipMemcpyAsync(hbuf, dcnt, 8, hipMemcpyDeviceToHost, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  long long lb = std::max(hbuf[0], 1LL);
  long long nc = np > lb ? np - lb : 0;
  c->stat_candidates = nc;

  // candidate state in BUF_CAND: G, cnt, r2 (int), kept (u8), pend x2, hidx
  size_t coff = 0;
  auto ctake = [&](size_t bytes) {
    size_t r = coff;
    coff += (bytes + 255) & ~(size_t)255;
    return r;
  };
  size_t ncs = (size_t)std::max(nc, 1LL);
  long long nblk_c = (nc + PK_TILE - 1) / PK_TILE;
  // first-round order: bricks of 2^bsh x 2^bsh cells (16 x 16, wider on big
  // grids so that the brick histogram fits in LDS)
  int bsh = 4;
  while ((long long)((n0 + (1 << bsh) - 1) >> bsh) * ((n1 + (1 << bsh) - 1) >> bsh) > ORD_MAXB)
    bsh++;
  const int nby = (n1 + (1 << bsh) - 1) >> bsh;
  const int nbr = ((n0 + (1 << bsh) - 1) >> bsh) * nby;
  const long long nblk_r = std::max((nc + ORD_RB - 1) / ORD_RB, 1LL);
  size_t o_c = ctake(ncs * 4), o_r = ctake(ncs * 4), o_k = ctake(ncs),
         o_p0 = ctake(ncs * 4), o_p1 = ctake(ncs * 4), o_kl = ctake((ncs + 1) * 8),
         o_hid = ctake((ncs + 1) * 8), o_cc = ctake(ncs * 4), o_cd = ctake(ncs * sizeof(T)),
         o_bk = ctake(((size_t)nbr * nblk_r + 1) * 4), o_rf = ctake(ncs * 4),
         o_pb2 = ctake((size_t)(2 * std::max(nblk_c, 1LL) + 2) * 8), o_pc0 = ctake(ncs * 4),
         o_pc1 = ctake(ncs * 4), o_pd0 = ctake(ncs * sizeof(T)), o_pd1 = ctake(ncs * sizeof(T)),
         o_pg = ctake(ncs * 4), o_pn = ctake(ncs * 4), o_pr = ctake(ncs * 4),
         o_rt = ctake(RT_RING * 8);
  char* cand = (char*)ctx_buf(c, BUF_CAND, coff);
  int* ccell = (int*)(cand + o_cc);
  T* cden = (T*)(cand + o_cd);
  // eff (delta) and the free flags were written by the peak passes
  S.flag = flag;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, derr, 1, 0);
  EXH_LAUNCH_CHECK();

  GrowArgs<T> ga{};
  ga.n0 = n0;
  ga.n1 = n1;
  ga.n2 = n2;
  ga.delta = delta;
  ga.eff = eff;
  ga.flag = flag;
  ga.claim = claim;
  ga.sh = S.sh;
  ga.Bt = (const T*)(tabs + o_bt);
  ga.nB = nBt;
  ga.Bmin = Bmin;
  ga.negMtot = -(T)b->Mtot;
  ga.Nmin = o->Nmin;
  ga.pG = (int*)(cand + o_pg);
  ga.pcnt = (int*)(cand + o_pn);
  ga.pr2 = (int*)(cand + o_pr);
  ga.cnt = (int*)(cand + o_c);
  ga.r2 = (int*)(cand + o_r);
  ga.kept = (unsigned char*)(cand + o_k);
  ga.err = derr;
  int* pend = (int*)(cand + o_p0);
  int* pnext = (int*)(cand + o_p1);
  int* qcell = (int*)(cand + o_pc0);
  int* qcell_next = (int*)(cand + o_pc1);
  T* qden = (T*)(cand + o_pd0);
  T* qden_next = (T*)(cand + o_pd1);
  int* retry = (int*)(cand + o_rf);
  long long* pb2 = (long long*)(cand + o_pb2);
  {
    TimedScope ts(c, "halo_candidates");
    hipLaunchKernelGGL(k_cand_setup<T>, dim3(grid_for(nc, 256 * 4)), dim3(256), 0, c->stream,
                       pkey, pcell, np, lb, ccell, cden);
    EXH_LAUNCH_CHECK();
    if (nc > 0) {
      int* bk = (int*)(cand + o_bk);
      hipLaunchKernelGGL(k_brick_hist, dim3((unsigned)nblk_r), dim3(256), 0, c->stream, ccell, nc,
                         n1, n2, bsh, nby, nbr, (int)nblk_r, bk);
      EXH_LAUNCH_CHECK();
      exclusive_scan<int>(c, bk, bk, (long long)nbr * nblk_r);
      hipLaunchKernelGGL(k_brick_scatter<T>, dim3((unsigned)nblk_r), dim3(256), 0, c->stream,
                         ccell, cden, nc, n1, n2, bsh, nby, nbr, (int)nblk_r, bk, pend, qcell,
                         qden);
    }
    EXH_LAUNCH_CHECK();
  }
  if (phase) phase(2);
  // Rounds run in batches without a host read between them: each round's
  // pending count stays on the device (a ring of RT_RING counts), grids are
  // sized by the last count the host read (counts only shrink), and the
  // host reads the counts and the error flag once per batch.  A round with
  // no pending candidate launches blocks that return at once.
  long long* rt = (long long*)(cand + o_rt);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, (unsigned long long*)rt, 1,
                     (unsigned long long)nc);
  EXH_LAUNCH_CHECK();
  // Claims carry their round's tag (an epoch counted over the context's
  // calls) above the inverted rank, so the claims of earlier rounds and
  // calls never match a later round: the array is cleared only when it is
  // newly allocated, when the claim width changes, or before the epoch
  // wraps (every 255 rounds with 32-bit claims).
  const bool c32 = nc < (1LL << 24) && c->tune_halo_claims != 64;
  const int cmode = c32 ? 32 : 64;
  auto clear_claims = [&]() {
    const long long ncap = (long long)(c->cap[BUF_CLAIM] / 8);
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ncap, 256 * 8)), dim3(256), 0, c->stream, claim,
                       ncap, 0ull);
    EXH_LAUNCH_CHECK();
    c->claim_epoch = 0;
  };
  if (c->claim_gen != c->buf_gen[BUF_CLAIM] || c->claim_mode != cmode ||
      c->claim_epoch > 0xffffffffu - 2000000u) {
    clear_claims();
    c->claim_gen = c->buf_gen[BUF_CLAIM];
    c->claim_mode = cmode;
  }
  long long ub = nc;
  int round = 0, launched = 0;
  {
    TimedScope ts(c, "halo_rounds");
    while (ub > 0) {
      const int batch = launched == 0 ? 1 : RT_BATCH;
      const int first = launched;
      for (int q = 0; q < batch; q++) {
        const long long* dn = rt + launched % RT_RING;
        long long* dnext = rt + (launched + 1) % RT_RING;
        launched++;
        if (c32 && c->claim_epoch >= 255) clear_claims();
        const unsigned long long ep = ++c->claim_epoch;
        const unsigned long long tag_hi = c32 ? ep << 24 : ep << 32;
        ga.prank = pend;
        ga.pcell = qcell;
        ga.pden = qden;
        if (c32) {
          hipLaunchKernelGGL((k_grow<T, true>), dim3(grid_for(ub, 256)), dim3(256), 0, c->stream,
                             ga, dn, tag_hi);
          hipLaunchKernelGGL((k_commit<T, true>), dim3(grid_for(ub, 256)), dim3(256), 0,
                             c->stream, ga, dn, tag_hi, retry, tmax);
        } else {
          hipLaunchKernelGGL((k_grow<T, false>), dim3(grid_for(ub, 256)), dim3(256), 0, c->stream,
                             ga, dn, tag_hi);
          hipLaunchKernelGGL((k_commit<T, false>), dim3(grid_for(ub, 256)), dim3(256), 0,
                             c->stream, ga, dn, tag_hi, retry, tmax);
        }
        EXH_LAUNCH_CHECK();
        // candidates to retry, in this round's order
        const long long nbk = (ub + PK_TILE - 1) / PK_TILE;
        hipLaunchKernelGGL(k_nonneg_count, dim3((unsigned)nbk), dim3(256), 0, c->stream, retry,
                           dn, pb2);
        EXH_LAUNCH_CHECK();
        exclusive_scan<long long>(c, pb2, pb2 + nbk + 1, nbk);
        hipLaunchKernelGGL(k_retry_write<T>, dim3((unsigned)nbk), dim3(256), 0, c->stream, retry,
                           dn, pb2 + nbk + 1, nbk, qcell, qden, pnext, qcell_next, qden_next,
                           dnext);
        EXH_LAUNCH_CHECK();
        std::swap(pend, pnext);
        std::swap(qcell, qcell_next);
        std::swap(qden, qden_next);
      }
      long long* hr = (long long*)hbuf + 4;  // the ring after the error word
      EXH_CHECK_HIP(hipMemcpyAsync(hr, rt, RT_RING * 8, hipMemcpyDeviceToHost, c->stream));
      EXH_CHECK_HIP(hipMemcpyAsync((int*)hbuf + 1, derr, 4, hipMemcpyDeviceToHost, c->stream));
      EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
      int e = ((int*)hbuf)[1];
      EXH_REQUIRE(e == 0, EXH_ERR_RANGE,
                  "halo growth left the supported range (code %d): radius^2 beyond %d, count "
                  "beyond the barrier table, or a sphere wider than half the box",
                  e, stab - 1);
      for (int r = first; r < launched; r++)
        if (hr[r % RT_RING] > 0) round++;
      ub = hr[launched % RT_RING];
      EXH_REQUIRE(launched < 1000000, EXH_ERR_INTERNAL, "halo rounds did not converge");
    }
  }
  c->stat_rounds = round;

  // 4. catalogue
  long long* kl = (long long*)(cand + o_kl);
  long long* hidx = (long long*)(cand + o_hid);
  hipLaunchKernelGGL(k_kept_to_int, dim3(grid_for(nc, 256 * 8)), dim3(256), 0, c->stream,
                     ga.kept, nc, kl);
  EXH_LAUNCH_CHECK();
  exclusive_scan<long long>(c, kl, hidx, nc);
  EXH_CHECK_HIP(hipMemcpyAsync(hbuf, hidx + nc, 8, hipMemcpyDeviceToHost, c->stream));
  EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
  const long long nh = hbuf[0];
  S.hidx = hidx;
  size_t hs = (size_t)std::max(nh, 1LL);
  size_t hoff = 0;
  auto htake = [&](size_t bytes) {
    size_t r = hoff;
    hoff += (bytes + 255) & ~(size_t)255;
    return r;
  };
  size_t o_hx = htake(hs * 12), o_hc = htake(hs * 4), o_hr = htake(hs * 4),
         o_mh = htake(hs * sizeof(T)), o_u = htake(hs * 8), o_ph = htake(hs * 3 * sizeof(T)),
         o_vh = htake(hs * 3 * sizeof(T));
  // I(M) per sphere size for the masses: every index k_masses can read
  const int nsph = (int)sph.size();
  const int ntab = std::max(1, std::min(std::min(Ncells, nsph) + 1, nsph));
  const size_t o_itab = htake((size_t)ntab * 8), o_ibad = htake((size_t)ntab * 4);
  char* hal = (char*)ctx_buf(c, BUF_HALO, hoff);
  S.hx = (int*)(hal + o_hx);
  S.hcount = (int*)(hal + o_hc);
  S.hr2 = (int*)(hal + o_hr);
  S.Mh = hal + o_mh;
  S.posh = hal + o_ph;
  S.velh = hal + o_vh;
  hipLaunchKernelGGL(k_gather_halos<T>, dim3(grid_for(nc, 256)), dim3(256), 0, c->stream,
                     ga.kept, hidx, nc, ccell, ga.cnt, ga.r2, n1, n2, S.hx, S.hcount, S.hr2);
  EXH_LAUNCH_CHECK();
  // 5. masses
  if (nh > 0) {
    TimedScope ts(c, "halo_masses");
    uint32_t st[3];
    taus_seed_state(o->seed, st);
    double* u = (double*)(hal + o_u);
    hipLaunchKernelGGL(k_taus, dim3(grid_for((nh + 63) / 64, 256)), dim3(256), 0, c->stream,
                       (const unsigned*)(tabs + o_mats), st[0], st[1], st[2], nh, u);
    EXH_LAUNCH_CHECK();
    double* itab = (double*)(hal + o_itab);
    int* ibad = (int*)(hal + o_ibad);
    hipLaunchKernelGGL(k_mass_table<T>, dim3(grid_for(ntab, 256)), dim3(256), 0, c->stream,
                       (const int*)(tabs + o_sph), ntab, (T)b->Mcell,
                       (const double*)(tabs + o_mk), (const double*)(tabs + o_int),
                       (const double*)(tabs + o_mi), Nr, itab, ibad);
    hipLaunchKernelGGL(k_masses<T>, dim3(grid_for(nh, 256)), dim3(256), 0, c->stream, S.hcount, nh,
                       (const int*)(tabs + o_sph), nsph, Ncells, itab, ibad,
                       (const double*)(tabs + o_mk), (const double*)(tabs + o_int),
                       (const double*)(tabs + o_minv), Nr, u, (T*)S.Mh, derr);
    EXH_LAUNCH_CHECK();
    EXH_CHECK_HIP(hipMemcpyAsync(hbuf, derr, 4, hipMemcpyDeviceToHost, c->stream));
    EXH_CHECK_HIP(hipStreamSynchronize(c->stream));
    int e = ((int*)hbuf)[0];
    EXH_REQUIRE(e == 0, EXH_ERR_RANGE,
                "halo mass draw left the tabulated mass function (code %d): a halo's next "
                "sphere exceeds M_max = 1e16 Msun/h",
                e);
  }
  c->nh = nh;
  c->last_prec = b->precision;
  return nh;
}

template <typename T>
void run_halo_posh(exh_ctx* c, const exh_box* b, const T* S, const T* V, int lagrangian) {
  HaloState& st = c->halo;
  if (c->nh <= 0) return;
  TimedScope ts(c, "halo_posh");
  hipLaunchKernelGGL(k_halo_posh<T>, dim3(grid_for(c->nh, 256)), dim3(256), 0, c->stream, st.hx,
                     st.hcount, st.hr2, c->nh, st.sh, b->nd[0], b->nd[1], b->nd[2], S, V,
                     (T)b->Lc, (T)b->L[0], (T)b->L[1], (T)b->L[2], lagrangian, (T*)st.posh,
                     V ? (T*)st.velh : nullptr);
  EXH_LAUNCH_CHECK();
}

void run_halo_flag(exh_ctx* c, const exh_box* b, long long* flag_dev) {
  HaloState& st = c->halo;
  hipLaunchKernelGGL(k_flag_out, dim3(grid_for(b->ng, 256 * 8)), dim3(256), 0, c->stream, st.flag,
                     b->ng, st.hidx, flag_dev);
  EXH_LAUNCH_CHECK();
}

template long long run_find_halos<float>(exh_ctx*, const exh_box*, const float*,
                                        const exh_halo_opts*, const std::function<void(int)>&);
template long long run_find_halos<double>(exh_ctx*, const exh_box*, const double*,
                                         const exh_halo_opts*, const std::function<void(int)>&);
template void run_halo_posh<float>(exh_ctx*, const exh_box*, const float*, const float*, int);
template void run_halo_posh<double>(exh_ctx*, const exh_box*, const double*, const double*, int);

}  // namespace exh
