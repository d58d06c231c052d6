// Adaptive quadtree of the reference (flexmm/tree.py), built on the device.
//
// Every point gets a 31-level Morton key from the reference's own FP64
// midpoint recurrence (tree.py:169-176, 301-304: xmid = xmin + w/2, children
// start at xmid - w/2 or xmid, class 2*(x > xmid) + (y > ymid), tree.py:32-34),
// so the classification of every point at every level is the reference's
// bit for bit.  After a segmented MSD radix sort by key, the tree is a set of
// sorted code lists:
//   S_d   cells of depth d that are split (refinement: > ncut points,
//         tree.py:204; plus the minimal corner 2:1 balance, tree.py:529-541);
//   X_d   real leaves with a split same-level neighbour (tree.py:542-553,
//         385-397), which get four fake children;
//   P_d   = S_d u X_d, the nodes of depth d that have children.
// Level d+1 holds the children of P_d in Morton order.  Colleagues, parents,
// point ranges and M2L partners are binary searches in these sorted lists.
#include "tree.h"
#include "point_key.h"

#include <cstring>

namespace fmm {

namespace {

constexpr int BS = 256;

__device__ __forceinline__ int clz64(uint64_t v) { return v ? __clzll((long long)v) : 64; }

__device__ __forceinline__ int64_t bsearch_u64(const uint64_t* __restrict__ a, int64_t n,
                                               uint64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    uint64_t am = a[mid];
    if (am < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < n && a[lo] == v) ? lo : -1;
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* __restrict__ a, int64_t n,
                                                   uint64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

template <class T>
__device__ __forceinline__ int find_level(const T* __restrict__ off, int L, int64_t g) {
  int d = 0;
  while (d + 1 < L && g >= (int64_t)off[d + 1]) ++d;
  return d;
}

// --- points ----------------------------------------------------------------
// Morton keys of the points and an interleaved copy of their coordinates
// (the gather into tree order then reads one 16-byte record per point).
// One block of 1024 threads per chunk of kSortChunk points (4 per thread,
// strided by the block size; every load is issued before the first key is
// formed).  The key
// comes from the scaled-coordinate form where it provably equals the
// reference's recurrence (point_key.h) and from the recurrence itself for the
// few points within rounding reach of a midpoint.  With fr (more points than
// the sort's threshold) the block also counts its keys' top 8 bits for the
// sort's first round, and keys / vals are the sort's spare buffers.
constexpr int KEY_BS = 1024;
constexpr int kKeyPer = kSortChunk / KEY_BS;
static_assert(kKeyLevels == kMaxDepth, "keys cover every depth");
__global__ __launch_bounds__(KEY_BS) void k_point_keys(const double* __restrict__ x,
                                                   const double* __restrict__ y, int n,
                                                   double xmin0, double ymin0, double w0,
                                                   KeyScale ks, uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals,
                                                   double2* __restrict__ xy,
                                                   int* __restrict__ bad_blk,
                                                   int* __restrict__ cnt, int ncnt,
                                                   SortFirstRound fr, int with_hist) {
  static_assert(kKeyPer * KEY_BS == kSortChunk, "one chunk per block");
  __shared__ int hist[256];
  if (with_hist && threadIdx.x < 256) hist[threadIdx.x] = 0;
  const int i0 = blockIdx.x * kSortChunk + threadIdx.x;
  double px[kKeyPer], py[kKeyPer];
#pragma unroll
  for (int u = 0; u < kKeyPer; ++u) {
    const int i = i0 + u * KEY_BS;
    px[u] = i < n ? x[i] : 0.0;
    py[u] = i < n ? y[i] : 0.0;
  }
  if (with_hist) __syncthreads();
  bool bad = false;
#pragma unroll
  for (int u = 0; u < kKeyPer; ++u) {
    const int i = i0 + u * KEY_BS;
    if (i < n) {
      xy[i] = make_double2(px[u], py[u]);
      bad |= !(__builtin_isfinite(px[u]) && __builtin_isfinite(py[u]));
      bool ok;
      uint64_t key = point_key_fast(px[u], py[u], xmin0, ymin0, ks, &ok);
      if (!ok) key = point_key_recurrence(px[u], py[u], xmin0, ymin0, w0);
      keys[i] = key;
      vals[i] = (uint32_t)i;
      if (with_hist) atomicAdd(&hist[(int)(key >> 56)], 1);
    }
  }
  // every block writes its own flag (no zeroed counter to race with), and
  // block 0 zeroes the build's counters, which only later kernels touch
  const int any_bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) bad_blk[blockIdx.x] = any_bad;
  if (blockIdx.x == 0)
    for (int j = threadIdx.x; j < ncnt; j += KEY_BS) cnt[j] = 0;
  if (with_hist) sort_first_round_block(fr, blockIdx.x, hist);
}

// --- refinement: split cells from the sorted keys ----------------------------
// Position i starts the cells of depths s(i)+1.. (s = common depth with key
// i-1); such a cell has more than ncut points iff key i+ncut is still inside
// it, i.e. depth <= c(i) (common depth with key i+ncut).
// The same kernel counts the split cells per depth (position i starts one at
// each depth d with s < d <= c) and the split positions: each wave writes its
// ballot counts to its own LDS row, the block adds the rows and adds them to
// row b mod kSplitRows of `rows` ([kSplitRows][kMaxLevels + 1], zeroed; the
// host adds the rows): a thousand blocks adding to one set of counters
// serialised on those addresses (10 us of a 15 us kernel at 1M points).
// (Having the last block add the rows costs every block an agent-scope fence,
// which writes back its L2: 8.4 -> 32 us.)  A complete tree needs nothing
// more from the split positions.
constexpr int kSplitRows = 64;
// It also gathers the points into tree order (sx, sy from the interleaved
// copy xy through perm: one 16-byte gather per point), its loads in flight
// with the keys'.
__global__ __launch_bounds__(1024) void k_split_pos(const uint64_t* __restrict__ keys, int n,
                                                    int ncut, int* __restrict__ flag,
                                                    int2* __restrict__ sc, int* __restrict__ rows,
                                                    const uint32_t* __restrict__ perm,
                                                    const double2* __restrict__ xy,
                                                    double* __restrict__ sx,
                                                    double* __restrict__ sy) {
  __shared__ int lc[1024 / 64][kMaxLevels + 1];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  int s = 0, c = -1;
  if (i < n) {
    // the three keys are loaded together (clamped indices): behind their
    // conditions each load waited for the one before
    const int ip = i > 0 ? i - 1 : 0;
    const int ic = (int)min((int64_t)i + ncut, (int64_t)n - 1);
    const double2 pt = xy[perm[i]];
    const uint64_t k = keys[i], kp = keys[ip], kc = keys[ic];
    sx[i] = pt.x;
    sy[i] = pt.y;
    s = (i == 0) ? -1 : min(clz64(k ^ kp) >> 1, kMaxDepth);
    c = ((int64_t)i + ncut < n) ? min(clz64(k ^ kc) >> 1, kMaxDepth - 1) : -1;
    flag[i] = s < c;
    sc[i] = make_int2(s, c);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long fm = __ballot(s < c);
  if (fm) {   // wave-uniform: most waves hold no split position
    for (int d = 0; d < kMaxLevels; ++d) {
      const int cnt = (int)__popcll(__ballot(s < d && d <= c));
      if (lane == 0) lc[wv][d] = cnt;
    }
    if (lane == 0) lc[wv][kMaxLevels] = (int)__popcll(fm);
  } else if (lane <= kMaxLevels) {
    lc[wv][lane] = 0;
  }
  __syncthreads();
  if (threadIdx.x <= kMaxLevels) {
    int t = 0;
#pragma unroll
    for (int w = 0; w < 1024 / 64; ++w) t += lc[w][threadIdx.x];
    if (t != 0)
      atomicAdd(&rows[(blockIdx.x % kSplitRows) * (kMaxLevels + 1) + threadIdx.x], t);
  }
}

// The host's share of the split counts: per depth d <= kMaxLevels the sum of
// the kSplitRows rows, and the non-finite flag, into words 0 .. kMaxLevels + 1
// of the host mailbox.  With gate: *gate = whether they describe the complete
// tree of depth D (every cell of depths < D split, none deeper, all points
// finite), the test the host makes once it has the words.
__global__ __launch_bounds__(64) void k_split_counts(const int* __restrict__ rows,
                                                     const int* __restrict__ bad_blk, int nblk,
                                                     uint64_t* __restrict__ mbox, uint32_t seq,
                                                     int D, int* __restrict__ gate) {
  const int d = threadIdx.x;
  int sum = 0;
  if (d <= kMaxLevels)
    for (int r = 0; r < kSplitRows; ++r) sum += rows[r * (kMaxLevels + 1) + d];
  // the key kernel's per-block non-finite flags, or-ed over the wave
  int b = 0;
  for (int j = d; j < nblk; j += 64) b |= bad_blk[j];
  const int any_bad = __ballot(b != 0) != 0ull;
  if (d == kMaxLevels + 1) sum = any_bad;
  if (d <= kMaxLevels + 1) mailbox_put(mbox, d, seq, sum);
  if (gate) {
    bool ok = true;
    if (d < D) ok = sum == (1 << (2 * d));
    else if (d < kMaxLevels || d == kMaxLevels + 1) ok = sum == 0;
    const bool all = __ballot(!ok) == 0ull;
    if (d == 0) *gate = all ? 1 : 0;
  }
}

__global__ __launch_bounds__(BS) void k_compact_pos(const int* __restrict__ flag,
                                                    const int* __restrict__ rank,
                                                    const int2* __restrict__ sc, int n,
                                                    int* __restrict__ ppos, int2* __restrict__ psc) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n || !flag[i]) return;
  int r = rank[i];
  ppos[r] = i;
  psc[r] = sc[i];
}

__global__ __launch_bounds__(1024) void k_depth_lists(const int* __restrict__ ppos,
                                                     const int2* __restrict__ psc, int np,
                                                     const uint64_t* __restrict__ keys,
                                                     const int64_t* __restrict__ soff,
                                                     uint64_t* __restrict__ sref) {
  __shared__ int lds[1024 / 64];
  __shared__ int tot;
  const int d = blockIdx.x;
  uint64_t* out = sref + soff[d];
  int carry = 0;
  for (int base = 0; base < np; base += 1024) {
    int j = base + threadIdx.x;
    int pred = 0;
    uint64_t code = 0;
    if (j < np) {
      int2 v = psc[j];
      pred = (v.x < d) && (d <= v.y);
      if (pred && d > 0) code = keys[ppos[j]] >> (64 - 2 * d);
    }
    int ex = block_exclusive_scan<1024>(pred, lds, &tot);
    if (pred) out[carry + ex] = code;
    __syncthreads();
    carry += tot;
    __syncthreads();
  }
}

// --- 2:1 balance ---------------------------------------------------------------
// A leaf at depth d is split iff a same-level neighbour has a grandchild; the
// minimal balanced refinement is S_d = Sref_d u N(parents(S_{d+1})), with
// N the 3x3 neighbourhood.  This kernel counts neighbours missing from Sref,
// which is zero exactly when the refinement tree is already balanced.
__global__ __launch_bounds__(BS) void k_balance_check(const uint64_t* __restrict__ sref,
                                                      const int64_t* __restrict__ soff, int L,
                                                      int* __restrict__ rcount) {
  int64_t e = soff[2] + (int64_t)blockIdx.x * BS + threadIdx.x;
  if (e >= soff[L - 1]) return;
  int D = 2;
  while (e >= soff[D + 1]) ++D;
  uint64_t s = sref[e];
  if (e > soff[D] && (sref[e - 1] >> 2) == (s >> 2)) return;
  const int d = D - 1;
  uint64_t pc = s >> 2;
  uint32_t ix, iy;
  demorton(pc, &ix, &iy);
  const int64_t lim = (int64_t)1 << d;
  const uint64_t* lst = sref + soff[d];
  const int64_t nl = soff[d + 1] - soff[d];
  int miss = 0;
  for (int k = 0; k < 9; ++k) {
    int64_t nx = (int64_t)ix + k / 3 - 1, ny = (int64_t)iy + k % 3 - 1;
    if (nx < 0 || ny < 0 || nx >= lim || ny >= lim) continue;
    if (bsearch_u64(lst, nl, morton((uint32_t)nx, (uint32_t)ny)) < 0) ++miss;
  }
  if (miss) atomicAdd(rcount, miss);
}

__global__ __launch_bounds__(BS) void k_bal_candidates(const uint64_t* __restrict__ sref_d,
                                                       int64_t nref,
                                                       const uint64_t* __restrict__ s_next,
                                                       int64_t nnext, int d,
                                                       uint64_t* __restrict__ cand,
                                                       uint32_t* __restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int sh = 64 - 2 * d;
  if (i < nref) {
    cand[i] = sref_d[i] << sh;
    vals[i] = 0;
    return;
  }
  int64_t j = i - nref;
  if (j >= 9 * nnext) return;
  int64_t e = j / 9;
  int k = (int)(j % 9);
  uint64_t s = s_next[e];
  uint64_t out = ~0ull;
  if (e == 0 || (s_next[e - 1] >> 2) != (s >> 2)) {
    uint32_t ix, iy;
    demorton(s >> 2, &ix, &iy);
    int64_t lim = (int64_t)1 << d;
    int64_t nx = (int64_t)ix + k / 3 - 1, ny = (int64_t)iy + k % 3 - 1;
    if (nx >= 0 && ny >= 0 && nx < lim && ny < lim) out = morton((uint32_t)nx, (uint32_t)ny) << sh;
  }
  cand[i] = out;
  vals[i] = 0;
}

__global__ __launch_bounds__(BS) void k_unique_flags(const uint64_t* __restrict__ cand, int64_t n,
                                                     int* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  uint64_t c = cand[i];
  flag[i] = (c != ~0ull) && (i == 0 || cand[i - 1] != c);
}

__global__ __launch_bounds__(BS) void k_unique_scatter(const uint64_t* __restrict__ cand,
                                                       const int* __restrict__ flag,
                                                       const int* __restrict__ rank, int64_t n,
                                                       int d, uint64_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n || !flag[i]) return;
  out[rank[i]] = cand[i] >> (64 - 2 * d);
}

// Balance on bitmaps: bit c of depth d's bitmap is cell c (the depth-d Morton
// code); each depth's words start at a multiple of BS, so a block of BS words
// lies in one depth.  Marking is an idempotent OR, so the union needs neither
// sort nor unique, and the bitmap order is the code order.
constexpr int kBmDepth = 13;   // 4^13 bits = 8 MB at the deepest bitmap depth
struct BmOffsets {
  int64_t wo[kMaxLevels + 1];  // first word of depth d (depths 1 .. dmax)
};
__device__ __forceinline__ int bm_depth(const BmOffsets& o, int dmax, int64_t w) {
  int d = 1;
  while (d < dmax && w >= o.wo[d + 1]) ++d;
  return d;
}
__device__ __forceinline__ void bm_set(uint32_t* bm, uint64_t c) {
  atomicOr(bm + (c >> 5), 1u << (c & 31));
}
// the 3x3 neighbourhood (clipped to the domain) of parent cell pc at depth d;
// bits that fall in one word go out in one atomic (a neighbourhood spans at
// most four words)
__device__ __forceinline__ void bm_mark_nbrs(uint32_t* bm, uint64_t pc, int d) {
  uint32_t ix, iy;
  demorton(pc, &ix, &iy);
  const int64_t lim = (int64_t)1 << d;
  uint64_t wd[9];
  uint32_t bit[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int64_t nx = (int64_t)ix + k / 3 - 1, ny = (int64_t)iy + k % 3 - 1;
    const bool in = nx >= 0 && ny >= 0 && nx < lim && ny < lim;
    const uint64_t c = in ? morton((uint32_t)nx, (uint32_t)ny) : 0;
    wd[k] = c >> 5;
    bit[k] = in ? 1u << (c & 31) : 0u;
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    bool first = true;
#pragma unroll
    for (int j = 0; j < k; ++j) first = first && !(bit[j] && wd[j] == wd[k]);
    if (!first || !bit[k]) continue;
    uint32_t m = bit[k];
#pragma unroll
    for (int j = k + 1; j < 9; ++j) m |= wd[j] == wd[k] ? bit[j] : 0u;
    atomicOr(bm + wd[k], m);
  }
}
// Sref_d for depths 1 .. dmax into their bitmaps
__global__ __launch_bounds__(BS) void k_bm_mark(const uint64_t* __restrict__ sref,
                                                const int64_t* __restrict__ soff, int dmax,
                                                BmOffsets o, uint32_t* __restrict__ bm) {
  const int64_t e = soff[1] + (int64_t)blockIdx.x * BS + threadIdx.x;
  if (e >= soff[dmax + 1]) return;
  int d = 1;
  while (e >= soff[d + 1]) ++d;
  bm_set(bm + o.wo[d], sref[e]);
}
// N(parents) of the depth-(d+1) cells in bitmap bm1 into depth d's bitmap bm0,
// one thread per depth-d parent cell pc: its 4 children are bits
// 4(pc mod 8) .. +3 of word pc / 8
__global__ __launch_bounds__(BS) void k_bm_spread(const uint32_t* __restrict__ bm1, int d,
                                                  uint32_t* __restrict__ bm0) {
  const int64_t pc = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (pc >= (int64_t)1 << (2 * d)) return;
  if (!((bm1[pc >> 3] >> (4 * (pc & 7))) & 0xFu)) return;
  bm_mark_nbrs(bm0, (uint64_t)pc, d);
}
// the same with one thread per word of bm1 (8 parents), for the deep and
// sparse bitmaps where a thread per cell would mostly find nothing
__global__ __launch_bounds__(BS) void k_bm_spread_w(const uint32_t* __restrict__ bm1, int d,
                                                    uint32_t* __restrict__ bm0) {
  const int64_t w = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (w >= ((int64_t)1 << (2 * d)) / 8) return;
  const uint32_t v = bm1[w];
  if (!v) return;
  for (int q = 0; q < 8; ++q)
    if ((v >> (4 * q)) & 0xFu) bm_mark_nbrs(bm0, (uint64_t)w * 8 + q, d);
}
// the same from a sorted list of depth-(d+1) codes
__global__ __launch_bounds__(BS) void k_bm_spread_list(const uint64_t* __restrict__ lst,
                                                       int64_t n, int d,
                                                       uint32_t* __restrict__ bm0) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = lst[i];
  if (i > 0 && (lst[i - 1] >> 2) == (c >> 2)) return;
  bm_mark_nbrs(bm0, c >> 2, d);
}
// popcount per word (for the scan) and per depth
__global__ __launch_bounds__(BS) void k_bm_count(const uint32_t* __restrict__ bm, BmOffsets o,
                                                 int dmax, int* __restrict__ pc,
                                                 int* __restrict__ dcnt) {
  __shared__ int red[BS / 64];
  const int64_t w = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int c = __popc(bm[w]);
  pc[w] = c;
  int v = c;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < BS / 64; ++i) t += red[i];
    if (t) atomicAdd(dcnt + bm_depth(o, dmax, (int64_t)blockIdx.x * BS), t);
  }
}
// set bits in word order -> codes, at out[pos[w] ...]
__global__ __launch_bounds__(BS) void k_bm_scatter(const uint32_t* __restrict__ bm, BmOffsets o,
                                                   int dmax, const int* __restrict__ pos,
                                                   uint64_t* __restrict__ out) {
  const int64_t w = (int64_t)blockIdx.x * BS + threadIdx.x;
  uint32_t v = bm[w];
  if (!v) return;
  const int d = bm_depth(o, dmax, w);
  const uint64_t base = (uint64_t)(w - o.wo[d]) * 32;
  int p = pos[w];
  while (v) {
    const int b = __builtin_ctz(v);
    out[p++] = base + (uint64_t)b;
    v &= v - 1;
  }
}

// --- nodes with children (split or X-list) -----------------------------------
// Real nodes of level d are the children of S_{d-1} (the root for d = 0).
__device__ __forceinline__ void real_flags_one(const uint64_t* __restrict__ s_all,
                                               const int64_t* __restrict__ soff,
                                               const int64_t* __restrict__ roff, int L, int64_t r,
                                               uint64_t* __restrict__ rcode, int* __restrict__ hasc,
                                               int* __restrict__ kind, int* lc) {
  int d = 0;
  while (r >= roff[d + 1]) ++d;
  int64_t k = r - roff[d];
  uint64_t code = d == 0 ? 0 : ((s_all[soff[d - 1] + k / 4] << 2) | (uint64_t)(k & 3));
  int inS = 0, isX = 0;
  if (d <= L - 2) {
    const uint64_t* lst = s_all + soff[d];
    const int64_t nl = soff[d + 1] - soff[d];
    inS = bsearch_u64(lst, nl, code) >= 0;
    if (!inS && d >= 1) {
      uint32_t ix, iy;
      demorton(code, &ix, &iy);
      const int64_t lim = (int64_t)1 << d;
      for (int q = 0; q < 9 && !isX; ++q) {
        if (q == 4) continue;
        int64_t nx = (int64_t)ix + q / 3 - 1, ny = (int64_t)iy + q % 3 - 1;
        if (nx < 0 || ny < 0 || nx >= lim || ny >= lim) continue;
        isX = bsearch_u64(lst, nl, morton((uint32_t)nx, (uint32_t)ny)) >= 0;
      }
    }
  }
  rcode[r] = code;
  hasc[r] = inS || isX;
  kind[r] = isX;
  if (inS || isX) atomicAdd(&lc[d], 1);
}

__global__ __launch_bounds__(BS) void k_real_flags(const uint64_t* __restrict__ s_all,
                                                   const int64_t* __restrict__ soff,
                                                   const int64_t* __restrict__ roff, int L,
                                                   uint64_t* __restrict__ rcode,
                                                   int* __restrict__ hasc, int* __restrict__ kind,
                                                   int* __restrict__ pcount) {
  __shared__ int lc[kMaxLevels];
  if (threadIdx.x < kMaxLevels) lc[threadIdx.x] = 0;
  __syncthreads();
  int64_t r = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (r < roff[L]) real_flags_one(s_all, soff, roff, L, r, rcode, hasc, kind, lc);
  __syncthreads();
  if (threadIdx.x < kMaxLevels && lc[threadIdx.x]) atomicAdd(&pcount[threadIdx.x], lc[threadIdx.x]);
}

__global__ __launch_bounds__(BS) void k_p_scatter(const uint64_t* __restrict__ rcode,
                                                  const int* __restrict__ hasc,
                                                  const int* __restrict__ kind,
                                                  const int* __restrict__ rank, int64_t nr,
                                                  uint64_t* __restrict__ p_code,
                                                  int* __restrict__ p_kind) {
  int64_t r = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (r >= nr || !hasc[r]) return;
  int q = rank[r];
  p_code[q] = rcode[r];
  p_kind[q] = kind[r];
}

// --- final node arrays ---------------------------------------------------------
// Node index of the level-d cell with Morton code c, or -1: direct on a full
// level, a table lookup on the other levels up to depth kIdxDepth (one int per
// cell of the level, -1 where no node), else a binary search in the level's
// sorted codes.
constexpr int kIdxDepth = 11;   // tables of at most 4^12/3 ints (22 MB)
struct NodeIdx {
  const int* tab;
  int64_t off[kMaxLevels];      // level d's table in tab, or -1
};
__device__ __forceinline__ int node_of(const uint64_t* __restrict__ code,
                                       const int* __restrict__ levoff, const NodeIdx& ni, int d,
                                       uint64_t c) {
  const int64_t nl = levoff[d + 1] - levoff[d];
  if (d < 16 && nl == ((int64_t)1 << (2 * d))) return levoff[d] + (int)c;
  if (ni.off[d] >= 0) return ni.tab[ni.off[d] + (int64_t)c];
  const int64_t pos = bsearch_u64(code + levoff[d], nl, c);
  return pos >= 0 ? levoff[d] + (int)pos : -1;
}
__global__ __launch_bounds__(BS) void k_node_index(const uint64_t* __restrict__ code,
                                                   const int* __restrict__ nlev, int nn,
                                                   NodeIdx ni, int* __restrict__ tab) {
  const int g = blockIdx.x * BS + threadIdx.x;
  if (g >= nn) return;
  const int64_t o = ni.off[nlev[g]];
  if (o >= 0) tab[o + (int64_t)code[g]] = g;
}

// Source ranges of all nodes from the sorted keys: point i starts its cells
// of depths s(i)+1.. and ends those of depths s(i+1)+1.. (s = common depth
// with the previous key).  Cells without points keep -1 (fixed below).  A
// cell exists only below an existing parent, so each walk stops at the first
// depth without one.
__global__ __launch_bounds__(BS) void k_cell_ranges(const uint64_t* __restrict__ keys,
                                                    const int2* __restrict__ sc, int n,
                                                    const uint64_t* __restrict__ code,
                                                    const int* __restrict__ levoff, int L,
                                                    NodeIdx ni, int* __restrict__ bot,
                                                    int* __restrict__ top) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const int s0 = sc[i].x;
  const int s1 = i + 1 < n ? sc[i + 1].x : -1;
  if (s0 >= L - 1 && s1 >= L - 1) return;
  const uint64_t k = keys[i];
  for (int d = max(1, s0 + 1); d <= L - 1; ++d) {
    const int g = node_of(code, levoff, ni, d, k >> (64 - 2 * d));
    if (g < 0) break;
    bot[g] = i;
  }
  for (int d = max(1, s1 + 1); d <= L - 1; ++d) {
    const int g = node_of(code, levoff, ni, d, k >> (64 - 2 * d));
    if (g < 0) break;
    top[g] = i + 1;
  }
}

__global__ __launch_bounds__(BS) void k_cell_ranges_init(int nn, int npts, int* __restrict__ bot,
                                                         int* __restrict__ top) {
  const int g = blockIdx.x * BS + threadIdx.x;
  if (g >= nn) return;
  bot[g] = g == 0 ? 0 : -1;
  top[g] = g == 0 ? npts : -1;
}

__global__ __launch_bounds__(BS) void k_nodes_a(const uint64_t* __restrict__ p_code,
                                                const int* __restrict__ p_kind,
                                                const int64_t* __restrict__ poff,
                                                const int* __restrict__ levoff, int L,
                                                uint64_t* __restrict__ code,
                                                int* __restrict__ nlev, int* __restrict__ flags) {
  int g = blockIdx.x * BS + threadIdx.x;
  if (g >= levoff[L]) return;
  int d = find_level(levoff, L, g);
  int k = g - levoff[d];
  uint64_t c = 0;
  int fake = 0;
  if (d > 0) {
    int64_t pj = poff[d - 1] + k / 4;
    c = (p_code[pj] << 2) | (uint64_t)(k & 3);
    fake = p_kind[pj] == 1;
  }
  code[g] = c;
  nlev[g] = d;
  flags[g] = fake ? NF_FAKE : 0;
}

__global__ __launch_bounds__(64) void k_nodes_b(
    const uint64_t* __restrict__ p_code, const int* __restrict__ p_kind,
    const int64_t* __restrict__ poff, const int* __restrict__ levoff, int L,
    const uint64_t* __restrict__ keys, int npts, double xmin0, double ymin0, double w0,
    NodeIdx ni, const uint64_t* __restrict__ code, int* __restrict__ flags, int* __restrict__ bot,
    int* __restrict__ top, int* __restrict__ parent, int* __restrict__ child,
    int* __restrict__ coll, double* __restrict__ xmid, double* __restrict__ ymid) {
  int g = blockIdx.x * 64 + threadIdx.x;
  if (g >= levoff[L]) return;
  const int d = find_level(levoff, L, g);
  const uint64_t c = code[g];
  const int fake = flags[g] & NF_FAKE;
  // children
  int ch = -1, xl = 0;
  if (!fake && d <= L - 2) {
    int64_t pos = bsearch_u64(p_code + poff[d], poff[d + 1] - poff[d], c);
    if (pos >= 0) {
      ch = levoff[d + 1] + 4 * (int)pos;
      xl = p_kind[poff[d] + pos] == 1;
    }
  }
  const int leaf = !fake && (ch < 0 || xl);
  // point range of the cell (k_cell_ranges); a cell without points sits at
  // the position where its keys would be
  int b = bot[g], t = top[g];
  if (b < 0) {
    b = (int)lower_bound_u64(keys, npts, c << (64 - 2 * d));
    t = b;
  }
  const int has_src = t > b;
  int f = fake ? NF_FAKE : 0;
  if (leaf) f |= NF_LEAF;
  if (xl) f |= NF_XLIST;
  if (ch >= 0) f |= NF_CHILDREN;
  if (has_src) f |= NF_SOURCE;
  if (((leaf && !xl) || fake) && has_src) f |= NF_UP;   // tree.py:557-560
  if ((!leaf || xl) && !fake) f |= NF_DOWN;             // tree.py:562 (fake nodes have leaf=False)
  flags[g] = f;
  bot[g] = b;
  top[g] = t;
  child[g] = ch;
  // centre by the reference recurrence
  double xm0 = xmin0, ym0 = ymin0, w = w0;
  for (int l = 1; l <= d; ++l) {
    const double h = 0.5 * w;
    const double xm = xm0 + h, ym = ym0 + h;
    const int dig = (int)((c >> (2 * (d - l))) & 3);
    xm0 = (dig & 2) ? xm : xm - h;
    ym0 = (dig & 1) ? ym : ym - h;
    w = h;
  }
  xmid[g] = xm0 + 0.5 * w;
  ymid[g] = ym0 + 0.5 * w;
  // colleagues and parent
  uint32_t ix, iy;
  demorton(c, &ix, &iy);
  const int64_t lim = (int64_t)1 << d;
  for (int q = 0; q < 9; ++q) {
    int64_t nx = (int64_t)ix + q / 3 - 1, ny = (int64_t)iy + q % 3 - 1;
    int r = -1;
    if (nx >= 0 && ny >= 0 && nx < lim && ny < lim) {
      r = (q == 4) ? g : node_of(code, levoff, ni, d, morton((uint32_t)nx, (uint32_t)ny));
    }
    coll[(int64_t)g * 9 + q] = r;
  }
  const int par = d > 0 ? node_of(code, levoff, ni, d - 1, c >> 2) : -1;
  parent[g] = par;
}

// Source ranges of every node's nine colleagues, for the evaluation kernels
// (one load per colleague instead of a colleague index and then its range).
__global__ __launch_bounds__(BS) void k_coll_ranges(int nslots, const int* __restrict__ coll,
                                                   const int* __restrict__ bot,
                                                   const int* __restrict__ top,
                                                   int2* __restrict__ crange) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= nslots) return;
  const int c = coll[i];
  crange[i] = c >= 0 ? make_int2(bot[c], top[c]) : make_int2(0, 0);
}

// --- density rows: sibling groups (levels >= 2) holding a source ------------------
__global__ __launch_bounds__(BS) void k_group_flags(const int* __restrict__ flags, int g0, int ngroups,
                                                    const int* __restrict__ levoff, int L,
                                                    int* __restrict__ gflag,
                                                    int* __restrict__ lcount) {
  __shared__ int lc[kMaxLevels];
  if (threadIdx.x < kMaxLevels) lc[threadIdx.x] = 0;
  __syncthreads();
  int q = blockIdx.x * BS + threadIdx.x;
  if (q < ngroups) {
    int g = g0 + 4 * q;
    int any = 0;
    for (int j = 0; j < 4; ++j) any |= flags[g + j] & NF_SOURCE;
    gflag[q] = any ? 1 : 0;
    if (any) atomicAdd(&lc[find_level(levoff, L, g)], 4);
  }
  __syncthreads();
  if (threadIdx.x < kMaxLevels && lc[threadIdx.x]) atomicAdd(&lcount[threadIdx.x], lc[threadIdx.x]);
}

__global__ __launch_bounds__(BS) void k_rows(const int* __restrict__ gflag,
                                             const int* __restrict__ grank, int g0, int nnodes,
                                             int* __restrict__ row, int* __restrict__ rownode) {
  int g = blockIdx.x * BS + threadIdx.x;
  if (g >= nnodes) return;
  int r = -1;
  if (g >= g0) {
    int q = (g - g0) / 4;
    if (gflag[q]) r = 4 * grank[q] + (g - g0) % 4;
  }
  row[g] = r;
  if (r >= 0) rownode[r] = g;
}

// Per multipole row: the row of its node's first child (-1 without children)
// and the node's flags, so that the upward pass reads one record per row.
__global__ __launch_bounds__(BS) void k_row_info(const int* __restrict__ rownode, int nrows,
                                                 const int* __restrict__ flags,
                                                 const int* __restrict__ child,
                                                 const int* __restrict__ row,
                                                 int2* __restrict__ info) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= nrows) return;
  const int g = rownode[r];
  const int f = flags[g];
  info[r] = make_int2((f & NF_CHILDREN) ? row[child[g]] : -1, f);
}

// --- M2L partners: children of the parent's colleagues, outside the 3x3 --------
// (tree.py:399-435).  For child position (cx, cy) the parent's 3x3 covers
// offsets [-2-cx, 3-cx] x [-2-cy, 3-cy].
struct SlotOffsets {
  int ox[kNumM2L], oy[kNumM2L];
};
__constant__ SlotOffsets c_slot_off;

__global__ __launch_bounds__(BS) void k_m2l_lists(const uint64_t* __restrict__ code,
                                                  const int* __restrict__ flags,
                                                  const int* __restrict__ row,
                                                  const int* __restrict__ levoff, int L,
                                                  NodeIdx ni, int* __restrict__ m2l) {
  const int64_t tid = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int g = levoff[2] + (int)(tid / kNumM2L);
  const int s = (int)(tid % kNumM2L);
  if (g >= levoff[L]) return;
  int r = -1;
  if (!(flags[g] & NF_FAKE)) {
    const int d = find_level(levoff, L, g);
    uint32_t ix, iy;
    demorton(code[g], &ix, &iy);
    const int cx = ix & 1, cy = iy & 1;
    const int ox = c_slot_off.ox[s], oy = c_slot_off.oy[s];
    if (ox >= -2 - cx && ox <= 3 - cx && oy >= -2 - cy && oy <= 3 - cy) {
      const int64_t lim = (int64_t)1 << d;
      const int64_t nx = (int64_t)ix + ox, ny = (int64_t)iy + oy;
      if (nx >= 0 && ny >= 0 && nx < lim && ny < lim) {
        const int q = node_of(code, levoff, ni, d, morton((uint32_t)nx, (uint32_t)ny));
        if (q >= 0) r = row[q];
      }
    }
  }
  m2l[(int64_t)g * kNumM2L + s] = r;
}

__global__ __launch_bounds__(BS) void k_flag_bit(const int* __restrict__ flags, int n, int bit,
                                                 int* __restrict__ out) {
  int g = blockIdx.x * BS + threadIdx.x;
  if (g >= n) return;
  out[g] = (flags[g] & bit) ? 1 : 0;
}

// S2M boxes: compute_upwards nodes that own a multipole row (levels >= 2;
// the upward pass stops at level 2, fmm.py:98).
__global__ __launch_bounds__(BS) void k_up_flag(const int* __restrict__ flags,
                                                const int* __restrict__ row, int n,
                                                int* __restrict__ out) {
  int g = blockIdx.x * BS + threadIdx.x;
  if (g >= n) return;
  out[g] = ((flags[g] & NF_UP) && row[g] >= 0) ? 1 : 0;
}

__global__ __launch_bounds__(BS) void k_list_scatter(const int* __restrict__ flag,
                                                     const int* __restrict__ rank, int n,
                                                     int* __restrict__ list,
                                                     int* __restrict__ idx_of) {
  int g = blockIdx.x * BS + threadIdx.x;
  if (g >= n) return;
  if (flag[g]) {
    list[rank[g]] = g;
    if (idx_of) idx_of[g] = rank[g];
  } else if (idx_of) {
    idx_of[g] = -1;
  }
}

__global__ __launch_bounds__(BS) void k_locate(const double* __restrict__ tx,
                                               const double* __restrict__ ty, int64_t nt,
                                               const int* __restrict__ flags,
                                               const int* __restrict__ child,
                                               const double* __restrict__ xmid,
                                               const double* __restrict__ ymid,
                                               int* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= nt) return;
  const double x = tx[i], y = ty[i];
  int g = 0;
  for (int it = 0; it < kMaxLevels && !(flags[g] & NF_LEAF); ++it)
    g = child[g] + 2 * (x > xmid[g]) + (y > ymid[g]);   // tree.py:720-725
  out[i] = g;
}

inline unsigned grid_for(int64_t n) { return (unsigned)(n <= 0 ? 1 : ceil_div(n, BS)); }

// --- complete trees --------------------------------------------------------------
// When the split cells of every depth d < D are all 4^d cells of that depth
// and no cell of depth D splits, the reference's refinement (split iff more
// than ncut points, tree.py:204) has built the complete quadtree of depth D:
// its leaves all sit on level D, so no 2:1 balance split and no X-list node
// exists (both need leaves on different levels), and every sibling group of
// the levels >= 2 holds a source (its parent has more than ncut).  Uniform
// data give such trees.  After the sort and the split-cell counts that every
// build reads back, such a tree takes closed forms of all node arrays in one
// kernel instead of the split lists, their scans and the two further read-backs.
constexpr int kFullMaxDepth = 15;   // depth-D cell index tables of at most 4^15 ints

// first sorted point of each nonempty depth-D cell, -1 for the empty ones:
// the point that starts a cell also marks the empty cells between the
// previous point's cell and its own (and the last point those after it).  In
// a complete tree every depth D-1 cell has points, so such a run is at most
// six cells.
__global__ __launch_bounds__(BS) void k_full_starts(const uint64_t* __restrict__ keys,
                                                    const int2* __restrict__ sc, int n, int D,
                                                    int* __restrict__ start,
                                                    const int* __restrict__ gate) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n || (gate && *gate == 0)) return;
  const int2 si = sc[i];
  // loaded with sc, not behind its test
  const uint64_t k = keys[i], kp = keys[i > 0 ? i - 1 : 0];
  const int64_t cell = (int64_t)(k >> (64 - 2 * D));
  if (si.x < D) {
    for (int64_t e = i > 0 ? (int64_t)(kp >> (64 - 2 * D)) + 1 : 0; e < cell; ++e) start[e] = -1;
    start[cell] = i;
  }
  if (i == n - 1)
    for (int64_t e = cell + 1; e < ((int64_t)1 << (2 * D)); ++e) start[e] = -1;
}

// points of a depth-d cell c: [bot, top) from the depth-D starts (a cell
// without points sits where its keys would be)
__device__ __forceinline__ int full_pos(const int* __restrict__ start,
                                        const uint64_t* __restrict__ keys, int n, int D, int d,
                                        uint64_t c) {
  const uint64_t leaf = c << (2 * (D - d));
  if (leaf >= ((uint64_t)1 << (2 * D))) return n;
  const int v = start[leaf];
  return v >= 0 ? v : (int)lower_bound_u64(keys, n, leaf << (64 - 2 * D));
}

__host__ __device__ __forceinline__ int full_levoff(int d) { return (int)((((int64_t)1 << (2 * d)) - 1) / 3); }

// Every node array of the complete tree of depth D (levels 0 .. D, level d =
// the 4^d cells in Morton order from levoff(d) = (4^d - 1) / 3).  Sixteen
// lanes per node: lane 0 writes the node's own fields, lanes 0..8 its nine
// colleague slots and all sixteen its 40 M2L slots (j, j + 16, j + 32), so
// the per-node arrays are written in contiguous runs.
constexpr int kFullLanes = 16;
__global__ __launch_bounds__(BS) void k_full_nodes(
    int D, int nn, const int* __restrict__ start, const uint64_t* __restrict__ keys, int npts,
    double xmin0, double ymin0, double w0, uint64_t* __restrict__ code, int* __restrict__ nlev,
    int* __restrict__ flags, int* __restrict__ bot, int* __restrict__ top,
    int* __restrict__ parent, int* __restrict__ child, int* __restrict__ coll,
    int2* __restrict__ crange, double* __restrict__ xmid, double* __restrict__ ymid,
    int* __restrict__ row, int* __restrict__ m2l, int* __restrict__ evalid,
    int* __restrict__ evalnode, int* __restrict__ upnodes, int2* __restrict__ rowinfo,
    int* __restrict__ levoff, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t tid = (int64_t)blockIdx.x * BS + threadIdx.x;
  const int g = (int)(tid / kFullLanes), j = (int)(tid % kFullLanes);
  if (tid <= D + 1) levoff[tid] = full_levoff((int)tid);   // levels 0 .. D and the end
  if (g >= nn) return;
  int d = 0;
  while (d < D && g >= full_levoff(d + 1)) ++d;
  const int lo = full_levoff(d);
  const uint64_t c = (uint64_t)(g - lo);
  const bool leaf = d == D;
  const int g0 = full_levoff(2);
  uint32_t ix, iy;
  demorton(c, &ix, &iy);
  const int64_t lim = (int64_t)1 << d;
  if (j == 0) {
    const int b = full_pos(start, keys, npts, D, d, c), t = full_pos(start, keys, npts, D, d, c + 1);
    int f = leaf ? NF_LEAF : (NF_CHILDREN | NF_DOWN);   // tree.py:557-562
    if (t > b) f |= NF_SOURCE | (leaf ? NF_UP : 0);
    code[g] = c;
    nlev[g] = d;
    flags[g] = f;
    bot[g] = b;
    top[g] = t;
    parent[g] = d > 0 ? full_levoff(d - 1) + (int)(c >> 2) : -1;
    const int ch = leaf ? -1 : full_levoff(d + 1) + 4 * (int)c;
    child[g] = ch;
    // centre by the reference recurrence (as k_nodes_b)
    double xm0 = xmin0, ym0 = ymin0, w = w0;
    for (int l = 1; l <= d; ++l) {
      const double h = 0.5 * w;
      const double xm = xm0 + h, ym = ym0 + h;
      const int dig = (int)((c >> (2 * (d - l))) & 3);
      xm0 = (dig & 2) ? xm : xm - h;
      ym0 = (dig & 1) ? ym : ym - h;
      w = h;
    }
    xmid[g] = xm0 + 0.5 * w;
    ymid[g] = ym0 + 0.5 * w;
    // rows: every sibling group of levels >= 2 holds sources, so row = g - g0
    const int r = d >= 2 ? g - g0 : -1;
    row[g] = r;
    if (r >= 0) rowinfo[r] = make_int2(ch >= 0 ? ch - g0 : -1, f);
    evalid[g] = leaf ? (int)c : -1;
    if (leaf) {
      evalnode[c] = g;
      upnodes[c] = g;   // every leaf: an empty one has S2M check values 0 (no sources)
    }
  }
  if (j < 9) {
    const int64_t nx = (int64_t)ix + j / 3 - 1, ny = (int64_t)iy + j % 3 - 1;
    int r = -1;
    int2 rg = make_int2(0, 0);
    if (nx >= 0 && ny >= 0 && nx < lim && ny < lim) {
      const uint64_t cc = morton((uint32_t)nx, (uint32_t)ny);
      r = lo + (int)cc;
      rg = make_int2(full_pos(start, keys, npts, D, d, cc),
                     full_pos(start, keys, npts, D, d, cc + 1));
    }
    coll[(int64_t)g * 9 + j] = r;
    crange[(int64_t)g * 9 + j] = rg;
  }
  if (d >= 2) {
    const int cx = ix & 1, cy = iy & 1;
    for (int s = j; s < kNumM2L; s += kFullLanes) {
      const int ox = c_slot_off.ox[s], oy = c_slot_off.oy[s];
      int rr = -1;
      if (ox >= -2 - cx && ox <= 3 - cx && oy >= -2 - cy && oy <= 3 - cy) {
        const int64_t nx = (int64_t)ix + ox, ny = (int64_t)iy + oy;
        if (nx >= 0 && ny >= 0 && nx < lim && ny < lim)
          rr = lo + (int)morton((uint32_t)nx, (uint32_t)ny) - g0;
      }
      m2l[(int64_t)g * kNumM2L + s] = rr;
    }
  }
}

}  // namespace

// Node arrays of a complete tree of depth D from the sorted points (keys and
// the common depths sc of neighbouring keys); see k_full_nodes.
// The complete tree of depth D, in three parts: its buffers, its node arrays
// (k_full_starts, k_full_nodes; with gate, the kernels do nothing unless
// *gate is set when they run) and the host's level tables.
static void tree_full_buffers(Tree& t, int D) {
  const int NN = full_levoff(D + 1);
  const int64_t n_eval = (int64_t)1 << (2 * D);
  const int64_t n_rows = NN - full_levoff(2);
  t.levoff.ensure((D + 2) * sizeof(int));   // written by k_full_nodes
  t.w_flag.ensure((n_eval + 1) * sizeof(int));
  t.code.ensure(NN * sizeof(uint64_t));
  t.nlevel.ensure(NN * sizeof(int));
  t.bot.ensure(NN * sizeof(int));
  t.top.ensure(NN * sizeof(int));
  t.parent.ensure(NN * sizeof(int));
  t.child.ensure(NN * sizeof(int));
  t.flags.ensure(NN * sizeof(int));
  t.coll.ensure(NN * 9 * sizeof(int));
  t.crange.ensure(((int64_t)NN * 9 + 1) * sizeof(int2));
  t.xmid.ensure(NN * sizeof(double));
  t.ymid.ensure(NN * sizeof(double));
  t.row.ensure(NN * sizeof(int));
  t.m2l.ensure((int64_t)NN * kNumM2L * sizeof(int));
  t.evalid.ensure(NN * sizeof(int));
  t.evalnode.ensure((n_eval + 1) * sizeof(int));
  t.upnodes.ensure((n_eval + 1) * sizeof(int));
  t.rowinfo.ensure((n_rows + 1) * sizeof(int2));
}

static void tree_full_kernels(Tree& t, int D, int npts, hipStream_t s, const int* gate) {
  const int NN = full_levoff(D + 1);
  int* start = t.w_flag.as<int>();
  hipLaunchKernelGGL(k_full_starts, dim3(grid_for(npts)), dim3(BS), 0, s, t.keys.as<uint64_t>(),
                     t.w_a.as<int2>(), npts, D, start, gate);
  FMM_LAUNCH_CHECK();
  // levels 0 and 1 have no M2L lists (their entries are never read), as in
  // the general build
  hipLaunchKernelGGL(k_full_nodes, dim3(grid_for((int64_t)NN * kFullLanes)), dim3(BS), 0, s, D, NN,
                     start,
                     t.keys.as<uint64_t>(), npts, t.xmin, t.ymin, t.w0, t.code.as<uint64_t>(),
                     t.nlevel.as<int>(), t.flags.as<int>(), t.bot.as<int>(), t.top.as<int>(),
                     t.parent.as<int>(), t.child.as<int>(), t.coll.as<int>(), t.crange.as<int2>(),
                     t.xmid.as<double>(), t.ymid.as<double>(), t.row.as<int>(), t.m2l.as<int>(),
                     t.evalid.as<int>(), t.evalnode.as<int>(), t.upnodes.as<int>(),
                     t.rowinfo.as<int2>(), t.levoff.as<int>(), gate);
  FMM_LAUNCH_CHECK();
}

static void tree_full_host(Tree& t, int D) {
  const int L = D + 1;
  t.L = L;
  t.lev.assign(L, TreeLevel());
  for (int d = 0; d < L; ++d) {
    t.lev[d].node_off = full_levoff(d);
    t.lev[d].n_nodes = (int64_t)1 << (2 * d);
    t.lev[d].width = std::ldexp(t.w0, -d);
  }
  const int NN = full_levoff(L);
  t.n_nodes = NN;
  const int g0 = full_levoff(2);
  t.n_rows = NN - g0;
  t.row_off.assign(L + 1, 0);
  for (int d = 0; d <= L; ++d) t.row_off[d] = d <= 2 ? 0 : full_levoff(d) - g0;
  t.n_eval = (int64_t)1 << (2 * D);
  t.n_up = t.n_eval;
  t.nidx_off.assign(kMaxLevels, -1);
  t.complete = true;
  t.full_D = D;
  ++t.build_id;
  t.built = true;
}

// Device counters c[0 .. n) into the host mailbox, for read_counts.
__global__ __launch_bounds__(128) void k_counts_to_mailbox(const int* __restrict__ c, int n,
                                                           uint64_t* __restrict__ mbox,
                                                           uint32_t seq) {
  for (int i = threadIdx.x; i < n; i += 128) mailbox_put(mbox, i, seq, c[i]);
}

// host[0 .. n) = the device counters dev[0 .. n) once the work queued on s
// before this call has run, without draining s (HostMailbox).
static void read_counts(Tree& t, const int* dev, int n, int* host, hipStream_t s) {
  HostMailbox& mb = t.pw.mbox;
  mb.ensure();
  const uint32_t seq = mb.next();
  hipLaunchKernelGGL(k_counts_to_mailbox, dim3(1), dim3(128), 0, s, dev, n, mb.dev, seq);
  FMM_LAUNCH_CHECK();
  int64_t v[HostMailbox::kWords];
  mb.take(n, seq, s, v);
  for (int i = 0; i < n; ++i) host[i] = (int)v[i];
}

// ---------------------------------------------------------------------------
void tree_build(Tree& t, const double* x, const double* y, int64_t n, double xmin, double ymin,
                double w0, int ncut, hipStream_t s) {
  FMM_REQUIRE(n >= 1 && n < ((int64_t)1 << 31) - 1, FMM_E_LIMIT,
              "tree: number of sources %lld outside [1, 2^31-2]", (long long)n);
  FMM_REQUIRE(ncut >= 1, FMM_E_ARG, "tree: Ncutoff must be >= 1 (got %d)", ncut);
  FMM_REQUIRE(w0 > 0.0 && std::isfinite(w0), FMM_E_ARG, "tree: root width must be positive");
  t.built = false;
  t.complete = false;
  t.n = n;
  t.ncut = ncut;
  t.xmin = xmin;
  t.ymin = ymin;
  t.w0 = w0;
  const int nn = (int)n;

  // -- sorted points ----------------------------------------------------------
  t.keys.ensure(n * sizeof(uint64_t));
  t.perm.ensure(n * sizeof(uint32_t));
  t.sx.ensure(n * sizeof(double));
  t.sy.ensure(n * sizeof(double));
  uint64_t* keys = t.keys.as<uint64_t>();
  uint32_t* perm = t.perm.as<uint32_t>();
  t.xy.ensure(n * sizeof(double2));
  // -- counters (zeroed before the first kernel, which flags non-finite points)
  enum { CNT_R = 33, CNT_PC = 34, CNT_ROWS = 66, CNT_EVAL = 67, CNT_UP = 68, CNT_UNIQ = 69,
         CNT_LR = 72, CNT_BAL = 104, CNT_BAD = 136, CNT_N = 137,
         CNT_SPLIT = CNT_N, CNT_ALL = CNT_SPLIT + kSplitRows * (kMaxLevels + 1) };
  t.w_cnt.ensure(CNT_ALL * sizeof(int));
  t.h_cnt.ensure(CNT_ALL * sizeof(int));
  int* cnt = t.w_cnt.as<int>();
  int* hc = t.h_cnt.as<int>();
  const int nkey_blk = (int)ceil_div(n, kSortChunk);
  t.w_bad.ensure(nkey_blk * sizeof(int));
  const int sort_T = ncut < kLocalSortMax ? ncut : kLocalSortMax;
  const bool first_round = n > sort_T;
  const SortFirstRound fr = first_round ? sort_first_round(t.pw, n) : SortFirstRound{};
  hipLaunchKernelGGL(k_point_keys, dim3((unsigned)ceil_div(n, kSortChunk)), dim3(KEY_BS), 0, s, x, y,
                     nn, xmin, ymin, w0, key_scale(xmin, ymin, w0),
                     first_round ? fr.keys : keys, first_round ? fr.vals : perm,
                     t.xy.as<double2>(), t.w_bad.as<int>(), cnt, (int)CNT_ALL, fr,
                     first_round ? 1 : 0);
  FMM_LAUNCH_CHECK();
  // Segments of at most Ncutoff points that end in the last radix round sit
  // at the deepest radix depth D, which no cell of the tree goes below (2:1
  // balance splits lie at least two levels above the deepest leaf, X-list
  // children at most at depth D): the radix order within them stands.
  // Segments that end earlier are sorted, since balance and X-list children
  // can split their cells further.
  sort_pairs_u64(keys, perm, n, sort_T, t.pw, s, 64, true, first_round);

  // Small host->device uploads go through distinct regions of a pinned buffer
  // so that the host never rewrites memory a queued copy still reads.
  constexpr int kUpRegion = kMaxLevels + 2;
  t.h_up.ensure(8 * kUpRegion * sizeof(int64_t));
  int64_t* hup = t.h_up.as<int64_t>();
  int up_region = 0;
  auto upload = [&](void* dst, const void* src, size_t bytes) {
    FMM_REQUIRE(up_region < 8 && bytes <= kUpRegion * sizeof(int64_t), FMM_E_INTERNAL,
                "tree: upload staging exhausted");
    int64_t* slot = hup + (int64_t)up_region * kUpRegion;
    ++up_region;
    std::memcpy(slot, src, bytes);
    FMM_HIP(hipMemcpyAsync(dst, slot, bytes, hipMemcpyHostToDevice, s));
  };

  // -- split cells of the refinement --------------------------------------------
  t.w_flag.ensure(n * sizeof(int));
  t.w_rank.ensure(n * sizeof(int));
  t.w_a.ensure(n * sizeof(int2));
  int* flag = t.w_flag.as<int>();
  int* rank = t.w_rank.as<int>();
  int2* sc = t.w_a.as<int2>();
  hipLaunchKernelGGL(k_split_pos, dim3((unsigned)ceil_div(n, 1024)), dim3(1024), 0, s, keys, nn,
                     ncut, flag, sc, cnt + CNT_SPLIT, perm, t.xy.as<double2>(), t.sx.as<double>(),
                     t.sy.as<double>());
  FMM_LAUNCH_CHECK();
  // per depth, then the split positions; then the non-finite flag
  int64_t dcount[kMaxLevels + 2];
  // A complete tree of the previous build's depth is built at once, gated on
  // the device by the split counts (its kernels then follow without waiting
  // for the host to read them); that needs its buffers to leave the split
  // flags in w_flag (still needed otherwise) where they are.
  const int specD = t.full_D;
  const bool spec = t.allow_full && specD >= 2 && specD <= kFullMaxDepth &&
                    t.w_flag.bytes >= ((((size_t)1 << (2 * specD)) + 1) * sizeof(int));
  {
    HostMailbox& mb = t.pw.mbox;
    mb.ensure();
    const uint32_t seq = mb.next();
    int* gate = nullptr;
    if (spec) {
      t.w_gate.ensure(sizeof(int));
      gate = t.w_gate.as<int>();
      tree_full_buffers(t, specD);
    }
    hipLaunchKernelGGL(k_split_counts, dim3(1), dim3(64), 0, s, cnt + CNT_SPLIT,
                       t.w_bad.as<int>(), nkey_blk, mb.dev, seq, specD, gate);
    FMM_LAUNCH_CHECK();
    if (spec) tree_full_kernels(t, specD, nn, s, gate);
    mb.take(kMaxLevels + 2, seq, s, dcount);
  }
  FMM_REQUIRE(dcount[kMaxLevels + 1] == 0, FMM_E_ARG,
              "tree: the source coordinates hold a NaN or infinite value");
  const int np = (int)dcount[kMaxLevels];
  int maxd = -1;
  for (int d = 0; d < kMaxLevels; ++d)
    if (dcount[d] > 0) maxd = d;
  const int L = maxd + 2;   // deepest split depth + children level + 1
  t.L = L;
  FMM_REQUIRE(L <= kMaxLevels, FMM_E_INTERNAL, "tree: %d levels exceed the depth cap", L);

  // a complete tree (every cell of the depths above the deepest split cells
  // split): closed forms of the node arrays (k_full_nodes)
  if (t.allow_full && L >= 3 && L - 1 <= kFullMaxDepth) {
    bool complete = true;
    for (int d = 0; d <= L - 2 && complete; ++d) complete = dcount[d] == ((int64_t)1 << (2 * d));
    if (complete) {
      if (!(spec && L - 1 == specD)) {   // else built already
        tree_full_buffers(t, L - 1);
        tree_full_kernels(t, L - 1, nn, s, nullptr);
      }
      tree_full_host(t, L - 1);
      return;
    }
  }
  t.full_D = 0;

  // compacted split positions and the split-cell lists per depth (S_ref)
  exclusive_scan_i32(flag, rank, n, nullptr, t.pw, s);
  t.w_c.ensure((int64_t)(n + 2) * (sizeof(int) + sizeof(int2)) + 16);
  int* ppos = t.w_c.as<int>();
  int2* psc = reinterpret_cast<int2*>(ppos + n + (n & 1));
  hipLaunchKernelGGL(k_compact_pos, dim3(grid_for(n)), dim3(BS), 0, s, flag, rank, sc, nn, ppos,
                     psc);
  FMM_LAUNCH_CHECK();
  t.s_off.assign(L + 1, 0);
  for (int d = 0; d < L; ++d) t.s_off[d + 1] = t.s_off[d] + (d <= L - 2 ? dcount[d] : 0);
  const int64_t ns_total = t.s_off[L];
  t.w_b.ensure((L + 1) * sizeof(int64_t) * 3 + 64);
  int64_t* d_soff = t.w_b.as<int64_t>();
  t.sref_all.ensure((ns_total + 1) * sizeof(uint64_t));
  if (np > 0) {
    upload(d_soff, t.s_off.data(), (L + 1) * sizeof(int64_t));
    hipLaunchKernelGGL(k_depth_lists, dim3(L - 1), dim3(1024), 0, s, ppos, psc, np, keys, d_soff,
                       t.sref_all.as<uint64_t>());
    FMM_LAUNCH_CHECK();
  }

  // -- 2:1 balance ---------------------------------------------------------------
  bool balanced = true;
  if (L >= 4) {
    int64_t ne = t.s_off[L - 1] - t.s_off[2];
    // the check looks for neighbours of split cells' parents (depths 1 .. L-3)
    // that are not split themselves; with every cell of those depths split
    // there is none
    bool full = true;
    for (int d = 1; d <= L - 3 && full; ++d)
      full = d < 31 && t.s_off[d + 1] - t.s_off[d] == ((int64_t)1 << (2 * d));
    if (ne > 0 && !full) {
      hipLaunchKernelGGL(k_balance_check, dim3(grid_for(ne)), dim3(BS), 0, s,
                         t.sref_all.as<uint64_t>(), d_soff, L, cnt + CNT_R);
      FMM_LAUNCH_CHECK();
      read_counts(t, cnt + CNT_R, 1, hc + CNT_R, s);
      balanced = hc[CNT_R] == 0;
    }
  }
  std::vector<int64_t> soff = t.s_off;
  if (balanced) {
    t.s_all.ensure((ns_total + 1) * sizeof(uint64_t));
    if (ns_total > 0)
      FMM_HIP(hipMemcpyAsync(t.s_all.ptr, t.sref_all.ptr, ns_total * sizeof(uint64_t),
                             hipMemcpyDeviceToDevice, s));
  } else {
    // Bottom-up: S_d = Sref_d u N(parents(S_{d+1})) for d = L-3 .. 1, with
    // S_0 = Sref_0 and S_{L-2} = Sref_{L-2}.  Depths above kBmDepth (very deep
    // trees) build sorted candidate lists; the others are bitmaps.
    const int DM = std::min(L - 3, kBmDepth);
    std::vector<int64_t> sz(L, 0);
    for (int d = 0; d < L; ++d) sz[d] = t.s_off[d + 1] - t.s_off[d];
    const uint64_t* sref = t.sref_all.as<uint64_t>();
    std::vector<DevBuf>& lists = t.bal_lists;
    if ((int)lists.size() < L) lists.resize(L);
    for (int d = L - 3; d > DM; --d) {
      const uint64_t* nxt = d + 1 == L - 2 ? sref + t.s_off[L - 2] : lists[d + 1].as<uint64_t>();
      const int64_t nref = sz[d];
      const int64_t nnext = sz[d + 1];
      const int64_t nc = nref + 9 * nnext;
      t.bal_k.ensure((nc + 1) * sizeof(uint64_t));
      t.bal_v.ensure((nc + 1) * sizeof(uint32_t));
      t.bal_f.ensure((nc + 1) * 2 * sizeof(int));
      uint64_t* candk = t.bal_k.as<uint64_t>();
      int* fl = t.bal_f.as<int>();
      int* rk = fl + (nc + 1);
      hipLaunchKernelGGL(k_bal_candidates, dim3(grid_for(nc)), dim3(BS), 0, s,
                         sref + t.s_off[d], nref, nxt, nnext, d, candk, t.bal_v.as<uint32_t>());
      FMM_LAUNCH_CHECK();
      sort_pairs_u64(candk, t.bal_v.as<uint32_t>(), nc, 1024, t.pw, s, 2 * d);
      hipLaunchKernelGGL(k_unique_flags, dim3(grid_for(nc)), dim3(BS), 0, s, candk, nc, fl);
      FMM_LAUNCH_CHECK();
      exclusive_scan_i32(fl, rk, nc, cnt + CNT_UNIQ, t.pw, s);
      read_counts(t, cnt + CNT_UNIQ, 1, hc + CNT_UNIQ, s);
      sz[d] = hc[CNT_UNIQ];
      lists[d].ensure((sz[d] + 1) * sizeof(uint64_t));
      hipLaunchKernelGGL(k_unique_scatter, dim3(grid_for(nc)), dim3(BS), 0, s, candk, fl, rk, nc,
                         d, lists[d].as<uint64_t>());
      FMM_LAUNCH_CHECK();
    }
    BmOffsets bo{};
    int64_t nwords = 0;
    for (int d = 1; d <= DM; ++d) {
      bo.wo[d] = nwords;
      const int64_t w = std::max<int64_t>(1, ((int64_t)1 << (2 * d)) / 32);
      nwords += ceil_div(w, BS) * BS;
    }
    bo.wo[DM + 1] = nwords;
    if (DM >= 1) {
      t.bal_bm.ensure(nwords * sizeof(uint32_t));
      t.bal_pc.ensure(nwords * sizeof(int));
      uint32_t* bm = t.bal_bm.as<uint32_t>();
      FMM_HIP(hipMemsetAsync(bm, 0, nwords * sizeof(uint32_t), s));
      const int64_t nmark = t.s_off[DM + 1] - t.s_off[1];
      if (nmark > 0) {
        hipLaunchKernelGGL(k_bm_mark, dim3(grid_for(nmark)), dim3(BS), 0, s, sref, d_soff, DM, bo,
                           bm);
        FMM_LAUNCH_CHECK();
      }
      for (int d = DM; d >= 1; --d) {
        if (d + 1 <= DM) {
          const int64_t npar = (int64_t)1 << (2 * d);   // depth-d cells
          if (npar <= ((int64_t)1 << 22))
            hipLaunchKernelGGL(k_bm_spread, dim3(grid_for(npar)), dim3(BS), 0, s,
                               bm + bo.wo[d + 1], d, bm + bo.wo[d]);
          else
            hipLaunchKernelGGL(k_bm_spread_w, dim3(grid_for(npar / 8)), dim3(BS), 0, s,
                               bm + bo.wo[d + 1], d, bm + bo.wo[d]);
        } else {
          const uint64_t* lst =
              d + 1 == L - 2 ? sref + t.s_off[L - 2] : lists[d + 1].as<uint64_t>();
          if (sz[d + 1] == 0) continue;
          hipLaunchKernelGGL(k_bm_spread_list, dim3(grid_for(sz[d + 1])), dim3(BS), 0, s, lst,
                             sz[d + 1], d, bm + bo.wo[d]);
        }
        FMM_LAUNCH_CHECK();
      }
      hipLaunchKernelGGL(k_bm_count, dim3(nwords / BS), dim3(BS), 0, s, bm, bo, DM,
                         t.bal_pc.as<int>(), cnt + CNT_BAL);
      FMM_LAUNCH_CHECK();
      exclusive_scan_i32(t.bal_pc.as<int>(), t.bal_pc.as<int>(), nwords, nullptr, t.pw, s);
      read_counts(t, cnt + CNT_BAL, kMaxLevels, hc + CNT_BAL, s);
    }
    for (int d = 1; d <= DM; ++d) sz[d] = hc[CNT_BAL + d];
    soff.assign(L + 1, 0);
    for (int d = 0; d < L; ++d) soff[d + 1] = soff[d] + (d <= L - 2 ? sz[d] : 0);
    t.s_all.ensure((soff[L] + 1) * sizeof(uint64_t));
    uint64_t* sall = t.s_all.as<uint64_t>();
    if (sz[0] > 0)
      FMM_HIP(hipMemcpyAsync(sall, sref, sz[0] * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    if (DM >= 1 && soff[DM + 1] > soff[1]) {
      hipLaunchKernelGGL(k_bm_scatter, dim3(nwords / BS), dim3(BS), 0, s, t.bal_bm.as<uint32_t>(),
                         bo, DM, t.bal_pc.as<int>(), sall + soff[1]);
      FMM_LAUNCH_CHECK();
    }
    for (int d = DM + 1; d <= L - 3; ++d)
      if (sz[d] > 0)
        FMM_HIP(hipMemcpyAsync(sall + soff[d], lists[d].ptr, sz[d] * sizeof(uint64_t),
                               hipMemcpyDeviceToDevice, s));
    if (L - 2 >= 1 && sz[L - 2] > 0)
      FMM_HIP(hipMemcpyAsync(sall + soff[L - 2], sref + t.s_off[L - 2],
                             sz[L - 2] * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  }
  t.s_off = soff;

  // -- nodes with children (P_d = S_d u X_d) ----------------------------------------
  std::vector<int64_t> roff(L + 1, 0);
  roff[1] = 1;
  for (int d = 1; d < L; ++d) roff[d + 1] = roff[d] + 4 * (soff[d] - soff[d - 1]);
  const int64_t nr = roff[L];
  int64_t* d_soff2 = d_soff;
  int64_t* d_roff = d_soff + (L + 1);
  int64_t* d_poff = d_soff + 2 * (L + 1);
  upload(d_soff2, soff.data(), (L + 1) * sizeof(int64_t));
  upload(d_roff, roff.data(), (L + 1) * sizeof(int64_t));
  t.w_c.ensure(nr * (sizeof(uint64_t) + 3 * sizeof(int)) + 64);
  uint64_t* rcode = t.w_c.as<uint64_t>();
  int* hasc = reinterpret_cast<int*>(rcode + nr);
  int* kind = hasc + nr;
  int* rrank = kind + nr;
  hipLaunchKernelGGL(k_real_flags, dim3(grid_for(nr)), dim3(BS), 0, s, t.s_all.as<uint64_t>(),
                     d_soff2, d_roff, L, rcode, hasc, kind, cnt + CNT_PC);
  FMM_LAUNCH_CHECK();
  exclusive_scan_i32(hasc, rrank, nr, nullptr, t.pw, s);
  read_counts(t, cnt + CNT_PC, kMaxLevels, hc + CNT_PC, s);
  t.p_off.assign(L + 1, 0);
  for (int d = 0; d < L; ++d) t.p_off[d + 1] = t.p_off[d] + hc[CNT_PC + d];
  FMM_REQUIRE(hc[CNT_PC + L - 1] == 0, FMM_E_INTERNAL, "tree: deepest level has children");
  const int64_t np_total = t.p_off[L];
  t.p_code.ensure((np_total + 1) * sizeof(uint64_t));
  t.p_kind.ensure((np_total + 1) * sizeof(int));
  hipLaunchKernelGGL(k_p_scatter, dim3(grid_for(nr)), dim3(BS), 0, s, rcode, hasc, kind, rrank, nr,
                     t.p_code.as<uint64_t>(), t.p_kind.as<int>());
  FMM_LAUNCH_CHECK();
  upload(d_poff, t.p_off.data(), (L + 1) * sizeof(int64_t));

  // -- final node arrays ------------------------------------------------------------
  t.lev.assign(L, TreeLevel());
  std::vector<int> levoff(L + 1, 0);
  int64_t off = 0;
  for (int d = 0; d < L; ++d) {
    int64_t nd = d == 0 ? 1 : 4 * (t.p_off[d] - t.p_off[d - 1]);
    t.lev[d].node_off = off;
    t.lev[d].n_nodes = nd;
    t.lev[d].width = std::ldexp(w0, -d);
    levoff[d] = (int)off;
    off += nd;
  }
  levoff[L] = (int)off;
  FMM_REQUIRE(off < ((int64_t)1 << 31) / kNumM2L, FMM_E_LIMIT, "tree: too many nodes (%lld)",
              (long long)off);
  const int NN = (int)off;
  t.n_nodes = off;
  t.levoff.ensure((L + 1) * sizeof(int));
  upload(t.levoff.ptr, levoff.data(), (L + 1) * sizeof(int));
  int* d_levoff = t.levoff.as<int>();
  t.code.ensure(off * sizeof(uint64_t));
  t.nlevel.ensure(off * sizeof(int));
  t.bot.ensure(off * sizeof(int));
  t.top.ensure(off * sizeof(int));
  t.parent.ensure(off * sizeof(int));
  t.child.ensure(off * sizeof(int));
  t.flags.ensure(off * sizeof(int));
  t.coll.ensure(off * 9 * sizeof(int));
  t.xmid.ensure(off * sizeof(double));
  t.ymid.ensure(off * sizeof(double));
  t.row.ensure(off * sizeof(int));
  t.m2l.ensure(off * kNumM2L * sizeof(int));
  t.evalid.ensure(off * sizeof(int));
  hipLaunchKernelGGL(k_nodes_a, dim3(grid_for(off)), dim3(BS), 0, s, t.p_code.as<uint64_t>(),
                     t.p_kind.as<int>(), d_poff, d_levoff, L, t.code.as<uint64_t>(),
                     t.nlevel.as<int>(), t.flags.as<int>());
  FMM_LAUNCH_CHECK();
  NodeIdx ni{};
  int64_t ntab = 0;
  for (int d = 0; d < kMaxLevels; ++d) ni.off[d] = -1;
  for (int d = 1; d < L && d <= kIdxDepth; ++d) {
    if (t.lev[d].n_nodes == ((int64_t)1 << (2 * d))) continue;   // full: direct
    ni.off[d] = ntab;
    ntab += (int64_t)1 << (2 * d);
  }
  t.nidx_off.assign(ni.off, ni.off + kMaxLevels);
  if (ntab > 0) {
    t.nidx.ensure(ntab * sizeof(int));
    FMM_HIP(hipMemsetAsync(t.nidx.ptr, 0xFF, ntab * sizeof(int), s));
    ni.tab = t.nidx.as<int>();
    hipLaunchKernelGGL(k_node_index, dim3(grid_for(off)), dim3(BS), 0, s, t.code.as<uint64_t>(),
                       t.nlevel.as<int>(), NN, ni, t.nidx.as<int>());
    FMM_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_cell_ranges_init, dim3(grid_for(off)), dim3(BS), 0, s, NN, nn,
                     t.bot.as<int>(), t.top.as<int>());
  FMM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cell_ranges, dim3(grid_for(n)), dim3(BS), 0, s, keys, sc, nn,
                     t.code.as<uint64_t>(), d_levoff, L, ni, t.bot.as<int>(), t.top.as<int>());
  FMM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_nodes_b, dim3(ceil_div(off, 64)), dim3(64), 0, s, t.p_code.as<uint64_t>(),
                     t.p_kind.as<int>(), d_poff, d_levoff, L, keys, nn, xmin, ymin, w0, ni,
                     t.code.as<uint64_t>(), t.flags.as<int>(), t.bot.as<int>(), t.top.as<int>(),
                     t.parent.as<int>(), t.child.as<int>(), t.coll.as<int>(), t.xmid.as<double>(),
                     t.ymid.as<double>());
  FMM_LAUNCH_CHECK();
  t.crange.ensure(((int64_t)off * 9 + 1) * sizeof(int2));
  hipLaunchKernelGGL(k_coll_ranges, dim3(grid_for((int64_t)off * 9)), dim3(BS), 0, s,
                     (int)(off * 9), t.coll.as<int>(), t.bot.as<int>(), t.top.as<int>(),
                     t.crange.as<int2>());
  FMM_LAUNCH_CHECK();

  // -- density rows ------------------------------------------------------------------
  t.w_flag.ensure((off + 1) * sizeof(int));
  t.w_rank.ensure((off + 1) * sizeof(int));
  flag = t.w_flag.as<int>();
  rank = t.w_rank.as<int>();
  const int g0 = L > 2 ? levoff[2] : NN;
  const int ngroups = (NN - g0) / 4;
  if (ngroups > 0) {
    hipLaunchKernelGGL(k_group_flags, dim3(grid_for(ngroups)), dim3(BS), 0, s, t.flags.as<int>(),
                       g0, ngroups, d_levoff, L, flag, cnt + CNT_LR);
    FMM_LAUNCH_CHECK();
    exclusive_scan_i32(flag, rank, ngroups, cnt + CNT_ROWS, t.pw, s);
  }
  t.rownode.ensure((4 * (int64_t)ngroups + 1) * sizeof(int));
  hipLaunchKernelGGL(k_rows, dim3(grid_for(off)), dim3(BS), 0, s, flag, rank, g0, NN,
                     t.row.as<int>(), t.rownode.as<int>());
  FMM_LAUNCH_CHECK();
  if (L > 2) {
    hipLaunchKernelGGL(k_m2l_lists, dim3(grid_for((int64_t)(NN - g0) * kNumM2L)), dim3(BS), 0, s,
                       t.code.as<uint64_t>(), t.flags.as<int>(), t.row.as<int>(), d_levoff, L,
                       ni, t.m2l.as<int>());
    FMM_LAUNCH_CHECK();
  }

  // -- evaluation boxes and S2M boxes ---------------------------------------------------
  t.w_a.ensure((off + 1) * 2 * sizeof(int));
  int* f2 = t.w_a.as<int>();
  int* r2 = f2 + off;
  hipLaunchKernelGGL(k_flag_bit, dim3(grid_for(off)), dim3(BS), 0, s, t.flags.as<int>(), NN,
                     (int)NF_LEAF, flag);
  FMM_LAUNCH_CHECK();
  exclusive_scan_i32(flag, rank, off, cnt + CNT_EVAL, t.pw, s);
  hipLaunchKernelGGL(k_up_flag, dim3(grid_for(off)), dim3(BS), 0, s, t.flags.as<int>(),
                     t.row.as<int>(), NN, f2);
  FMM_LAUNCH_CHECK();
  exclusive_scan_i32(f2, r2, off, cnt + CNT_UP, t.pw, s);
  read_counts(t, cnt + CNT_ROWS, CNT_N - CNT_ROWS, hc + CNT_ROWS, s);
  t.n_rows = ngroups > 0 ? 4 * (int64_t)hc[CNT_ROWS] : 0;
  t.row_off.assign(L + 1, 0);
  for (int d = 0; d < L; ++d) t.row_off[d + 1] = t.row_off[d] + hc[CNT_LR + d];
  FMM_REQUIRE(t.row_off[L] == t.n_rows, FMM_E_INTERNAL, "tree: row count mismatch");
  t.n_eval = hc[CNT_EVAL];
  t.n_up = hc[CNT_UP];
  t.rowinfo.ensure((t.n_rows + 1) * sizeof(int2));
  if (t.n_rows > 0) {
    hipLaunchKernelGGL(k_row_info, dim3(grid_for(t.n_rows)), dim3(BS), 0, s, t.rownode.as<int>(),
                       (int)t.n_rows, t.flags.as<int>(), t.child.as<int>(), t.row.as<int>(),
                       t.rowinfo.as<int2>());
    FMM_LAUNCH_CHECK();
  }
  t.evalnode.ensure((t.n_eval + 1) * sizeof(int));
  t.upnodes.ensure((t.n_up + 1) * sizeof(int));
  hipLaunchKernelGGL(k_list_scatter, dim3(grid_for(off)), dim3(BS), 0, s, flag, rank, NN,
                     t.evalnode.as<int>(), t.evalid.as<int>());
  FMM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_list_scatter, dim3(grid_for(off)), dim3(BS), 0, s, f2, r2, NN,
                     t.upnodes.as<int>(), (int*)nullptr);
  FMM_LAUNCH_CHECK();
  ++t.build_id;
  t.built = true;
}

void tree_init_device() {
  SlotOffsets so;
  for (int q = 0; q < kNumM2L; ++q) m2l_slot_offset(q, &so.ox[q], &so.oy[q]);
  FMM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_slot_off), &so, sizeof so));
}

void tree_locate(const Tree& t, const double* tx, const double* ty, int64_t nt, int* out_node,
                 hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_locate, dim3(grid_for(nt)), dim3(BS), 0, s, tx, ty, nt, t.flags.as<int>(),
                     t.child.as<int>(), t.xmid.as<double>(), t.ymid.as<double>(), out_node);
  FMM_LAUNCH_CHECK();
}

}  // namespace fmm
