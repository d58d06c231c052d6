// Device-wide primitives: exclusive scan and segmented MSD radix sort.
#include "prims.h"

namespace fmm {

namespace {

constexpr int SCAN_BS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BS * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_BS) void k_scan_block_sums(const int* __restrict__ in, int64_t n,
                                                            int* __restrict__ sums) {
  __shared__ int lds[SCAN_BS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int acc = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    int64_t idx = base + (int64_t)i * SCAN_BS + threadIdx.x;
    if (idx < n) acc += in[idx];
  }
  __shared__ int tot;
  (void)block_exclusive_scan<SCAN_BS>(acc, lds, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_sums(int* __restrict__ sums, int64_t nb,
                                                   int* __restrict__ total) {
  __shared__ int lds[1024 / 64];
  __shared__ int tot;
  int carry = 0;
  for (int64_t base = 0; base < nb; base += 1024) {
    int64_t i = base + threadIdx.x;
    int v = i < nb ? sums[i] : 0;
    int ex = block_exclusive_scan<1024>(v, lds, &tot);
    if (i < nb) sums[i] = carry + ex;
    __syncthreads();
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_apply(const int* in, int* out, int64_t n,
                                                       const int* __restrict__ sums) {
  __shared__ int lds[SCAN_BS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int acc = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    int64_t idx = base + i;
    v[i] = idx < n ? in[idx] : 0;
    acc += v[i];
  }
  int ex = block_exclusive_scan<SCAN_BS>(acc, lds, nullptr) + sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; ++i) {
    int64_t idx = base + i;
    if (idx < n) out[idx] = ex;
    ex += v[i];
  }
}

// n <= kScanSmallMax: one block of 1024 threads; element j * 1024 + t is
// loaded by thread t (coalesced, all loads in flight together), each of the
// J = ceil(n / 1024) tiles is scanned by wave scans, and the wave and tile
// totals are combined through LDS with one barrier.
constexpr int kScanSmallTiles = 32;
constexpr int kScanSmallMax = kScanSmallTiles * 1024;
__global__ __launch_bounds__(1024) void k_scan_small(const int* in, int* out, int n,
                                                    int* __restrict__ total) {
  __shared__ int wsum[kScanSmallTiles][16];
  __shared__ int woff[kScanSmallTiles][16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int J = (n + 1023) >> 10;
  int v[kScanSmallTiles], inc[kScanSmallTiles];
#pragma unroll
  for (int j = 0; j < kScanSmallTiles; ++j) {
    const int idx = j * 1024 + t;
    v[j] = (j < J && idx < n) ? in[idx] : 0;
  }
#pragma unroll
  for (int j = 0; j < kScanSmallTiles; ++j) {
    if (j < J) {
      inc[j] = wave_inclusive_scan(v[j]);
      if (lane == 63) wsum[j][wid] = inc[j];
    }
  }
  __syncthreads();
  if (wid == 0) {
    // lane j < J: offsets of the 16 waves inside tile j, then the tile prefix
    int tile_tot = 0;
    if (lane < J)
      for (int w = 0; w < 16; ++w) tile_tot += wsum[lane][w];
    const int tile_off = wave_inclusive_scan(lane < J ? tile_tot : 0) - (lane < J ? tile_tot : 0);
    if (lane < J) {
      int run = tile_off;
      for (int w = 0; w < 16; ++w) {
        woff[lane][w] = run;
        run += wsum[lane][w];
      }
      if (lane == J - 1 && total) *total = run;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanSmallTiles; ++j) {
    const int idx = j * 1024 + t;
    if (j < J && idx < n) out[idx] = woff[j][wid] + inc[j] - v[j];
  }
}

}  // namespace

// 16-byte stores for the aligned body (one int per thread ran at a third of
// the bandwidth), single ints for the up to 3 before it and after it.
__global__ __launch_bounds__(256) void k_fill_i32(int* __restrict__ p, int64_t n, int v) {
  const int64_t head = min((int64_t)(((16 - ((uintptr_t)p & 15)) & 15) >> 2), n);
  const int64_t nb = (n - head) >> 2;
  const int64_t t0 = head + 4 * nb;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < nb) reinterpret_cast<int4*>(p + head)[i] = make_int4(v, v, v, v);
  if (i < head) p[i] = v;
  if (i < n - t0) p[t0 + i] = v;
}

void fill_i32(int* p, int64_t n, int v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(ceil_div(n / 4 + 4, 256)), dim3(256), 0, s, p, n, v);
  FMM_LAUNCH_CHECK();
}

int64_t scan_partials_i32(int64_t n) { return n <= kScanSmallMax ? 0 : ceil_div(n, SCAN_TILE); }

void exclusive_scan_i32(const int* in, int* out, int64_t n, int* total_dev, PrimWork& w,
                        hipStream_t s, int* partials) {
  if (n <= 0) {
    if (total_dev) fill_i32(total_dev, 1, 0, s);
    return;
  }
  if (n <= kScanSmallMax) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, in, out, (int)n, total_dev);
    FMM_LAUNCH_CHECK();
    return;
  }
  int64_t nb = ceil_div(n, SCAN_TILE);
  int* sums = partials;
  if (!sums) {
    w.blocksums.ensure(nb * sizeof(int));
    sums = w.blocksums.as<int>();
  }
  hipLaunchKernelGGL(k_scan_block_sums, dim3(nb), dim3(SCAN_BS), 0, s, in, n, sums);
  FMM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nb, total_dev);
  FMM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_BS), 0, s, in, out, n, sums);
  FMM_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Segmented MSD radix sort.
namespace {

using namespace sort_detail;
constexpr int CHUNK = kSortChunk;
constexpr int SORT_BS = 256;
constexpr int kChunkPer = CHUNK / SORT_BS;   // keys per thread of a chunk block
constexpr int kWaveSortMax = 64;

// nseg_host < 0: the round's segments are the C_NEXT of the round before, as
// the device holds it (a round launched before the host read the counts).
__global__ __launch_bounds__(1024) void k_make_chunks(const int2* __restrict__ seg, int nseg_host,
                                                     int4* __restrict__ chunks,
                                                     int* __restrict__ segChunk0,
                                                     int* __restrict__ counters) {
  __shared__ int lds[1024 / 64];
  __shared__ int tot;
  __shared__ int s_nseg;
  // counters of this round's segment classification (k_seg_offsets adds to
  // them, after this kernel), and the ones it starts from
  if (threadIdx.x == 0) {
    const int ns = nseg_host >= 0 ? nseg_host : counters[C_NEXT];
    s_nseg = ns;
    counters[C_SEGS] = ns;
    counters[C_ACTIVE_PREV] = counters[C_ACTIVE];
    counters[C_FINAL_PREV] = counters[C_FINAL];
    counters[C_BIG_PREV] = counters[C_FINAL_BIG];
    counters[C_NEXT] = 0;
    counters[C_ACTIVE] = 0;
  }
  __syncthreads();
  const int nseg = s_nseg;
  int carry = 0;
  for (int base = 0; base < nseg; base += 1024) {
    int si = base + threadIdx.x;
    int2 sg = si < nseg ? seg[si] : make_int2(0, 0);
    int cnt = si < nseg ? (int)ceil_div(sg.y, CHUNK) : 0;
    int ex = block_exclusive_scan<1024>(cnt, lds, &tot);
    if (si < nseg) segChunk0[si] = carry + ex;
    __syncthreads();
    carry += tot;
    __syncthreads();
  }
  __threadfence_block();
  __syncthreads();
  // chunk c belongs to the last segment whose first chunk is <= c
  for (int c = threadIdx.x; c < carry; c += 1024) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (segChunk0[mid] <= c) lo = mid;
      else hi = mid - 1;
    }
    const int2 sg = seg[lo];
    const int k = c - segChunk0[lo];
    chunks[c] = make_int4(lo, sg.x + k * CHUNK, min(CHUNK, sg.y - k * CHUNK), 0);
  }
  if (threadIdx.x == 0) counters[C_CHUNKS] = carry;
}

__global__ __launch_bounds__(SORT_BS) void k_chunk_hist(const uint64_t* __restrict__ keys,
                                                       const int4* __restrict__ chunks,
                                                       const int* __restrict__ counters, int shift,
                                                       int* __restrict__ chunkHist) {
  __shared__ int hist[256];
  const int b = blockIdx.x;
  if (b >= counters[C_CHUNKS]) return;
  int4 ch = chunks[b];
  hist[threadIdx.x] = 0;
  // all loads first: the LDS atomics would otherwise wait for each in turn
  uint64_t k[kChunkPer];
#pragma unroll
  for (int u = 0; u < kChunkPer; ++u) {
    const int i = threadIdx.x + u * SORT_BS;
    k[u] = i < ch.z ? keys[ch.y + i] : 0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kChunkPer; ++u)
    if (threadIdx.x + u * SORT_BS < ch.z) atomicAdd(&hist[(int)((k[u] >> shift) & 255u)], 1);
  __syncthreads();
  chunkHist[(int64_t)b * 256 + threadIdx.x] = hist[threadIdx.x];
}

// One block per segment, 1024 threads: digit d = tid % 256, part = tid / 256
// scans a quarter of the segment's chunks.  Output: chunkHist[c][d] = the
// destination of chunk c's first element of digit d.  The digit segments of
// the next round and the final ones are appended with one atomic per block.
constexpr int SEG_BS = 1024;

__global__ __launch_bounds__(SEG_BS) void k_seg_offsets(
    const int2* __restrict__ seg, const int* __restrict__ segChunk0, int* __restrict__ chunkHist,
    int threshold, int last_round, int2* __restrict__ segNext, int2* __restrict__ fin,
    int finBigOff, int* __restrict__ counters) {
  __shared__ int lds[SEG_BS / 64];
  __shared__ int part_sum[4][256];
  __shared__ int bases[3];
  const int si = blockIdx.x;
  if (si >= counters[C_SEGS]) return;   // block-uniform: a grid sized by a bound
  const int d = threadIdx.x & 255, part = threadIdx.x >> 8;
  const int2 sg = seg[si];
  const int c0 = segChunk0[si];
  const int nc = (int)ceil_div(sg.y, CHUNK);
  const int q = (nc + 3) >> 2;
  const int cb = c0 + min(nc, part * q), ce = c0 + min(nc, (part + 1) * q);
  // A part of at most R chunks (segments of up to 4 R chunks) is read in one
  // batch and kept in registers for the second pass; larger ones are read in
  // batches of B, twice.  The kernel's time is the latency of these reads.
  constexpr int R = 64, B = 16;
  const bool kept = q > B && ce - cb <= R;
  int hk[R];
  int sum = 0;
  int c = cb;
  if (kept) {
#pragma unroll
    for (int u = 0; u < R; ++u) hk[u] = cb + u < ce ? chunkHist[(int64_t)(cb + u) * 256 + d] : 0;
#pragma unroll
    for (int u = 0; u < R; ++u) sum += hk[u];
  } else {
    for (; c + B <= ce; c += B) {
      int h[B];
#pragma unroll
      for (int u = 0; u < B; ++u) h[u] = chunkHist[(int64_t)(c + u) * 256 + d];
#pragma unroll
      for (int u = 0; u < B; ++u) sum += h[u];
    }
    for (; c < ce; ++c) sum += chunkHist[(int64_t)c * 256 + d];
  }
  part_sum[part][d] = sum;
  __syncthreads();
  const int run = part_sum[0][d] + part_sum[1][d] + part_sum[2][d] + part_sum[3][d];
  // digit starts: exclusive scan of the totals over the 256 digits
  const int dstart = block_exclusive_scan<SEG_BS>(part == 0 ? run : 0, lds, nullptr);
  __shared__ int dstart_s[256];
  if (part == 0) dstart_s[d] = dstart;
  __syncthreads();
  int pre = 0;
  for (int pp = 0; pp < part; ++pp) pre += part_sum[pp][d];
  int pos = sg.x + dstart_s[d] + pre;
  // bit 31 of a destination: the digit's segment continues to the next round
  const int cont = (run > threshold && !last_round) ? (int)0x80000000u : 0;
  if (kept) {
#pragma unroll
    for (int u = 0; u < R; ++u) {
      if (cb + u < ce) chunkHist[(int64_t)(cb + u) * 256 + d] = pos | cont;
      pos += hk[u];
    }
  }
  c = kept ? ce : cb;
  for (; c + B <= ce; c += B) {
    int h[B];
#pragma unroll
    for (int u = 0; u < B; ++u) h[u] = chunkHist[(int64_t)(c + u) * 256 + d];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      chunkHist[(int64_t)(c + u) * 256 + d] = pos | cont;
      pos += h[u];
    }
  }
  for (; c < ce; ++c) {
    const int h = chunkHist[(int64_t)c * 256 + d];
    chunkHist[(int64_t)c * 256 + d] = pos | cont;
    pos += h;
  }
  // classify the digit segments and append them with one atomic per list
  const int start = sg.x + dstart_s[d];
  const bool nxt = part == 0 && run > threshold && !last_round;
  const bool fsm = part == 0 && !nxt && run >= 2 && run <= kWaveSortMax;
  const bool fbg = part == 0 && !nxt && run > kWaveSortMax && run <= kLocalSortMax;
  // counts of the three kinds packed in one scan (each at most 256)
  const int code = (nxt ? 1 : 0) | (fsm ? 1 << 10 : 0) | (fbg ? 1 << 20 : 0);
  __shared__ int tot;
  const int ex = block_exclusive_scan<SEG_BS>(code, lds, &tot);
  if (threadIdx.x == 0) {
    const int tn = tot & 1023, tf = (tot >> 10) & 1023, tb = (tot >> 20) & 1023;
    bases[0] = tn ? atomicAdd(&counters[C_NEXT], tn) : 0;
    bases[1] = tf ? atomicAdd(&counters[C_FINAL], tf) : 0;
    bases[2] = tb ? atomicAdd(&counters[C_FINAL_BIG], tb) : 0;
  }
  int act = nxt ? run : 0;
  __syncthreads();
  if (nxt) segNext[bases[0] + (ex & 1023)] = make_int2(start, run);
  if (fsm) fin[bases[1] + ((ex >> 10) & 1023)] = make_int2(start, run);
  if (fbg) fin[finBigOff - 1 - (bases[2] + ((ex >> 20) & 1023))] = make_int2(start, run);
  // active elements of the next round: one atomic per block
  __shared__ int act_tot;
  (void)block_exclusive_scan<SEG_BS>(act, lds, &act_tot);
  if (threadIdx.x == 0 && act_tot) atomicAdd(&counters[C_ACTIVE], act_tot);
}

// Scatter of one chunk by its digit, staged through LDS: the chunk is first
// grouped by digit in LDS (rank within the digit from an LDS atomic, digit
// starts from a scan of the counts), then written out in that order, so
// consecutive lanes store consecutive places of one digit's run instead of 64
// unrelated places per store instruction.
__global__ __launch_bounds__(SORT_BS) void k_chunk_scatter(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
    const int4* __restrict__ chunks, const int* __restrict__ counters, int shift,
    const int* __restrict__ chunkHist, uint64_t* __restrict__ keysB, uint32_t* __restrict__ valsB,
    uint64_t* __restrict__ kfin, uint32_t* __restrict__ vfin, uint64_t* __restrict__ mbox,
    uint32_t seq) {
  __shared__ int gdst[256];    // destination of the chunk's first element of digit d
  __shared__ int lstart[256];  // count, then start, of digit d inside the chunk
  __shared__ int scan_lds[SORT_BS / 64];
  __shared__ uint64_t sk[CHUNK];
  __shared__ uint32_t sv[CHUNK];
  const int b = blockIdx.x;
  // the round's counters are final (k_seg_offsets ran before): block 0 hands
  // them to the host at once, so the host can queue the next round's kernels
  // while this one scatters
  if (b == 0 && threadIdx.x < C_COUNT) mailbox_put(mbox, threadIdx.x, seq, counters[threadIdx.x]);
  if (b >= counters[C_CHUNKS]) return;
  const int4 ch = chunks[b];
  gdst[threadIdx.x] = chunkHist[(int64_t)b * 256 + threadIdx.x];
  lstart[threadIdx.x] = 0;
  uint64_t k[kChunkPer];
  uint32_t v[kChunkPer];
#pragma unroll
  for (int u = 0; u < kChunkPer; ++u) {
    const int i = threadIdx.x + u * SORT_BS;
    if (i < ch.z) {
      k[u] = keys[ch.y + i];
      v[u] = vals[ch.y + i];
    }
  }
  __syncthreads();
  int r[kChunkPer];
#pragma unroll
  for (int u = 0; u < kChunkPer; ++u)
    if (threadIdx.x + u * SORT_BS < ch.z) r[u] = atomicAdd(&lstart[(int)((k[u] >> shift) & 255u)], 1);
  __syncthreads();
  const int start = block_exclusive_scan<SORT_BS>(lstart[threadIdx.x], scan_lds, nullptr);
  lstart[threadIdx.x] = start;
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kChunkPer; ++u) {
    if (threadIdx.x + u * SORT_BS < ch.z) {
      const int j = lstart[(int)((k[u] >> shift) & 255u)] + r[u];
      sk[j] = k[u];
      sv[j] = v[u];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kChunkPer; ++u) {
    const int j = threadIdx.x + u * SORT_BS;
    if (j < ch.z) {
      const uint64_t kj = sk[j];
      const int d = (int)((kj >> shift) & 255u);
      const int gd = gdst[d];
      const int pos = (gd & 0x7fffffff) + (j - lstart[d]);
      if (kfin && gd >= 0) {   // the digit's segment finished: its final place
        kfin[pos] = kj;
        vfin[pos] = sv[j];
      } else {
        keysB[pos] = kj;
        valsB[pos] = sv[j];
      }
    }
  }
}

__global__ __launch_bounds__(SORT_BS) void k_chunk_copy(const uint64_t* __restrict__ keysB,
                                                       const uint32_t* __restrict__ valsB,
                                                       const int4* __restrict__ chunks,
                                                       const int* __restrict__ counters,
                                                       uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
  const int b = blockIdx.x;
  if (b >= counters[C_CHUNKS]) return;
  int4 ch = chunks[b];
  for (int i = threadIdx.x; i < ch.z; i += SORT_BS) {
    keys[ch.y + i] = keysB[ch.y + i];
    vals[ch.y + i] = valsB[ch.y + i];
  }
}

__global__ void k_sort_init(int2* __restrict__ seg0, int2* __restrict__ fin, int finBigOff,
                            int* __restrict__ counters, int n, int T) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i < C_COUNT; ++i) counters[i] = 0;
  if (n > T) {
    seg0[0] = make_int2(0, n);
  } else if (n <= kWaveSortMax) {
    fin[0] = make_int2(0, n);
    counters[C_FINAL] = 1;
  } else {
    fin[finBigOff - 1] = make_int2(0, n);
    counters[C_FINAL_BIG] = 1;
  }
}

// Wave-level bitonic sort of segments of at most 64 (key, value) pairs held
// one per lane; 4 segments per block.
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  const int lo = __shfl((int)(uint32_t)v, src, 64);
  const int hi = __shfl((int)(uint32_t)(v >> 32), src, 64);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Segments of at most 64 pairs: one wave per segment, each lane computes the
// rank of its (key, value) among the segment's pairs by broadcasting every
// pair in turn (equal pairs are ranked by lane, so the ranks are a
// permutation), then stores its pair at that rank.
__global__ __launch_bounds__(256) void k_wave_sort(uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals,
                                                  const int2* __restrict__ fin, int nfin) {
  const int si = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (si >= nfin) return;   // wave-uniform
  const int lane = threadIdx.x & 63;
  const int2 sg = fin[si];
  const int cnt = __builtin_amdgcn_readfirstlane(sg.y);
  uint64_t k = ~0ull;
  uint32_t v = 0xffffffffu;
  if (lane < cnt) {
    k = keys[sg.x + lane];
    v = vals[sg.x + lane];
  }
  const uint32_t klo = (uint32_t)k, khi = (uint32_t)(k >> 32);
  int rank = 0;
  for (int j = 0; j < cnt; ++j) {
    const uint32_t jlo = (uint32_t)__builtin_amdgcn_readlane((int)klo, j);
    const uint32_t jhi = (uint32_t)__builtin_amdgcn_readlane((int)khi, j);
    const uint32_t jv = (uint32_t)__builtin_amdgcn_readlane((int)v, j);
    const uint64_t kj = ((uint64_t)jhi << 32) | jlo;
    rank += (kj < k || (kj == k && (jv < v || (jv == v && j < lane)))) ? 1 : 0;
  }
  if (lane < cnt) {
    keys[sg.x + rank] = k;
    vals[sg.x + rank] = v;
  }
}

__device__ __forceinline__ bool pair_gt(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb) {
  return ka > kb || (ka == kb && va > vb);
}

__global__ __launch_bounds__(SORT_BS) void k_local_sort(uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals,
                                                       const int2* __restrict__ fin) {
  __shared__ uint64_t sk[kLocalSortMax];
  __shared__ uint32_t sv[kLocalSortMax];
  int2 sg = fin[blockIdx.x];
  int P = 1;
  while (P < sg.y) P <<= 1;
  for (int i = threadIdx.x; i < P; i += SORT_BS) {
    if (i < sg.y) {
      sk[i] = keys[sg.x + i];
      sv[i] = vals[sg.x + i];
    } else {
      sk[i] = ~0ull;
      sv[i] = 0xffffffffu;
    }
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += SORT_BS) {
        int ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          uint64_t ka = sk[i], kb = sk[ixj];
          uint32_t va = sv[i], vb = sv[ixj];
          bool gt = pair_gt(ka, va, kb, vb);
          if (gt == up) {
            sk[i] = kb;
            sk[ixj] = ka;
            sv[i] = vb;
            sv[ixj] = va;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < sg.y; i += SORT_BS) {
    keys[sg.x + i] = sk[i];
    vals[sg.x + i] = sv[i];
  }
}

}  // namespace

// Work buffers of the first round for n keys (sort_pairs_u64 sizes them the
// same way, so it finds them in place).
constexpr int kSecondRoundSegs = 256;   // segments a one-segment round can leave

static void first_round_buffers(PrimWork& w, int64_t n) {
  const int64_t segcap = n / 2 + 2;
  const int64_t cbound = kSecondRoundSegs + n / CHUNK + 1;   // either of the first two rounds
  w.keysB.ensure(n * sizeof(uint64_t));
  w.valsB.ensure(n * sizeof(uint32_t));
  w.seg[0].ensure(segcap * sizeof(int2));
  w.seg[1].ensure(segcap * sizeof(int2));
  w.counters.ensure(C_COUNT * sizeof(int));
  w.chunks.ensure(cbound * sizeof(int4));
  w.chunkHist.ensure(cbound * 256 * sizeof(int));
  w.segChunk0.ensure(kSecondRoundSegs * sizeof(int));
}

SortFirstRound sort_first_round(PrimWork& w, int64_t n) {
  FMM_REQUIRE(n >= 2 && n < (int64_t)1 << 31, FMM_E_LIMIT, "sort: n = %lld outside [2, 2^31-1]",
              (long long)n);
  first_round_buffers(w, n);
  w.keysC.ensure(n * sizeof(uint64_t));
  w.valsC.ensure(n * sizeof(uint32_t));
  SortFirstRound f;
  f.keys = w.keysB.as<uint64_t>();
  f.vals = w.valsB.as<uint32_t>();
  f.chunks = w.chunks.as<int4>();
  f.chunk_hist = w.chunkHist.as<int>();
  f.seg_chunk0 = w.segChunk0.as<int>();
  f.seg0 = w.seg[0].as<int2>();
  f.counters = w.counters.as<int>();
  f.n = (int)n;
  return f;
}

bool sort_pairs_u64(uint64_t* keys, uint32_t* vals, int64_t n, int threshold, PrimWork& w,
                    hipStream_t s, int key_bits, bool prefix_only, bool first_round_ready) {
  if (n <= 1) return prefix_only;
  FMM_REQUIRE(n < (int64_t)1 << 31, FMM_E_LIMIT, "sort: n = %lld exceeds 2^31-1", (long long)n);
  int T = threshold < 1 ? 1 : threshold;
  if (T > kLocalSortMax) T = kLocalSortMax;
  FMM_REQUIRE(!first_round_ready || n > T, FMM_E_INTERNAL,
              "sort: a prepared first round needs more than %d keys", T);
  first_round_buffers(w, n);
  const int64_t segcap = n / 2 + 2;
  w.keysB.ensure(n * sizeof(uint64_t));
  w.valsB.ensure(n * sizeof(uint32_t));
  w.seg[0].ensure(segcap * sizeof(int2));
  w.seg[1].ensure(segcap * sizeof(int2));
  w.fin.ensure(segcap * sizeof(int2));
  w.counters.ensure(C_COUNT * sizeof(int));
  w.mbox.ensure();
  int* cnt = w.counters.as<int>();
  int64_t hc[C_COUNT];
  const int finBigOff = (int)segcap;
  if (!first_round_ready) {
    hipLaunchKernelGGL(k_sort_init, dim3(1), dim3(64), 0, s, w.seg[0].as<int2>(),
                       w.fin.as<int2>(), finBigOff, cnt, (int)n, T);
    FMM_LAUNCH_CHECK();
  }

  int nseg = 0;
  int64_t active = 0;
  int nfinal = 0, nbig = 0;
  int cur = 0;
  if (n > T) {
    nseg = 1;
    active = n;
  } else if (n <= kWaveSortMax) {
    nfinal = 1;
  } else {
    nbig = 1;
  }
  const int rounds = key_bits < 8 ? 1 : (key_bits > 64 ? 8 : (key_bits + 7) / 8);
  // final segments (small, big) found before the last round run: the small
  // ones fill fin from the front, the big ones from finBigOff downwards, so
  // the segments of the last round are fin[nfin_early, nfinal) and
  // fin[finBigOff - nbig, finBigOff - nbig_early)
  int nfin_early = nfinal, nbig_early = nbig;
  bool ran = false;
  // The round's input (kin, vin) and scatter target (kout, vout): a round whose
  // segments all continue leaves its output in place for the next round to
  // read (ping-pong); otherwise its chunks are copied back into keys / vals,
  // where every finished segment must end.  A prepared sort alternates
  // between the two spare pairs and sends finished elements to keys / vals
  // (kfin) as it scatters.
  const bool ready = first_round_ready;
  uint64_t* kin = ready ? w.keysB.as<uint64_t>() : keys;
  uint32_t* vin = ready ? w.valsB.as<uint32_t>() : vals;
  uint64_t* kout = ready ? w.keysC.as<uint64_t>() : w.keysB.as<uint64_t>();
  uint32_t* vout = ready ? w.valsC.as<uint32_t>() : w.valsB.as<uint32_t>();
  uint64_t* kfin = ready ? keys : nullptr;
  uint32_t* vfin = ready ? vals : nullptr;
  // After a prepared first round (one segment, so at most kSecondRoundSegs
  // segments and n keys left) the second round is launched at once, on grids
  // sized by those bounds; its kernels take the counts from the device.
  const bool early_second = first_round_ready && rounds > 1;
  for (int r = 0; r < rounds && nseg > 0; ++r) {
    const bool on_device_counts = early_second && r == 1;
    nfin_early = nfinal;
    nbig_early = nbig;
    ran = true;
    int64_t active_in = active;
    const int shift = 56 - 8 * r;
    const int64_t cbound =
        on_device_counts ? kSecondRoundSegs + n / CHUNK + 1 : nseg + active / CHUNK + 1;
    const int nseg_grid = on_device_counts ? kSecondRoundSegs : nseg;
    w.chunks.ensure(cbound * sizeof(int4));
    w.chunkHist.ensure(cbound * 256 * sizeof(int));
    w.segChunk0.ensure((int64_t)nseg_grid * sizeof(int));
    int2* segCur = w.seg[cur].as<int2>();
    int2* segNext = w.seg[cur ^ 1].as<int2>();
    if (!(first_round_ready && r == 0)) {
      hipLaunchKernelGGL(k_make_chunks, dim3(1), dim3(1024), 0, s, segCur,
                         on_device_counts ? -1 : nseg, w.chunks.as<int4>(),
                         w.segChunk0.as<int>(), cnt);
      FMM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_chunk_hist, dim3(cbound), dim3(SORT_BS), 0, s, kin,
                         w.chunks.as<int4>(), cnt, shift, w.chunkHist.as<int>());
      FMM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_seg_offsets, dim3(nseg_grid), dim3(SEG_BS), 0, s, segCur,
                       w.segChunk0.as<int>(), w.chunkHist.as<int>(), T, r == rounds - 1 ? 1 : 0, segNext,
                       w.fin.as<int2>(), finBigOff, cnt);
    FMM_LAUNCH_CHECK();
    const uint32_t seq = w.mbox.next();
    hipLaunchKernelGGL(k_chunk_scatter, dim3(cbound), dim3(SORT_BS), 0, s, kin, vin,
                       w.chunks.as<int4>(), cnt, shift, w.chunkHist.as<int>(), kout, vout,
                       kfin, vfin, w.mbox.dev, seq);
    FMM_LAUNCH_CHECK();
    cur ^= 1;
    if (early_second && r == 0) {
      std::swap(kin, kout);
      std::swap(vin, vout);
      continue;
    }
    w.mbox.take(C_COUNT, seq, s, hc);
    nseg = (int)hc[C_NEXT];
    active = hc[C_ACTIVE];
    nfinal = hc[C_FINAL];
    nbig = hc[C_FINAL_BIG];
    if (on_device_counts) {   // the first round's results, as this round began
      active_in = hc[C_ACTIVE_PREV];
      nfin_early = hc[C_FINAL_PREV];
      nbig_early = hc[C_BIG_PREV];
    }
    // Elements that left the active set finished in this round, in segments of
    // any length (one-element segments are on no list, but their place in
    // kout is final too): they must reach keys / vals before a later round
    // writes its output there instead.
    const bool finished_some = active < active_in;
    const bool more = nseg > 0 && r + 1 < rounds;
    if (!ready && kout != keys && (finished_some || !more)) {
      hipLaunchKernelGGL(k_chunk_copy, dim3(cbound), dim3(SORT_BS), 0, s, kout, vout,
                         w.chunks.as<int4>(), cnt, keys, vals);
      FMM_LAUNCH_CHECK();
      kin = keys;
      vin = vals;
      kout = w.keysB.as<uint64_t>();
      vout = w.valsB.as<uint32_t>();
    } else {
      std::swap(kin, kout);
      std::swap(vin, vout);
    }
  }
  FMM_REQUIRE(nfinal + nbig <= segcap, FMM_E_INTERNAL, "sort: segment list overflow");
  // prefix_only: the segments of the last round are left in radix order
  const bool skip_last = prefix_only && ran && nseg == 0;
  const int nws = skip_last ? nfin_early : nfinal;
  const int nls = skip_last ? nbig_early : nbig;
  if (nws > 0) {
    hipLaunchKernelGGL(k_wave_sort, dim3(ceil_div(nws, 4)), dim3(256), 0, s, keys, vals,
                       w.fin.as<int2>(), nws);
    FMM_LAUNCH_CHECK();
  }
  if (nls > 0) {
    hipLaunchKernelGGL(k_local_sort, dim3(nls), dim3(SORT_BS), 0, s, keys, vals,
                       w.fin.as<int2>() + (finBigOff - nls));
    FMM_LAUNCH_CHECK();
  }
  return skip_last && (nws < nfinal || nls < nbig);
}

}  // namespace fmm
