// Device-wide primitives used by the tree build: exclusive scan, compaction and
// a segmented MSD radix sort of (u64 key, u32 value) pairs.
#pragma once

#include "common.h"

namespace fmm {

// ---------------------------------------------------------------------------
// Wave / block helpers (wave64).
__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exclusive scan over a block of BS threads; lds must hold BS/64 ints.
template <int BS>
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  int inc = wave_inclusive_scan(v);
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    int w = (lane < BS / 64) ? lds[lane] : 0;
    int winc = wave_inclusive_scan(w);
    if (lane < BS / 64) lds[lane] = winc - w;
    if (lane == BS / 64 - 1 && total) *total = winc;
  }
  __syncthreads();
  int res = inc - v + lds[wid];
  __syncthreads();
  return res;
}

// ---------------------------------------------------------------------------
// Host-side work area for the primitives (grown on demand, reused).
struct PrimWork {
  DevBuf blocksums;     // scan partials
  DevBuf counters;      // small device counters
  HostBuf hcounters;    // pinned mirror of the counters
  // MSD sort
  DevBuf keysB, valsB;
  DevBuf seg[2];        // (start, len) int2
  DevBuf fin;           // final segments (start, len)
  DevBuf chunks;        // int4 (seg, start, len, -)
  DevBuf chunkHist;     // [chunks][256]
  DevBuf segChunk0;     // first chunk of each segment
};

// out[i] = sum_{j<i} in[i]; *total_dev (optional) = sum of all in.
// in and out may alias.
void exclusive_scan_i32(const int* in, int* out, int64_t n, int* total_dev, PrimWork& w,
                        hipStream_t s);

// Sort (keys, vals) of length n by (key, val) ascending, in place.  Segments
// whose length drops to <= threshold are finished by an LDS bitonic sort; the
// radix part looks at 8 key bits per round, most significant first.  Keys
// that agree in all 64 bits keep an unspecified order when more than
// kLocalSortMax of them coincide.
constexpr int kLocalSortMax = 2048;
void sort_pairs_u64(uint64_t* keys, uint32_t* vals, int64_t n, int threshold, PrimWork& w,
                    hipStream_t s);

}  // namespace fmm
