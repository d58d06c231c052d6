// Device-wide primitives used by the tree build: exclusive scan, compaction and
// a segmented MSD radix sort of (u64 key, u32 value) pairs.
#pragma once

#include "common.h"

namespace fmm {

// ---------------------------------------------------------------------------
// Wave / block helpers (wave64).
// Inclusive scan with DPP moves (no LDS traffic): Hillis-Steele within each
// row of 16 lanes (row_shr 1, 2, 4, 8), then row_bcast:15 adds the last lane
// of rows 0 and 2 to rows 1 and 3, and row_bcast:31 adds the last lane of
// row 1 to rows 2 and 3.  Lanes without a source (or outside the row mask)
// receive the "old" operand 0.
__device__ __forceinline__ int wave_inclusive_scan(int v) {
  constexpr int kRowShr = 0x110, kRowBcast15 = 0x142, kRowBcast31 = 0x143;
  v += __builtin_amdgcn_update_dpp(0, v, kRowShr + 1, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kRowShr + 2, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kRowShr + 4, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kRowShr + 8, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kRowBcast15, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, kRowBcast31, 0xc, 0xf, false);
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
  HostMailbox mbox;     // the counters the host reads between rounds
  // MSD sort
  DevBuf keysB, valsB;
  DevBuf keysC, valsC;  // second spare pair of a prepared first round
  DevBuf seg[2];        // (start, len) int2
  DevBuf fin;           // final segments (start, len)
  DevBuf chunks;        // int4 (seg, start, len, -)
  DevBuf chunkHist;     // [chunks][256]
  DevBuf segChunk0;     // first chunk of each segment
};

// out[i] = sum_{j<i} in[i]; *total_dev (optional) = sum of all in.
// in and out may alias.  partials (optional): caller-owned space for
// scan_partials_i32(n) ints; without it the block partials live in w.
void exclusive_scan_i32(const int* in, int* out, int64_t n, int* total_dev, PrimWork& w,
                        hipStream_t s, int* partials = nullptr);
int64_t scan_partials_i32(int64_t n);

// p[0..n) = v with a fill kernel (no memset node, so the call can sit in a
// captured HIP graph).
void fill_i32(int* p, int64_t n, int v, hipStream_t s);

// Sort (keys, vals) of length n by (key, val) ascending, in place.  Segments
// whose length drops to <= threshold are finished by an LDS sort; the radix
// part looks at 8 key bits per round, most significant first, over the top
// key_bits bits (the others must be equal in all keys).  Keys that agree in
// all their bits keep an unspecified order when more than kLocalSortMax of
// them coincide.
// prefix_only: leave the segments that reach the threshold in the last radix
// round run in their radix (input) order instead of sorting them on the
// remaining bits, and return true if any was left so.  The quadtree uses
// this: those segments lie at the deepest radix depth D, every cell at depth
// D holds at most Ncutoff points, so no cell of the tree below depth D exists
// and none needs its points grouped; segments that ended earlier are sorted.
constexpr int kLocalSortMax = 2048;
// first_round_ready: the caller's kernel wrote the first round's records and
// its keys and values into the sort's spare buffers (sort_first_round
// below), so the sort starts at that round's segment offsets; it requires
// n > threshold.  Such a sort never reads keys / vals: every round scatters
// the elements of finished segments straight to their places there and the
// others to the other spare buffer, so no round copies anything back.  The
// second round is launched on the first round's counts as the device holds
// them, without waiting for the host to read them.
bool sort_pairs_u64(uint64_t* keys, uint32_t* vals, int64_t n, int threshold, PrimWork& w,
                    hipStream_t s, int key_bits = 64, bool prefix_only = false,
                    bool first_round_ready = false);

// Layout of the sort's device counters and chunks.
namespace sort_detail {
// C_SEGS: segments of the current round; C_*_PREV: C_ACTIVE, C_FINAL and
// C_FINAL_BIG as they stood when the round began
enum { C_NEXT = 0, C_FINAL = 1, C_CHUNKS = 2, C_ACTIVE = 3, C_FINAL_BIG = 4, C_SEGS = 5,
       C_ACTIVE_PREV = 6, C_FINAL_PREV = 7, C_BIG_PREV = 8, C_COUNT = 9 };
}
constexpr int kSortChunk = 4096;   // keys per chunk of a radix round

// The first radix round's records, written by a key-producing kernel with one
// block of at least 256 threads per chunk of kSortChunk keys (chunk b = keys
// [b kSortChunk, min(n, (b + 1) kSortChunk))): each block passes the counts of
// its keys' top 8 bits (bits 56..63) to sort_first_round_block.  This saves
// the sort three launches and a pass over the keys.
struct SortFirstRound {
  uint64_t* keys = nullptr;   // where the key kernel writes the keys and values
  uint32_t* vals = nullptr;
  int4* chunks = nullptr;
  int* chunk_hist = nullptr;
  int* seg_chunk0 = nullptr;
  int2* seg0 = nullptr;
  int* counters = nullptr;
  int n = 0;
};
// Grows the sort's work buffers for n keys (as sort_pairs_u64 would) and
// returns where the records go.
SortFirstRound sort_first_round(PrimWork& w, int64_t n);

// Every thread of block b (at least 256 threads) calls this once hist (256
// counts in LDS) is complete.
__device__ inline void sort_first_round_block(const SortFirstRound& f, int b, const int* hist) {
  const int t = threadIdx.x;
  if (t < 256) f.chunk_hist[(int64_t)b * 256 + t] = hist[t];
  if (t == 0) {
    const int start = b * kSortChunk;
    f.chunks[b] = make_int4(0, start, min(kSortChunk, f.n - start), 0);
    if (b == 0) {
      for (int i = 0; i < sort_detail::C_COUNT; ++i) f.counters[i] = 0;
      f.counters[sort_detail::C_CHUNKS] = (int)ceil_div(f.n, kSortChunk);
      f.counters[sort_detail::C_SEGS] = 1;
      f.seg0[0] = make_int2(0, f.n);
      f.seg_chunk0[0] = 0;
    }
  }
}

}  // namespace fmm
