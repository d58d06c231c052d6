// Adaptive quadtree on the device (see tree.hip).
#pragma once

#include "prims.h"

namespace fmm {

constexpr int kMaxDepth = 31;    // cells exist down to depth 31; only depth <= 30 splits
constexpr int kMaxLevels = kMaxDepth + 1;
constexpr int kNumM2L = 40;      // the 7x7 ring offsets minus the 3x3 neighbourhood

// Node flag bits (per node).
enum NodeFlag : int {
  NF_LEAF = 1,       // reference `leaf`: a real node without real children (X-list nodes included)
  NF_FAKE = 2,       // reference `fake_leaf`: child of an X-list node
  NF_XLIST = 4,      // reference `Xlist`
  NF_UP = 8,         // reference `compute_upwards`
  NF_DOWN = 16,      // reference `compute_downwards`
  NF_CHILDREN = 32,  // has children (real or fake): children index >= 0
  NF_SOURCE = 64,    // holds at least one source
};

struct TreeLevel {
  int64_t node_off = 0;
  int64_t n_nodes = 0;
  double width = 0.0;
};

struct Tree {
  bool built = false;
  bool allow_full = true; // complete trees take the closed-form build (tree.hip)
  bool complete = false;  // the last build was such a tree: every leaf at depth L - 1, and
                          // leaf (depth L - 1, code c) is evaluation box c
  int64_t build_id = 0;   // incremented by every build
  int full_D = 0;         // depth of the last build's complete tree (0: it was not one)
  int64_t n = 0;           // number of sources
  int ncut = 0;
  int L = 0;               // number of levels
  double xmin = 0, ymin = 0, w0 = 0;
  std::vector<TreeLevel> lev;
  int64_t n_nodes = 0, n_rows = 0, n_eval = 0, n_up = 0;

  // sorted points (tree order)
  DevBuf keys;     // u64 Morton key (31 levels, left aligned)
  DevBuf perm;     // u32: tree position -> input index
  DevBuf sx, sy;   // f64 sorted coordinates

  // nodes, level-major, Morton order inside a level
  DevBuf code;     // u64
  DevBuf nlevel;   // i32
  DevBuf bot, top; // i32 point range
  DevBuf parent;   // i32 global
  DevBuf child;    // i32 global index of first child or -1
  DevBuf flags;    // i32 NodeFlag bits
  DevBuf coll;     // i32[9] global colleagues or -1; slot (dx+1)*3+(dy+1)
  DevBuf crange;   // int2[9]: colleague q's source range [bot, top); (0, 0) when absent
  DevBuf xmid, ymid;
  DevBuf row;      // i32 multipole row (levels >= 2, sibling groups holding sources) or -1
  DevBuf rownode;  // i32[n_rows] node of each row
  DevBuf rowinfo;  // int2[n_rows]: row of the node's first child (or -1), node flags
  std::vector<int64_t> row_off;  // first row of each level (size L+1)
  DevBuf m2l;      // i32[40] multipole rows of the M2L partners or -1 (real nodes, level >= 2)
  DevBuf evalid;   // i32 index among evaluation boxes (NF_LEAF) or -1
  DevBuf evalnode; // i32[n_eval]
  DevBuf upnodes;  // i32[n_up] nodes with NF_UP
  DevBuf levoff;   // i32[L+1] node offsets on the device

  // work areas
  PrimWork pw;
  DevBuf w_flag, w_rank, w_a, w_b, w_c, w_cnt;
  DevBuf w_gate;   // i32: the split counts confirm the guessed complete tree
  DevBuf w_bad;    // i32 per key block: it holds a non-finite coordinate
  HostBuf h_cnt;
  HostBuf h_up;    // pinned staging for small host->device uploads
  std::vector<int64_t> s_off;   // host offsets of the split-cell lists per depth
  DevBuf s_all, sref_all;       // u64 split-cell codes per depth (concatenated)
  DevBuf bal_bm, bal_pc;        // balance: cell bitmaps per depth, popcount scan
  DevBuf nidx;                  // dense node index of the non-full shallow levels
  std::vector<int64_t> nidx_off; // level d's table in nidx (per cell code), or -1
  DevBuf xy;                    // interleaved input coordinates (input order)
  DevBuf bal_k, bal_v, bal_f;   // balance below the bitmap depths: candidates
  std::vector<DevBuf> bal_lists;
  DevBuf p_code;                // u64 codes of nodes with children, per level (concatenated)
  DevBuf p_kind;                // i32 0 = split, 1 = X-list
  std::vector<int64_t> p_off;
};

// Build the tree of n sources (device pointers x, y) in the square root box
// [xmin, xmin + w0] x [ymin, ymin + w0] with leaf capacity ncut.  Synchronises
// the stream a few times to size the per-level arrays.
void tree_build(Tree& t, const double* x, const double* y, int64_t n, double xmin, double ymin,
                double w0, int ncut, hipStream_t s);

// Upload the constant tables of the tree kernels to the current device (once
// per device, before the first tree_build on it).
void tree_init_device();

// Locate targets: out_node[i] = global index of the leaf (NF_LEAF) reached by
// the reference descent (x > xmid -> +2, y > ymid -> +1).
void tree_locate(const Tree& t, const double* tx, const double* ty, int64_t nt, int* out_node,
                 hipStream_t s);

// Slot order: (ii, jj) over 0..6 x 0..6 row-major, skipping |ii-3|<=1 && |jj-3|<=1;
// the offset of slot (ii, jj) is (ii-3, jj-3).
__host__ __device__ inline void m2l_slot_offset(int slot, int* dx, int* dy) {
  int k = 0;
  for (int ii = 0; ii < 7; ++ii)
    for (int jj = 0; jj < 7; ++jj) {
      if (ii >= 2 && ii <= 4 && jj >= 2 && jj <= 4) continue;
      if (k == slot) {
        *dx = ii - 3;
        *dy = jj - 3;
        return;
      }
      ++k;
    }
  *dx = 0;
  *dy = 0;
}

}  // namespace fmm
