// KIFMM passes on the device and the C-ABI of libflexmm_amd.
//
//   upward    S2M (check potential of leaf sources on the large circle) and
//             M2M (children's equivalent densities to the parent's check
//             surface), then the pseudo-inverse apply q = V diag(1/s) U^T b;
//             fmm.py:73-124, helpers.py:159-173, 240-248.
//   downward  L2L (parent's local density to the child's check surface), the
//             M2L translations of the 40-offset ring, then the same
//             pseudo-inverse; fmm.py:125-166.
//   evaluate  target location, targets grouped by leaf, L2T + P2P fused in one
//             pass; fmm.py:182-199, helpers.py:180-235.
// Dense contractions run on FP64 MFMA (v_mfma_f64_16x16x4f64) with the
// operator block staged in LDS; expansions are rows of PD doubles (PD = 32 or
// 64, p padded with zeros).
#include "bessel.h"
#include "tree.h"

#include <cstdarg>
#include <cstring>
#include <exception>
#include <mutex>
#include <type_traits>

#include "../../include/flexmm_amd.h"

namespace fmm {

// ---------------------------------------------------------------------------
static thread_local std::string g_err;
void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}
void clear_error() { g_err.clear(); }
const char* last_error() { return g_err.c_str(); }

void LogTableHost::fill(double2* out) {
  for (int j = 0; j < kLogTableSize; ++j) {
    long double c = 1.0L + ((long double)j + 0.5L) / kLogTableSize;
    double T = (double)(2.0L / c);
    long double lg = -logl((long double)T);
    out[j] = make_double2(T, (double)lg);
  }
}

// ---------------------------------------------------------------------------
// Operator layout per level: OP_NBLK blocks of [PD][PD] doubles (row-major,
// zero padded), then the large-circle coordinates lx[PD], ly[PD].
enum {
  OP_PINV1 = 0,   // diag(1/s) U^T
  OP_PINV2 = 1,   // V
  OP_M2M = 2,     // 4 blocks: M2MC[:, c p:(c+1) p]
  OP_L2L = 6,     // 4 blocks: M2MC[:, c p:(c+1) p]^T
  OP_M2L = 10,    // 40 blocks in slot order
  OP_NBLK = 50,
};
// ... followed, for the mixed-precision M2L, by the 40 row-mean vectors c_s
// (FP64 [40][PD]) and the FP32 residuals R_s = M2L_s - c_s 1^T ([40][PD][PD]);
// then the SVD-compressed M2L (helpers.py:17-45): 64 slot ranks (as doubles;
// r > 0: compressed of rank r, 0: null slot, -1: dense) and 40 [PD][PD]
// blocks, rows 0 .. PD/2-1 = M1 = diag(s) V^T (r x p) and rows PD/2 .. PD-1 =
// M2 = U (p x r) two rows per block row: block[PD/2 + i/2][(i%2) PD/2 + j] =
// M2[i][j].
__host__ __device__ inline int64_t ops_off_circles(int PD) { return (int64_t)OP_NBLK * PD * PD; }
__host__ __device__ inline int64_t ops_off_means(int PD) { return ops_off_circles(PD) + 2 * PD; }
__host__ __device__ inline int64_t ops_off_r32(int PD) { return ops_off_means(PD) + 40 * PD; }
__host__ __device__ inline int64_t ops_off_svd_rank(int PD) {
  return ops_off_r32(PD) + 20 * (int64_t)PD * PD;
}
__host__ __device__ inline int64_t ops_off_svd(int PD) { return ops_off_svd_rank(PD) + 64; }
inline int64_t ops_per_level(int PD) { return ops_off_svd(PD) + 40 * (int64_t)PD * PD; }

// Position of entry k of a multipole row in the FP32 M2L's copy of it
// (k_m2l_waves_f32): lane group kq = (k mod 8) / 2 reads KS = PD / 4
// consecutive floats, in the order of its MFMA steps 2 m + h, k = 8 m + 2 kq + h.
template <int PD>
__host__ __device__ __forceinline__ int m32_pos(int k) {
  return ((k & 7) >> 1) * (PD / 4) + 2 * (k >> 3) + (k & 1);
}

// Stores this lane's entries 16 rb + 4 v + kq (v = 0 .. 3) of a multipole row
// into its FP32 copy row32, with the lane's partner l ^ 16 (lane group
// kq ^ 1) as one 16-byte store each: the even lane group writes the v = 0, 2
// entries of both, the odd one the v = 1, 3 entries (positions m32_pos).
// Every lane of the wave takes part in the exchange; store says whether this
// lane writes.
template <int PD>
__device__ __forceinline__ void store_m32(float* __restrict__ row32, int rb, const double (&v)[4],
                                          bool store) {
  const int kq = (threadIdx.x & 63) >> 4;
  const bool odd = kq & 1;
  const float f0 = (float)v[0], f1 = (float)v[1], f2 = (float)v[2], f3 = (float)v[3];
  const float rx = __shfl_xor(odd ? f0 : f1, 16, 64);
  const float ry = __shfl_xor(odd ? f2 : f3, 16, 64);
  const float4 out = odd ? make_float4(rx, f1, ry, f3) : make_float4(f0, rx, f2, ry);
  if (store)
    *reinterpret_cast<float4*>(row32 + (PD / 4) * ((kq >> 1) + (odd ? 2 : 0)) + 4 * rb) = out;
}

// Chebyshev (black-box) expansions, bbfmm.py:95-231: order pc <= kMaxCheb,
// p = pc^2 coefficients per box, index i pc + j for T_i(x) T_j(y), with
// (x, y) = (point - box centre) * (2 / w).  The level's 2/w sits in the first
// entry of the operator block's circle area (no circles are used), and the
// pseudo-inverse blocks are the identity (fmm.py:58-61 of flexmm2d).
constexpr int kMaxCheb = 8;

constexpr double kInv4Pi = 0.07957747154594767;   // 1/(4 pi)
constexpr double kInv2Pi = 0.15915494309189535;   // 1/(2 pi)
constexpr int LDA = 68;   // LDS row stride of an A block: 68 = 4 (mod 32) spreads an MFMA A
                          // fragment (16 rows x 4 columns) over distinct LDS banks

typedef double d4 __attribute__((ext_vector_type(4)));

struct SlotTable {
  int ox[kNumM2L], oy[kNumM2L];
};
__constant__ SlotTable c_slots;

struct Plan {
  int kernel_id = KER_LAPLACE;
  double kparam = 0.0;
  int p = 0, PD = 0, precision = PREC_FP64;
  int cheb = 0;        // Chebyshev order of a black-box plan (p = cheb^2); 0: KIFMM
  int m2l_mode = FMM2D_M2L_DENSE;   // FP64 M2L: every slot dense, or compressed as the reference
  Tree tree;
  DevBuf ops;
  int ops_levels = 0;
  std::vector<char> ops_set;
  DevBuf logtab;
  DevBuf tau_s, src, check, mult, le, ple, qs;
  DevBuf mult32;       // FP32 M2L: the multipoles rounded, in operand order (m32_pos)
  DevBuf srcf, tsum;   // mixed precision: leaf-relative FP32 sources, per-leaf charge sums
  DevBuf dip;          // dipole moments of the sources in tree order (double layer)
  DevBuf sdip;         // Stokes double layer: (n.sigma, a, b, 0) per source in tree order
  bool have_dip = false;
  int nd = 1;          // densities of the current expansions
  DevBuf m2lperm;      // M2L column order of an adaptive tree
  int64_t perm_tree = -1;   // tree build the column order belongs to
  bool have_srcf = false;
  bool have_exp = false;
  bool up_done = false;
  // targets
  DevBuf tnode, tcnt, toff, tcur, tperm, tsx, tsy, icnt, ioff, items, ntot;
  PrimWork pw;
  HostBuf hbuf;
  int device = 0;
  // end of the last call's work on its stream (waited for by plan destroy)
  hipEvent_t done_ev = nullptr;
  bool done_recorded = false;
  // the last evaluation's per-box target counts (work_counts)
  const int* last_tcnt = nullptr;
  // optional phase timing with HIP events on the caller's stream: one
  // (start, end) pair per run of a phase, kept until phase_times reads them
  bool profile = false;
  std::vector<hipEvent_t> ev[FMM2D_NUM_PHASES];   // start, end, start, end ...
  int runs[FMM2D_NUM_PHASES] = {};
  bool open_run[FMM2D_NUM_PHASES] = {};
  ~Plan() {
    for (auto& v : ev)
      for (auto e : v)
        if (e) (void)hipEventDestroy(e);
    if (done_ev) (void)hipEventDestroy(done_ev);
  }
  void mark(int phase, bool end, hipStream_t s) {
    if (!profile) return;
    auto& v = ev[phase];
    const size_t k = 2 * (size_t)runs[phase] + (end ? 1 : 0);
    if (end && !open_run[phase]) return;
    while (v.size() <= k) {
      hipEvent_t e = nullptr;
      FMM_HIP(hipEventCreate(&e));
      v.push_back(e);
    }
    FMM_HIP(hipEventRecord(v[k], s));
    open_run[phase] = !end;
    if (end) ++runs[phase];
  }
};

// Binds a C-ABI call on a plan to its stream: buffers that grow during the
// call wait for that stream (not the whole device) before freeing their old
// block, and the end of the call is recorded so that destroying the plan
// waits for exactly the work queued on it.
struct PlanCall {
  Plan& P;
  hipStream_t s, prev;
  bool record;
  PlanCall(Plan& plan, void* stream, bool record_end = true)
      : P(plan), s((hipStream_t)stream), prev(g_op_stream), record(record_end) {
    g_op_stream = s;
  }
  ~PlanCall() {
    g_op_stream = prev;
    if (record && P.done_ev && hipEventRecord(P.done_ev, s) == hipSuccess) P.done_recorded = true;
  }
  PlanCall(const PlanCall&) = delete;
  PlanCall& operator=(const PlanCall&) = delete;
};

namespace {

__device__ __forceinline__ void load_logtab(double2* tab, const double2* __restrict__ g) {
  for (int i = threadIdx.x; i < kLogTableSize; i += blockDim.x) tab[i] = g[i];
}

// ---------------------------------------------------------------------------
// Sources in tree order packed as (x, y, tau, 0): one 32-byte record per
// source, read by wave-uniform (scalar) loads in the S2M and P2P loops.
__global__ void k_pack_sources(const double* __restrict__ tau, const uint32_t* __restrict__ perm,
                               const double* __restrict__ sx, const double* __restrict__ sy,
                               int64_t n, double* __restrict__ tau_s,
                               double4* __restrict__ packed) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double t = tau[perm[i]];
  tau_s[i] = t;
  packed[i] = make_double4(sx[i], sy[i], t, 0.0);
}

// 1/x to a few ulp: hardware reciprocal and two Newton steps (x > 0, normal)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Single + double layer sources (examples/pure_kifmm_laplace_dipoles.py:32-40):
// tau holds the single-layer charges [n] then the dipole strengths [n], dvec
// the dipole directions, x [n] then y [n].  Packed in tree order as
// (x, y, charge, 0) and the dipole moment strength * direction.
__global__ void k_pack_dipole(const double* __restrict__ tau, const double* __restrict__ dvec,
                              const uint32_t* __restrict__ perm, const double* __restrict__ sx,
                              const double* __restrict__ sy, int64_t n,
                              double* __restrict__ tau_s, double4* __restrict__ packed,
                              double2* __restrict__ dip) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  const double t = tau[j], st = tau[n + j];
  tau_s[i] = t;
  packed[i] = make_double4(sx[i], sy[i], t, 0.0);
  dip[i] = make_double2(st * dvec[j], st * dvec[n + j]);
}

// Accumulate w * log(|t - s|^2) = w * (e ln2 + lt) as (ae += w e, at += w lt).
struct LogAcc {
  double ae = 0.0, at = 0.0;
  __device__ __forceinline__ void add(double w, double d2, const double2* __restrict__ tab) {
    double e, lt;
    log_split(d2, tab, e, lt);
    ae = __builtin_fma(w, e, ae);
    at = __builtin_fma(w, lt, at);
  }
  __device__ __forceinline__ double value() const { return __builtin_fma(ae, kLn2, at); }
};

// K0(x) and k K1(x)/r at x = k r from d2 = r^2.  On the ascending-series
// branch (x <= 2, i.e. q = k^2 d2 / 4 <= 1) neither a square root nor the
// library log is needed: q is (k^2/4) d2, ln(x/2) = ln(d2)/2 + ln(k/2) with
// ln(d2) from the table log, and
//   k K1(x)/r = 1/d2 + (k^2/4) (2 ln(x/2) I1' - S1'),
// I1' = sum q^j/(j!(j+1)!), S1' the psi sum of bessel_k01 (x I1' / 2 = I1).
// The x > 2 branch is bessel_k0_large / bessel_k01_large at x = k sqrt(d2).
struct HelmConsts {
  double k, k2q, lnk2;   // k, k^2/4, ln(k/2)
};
__device__ __forceinline__ HelmConsts helm_consts(double k) {
  return HelmConsts{k, 0.25 * k * k, log(0.5 * k)};
}
__device__ __forceinline__ double helm_lnx2(double d2, const double2* __restrict__ tab,
                                            const HelmConsts& h) {
  double e, t;
  log_split(d2, tab, e, t);
  return __builtin_fma(0.5, __builtin_fma(e, kLn2, t), h.lnk2);
}
__device__ __forceinline__ double helm_k0(double d2, const double2* __restrict__ tab,
                                          const HelmConsts& h) {
  const double q = h.k2q * d2;
  if (q <= 1.0) {
    const BesselSeries& c = bessel_series();
    constexpr int n = kBesselSeriesTerms - 1;
    double i0 = c.i0[n], s = c.s0[n];
#pragma unroll
    for (int j = n - 1; j >= 0; --j) {
      i0 = __builtin_fma(i0, q, c.i0[j]);
      s = __builtin_fma(s, q, c.s0[j]);
    }
    return __builtin_fma(-(helm_lnx2(d2, tab, h) + kEulerGamma), i0, s);
  }
  return bessel_k0_large(h.k * __builtin_sqrt(d2));
}
// K0 and k K1(x)/r; id2 = 1/d2
__device__ __forceinline__ void helm_k01(double d2, double id2, const double2* __restrict__ tab,
                                         const HelmConsts& h, double& k0, double& gk1) {
  const double q = h.k2q * d2;
  if (q <= 1.0) {
    // K0's two sums take their coefficients from constant memory (scalar
    // operands), K1's are compile-time values in vector registers: fewer
    // register moves per pair than all in registers (gradient evaluation of
    // 2.2e8 pairs 1.33 -> 1.22 ms), and all in constant memory exceeds the
    // scalar registers
    const BesselSeries& c = kBesselSeries;
    const BesselSeries& cm = bessel_series();
    constexpr int n = kBesselSeriesTerms - 1;
    double i0 = cm.i0[n], s0 = cm.s0[n], i1 = c.i1[n], s1 = c.s1[n];
#pragma unroll
    for (int j = n - 1; j >= 0; --j) {
      i0 = __builtin_fma(i0, q, cm.i0[j]);
      s0 = __builtin_fma(s0, q, cm.s0[j]);
      i1 = __builtin_fma(i1, q, c.i1[j]);
      s1 = __builtin_fma(s1, q, c.s1[j]);
    }
    const double lg = helm_lnx2(d2, tab, h);
    k0 = __builtin_fma(-(lg + kEulerGamma), i0, s0);
    gk1 = __builtin_fma(h.k2q, __builtin_fma(2.0 * lg, i1, -s1), id2);
    return;
  }
  const double r = __builtin_sqrt(d2);
  double k1;
  bessel_k01_large(h.k * r, k0, k1);
  gk1 = h.k * k1 / r;
}


// Potential accumulators of the supported kernels: sum_j w_j K(d_j^2).
//   Laplace           K = -log(d^2)/(4 pi)   (examples/pure_kifmm_laplace.py:24-29)
//   modified Helmholtz K = K0(k d)/(2 pi)     (pure_kifmm_modified_helmholtz_gradient.py:36-42)
template <int KER>
struct PotAcc;
template <>
struct PotAcc<KER_LAPLACE> {
  LogAcc a;
  __device__ __forceinline__ void add(double w, double d2, const double2* __restrict__ tab,
                                      const HelmConsts&) {
    a.add(w, d2, tab);
  }
  __device__ __forceinline__ double result() const { return -kInv4Pi * a.value(); }
};

template <>
struct PotAcc<KER_MODHELM> {
  double s = 0.0;
  __device__ __forceinline__ void add(double w, double d2, const double2* __restrict__ tab,
                                      const HelmConsts& h) {
    s = __builtin_fma(w, helm_k0(d2, tab, h), s);
  }
  __device__ __forceinline__ double result() const { return kInv2Pi * s; }
};

// S2M: one wave per leaf (kS2MWaves leaves per block share the log table),
// lane k = check point k on the large circle (helpers.py:159-173; the source
// is taken relative to the box centre).
constexpr int kS2MWaves = 4;
// DIP: sources also carry a dipole moment q (the double layer of the Laplace
// kernel, (q.d)/d^2/(2 pi) with d = check point - source).
template <int PD, int KER, bool DIP>
__global__ __launch_bounds__(64 * kS2MWaves) void k_s2m(int nup, const int* __restrict__ upnodes,
                                            const int* __restrict__ nlevel,
                                            const int* __restrict__ bot, const int* __restrict__ top,
                                            const int* __restrict__ row,
                                            const double* __restrict__ xmid,
                                            const double* __restrict__ ymid,
                                            const double4* __restrict__ src,
                                            const double* __restrict__ ops, int64_t opsz, int p,
                                            const double2* __restrict__ logtab, double kp,
                                            const double2* __restrict__ dip,
                                            double* __restrict__ check) {
  __shared__ double2 tab[kLogTableSize];
  load_logtab(tab, logtab);
  __syncthreads();
  const int gi = blockIdx.x * kS2MWaves + (int)(threadIdx.x >> 6);
  if (gi >= nup) return;
  const int g = upnodes[gi];
  const int d = nlevel[g];
  const int k = threadIdx.x & 63;
  const int r = row[g];
  if (r < 0 || d < 2) return;   // wave-uniform
  const double* lx = ops + d * opsz + (int64_t)OP_NBLK * PD * PD;
  const double* ly = lx + PD;
  // check point k in absolute coordinates (box centre + large-circle offset)
  // lanes k >= p repeat check point 0 (their values are not stored), so that
  // they take the same branch of the kernel evaluation as the active lanes
  const double ax = lx[k < p ? k : 0] + xmid[g];
  const double ay = ly[k < p ? k : 0] + ymid[g];
  const int b = __builtin_amdgcn_readfirstlane(bot[g]);
  const int t = __builtin_amdgcn_readfirstlane(top[g]);
  PotAcc<KER> acc;
  const HelmConsts hc = helm_consts(kp);
  double dsum = 0.0;   // double layer
  int j = b;
  for (; j + 4 <= t; j += 4) {
    double4 s4[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) s4[u] = src[j + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double dx = ax - s4[u].x;
      const double dy = ay - s4[u].y;
      const double d2 = __builtin_fma(dy, dy, dx * dx);
      acc.add(s4[u].z, d2, tab, hc);
      if constexpr (DIP) {
        const double2 q = dip[j + u];
        dsum = __builtin_fma(__builtin_fma(q.x, dx, q.y * dy), rcp_nr(d2), dsum);
      }
    }
  }
  for (; j < t; ++j) {
    const double4 sj = src[j];
    const double dx = ax - sj.x;
    const double dy = ay - sj.y;
    const double d2 = __builtin_fma(dy, dy, dx * dx);
    acc.add(sj.z, d2, tab, hc);
    if constexpr (DIP) {
      const double2 q = dip[j];
      dsum = __builtin_fma(__builtin_fma(q.x, dx, q.y * dy), rcp_nr(d2), dsum);
    }
  }
  if (k < PD) check[(int64_t)r * PD + k] = k < p ? acc.result() + kInv2Pi * dsum : 0.0;
}

// Chebyshev moments (source2pMs, bbfmm.py:173-184; anterpolate2 :11-24):
// m[i pc + j] = sum_s T_i(x_s) T_j(y_s) tau_s, (x, y) = (s - centre) * (2 / w).
// One 64-lane block per box: for each chunk of 64 sources, lane l writes
// T_i(x) tau and T_j(y) of source l to LDS (three-term recurrence), then lane
// i pc + j sums the products over the chunk.
template <int PD>
__global__ __launch_bounds__(64) void k_s2m_cheb(const int* __restrict__ upnodes,
                                                 const int* __restrict__ nlevel,
                                                 const int* __restrict__ bot,
                                                 const int* __restrict__ top,
                                                 const int* __restrict__ row,
                                                 const double* __restrict__ xmid,
                                                 const double* __restrict__ ymid,
                                                 const double4* __restrict__ src,
                                                 const double* __restrict__ ops, int64_t opsz,
                                                 int pc, double* __restrict__ check) {
  __shared__ double tx[64][kMaxCheb + 1];
  __shared__ double ty[64][kMaxCheb + 1];
  const int g = upnodes[blockIdx.x];
  const int d = nlevel[g];
  const int r = row[g];
  if (r < 0 || d < 2) return;   // block-uniform
  const int k = threadIdx.x;
  const double wadj = ops[d * opsz + ops_off_circles(PD)];
  const double cx = xmid[g], cy = ymid[g];
  const int b = bot[g], e = top[g];
  const int ci = k / pc, cj = k - ci * pc;
  const bool live = k < pc * pc;
  double acc = 0.0;
  for (int c0 = b; c0 < e; c0 += 64) {
    const int sidx = c0 + k;
    if (sidx < e) {
      const double4 sj = src[sidx];
      const double x = (sj.x - cx) * wadj, y = (sj.y - cy) * wadj;
      double a0 = sj.z, a1 = x * sj.z, b0 = 1.0, b1 = y;
      tx[k][0] = a0;
      tx[k][1] = a1;
      ty[k][0] = b0;
      ty[k][1] = b1;
#pragma unroll
      for (int m = 2; m < kMaxCheb; ++m) {
        const double a2 = __builtin_fma(2.0 * x, a1, -a0);
        const double b2 = __builtin_fma(2.0 * y, b1, -b0);
        tx[k][m] = a2;
        ty[k][m] = b2;
        a0 = a1;
        a1 = a2;
        b0 = b1;
        b1 = b2;
      }
    }
    __syncthreads();
    const int n = min(64, e - c0);
    if (live)
      for (int q = 0; q < n; ++q) acc = __builtin_fma(tx[q][ci], ty[q][cj], acc);
    __syncthreads();
  }
  if (k < PD) check[(int64_t)r * PD + k] = live ? acc : 0.0;
}

// ---------------------------------------------------------------------------
// MFMA helpers.  For v_mfma_f64_16x16x4f64 lane l holds A[l%16][l/16],
// B[l/16][l%16] and D[4v + l/16][l%16] (v = 0..3).  A PD-row result therefore
// sits in lane l as rows 4m + l/16 (m = 4 rb + v), which is exactly the
// B-operand layout of the next product: results chain without shuffles.
//
// A block size: 4 waves, 16 columns (boxes) per wave, every wave computes all
// PD rows of its 16 columns.  A sequence of "phases" acc += A_i * B_i runs
// through a double-buffered LDS stage for A_i: while phase i multiplies, the
// 16-byte global loads of A_{i+1} and the B gathers of phase i+1 are in
// flight, and one barrier per phase separates the buffers.
template <int PD>
struct Mfma {
  static constexpr int RB = PD / 16;     // 16-row blocks
  static constexpr int KS = PD / 4;      // k-steps of 4
  static constexpr int LDA = PD + 4;     // = 4 (mod 32): the 16 rows x 4 columns of an A
                                         // fragment fall into distinct LDS banks
};

typedef double dv2 __attribute__((ext_vector_type(2)));

// Register stage of one PD x PD operator block, loaded by all NT threads of
// the block with 16-byte loads.
template <int PD, int NT = 256>
struct AStage {
  static constexpr int N2 = PD * PD / 2 / NT;
  dv2 v[N2];
  __device__ __forceinline__ void load(const double* __restrict__ G) {
    const dv2* G2 = reinterpret_cast<const dv2*>(G);
#pragma unroll
    for (int i = 0; i < N2; ++i) v[i] = G2[threadIdx.x + i * NT];
  }
  __device__ __forceinline__ void store(double* __restrict__ A) const {
#pragma unroll
    for (int i = 0; i < N2; ++i) {
      const int idx = 2 * (threadIdx.x + i * NT);
      const int r = idx / PD, c = idx - r * PD;
      *reinterpret_cast<dv2*>(A + r * Mfma<PD>::LDA + c) = v[i];
    }
  }
};

// acc[rb] += A[rows of row block rb0 + rb] * B, rb < NRB
template <int PD, int NRB = PD / 16>
__device__ __forceinline__ void mfma_block(d4 (&acc)[NRB], const double* __restrict__ A,
                                           const double (&b)[PD / 4], int rb0 = 0) {
  constexpr int KS = Mfma<PD>::KS, LDA = Mfma<PD>::LDA;
  const int l = threadIdx.x & 63;
  const double* a0 = A + (rb0 * 16 + (l & 15)) * LDA + (l >> 4);
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
      acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[rb * 16 * LDA + ks * 4], b[ks], acc[rb], 0,
                                                     0, 0);
}

template <int PD>
__device__ __forceinline__ void zero_acc(d4 (&acc)[PD / 16]) {
#pragma unroll
  for (int rb = 0; rb < PD / 16; ++rb) acc[rb] = d4{0.0, 0.0, 0.0, 0.0};
}

// acc (D layout) from / to a row of PD doubles
template <int PD>
__device__ __forceinline__ void load_acc(d4 (&acc)[PD / 16], const double* __restrict__ rowp) {
  const int kq = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int rb = 0; rb < PD / 16; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[rb][v] = rowp ? rowp[16 * rb + 4 * v + kq] : 0.0;
}
template <int PD>
__device__ __forceinline__ void store_acc(const d4 (&acc)[PD / 16], double* __restrict__ rowp) {
  const int kq = (threadIdx.x & 63) >> 4;
#pragma unroll
  for (int rb = 0; rb < PD / 16; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) rowp[16 * rb + 4 * v + kq] = acc[rb][v];
}

// One phase: operator block A (PD x PD, global) and either a per-lane source
// row (B column = PD doubles, or null for zeros) or the accumulator itself
// (chain: acc <- A * acc).
struct Phase {
  const double* A;
  const double* brow;
  bool chain;
};

// `next()` returns the phases in order (it may keep state).
template <int PD, class Next>
__device__ __forceinline__ void run_phases(int nph, d4 (&acc)[PD / 16], double* __restrict__ buf0,
                                           double* __restrict__ buf1, Next& next) {
  constexpr int KS = Mfma<PD>::KS;
  if (nph <= 0) return;
  const int kq = (threadIdx.x & 63) >> 4;
  AStage<PD> st;
  double bn[KS];
  Phase ps = next();
  st.load(ps.A);
  bool chain_n = ps.chain;
  bool any_n = ps.chain || __ballot(ps.brow != nullptr) != 0;
  if (!ps.chain) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) bn[ks] = ps.brow ? ps.brow[4 * ks + kq] : 0.0;
  }
  for (int i = 0; i < nph; ++i) {
    double* A = (i & 1) ? buf1 : buf0;
    st.store(A);
    const bool chain = chain_n, any = any_n;
    double bc[KS];
    if (chain) {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) bc[ks] = acc[ks >> 2][ks & 3];
      zero_acc<PD>(acc);
    } else {
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) bc[ks] = bn[ks];
    }
    __syncthreads();
    if (i + 1 < nph) {
      ps = next();
      st.load(ps.A);
      chain_n = ps.chain;
      any_n = ps.chain || __ballot(ps.brow != nullptr) != 0;
      if (!ps.chain) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) bn[ks] = ps.brow ? ps.brow[4 * ks + kq] : 0.0;
      }
    }
    if (any) mfma_block<PD>(acc, A, bc);
  }
}

// ---------------------------------------------------------------------------
// Upward, one level: rows [row0, row0 + nrows); 64 rows (boxes) per block.
// check = S2M result (leaves) or M2M of the four children (parents), then
// multipole = pinv2 * (pinv1 * check)  (fmm.py:98-122).
template <int PD>
__global__ __launch_bounds__(256) void k_mult_level(int64_t row0, int nrows, int with_m2m,
                                                   const int2* __restrict__ rowinfo,
                                                   const double* __restrict__ check,
                                                   const double* __restrict__ opsL,
                                                   double* __restrict__ mult,
                                                   double* __restrict__ qs, int64_t dstride,
                                                   float* __restrict__ m32) {
  __shared__ __attribute__((aligned(16))) double buf[2][PD * Mfma<PD>::LDA];
  // density blockIdx.y: its rows sit dstride rows further in check, mult and qs
  check += (int64_t)blockIdx.y * dstride * PD;
  mult += (int64_t)blockIdx.y * dstride * PD;
  qs += (int64_t)blockIdx.y * dstride;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rl = blockIdx.x * 64 + w * 16 + (l & 15);
  const bool active = rl < nrows;
  const int64_t r = row0 + rl;
  const int2 ri = active ? rowinfo[r] : make_int2(-1, 0);
  const int f = ri.y;
  const int crow = ri.x;
  d4 acc[PD / 16];
  load_acc<PD>(acc, (active && (f & NF_UP)) ? check + r * PD : nullptr);
  const int nm = with_m2m ? 4 : 0;
  int i = 0;
  auto next = [&]() -> Phase {
    const int k = i++;
    if (k < nm)
      return Phase{opsL + (int64_t)(OP_M2M + k) * PD * PD,
                   crow >= 0 ? mult + (int64_t)(crow + k) * PD : nullptr, false};
    return Phase{opsL + (int64_t)(OP_PINV1 + (k - nm)) * PD * PD, nullptr, true};
  };
  run_phases<PD>(nm + 2, acc, buf[0], buf[1], next);
  // total equivalent charge of the box (used by the mixed-precision M2L)
  double part = 0.0;
#pragma unroll
  for (int rb = 0; rb < PD / 16; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) part += acc[rb][v];
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  if (active) {
    store_acc<PD>(acc, mult + r * PD);
    if ((l >> 4) == 0) qs[r] = part;
  }
  if (m32) {
    float* row32 = m32 + ((int64_t)blockIdx.y * dstride + r) * PD;
#pragma unroll
    for (int rb = 0; rb < PD / 16; ++rb) {
      const double v[4] = {acc[rb][0], acc[rb][1], acc[rb][2], acc[rb][3]};
      store_m32<PD>(row32, rb, v, active);
    }
  }
}

// ---------------------------------------------------------------------------
// M2L of all levels in one launch (the translations only read multipoles).
// Blocks of a level are tiled by child position c so that all columns of a
// block share the 27 offsets that can occur for c; block (tile, c) handles the
// c-th child of sibling groups [64 tile, 64 tile + 64).  Offsets absent from
// every column of a block are skipped.
struct LevelTable {
  int nlev;
  int d[kMaxLevels];
  int g0[kMaxLevels];
  int ngroups[kMaxLevels];
  int boff[kMaxLevels + 1];
  // optional column order (adaptive trees): sibling group of column qq of
  // child position c at level li is perm[poff[li] + c * ngroups[li] + qq]
  const int* perm;
  int poff[kMaxLevels];
  // > 1: each tile runs as `split` neighbouring blocks, each taking an equal
  // share of its phases and adding into a zeroed PLE with FP64 atomics; used
  // when few tiles would leave CUs idle.  Two addends onto 0 give the same sum
  // in either order; more parts make the sum's rounding depend on the order.
  int split;
};

struct ValidSlots {
  int n[4];
  int slot[4][kNumM2L];
};
__constant__ ValidSlots c_valid;
// zeros: the multipole row of an absent M2L source (FP32 M2L), so that its
// operand loads need no condition
__device__ double c_zero_row[64] = {};   // never written

// M2L block: 256 threads, 64 target boxes of one level and one child position
// c (4 waves x 16 columns, every wave all PD rows), phases = the slots that
// any of the 64 boxes uses.  Per phase the operator block is staged in LDS
// (double buffered) and each lane gathers its box's source multipole with
// 16-byte loads; the k index of an MFMA step is permuted so that one 16-byte
// load feeds two steps: step 2m + h of lane group kq multiplies k = 8m + 2kq
// + h (A is read with the same permutation, which leaves the product
// unchanged).  Loads of phase i + 1 are issued before the barrier of phase i.
// With a row split RS, a block holds 4 RS waves: wave w takes columns
// 16 (w mod 4) .. + 15 and the row blocks of half (or quarter) w / 4, which
// puts RS times more waves on the SIMDs for the same tiles (a 1M-point tree
// gives about 1.4 waves per SIMD with RS = 1).
constexpr int kM2LThreads = 256;   // threads per block and row split
constexpr int kM2LCols = 64;
constexpr int kM2LRowSplit = 1;   // blocks of 4 waves, each wave all PD rows of 16 boxes
constexpr int kM2LSplitF64 = 2;   // FP64 M2L: blocks per tile when tiles are few
// Dense FP64 M2L in narrow tiles (k_m2l_waves_f64) up to this padded
// dimension: at PD 32 it beats the LDS-staged tiles (100K points, p 32:
// 0.034 -> 0.028 ms); at PD 64 each wave re-reading 32 KB per slot from L2
// costs more than the shared stage (1M points, p 60: 0.151 -> 0.189 ms).
constexpr int kM2LWavesF64MaxPD = 32;

struct M2LSetup {
  int li, c, col, kq, g, nph, rh;
  bool active;
};

// Block prologue shared by the FP64 and FP32 M2L: decode the block, collect
// the 64 boxes' interaction lists in LDS and the list of used slots.
__device__ __forceinline__ M2LSetup m2l_setup(const LevelTable& lt, const int* __restrict__ flags,
                                              const int* __restrict__ m2l,
                                              int (&srcs)[27][kM2LCols], int (&pslot)[28],
                                              unsigned long long& smask) {
  M2LSetup S;
  const int bid0 = xcd_block(blockIdx.x, gridDim.x);
  const int part = bid0 % lt.split;   // the parts of a tile are neighbours
  const int bid = bid0 / lt.split;
  int li = 0;
  while (li + 1 < lt.nlev && bid >= lt.boff[li + 1]) ++li;
  const int lb = bid - lt.boff[li];
  S.li = li;
  S.c = lb & 3;
  const int tile = lb >> 2;
  const int ngroups = lt.ngroups[li];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  S.col = (w & 3) * 16 + (l & 15);
  S.kq = l >> 4;
  S.rh = w >> 2;
  const int qq = tile * kM2LCols + S.col;
  const int q = (lt.perm && qq < ngroups) ? lt.perm[lt.poff[li] + S.c * ngroups + qq] : qq;
  S.g = lt.g0[li] + 4 * q + S.c;
  S.active = q < ngroups && !(flags[q < ngroups ? S.g : lt.g0[li]] & NF_FAKE);
  const int* lst = m2l + (int64_t)S.g * kNumM2L;
  const int nv = c_valid.n[S.c];
  if (threadIdx.x == 0) smask = 0ull;
  __syncthreads();
  unsigned long long wm = 0ull;
  for (int i = 0; i < nv; ++i) {
    const int src = S.active ? lst[c_valid.slot[S.c][i]] : -1;
    if (S.kq == 0) srcs[i][S.col] = src;
    if (__ballot(src >= 0)) wm |= 1ull << i;
  }
  if (l == 0 && wm) atomicOr(&smask, wm);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = smask;
    int k = 0;
    while (m) {
      pslot[k++] = __builtin_ctzll(m);
      m &= m - 1ull;
    }
    if (lt.split > 1) {   // this block's share of the phases
      const int lo = k * part / lt.split, hi = k * (part + 1) / lt.split;
      for (int i = lo; i < hi; ++i) pslot[i - lo] = pslot[i];
      k = hi - lo;
    }
    pslot[27] = k;
  }
  __syncthreads();
  S.nph = pslot[27];
  return S;
}

// Column order of the M2L blocks on adaptive trees.  There the boxes of a
// level differ in which of their offsets carry a partner, and 64 neighbouring
// boxes together use most offsets although each uses few, so most MFMA
// columns would multiply absent partners.  Within chunks of kM2LOrderChunk
// sibling groups (in Morton order) of one level and child position, the
// boxes are sorted by their set of used offsets (ties by position); boxes
// without any (and fake nodes) go last.  One block per chunk: bitonic sort
// of (40-bit offset mask << 11 | index) in LDS.
constexpr int kM2LOrderChunk = 2048;
struct M2LOrderTable {
  int nlev;
  int g0[kMaxLevels];
  int ngroups[kMaxLevels];
  int nch[kMaxLevels];     // chunks per child position
  int boff[kMaxLevels + 1];
  int poff[kMaxLevels];
};
constexpr int kM2LOrderThreads = 1024;
__global__ __launch_bounds__(kM2LOrderThreads) void k_m2l_order(M2LOrderTable ot, const int* __restrict__ flags,
                                                   const int* __restrict__ m2l,
                                                   int* __restrict__ perm) {
  __shared__ uint64_t key[kM2LOrderChunk];
  const int b = blockIdx.x;
  int li = 0;
  while (li + 1 < ot.nlev && b >= ot.boff[li + 1]) ++li;
  const int lb = b - ot.boff[li];
  const int c = lb / ot.nch[li];
  const int chunk = lb % ot.nch[li];
  const int ng = ot.ngroups[li];
  const int q0 = chunk * kM2LOrderChunk;
  const int cnt = min(kM2LOrderChunk, ng - q0);
  static_assert(kNumM2L == 40 && kM2LOrderChunk <= 2048, "key layout: 40 + 11 bits");
#pragma unroll 2
  for (int i = threadIdx.x; i < kM2LOrderChunk; i += kM2LOrderThreads) {
    uint64_t k = ~0ull;
    if (i < cnt) {
      const int g = ot.g0[li] + 4 * (q0 + i) + c;
      const int4* lst = reinterpret_cast<const int4*>(m2l + (int64_t)g * kNumM2L);
      int4 e[kNumM2L / 4];
#pragma unroll
      for (int j = 0; j < kNumM2L / 4; ++j) e[j] = lst[j];
      uint64_t mask = 0;
#pragma unroll
      for (int j = 0; j < kNumM2L / 4; ++j)
        mask |= (uint64_t)((e[j].x >= 0) | (e[j].y >= 0) << 1 | (e[j].z >= 0) << 2 |
                           (e[j].w >= 0) << 3)
                << (4 * j);
      if (flags[g] & NF_FAKE) mask = 0;
      k = ((mask ? mask : (1ull << 52) - 1) << 11) | (uint64_t)i;
    }
    key[i] = k;
  }
  __syncthreads();
  for (int size = 2; size <= kM2LOrderChunk; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < kM2LOrderChunk / 2; t += kM2LOrderThreads) {
        // stride is a power of two: lo = 2 stride (t / stride) + t % stride
        const int lo = ((t & ~(stride - 1)) << 1) | (t & (stride - 1));
        const int hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t a = key[lo], e = key[hi];
        if ((a > e) == up) {
          key[lo] = e;
          key[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < cnt; i += kM2LOrderThreads)
    perm[ot.poff[li] + c * ng + q0 + i] = q0 + (int)(key[i] & 2047u);
}

// FP64 M2L.  use_svd: slots whose rank table entry is r > 0 are applied in
// the reference's compressed form, M2 (M1 q) (helpers.py:37-45), in two MFMA
// stages: T = M1 q (r rows, ceil(r/16) row blocks, K = PD), then
// acc += M2 T with K = 4 ceil(r/4); the first stage's result registers (lane
// (col, kq) holds rows 16 rb + 4 v + kq) are exactly the B operand of the
// second stage's k-steps j = 4 rb + v in natural k order, so no data moves
// between lanes.  Null slots (r = 0) add nothing, as in the reference.
template <int PD, int RS>
__global__ __launch_bounds__(kM2LThreads * RS) void k_m2l_all(LevelTable lt,
                                                        const int* __restrict__ flags,
                                                        const int* __restrict__ m2l,
                                                        const double* __restrict__ mult,
                                                        const double* __restrict__ ops,
                                                        int64_t opsz, double* __restrict__ ple,
                                                        int64_t rstride, int64_t nstride,
                                                        int use_svd) {
  // density blockIdx.y: multipole rows rstride rows further, local rows nstride nodes further
  mult += (int64_t)blockIdx.y * rstride * PD;
  ple += (int64_t)blockIdx.y * nstride * PD;
  constexpr int KS = Mfma<PD>::KS, RB = Mfma<PD>::RB / RS, LDA = Mfma<PD>::LDA;
  constexpr int H = PD / 2, R1 = H / 16, K2 = H / 4;
  __shared__ __attribute__((aligned(16))) double buf[2][PD * LDA];
  __shared__ int srcs[27][kM2LCols];
  __shared__ int pslot[28];
  __shared__ int srank[kNumM2L];
  __shared__ unsigned long long smask;
  const M2LSetup S = m2l_setup(lt, flags, m2l, srcs, pslot, smask);
  const int l = threadIdx.x & 63;
  const int rb0 = S.rh * RB;   // this wave's first row block
  const double* opsL = ops + (int64_t)lt.d[S.li] * opsz;
  if (threadIdx.x < kNumM2L)
    srank[threadIdx.x] = use_svd ? (int)opsL[ops_off_svd_rank(PD) + threadIdx.x] : -1;
  __syncthreads();
  d4 acc[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) acc[rb] = d4{0.0, 0.0, 0.0, 0.0};
  if (S.nph > 0) {
    AStage<PD, kM2LThreads * RS> st;
    double bn[KS], bc[KS];
    auto fetch = [&](int k) {
      const int i = pslot[k];
      const int slot = c_valid.slot[S.c][i];
      st.load(srank[slot] > 0 ? opsL + ops_off_svd(PD) + (int64_t)slot * PD * PD
                              : opsL + (int64_t)(OP_M2L + slot) * PD * PD);
      const int src = srcs[i][S.col];
      const double* brow = src >= 0 ? mult + (int64_t)src * PD + 2 * S.kq : nullptr;
#pragma unroll
      for (int m = 0; m < KS / 2; ++m) {
        const dv2 v = brow ? *reinterpret_cast<const dv2*>(brow + 8 * m) : dv2{0.0, 0.0};
        bn[2 * m] = v.x;
        bn[2 * m + 1] = v.y;
      }
    };
    fetch(0);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) bc[ks] = bn[ks];
    for (int i = 0; i < S.nph; ++i) {
      double* A = (i & 1) ? buf[1] : buf[0];
      st.store(A);
      const int r = srank[c_valid.slot[S.c][pslot[i]]];   // uniform in the block
      if (i + 1 < S.nph) fetch(i + 1);
      __syncthreads();
      const double* a0 = A + (l & 15) * LDA + 2 * S.kq;
      const double* ar = a0 + rb0 * 16 * LDA;   // this wave's rows
      if (r < 0) {
#pragma unroll
        for (int m = 0; m < KS / 2; ++m)
#pragma unroll
          for (int rb = 0; rb < RB; ++rb) {
            const dv2 a = *reinterpret_cast<const dv2*>(ar + rb * 16 * LDA + 8 * m);
            acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bc[2 * m], acc[rb], 0, 0, 0);
            acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bc[2 * m + 1], acc[rb], 0, 0, 0);
          }
      } else if (r > 0) {
        // stage 1: T = M1 q, row blocks rb1 < ceil(r / 16)
        const int r1b = (r + 15) >> 4;
        d4 t1[R1];
#pragma unroll
        for (int rb1 = 0; rb1 < R1; ++rb1) t1[rb1] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < KS / 2; ++m)
#pragma unroll
          for (int rb1 = 0; rb1 < R1; ++rb1)
            if (rb1 < r1b) {
              const dv2 a = *reinterpret_cast<const dv2*>(a0 + rb1 * 16 * LDA + 8 * m);
              t1[rb1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, bc[2 * m], t1[rb1], 0, 0, 0);
              t1[rb1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bc[2 * m + 1], t1[rb1], 0, 0, 0);
            }
        // stage 2: acc += M2 T, k-steps j < ceil(r / 4), k = 4 j + kq
        const int nk2 = (r + 3) >> 2;
        const double* a2 = A + H * LDA + S.kq;
#pragma unroll
        for (int j = 0; j < K2; ++j)
          if (j < nk2) {
            const double b = t1[j >> 2][j & 3];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
              const int row = 16 * (rb0 + rb) + (l & 15);
              const double a = a2[(row >> 1) * LDA + (row & 1) * H + 4 * j];
              acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[rb], 0, 0, 0);
            }
          }
      }
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) bc[ks] = bn[ks];
    }
  }
  if (S.active) {
    double* rowp = ple + (int64_t)S.g * PD + 16 * rb0;
    if (lt.split > 1) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) unsafeAtomicAdd(rowp + 16 * rb + 4 * v + S.kq, acc[rb][v]);
    } else {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) rowp[16 * rb + 4 * v + S.kq] = acc[rb][v];
    }
  }
}

// Mixed-precision M2L: PLE = sum_s [ R_s q_src (FP32 MFMA, FP32 accumulation
//                                                within the slot)
//                               + c_s * qs_src  (FP64) ], slots summed in FP64.
// FRAME (Stokeslet fields): the two fields D_i whose definition holds the
// position X relative to the root centre are translated in the target box's
// frame instead, D_i^(c) = D_i + c_i D_2 with c the target centre (formed in
// FP64 from the source's D_i and D_2 rows before rounding), so that the FP32
// operands carry no c_i D_2 part; the downward pass returns to the root
// frame in FP64 (k_local_level), where that part cancels exactly against
// X_i(t) D_2 at the targets.
// v_mfma_f32_16x16x4f32: lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[4(l/16) + v][l%16].
typedef float f4 __attribute__((ext_vector_type(4)));

// FP32 M2L in narrow tiles: block = 16 target boxes of one level and child
// position and NW waves; the slots that any of the 16 boxes uses are dealt
// round-robin to the waves, and each wave loads the residual block of its
// slot straight from L2 into registers, with no LDS stage and no barrier per
// slot.  The residuals are stored in the MFMA operand order (host packing,
// residual_fragment_order): one 16-byte load per lane feeds four MFMA steps
// and a wave's load reads 1 KB in one piece.  Step 2 m + h of lane group kq
// multiplies k = 8 m + 2 kq + h.  The source multipoles come from their FP32
// copy in the same order (m32_pos: a lane's 16 operands are 64 contiguous
// bytes), except in the Stokeslet frame fields, which form them in FP64 from
// two rows.  Each wave streams through its slots' residual pieces with a
// row block of them in flight and loads the next slot's multipoles at the
// start of the current one, so its loads overlap the MFMAs before them
// (measured at 1M points: 0.094 -> 0.077 ms).  The constant mode c_s * (sum of the
// source's equivalent density) is an FP64 MFMA product over the wave's slots
// at the end (rows permuted so that its output layout is the FP32 one).  The
// waves' FP64 sums are added in a fixed order through LDS.  Against 64-box
// tiles sharing an LDS stage this gives four times the waves (a 1M-point
// tree: ~5 per SIMD instead of ~1.4).
constexpr int kM2LWCols = 16;
constexpr int kM2LWWaves = 4;
constexpr int kM2LF32Waves = 4;

template <int PD, bool FRAME, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_m2l_waves_f32(
    LevelTable lt, const int* __restrict__ flags, const int* __restrict__ m2l,
    const double* __restrict__ mult, const float* __restrict__ m32,
    const double* __restrict__ qs, const double* __restrict__ ops, int64_t opsz,
    double* __restrict__ ple, int64_t rstride, int64_t nstride, const double* __restrict__ xmid,
    const double* __restrict__ ymid, double c0x, double c0y) {
  constexpr int RB = PD / 16, KS = PD / 4, NA = KS / 4;
  typedef float fv4 __attribute__((ext_vector_type(4)));
  // FRAME: fields 0 and 1 take the source's field-2 row (D_2) too
  const bool frame = FRAME && blockIdx.y < 2;
  const int64_t to2 = (int64_t)(2 - (int)blockIdx.y) * rstride;   // rows to field 2
  mult += (int64_t)blockIdx.y * rstride * PD;
  qs += (int64_t)blockIdx.y * rstride;
  ple += (int64_t)blockIdx.y * nstride * PD;
  __shared__ int srcs[27][16];
  __shared__ double qsl[27][16];   // source charge sums (0: no source)
  __shared__ int pslot[28];
  __shared__ int pslotv[27];   // the M2L slot (offset) of each used slot index
  __shared__ unsigned int smask;
  __shared__ double red[NW - 1][PD][17];
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  int li = 0;
  while (li + 1 < lt.nlev && bid >= lt.boff[li + 1]) ++li;
  const int lb = bid - lt.boff[li];
  const int c = lb & 3;
  const int tile = lb >> 2;
  const int ngroups = lt.ngroups[li];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = l & 15, kq = l >> 4;
  const int qq = tile * 16 + col;
  const int q = (lt.perm && qq < ngroups) ? lt.perm[lt.poff[li] + c * ngroups + qq] : qq;
  const int g = lt.g0[li] + 4 * q + c;
  const bool active = q < ngroups && !(flags[q < ngroups ? g : lt.g0[li]] & NF_FAKE);
  // target centre coordinate blockIdx.y in the root frame (frame fields)
  const double ct =
      (frame && active) ? (blockIdx.y == 0 ? xmid[g] - c0x : ymid[g] - c0y) : 0.0;
  // interaction lists: the 4 NW lane groups (w, kq) read slot indices
  // 4 w + kq + 4 NW j < 32 of their column
  constexpr int LG = 4 * NW, NJ = 32 / LG;
  static_assert(NJ * LG == 32, "lane groups cover the 32 slot indices");
  const int nv = c_valid.n[c];
  if (threadIdx.x == 0) smask = 0u;
  __syncthreads();
  unsigned int used = 0u;
  {
    const int* lst = m2l + (int64_t)g * kNumM2L;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int i = 4 * w + kq + LG * j;
      const int src = (active && i < nv) ? lst[c_valid.slot[c][i]] : -1;
      if (i < nv) {
        srcs[i][col] = src;
        qsl[i][col] = src < 0 ? 0.0 : frame ? __builtin_fma(ct, qs[to2 + src], qs[src]) : qs[src];
      }
      const unsigned long long b = __ballot(src >= 0);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if ((b >> (16 * k)) & 0xffffull) used |= 1u << (4 * w + k + LG * j);
    }
  }
  if (l == 0 && used) atomicOr(&smask, used);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int m = smask;
    int k = 0;
    while (m) {
      const int i = __builtin_ctz(m);
      pslot[k] = i;
      pslotv[k++] = c_valid.slot[c][i];
      m &= m - 1u;
    }
    pslot[27] = k;
  }
  __syncthreads();
  const int nph = pslot[27];
  const double* opsL = ops + (int64_t)lt.d[li] * opsz;
  const double* means = opsL + ops_off_means(PD);
  // residual piece j of row block rb of a slot: afr[(slot RB + rb) NA + j) 64]
  const fv4* afr = reinterpret_cast<const fv4*>(opsL + ops_off_r32(PD)) + l;
  d4 acc64[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) acc64[rb] = d4{0.0, 0.0, 0.0, 0.0};
  // this wave's slots: pslot[w + NW j], j < nmine
  const int nmine = w < nph ? (nph - 1 - w) / NW + 1 : 0;
  if (nmine > 0) {
    if constexpr (FRAME) {
      // Stokeslet fields: per slot, entries 8 m + 2 kq + {0, 1} of the
      // source rows formed in FP64 (the frame fields from two rows), then
      // each row block's residual pieces loaded before its steps
      for (int jj = 0; jj < nmine; ++jj) {
        const int k = w + NW * jj;
        const int slot = __builtin_amdgcn_readfirstlane(pslotv[k]);
        const int src = srcs[pslot[k]][col];
        const double* brow = src >= 0 ? mult + (int64_t)src * PD + 2 * kq : c_zero_row;
        fv4 bu[NA];
#pragma unroll
        for (int m = 0; m < KS / 2; ++m) {
          dv2 v = *reinterpret_cast<const dv2*>(brow + 8 * m);
          if (frame && src >= 0) {
            const dv2 v2 = *reinterpret_cast<const dv2*>(brow + to2 * PD + 8 * m);
            v.x = __builtin_fma(ct, v2.x, v.x);
            v.y = __builtin_fma(ct, v2.y, v.y);
          }
          bu[m >> 1][2 * (m & 1)] = (float)v.x;
          bu[m >> 1][2 * (m & 1) + 1] = (float)v.y;
        }
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          fv4 a[NA];
#pragma unroll
          for (int j = 0; j < NA; ++j) a[j] = afr[((int64_t)slot * RB * NA + rb * NA + j) * 64];
#pragma unroll
          for (int j = 0; j < NA; ++j) {
            f4 s0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][0], bu[j][0], f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            f4 s1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][1], bu[j][1], f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            s0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][2], bu[j][2], s0, 0, 0, 0);
            s1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][3], bu[j][3], s1, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) acc64[rb][v] += (double)(s0[v] + s1[v]);
          }
        }
      }
    } else {
      // operator pieces pc = rb NA + j of a slot (16 bytes per lane, four
      // MFMA steps each), NA of them in flight: consuming a piece issues the
      // load of the piece NA further on, across row block and slot
      // boundaries.  The multipole operands alternate between two buffers: a
      // slot's first NA pieces also load the next slot's operands into the
      // other one.  Absent sources read a zero row.
      constexpr int NP = RB * NA;
      const float* m32y = m32 + (int64_t)blockIdx.y * rstride * PD;
      auto bptr = [&](int src) -> const fv4* {
        return reinterpret_cast<const fv4*>(
            (src >= 0 ? m32y + (int64_t)src * PD : reinterpret_cast<const float*>(c_zero_row)) +
            kq * KS);
      };
      auto apiece = [&](int sl, int pc) -> const fv4* {
        return afr + ((int64_t)sl * NP + pc) * 64;
      };
      int slot = __builtin_amdgcn_readfirstlane(pslotv[w]);
      fv4 a[NA], b0[NA], b1[NA];
      // first loads in the order the loop consumes them, so that its waits
      // for the first pieces hold for every slot
#pragma unroll
      for (int h = 0; h < NA; ++h) {
        a[h] = *apiece(slot, h);
        __builtin_amdgcn_sched_barrier(0);
      }
      {
        const fv4* bp = bptr(srcs[pslot[w]][col]);
#pragma unroll
        for (int j = 0; j < NA; ++j) {
          b0[j] = bp[j];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      auto one_slot = [&](fv4 (&bu)[NA], fv4 (&bl)[NA], int jj) {
        // the next slot of this wave (the current one again after the last)
        const int kn = jj + 1 < nmine ? w + NW * (jj + 1) : w + NW * jj;
        const int slot_n = __builtin_amdgcn_readfirstlane(pslotv[kn]);
        const fv4* bpn = bptr(srcs[pslot[kn]][col]);
#pragma unroll
        for (int pc = 0; pc < NP; ++pc) {
          const int rb = pc / NA, j = pc % NA;
          // an FP32 run of four steps, added to the FP64 sums; two
          // accumulators keep consecutive MFMAs independent
          f4 s0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][0], bu[j][0], f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
          f4 s1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][1], bu[j][1], f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
          s0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][2], bu[j][2], s0, 0, 0, 0);
          s1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j][3], bu[j][3], s1, 0, 0, 0);
          // the piece just used is reloaded into the same registers
          a[j] = pc + NA < NP ? *apiece(slot, pc + NA) : *apiece(slot_n, pc + NA - NP);
          if (pc < NA) bl[pc] = bpn[pc];
#pragma unroll
          for (int v = 0; v < 4; ++v) acc64[rb][v] += (double)(s0[v] + s1[v]);
          // in program order: the scheduler would otherwise move each load
          // down to its use
          __builtin_amdgcn_sched_barrier(0);
        }
        slot = slot_n;
      };
      for (int jj = 0; jj < nmine; jj += 2) {
        one_slot(b0, b1, jj);
        if (jj + 1 < nmine) one_slot(b1, b0, jj + 1);
        // keeps the next slot's loads here, ahead of the loop's back edge
        // (otherwise they are merged with the first slot's into one load at
        // the top of the loop, which then waits for them at once)
        __asm__ volatile("" ::: "memory");
      }
    }
    // constant mode: acc64[rb] += sum over this wave's slots of
    // means[slot][row] * qsl[slot][col], four slots per FP64 MFMA step.  Lane
    // (kq, col) supplies A[col][kq] = means[slot_kq][16 rb + pi(col)] with
    // pi(4 a + b) = 4 b + a, so that D[4 v + kq][col] (the FP64 layout) is row
    // 16 rb + 4 kq + v, the row acc64[rb][v] holds.
    const int prow = 4 * (col & 3) + (col >> 2);
    for (int k0 = w; k0 < nph; k0 += 4 * NW) {
      const int kk = k0 + NW * kq;
      const bool ok = kk < nph;
      const int ks = ok ? kk : k0;
      const double bq = ok ? qsl[pslot[ks]][col] : 0.0;
      const double* mrow = means + (int64_t)pslotv[ks] * PD + prow;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
        acc64[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(mrow[16 * rb], bq, acc64[rb], 0, 0, 0);
    }
  }
  // waves 1 .. NW-1 hand their sums to wave 0, which adds them in wave order
  if (w > 0) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[w - 1][16 * rb + 4 * kq + v][col] = acc64[rb][v];
  }
  __syncthreads();
  if (w == 0 && active) {
    double* rowp = ple + (int64_t)g * PD;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double sum = acc64[rb][v];
#pragma unroll
        for (int ww = 0; ww < NW - 1; ++ww) sum += red[ww][16 * rb + 4 * kq + v][col];
        rowp[16 * rb + 4 * kq + v] = sum;
      }
  }
}

// FP64 M2L (dense form) in narrow tiles, as k_m2l_waves_f32: 16 boxes per
// block, their slots dealt to 4 waves, sums added in wave order at the end.
// The M2L blocks stay row-major: step 2 m + h of lane group kq multiplies
// k = 8 m + 2 kq + h, so a lane's A operands for two steps are one 16-byte
// load and the 4 lane groups of a row read 64 consecutive bytes (as for the
// multipoles).  Row blocks run one after another, each loading its operands
// just before its 16 steps, which keeps the kernel at 4 waves per SIMD.
// D layout of v_mfma_f64_16x16x4f64: lane l holds D[4 v + l/16][l%16].
template <int PD>
__global__ __launch_bounds__(64 * kM2LWWaves) __attribute__((amdgpu_waves_per_eu(4)))
void k_m2l_waves_f64(LevelTable lt, const int* __restrict__ flags, const int* __restrict__ m2l,
                     const double* __restrict__ mult, const double* __restrict__ ops,
                     int64_t opsz, double* __restrict__ ple, int64_t rstride, int64_t nstride) {
  constexpr int RB = PD / 16, KS = PD / 4, NW = kM2LWWaves;
  mult += (int64_t)blockIdx.y * rstride * PD;
  ple += (int64_t)blockIdx.y * nstride * PD;
  __shared__ int srcs[27][kM2LWCols];
  __shared__ int pslot[28];
  __shared__ unsigned int smask;
  __shared__ double red[NW - 1][PD][kM2LWCols + 1];
  const int bid = xcd_block(blockIdx.x, gridDim.x);
  int li = 0;
  while (li + 1 < lt.nlev && bid >= lt.boff[li + 1]) ++li;
  const int lb = bid - lt.boff[li];
  const int c = lb & 3;
  const int tile = lb >> 2;
  const int ngroups = lt.ngroups[li];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = l & 15, kq = l >> 4;
  const int qq = tile * kM2LWCols + col;
  const int q = (lt.perm && qq < ngroups) ? lt.perm[lt.poff[li] + c * ngroups + qq] : qq;
  const int g = lt.g0[li] + 4 * q + c;
  const bool active = q < ngroups && !(flags[q < ngroups ? g : lt.g0[li]] & NF_FAKE);
  const int nv = c_valid.n[c];
  const int* lst = m2l + (int64_t)g * kNumM2L;
  if (threadIdx.x == 0) smask = 0u;
  __syncthreads();
  unsigned int used = 0u;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int i = 4 * w + kq + 16 * j;
    const int src = (active && i < nv) ? lst[c_valid.slot[c][i]] : -1;
    if (i < nv) srcs[i][col] = src;
    const unsigned long long b = __ballot(src >= 0);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if ((b >> (16 * k)) & 0xffffull) used |= 1u << (4 * w + k + 16 * j);
  }
  if (l == 0 && used) atomicOr(&smask, used);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int m = smask;
    int k = 0;
    while (m) {
      pslot[k++] = __builtin_ctz(m);
      m &= m - 1u;
    }
    pslot[27] = k;
  }
  __syncthreads();
  const int nph = pslot[27];
  const double* opsL = ops + (int64_t)lt.d[li] * opsz;
  d4 acc[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) acc[rb] = d4{0.0, 0.0, 0.0, 0.0};
  for (int k = w; k < nph; k += NW) {
    const int i = pslot[k];
    const int slot = c_valid.slot[c][i];
    const int src = srcs[i][col];
    dv2 b[KS / 2];   // source row entries 8 m + 2 kq + {0, 1}
    if (src >= 0) {
      const double* brow = mult + (int64_t)src * PD + 2 * kq;
#pragma unroll
      for (int m = 0; m < KS / 2; ++m) b[m] = *reinterpret_cast<const dv2*>(brow + 8 * m);
    } else {
#pragma unroll
      for (int m = 0; m < KS / 2; ++m) b[m] = dv2{0.0, 0.0};
    }
    const double* arow =
        opsL + (int64_t)(OP_M2L + slot) * PD * PD + (int64_t)col * PD + 2 * kq;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      dv2 a[KS / 2];
#pragma unroll
      for (int m = 0; m < KS / 2; ++m)
        a[m] = *reinterpret_cast<const dv2*>(arow + (int64_t)rb * 16 * PD + 8 * m);
#pragma unroll
      for (int m = 0; m < KS / 2; ++m) {
        acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m].x, b[m].x, acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m].y, b[m].y, acc[rb], 0, 0, 0);
      }
    }
  }
  if (w > 0) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[w - 1][16 * rb + 4 * v + kq][col] = acc[rb][v];
  }
  __syncthreads();
  if (w == 0 && active) {
    double* rowp = ple + (int64_t)g * PD;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double sum = acc[rb][v];
#pragma unroll
        for (int ww = 0; ww < NW - 1; ++ww) sum += red[ww][16 * rb + 4 * v + kq][col];
        rowp[16 * rb + 4 * v + kq] = sum;
      }
  }
}

// ---------------------------------------------------------------------------
// Downward, one level: PLE += L2L(parent LE) (parent width's L2LC,
// fmm.py:142,162), LE = pinv2 * (pinv1 * PLE)  (fmm.py:156).  Tiled by child
// position c like the M2L.
// Return of the Stokeslet fields D_0, D_1 from the target-box frame of the
// FP32 M2L to the root frame: PLE_i -= c_i PLE_2, c the box centre relative to
// the root centre (xmid null: no change).
struct FrameFix {
  const double* xmid;
  const double* ymid;
  double c0x, c0y;
  // centre coordinate `field` (0: x, 1: y) of node g in the root frame
  __device__ __forceinline__ double centre(int g, int field) const {
    return field == 0 ? xmid[g] - c0x : ymid[g] - c0y;
  }
};

template <int PD>
__global__ __launch_bounds__(256) void k_local_level(int g0, int ngroups, int do_l2l,
                                                    const int* __restrict__ flags,
                                                    const int* __restrict__ parent,
                                                    const double* __restrict__ ple,
                                                    const double* __restrict__ opsL,
                                                    const double* __restrict__ opsParent,
                                                    double* __restrict__ le, int64_t nstride,
                                                    FrameFix fx) {
  const double* ple2 = ple + 2 * nstride * PD;   // field 2 (FrameFix)
  ple += (int64_t)blockIdx.y * nstride * PD;
  le += (int64_t)blockIdx.y * nstride * PD;
  __shared__ __attribute__((aligned(16))) double buf[2][PD * Mfma<PD>::LDA];
  const int c = blockIdx.x & 3, tile = blockIdx.x >> 2;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = tile * 64 + w * 16 + (l & 15);
  const int g = g0 + 4 * q + c;
  const bool active = q < ngroups && !(flags[q < ngroups ? g : g0] & NF_FAKE);
  d4 acc[PD / 16];
  load_acc<PD>(acc, active ? ple + (int64_t)g * PD : nullptr);
  if (fx.xmid && blockIdx.y < 2 && active) {
    d4 acc2[PD / 16];
    load_acc<PD>(acc2, ple2 + (int64_t)g * PD);
    const double ct = fx.centre(g, blockIdx.y);
#pragma unroll
    for (int rb = 0; rb < PD / 16; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[rb][v] = __builtin_fma(-ct, acc2[rb][v], acc[rb][v]);
  }
  const int pg = (active && do_l2l) ? parent[g] : -1;
  const int nl = do_l2l ? 1 : 0;
  int i = 0;
  auto next = [&]() -> Phase {
    const int k = i++;
    if (k < nl)
      return Phase{opsParent + (int64_t)(OP_L2L + c) * PD * PD,
                   pg >= 0 ? le + (int64_t)pg * PD : nullptr, false};
    return Phase{opsL + (int64_t)(OP_PINV1 + (k - nl)) * PD * PD, nullptr, true};
  };
  run_phases<PD>(nl + 2, acc, buf[0], buf[1], next);
  if (active) store_acc<PD>(acc, le + (int64_t)g * PD);
}

// ---------------------------------------------------------------------------
// Row-split variant of the two level kernels for levels with few boxes: a
// block holds 16 boxes and PD/16 waves, wave w computes rows 16w..16w+15, so
// one MFMA chain per phase is PD/16 times shorter than in the kernels above
// (which give a wave all PD rows of its 16 boxes).  A chained phase reads the
// previous result through a double-buffered LDS tile.
constexpr int kMaxSmallPhases = 6;   // 4 M2M + 2 pseudo-inverse factors

// Phases 0 .. nb-1 take their B column from row brow + i of `bbase`, where
// brow is the lane's box's source row (-1: none); these rows are loaded once
// per block, cooperatively and with 16-byte loads, into the LDS tile `btile`.
// Later phases are chained.  `opA(i)` is the operator of phase i.  A wave
// uses only its own 16 operator rows, so it loads them straight into
// registers, three phases ahead, with 16-byte loads: step 2 m + h of lane
// group kq multiplies k = 8 m + 2 kq + h.  (All six phases ahead was slower:
// the B tile fill then waits for every operator load issued before it.)  The B tiles are box-major in LDS
// (row padded to PD + 2), so that B arrives in the same order with one
// 16-byte LDS read per step pair.  (Staging the operator in LDS left every
// MFMA waiting on its own LDS read, with one block per CU to hide nothing.)
constexpr int kSmallPad = 2;
template <int PD, int NB, class OpFn>
__device__ __forceinline__ void run_small(
    int n, int nb, OpFn opA, const double* __restrict__ bbase, int brow, d4& acc,
    double (&bbuf)[2][16][PD + kSmallPad], double (&btile)[NB][16][PD + kSmallPad],
    int (&srow)[16]) {
  constexpr int KM = PD / 8, NT = 64 * (PD / 16), H = PD / 2;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = l & 15, kq = l >> 4;
  auto load_rows = [&](dv2 (&a)[KM], const double* __restrict__ A) {
    const double* r = A + (16 * w + col) * PD + 2 * kq;
#pragma unroll
    for (int m = 0; m < KM; ++m) a[m] = *reinterpret_cast<const dv2*>(r + 8 * m);
  };
  dv2 ar[3][KM];   // operator rows of phases i, i + 1, i + 2 (n <= kMaxSmallPhases)
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (i < n) load_rows(ar[i], opA(i));
  if (w == 0 && kq == 0) srow[col] = brow;
  __syncthreads();
  for (int e = threadIdx.x; e < NB * 16 * H; e += NT) {
    const int i = e / (16 * H), rem = e - i * 16 * H;
    const int box = rem / H, k2 = rem - box * H;
    const int rr = srow[box];
    dv2 v = dv2{0.0, 0.0};
    if (rr >= 0 && i < nb) v = *reinterpret_cast<const dv2*>(bbase + (int64_t)(rr + i) * PD + 2 * k2);
    *reinterpret_cast<dv2*>(&btile[i][box][2 * k2]) = v;
  }
#pragma unroll
  for (int i = 0; i < kMaxSmallPhases; ++i) {
    if (i >= n) break;
    const bool chain = i >= nb;
    if (chain) {
#pragma unroll
      for (int v = 0; v < 4; ++v) bbuf[i & 1][col][16 * w + 4 * v + kq] = acc[v];
      acc = d4{0.0, 0.0, 0.0, 0.0};
    }
    __syncthreads();   // also orders the btile fill before its first use
    dv2 b[KM];
    if (chain) {
#pragma unroll
      for (int m = 0; m < KM; ++m)
        b[m] = *reinterpret_cast<const dv2*>(&bbuf[i & 1][col][8 * m + 2 * kq]);
    } else {
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j == i) {
#pragma unroll
          for (int m = 0; m < KM; ++m)
            b[m] = *reinterpret_cast<const dv2*>(&btile[j][col][8 * m + 2 * kq]);
        }
    }
    __builtin_amdgcn_sched_barrier(0);   // every B read issued before the MFMA chain
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i % 3][m].x, b[m].x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i % 3][m].y, b[m].y, acc, 0, 0, 0);
    }
    if (i + 3 < n) load_rows(ar[i % 3], opA(i + 3));
  }
}

// Upward level (as k_mult_level): rows row0 .. row0 + nrows, 16 per unit;
// unit (bx, by) = rows 16 bx .. 16 bx + 15 of density by.
template <int PD>
__device__ __forceinline__ void mult_small_unit(
    int bx, int by, int64_t row0, int nrows, int with_m2m, const int2* __restrict__ rowinfo,
    const double* __restrict__ check, const double* __restrict__ opsL, double* __restrict__ mult,
    double* __restrict__ qs, int64_t dstride, float* __restrict__ m32) {
  __shared__ __attribute__((aligned(16))) double bbuf[2][16][PD + kSmallPad];
  __shared__ double qpart[PD / 16][16];
  check += (int64_t)by * dstride * PD;
  mult += (int64_t)by * dstride * PD;
  qs += (int64_t)by * dstride;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = l & 15, kq = l >> 4;
  const int rl = bx * 16 + col;
  const bool active = rl < nrows;
  const int64_t r = row0 + rl;
  const int2 ri = active ? rowinfo[r] : make_int2(-1, 0);
  const int f = ri.y;
  const int crow = ri.x;
  const double* src = (active && (f & NF_UP)) ? check + r * PD : nullptr;
  d4 acc;
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] = src ? src[16 * w + 4 * v + kq] : 0.0;
  const int nm = with_m2m ? 4 : 0;
  auto opA = [&](int k) -> const double* {
    return opsL + (int64_t)(k < nm ? OP_M2M + k : OP_PINV1 + (k - nm)) * PD * PD;
  };
  __shared__ __attribute__((aligned(16))) double btile[4][16][PD + kSmallPad];
  __shared__ int srow[16];
  run_small<PD, 4>(nm + 2, nm, opA, mult, crow, acc, bbuf, btile, srow);
  double part = acc[0] + acc[1] + acc[2] + acc[3];
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  if (kq == 0) qpart[w][col] = part;
  if (m32) {
    const double v[4] = {acc[0], acc[1], acc[2], acc[3]};
    store_m32<PD>(m32 + ((int64_t)by * dstride + r) * PD, w, v, active);
  }
  __syncthreads();
  if (active) {
    double* dst = mult + r * PD;
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[16 * w + 4 * v + kq] = acc[v];
    if (threadIdx.x < 16) {
      double q = 0.0;
#pragma unroll
      for (int ww = 0; ww < PD / 16; ++ww) q += qpart[ww][col];
      qs[r] = q;
    }
  }
}

template <int PD>
__global__ __launch_bounds__(64 * (PD / 16)) void k_mult_level_small(
    int64_t row0, int nrows, int with_m2m, const int2* __restrict__ rowinfo,
    const double* __restrict__ check, const double* __restrict__ opsL, double* __restrict__ mult,
    double* __restrict__ qs, int64_t dstride, float* __restrict__ m32) {
  mult_small_unit<PD>(blockIdx.x, blockIdx.y, row0, nrows, with_m2m, rowinfo, check, opsL, mult,
                      qs, dstride, m32);
}

// Downward level (as k_local_level): 16 boxes of child position c per unit;
// unit (bx, by) = child position bx mod 4 of sibling groups 16 (bx / 4) ..
// + 15, density by.
template <int PD>
__device__ __forceinline__ void local_small_unit(
    int bx, int by, int g0, int ngroups, int do_l2l, const int* __restrict__ flags,
    const int* __restrict__ parent, const double* __restrict__ ple,
    const double* __restrict__ opsL, const double* __restrict__ opsParent,
    double* __restrict__ le, int64_t nstride, const FrameFix& fx) {
  __shared__ __attribute__((aligned(16))) double bbuf[2][16][PD + kSmallPad];
  const double* ple2 = ple + 2 * nstride * PD;   // field 2 (FrameFix)
  ple += (int64_t)by * nstride * PD;
  le += (int64_t)by * nstride * PD;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = l & 15, kq = l >> 4;
  const int c = bx & 3;
  const int q = (bx >> 2) * 16 + col;
  const int g = g0 + 4 * q + c;
  const int gs = q < ngroups ? g : g0;   // a valid node for the independent loads
  const int fl = flags[gs];
  const int par = parent[gs];
  const bool active = q < ngroups && !(fl & NF_FAKE);
  d4 acc;
  const double* src = ple + (int64_t)gs * PD;
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] = src[16 * w + 4 * v + kq];
  if (fx.xmid && by < 2) {
    const double* src2 = ple2 + (int64_t)gs * PD;
    const double ct = fx.centre(gs, by);
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = __builtin_fma(-ct, src2[16 * w + 4 * v + kq], acc[v]);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] = active ? acc[v] : 0.0;
  const int pg = (active && do_l2l) ? par : -1;
  const int nl = do_l2l ? 1 : 0;
  auto opA = [&](int k) -> const double* {
    return k < nl ? opsParent + (int64_t)(OP_L2L + c) * PD * PD
                  : opsL + (int64_t)(OP_PINV1 + (k - nl)) * PD * PD;
  };
  __shared__ __attribute__((aligned(16))) double btile[1][16][PD + kSmallPad];
  __shared__ int srow[16];
  run_small<PD, 1>(nl + 2, nl, opA, le, pg, acc, bbuf, btile, srow);
  if (active) {
    double* dst = le + (int64_t)g * PD;
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[16 * w + 4 * v + kq] = acc[v];
  }
}

template <int PD>
__global__ __launch_bounds__(64 * (PD / 16)) void k_local_level_small(
    int g0, int ngroups, int do_l2l, const int* __restrict__ flags,
    const int* __restrict__ parent, const double* __restrict__ ple,
    const double* __restrict__ opsL, const double* __restrict__ opsParent,
    double* __restrict__ le, int64_t nstride, FrameFix fx) {
  local_small_unit<PD>(blockIdx.x, blockIdx.y, g0, ngroups, do_l2l, flags, parent, ple, opsL,
                       opsParent, le, nstride, fx);
}

// Levels with at most this many multipole rows (or 4 x this many boxes) use
// the row-split kernels.
constexpr int kSmallLevelRows = 4096;

// ---------------------------------------------------------------------------
// Targets: locate, count per evaluation box, scatter into box order.
// Node lookup by (depth, cell code) for target location: direct on a full
// level, a dense table on the other levels where the tree build made one,
// else a binary search in the level's sorted codes.
struct LocateTable {
  int L;
  int full_depth;   // a complete tree's leaf depth (evaluation box = leaf code), else -1
  double xmin, ymin, w0;
  int levoff[kMaxLevels + 1];
  int64_t tab_off[kMaxLevels];
  const int* tab;
  const uint64_t* code;
};
__device__ __forceinline__ int locate_node(const LocateTable& lt, int d, uint64_t c) {
  const int64_t nl = lt.levoff[d + 1] - lt.levoff[d];
  if (d < 16 && nl == ((int64_t)1 << (2 * d))) return lt.levoff[d] + (int)c;
  if (lt.tab_off[d] >= 0) return lt.tab[lt.tab_off[d] + (int64_t)c];
  const uint64_t* a = lt.code + lt.levoff[d];
  int64_t lo = 0, hi = nl;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return (lo < nl && a[lo] == c) ? lt.levoff[d] + (int)lo : -1;
}

// Each target gets its box and its rank inside the box.  The box is the first
// leaf on the target's root-to-leaf path (tree.py:712-727: descend while not
// leaf, child 2*(x > xmid) + (y > ymid)).  The cell codes along the path come
// from the reference's FP64 midpoint recurrence in registers (bit-identical to
// the stored node centres), so the node lookups of four depths are issued
// together instead of one dependent descent step per level.  Lanes of a wave
// that land in the same box share one atomic (targets arrive spatially
// coherent, so a wave touches few boxes).  In a complete tree the first leaf
// on the path is at the leaf depth D and its evaluation box is its code, so
// the recurrence alone gives the box.
__global__ __launch_bounds__(256) void k_locate_count(const double* __restrict__ tx,
                                                     const double* __restrict__ ty, int64_t nt,
                                                     LocateTable lt,
                                                     const int* __restrict__ flags,
                                                     const int* __restrict__ child,
                                                     const double* __restrict__ xmid,
                                                     const double* __restrict__ ymid,
                                                     const int* __restrict__ evalid,
                                                     int* __restrict__ tbox, int* __restrict__ trank,
                                                     int* __restrict__ tcnt) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool act = i < nt;
  int e = -1;
  if (act && lt.full_depth > 0) {
    const double x = tx[i], y = ty[i];
    double xmin = lt.xmin, ymin = lt.ymin, w = lt.w0;
    uint64_t c = 0;
    for (int d = 1; d <= lt.full_depth; ++d) {
      const double h = 0.5 * w;
      const double xm = xmin + h, ym = ymin + h;
      const int bx = x > xm, by = y > ym;
      c = (c << 2) | (uint64_t)(2 * bx + by);
      xmin = bx ? xm : xm - h;
      ymin = by ? ym : ym - h;
      w = h;
    }
    e = (int)c;
  } else if (act) {
    const double x = tx[i], y = ty[i];
    int g = 0;
    bool found = (flags[0] & NF_LEAF) != 0;
    if (!found) {
      double xmin = lt.xmin, ymin = lt.ymin, w = lt.w0;
      uint64_t c = 0;
      for (int d0 = 1; d0 < lt.L; d0 += 4) {
        int gg[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int d = d0 + u;
          gg[u] = -1;
          if (d < lt.L) {
            const double h = 0.5 * w;
            const double xm = xmin + h, ym = ymin + h;
            const int bx = x > xm, by = y > ym;
            c = (c << 2) | (uint64_t)(2 * bx + by);
            xmin = bx ? xm : xm - h;
            ymin = by ? ym : ym - h;
            w = h;
            gg[u] = locate_node(lt, d, c);
          }
        }
        int f[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) f[u] = gg[u] >= 0 ? flags[gg[u]] : 0;
        int hit = -1;
#pragma unroll
        for (int u = 3; u >= 0; --u)
          if (f[u] & NF_LEAF) hit = u;
        if (hit >= 0) {
          g = gg[hit];
          found = true;
          break;
        }
      }
    }
    // the path always ends in a leaf; the pointer descent is kept as a guard
    if (!found) {
      g = 0;
      for (int it = 0; it < kMaxLevels && !(flags[g] & NF_LEAF); ++it)
        g = child[g] + 2 * (x > xmid[g]) + (y > ymid[g]);
    }
    e = evalid[g];
  }
  // Group the wave's lanes by box with ballots first (no memory traffic),
  // then let every group's lowest lane add the group size with one atomic
  // instruction for all groups, so that targets spread over many boxes (in
  // random input order) cost one atomic round trip instead of one per box.
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long pending = __ballot(act);
  int leader = lane, inrank = 0, gsize = 0;
  while (pending) {
    const int ld = __builtin_ctzll(pending);
    const int key = __builtin_amdgcn_readlane(e, ld);
    const unsigned long long same = __ballot(act && e == key) & pending;
    if (same >> lane & 1ull) {
      leader = ld;
      inrank = __popcll(same & below);
    }
    if (lane == ld) gsize = __popcll(same);
    pending &= ~same;
  }
  int base = 0;
  if (act && lane == leader) base = atomicAdd(&tcnt[e], gsize);
  base = __shfl(base, leader, 64);
  if (act) {
    tbox[i] = e;
    trank[i] = base + inrank;
  }
}

__global__ __launch_bounds__(256) void k_target_scatter(const double* __restrict__ tx,
                                                       const double* __restrict__ ty, int64_t nt,
                                                       const int* __restrict__ tbox,
                                                       const int* __restrict__ trank,
                                                       const int* __restrict__ toff,
                                                       int* __restrict__ tperm,
                                                       double* __restrict__ tsx,
                                                       double* __restrict__ tsy) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nt) return;
  const int pos = toff[tbox[i]] + trank[i];
  tperm[pos] = (int)i;
  tsx[pos] = tx[i];
  tsy[pos] = ty[i];
}

constexpr int kTgtChunk = 64;

__global__ __launch_bounds__(256) void k_item_counts(const int* __restrict__ tcnt, int ne,
                                                    int* __restrict__ icnt) {
  int b = blockIdx.x * 256 + threadIdx.x;
  if (b < ne) icnt[b] = (tcnt[b] + kTgtChunk - 1) / kTgtChunk;
}

// Box offsets of a tree of at most kBoxOffsetsMax evaluation boxes, in one
// block: toff = exclusive scan of tcnt[0 .. ne] (tcnt[ne] = 0, so toff[ne] =
// nt) and ioff = exclusive scan of the boxes' item counts ceil(tcnt /
// kTgtChunk), with their total in *ntot: one pass over the counts for both
// (two scans and a count kernel otherwise).  Element j 1024 + t is thread t's
// (coalesced); tiles are scanned by wave scans and combined through LDS.
constexpr int kBoxTiles = 32;
constexpr int kBoxOffsetsMax = kBoxTiles * 1024 - 1;
__global__ __launch_bounds__(1024) void k_box_offsets(const int* __restrict__ tcnt, int ne,
                                                      int* __restrict__ toff,
                                                      int* __restrict__ ioff,
                                                      int* __restrict__ ntot) {
  __shared__ int wsum[2][kBoxTiles][16];
  __shared__ int woff[2][kBoxTiles][16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int n = ne + 1;
  const int J = (n + 1023) >> 10;
  int exv[kBoxTiles], exu[kBoxTiles];   // in-wave exclusive prefixes (loaded counts first)
#pragma unroll
  for (int j = 0; j < kBoxTiles; ++j) {
    const int idx = j * 1024 + t;
    exv[j] = (j < J && idx < ne) ? tcnt[idx] : 0;
  }
#pragma unroll
  for (int j = 0; j < kBoxTiles; ++j) {
    if (j < J) {
      const int v = exv[j], u = (v + kTgtChunk - 1) / kTgtChunk;
      const int iv = wave_inclusive_scan(v), iu = wave_inclusive_scan(u);
      if (lane == 63) {
        wsum[0][j][wid] = iv;
        wsum[1][j][wid] = iu;
      }
      exv[j] = iv - v;
      exu[j] = iu - u;
    }
  }
  __syncthreads();
  if (wid < 2) {   // wave a: offsets of array a (tile prefix, then the 16 waves of each tile)
    const int a = wid;
    int tile_tot = 0;
    if (lane < J)
      for (int w = 0; w < 16; ++w) tile_tot += wsum[a][lane][w];
    const int tile_off = wave_inclusive_scan(lane < J ? tile_tot : 0) - (lane < J ? tile_tot : 0);
    if (lane < J) {
      int run = tile_off;
      for (int w = 0; w < 16; ++w) {
        woff[a][lane][w] = run;
        run += wsum[a][lane][w];
      }
      if (a == 1 && lane == J - 1) *ntot = run;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kBoxTiles; ++j) {
    const int idx = j * 1024 + t;
    if (j < J && idx < n) {
      toff[idx] = woff[0][j][wid] + exv[j];
      ioff[idx] = woff[1][j][wid] + exu[j];
    }
  }
}

// Item = (evaluation node, first target, end target) of one chunk of at most
// kTgtChunk targets, in sorted-target positions.
__global__ __launch_bounds__(256) void k_make_items(const int* __restrict__ ioff, int ne,
                                                   const int* __restrict__ evalnode,
                                                   const int* __restrict__ toff,
                                                   int4* __restrict__ items) {
  int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= ne) return;
  const int o = ioff[b];
  const int g = evalnode[b], t0 = toff[b], t1 = toff[b + 1];
  for (int k = 0; t0 + k * kTgtChunk < t1; ++k)
    items[o + k] = make_int4(g, t0 + k * kTgtChunk, min(t1, t0 + (k + 1) * kTgtChunk), 0);
}

// ---------------------------------------------------------------------------
// L2T + P2P, one wave per (box, chunk of 64 targets); lane = target.
// Work items of the evaluation kernels, one per wave, kEvalWaves waves per
// block (the block shares one copy of the log table): XCD x (blocks b with
// b % 8 == x, dispatched in order) takes the x-th contiguous eighth of the
// items.  Items are in leaf (Morton) order, so the colleague sources that
// neighbouring leaves share stay in one XCD's L2.  The grid holds at least
// ceil(ni / 8) waves per XCD; surplus waves exit.  (A grid capped at 8
// blocks per CU, waves striding over several items, measured 8-20% slower.)
constexpr int kEvalWaves = 4;
constexpr int kStageSources = 64;   // LDS staging of few-target items: sources per wave and chunk
struct ItemRange {
  int first, last, step;
};
__device__ __forceinline__ ItemRange xcd_items(int ni) {
  const int x = blockIdx.x % kNumXcd;
  ItemRange r;
  r.first = (int)((int64_t)ni * x / kNumXcd) +
            (int)(blockIdx.x / kNumXcd) * kEvalWaves + (int)(threadIdx.x >> 6);
  r.last = min(r.first + 1, (int)((int64_t)ni * (x + 1) / kNumXcd));
  r.step = 1;
  return r;
}

template <int KER, int NOUT, bool CHECK>
struct Target {
  double x, y;
  HelmConsts kp;    // kernel parameter (Helmholtz k) and its derived constants
  PotAcc<KER> u;
  double gx = 0.0, gy = 0.0;
  // without self-exclusion: the smallest high word of d^2 seen; 0 means a
  // coincident pair (or one closer than 1e-154, below the normal range of d^2)
  uint32_t dmin = 1u;
  double dip = 0.0;   // double layer: sum (q.d)/d^2 (Laplace dipole sources)
  // a Laplace source with charge w and dipole moment (qx, qy)
  // (examples/pure_kifmm_laplace_dipoles.py:32-40).  GUARD false: the source
  // cannot coincide with the target (it lies outside the target's own box),
  // so the coincidence test and its bookkeeping are left out.
  template <bool GUARD = true>
  __device__ __forceinline__ void pair_dip(double sx, double sy, double w, double qx, double qy,
                                           const double2* __restrict__ tab) {
    const double dx = x - sx;
    const double dy = y - sy;
    const double d2 = __builtin_fma(dy, dy, dx * dx);
    const bool z = GUARD && d2 == 0.0;
    if constexpr (GUARD && CHECK) {
      w = z ? 0.0 : w;
      qx = z ? 0.0 : qx;
      qy = z ? 0.0 : qy;
    } else if constexpr (GUARD) {
      dmin = min(dmin, (uint32_t)(__builtin_bit_cast(uint64_t, d2) >> 32));
    }
    u.add(w, d2, tab, kp);
    dip = __builtin_fma(__builtin_fma(qx, dx, qy * dy), rcp_nr(z ? 1.0 : d2), dip);
  }
  // one source (or one equivalent-density point) at (sx, sy) with weight w;
  // GUARD as for pair_dip
  template <bool GUARD = true>
  __device__ __forceinline__ void pair(double sx, double sy, double w,
                                       const double2* __restrict__ tab) {
    const double dx = x - sx;
    const double dy = y - sy;
    double d2 = __builtin_fma(dy, dy, dx * dx);
    const bool z = GUARD && d2 == 0.0;
    if constexpr (GUARD && CHECK) w = z ? 0.0 : w;   // coincident source skipped (helpers.py:214)
    else if constexpr (GUARD) dmin = min(dmin, (uint32_t)(__builtin_bit_cast(uint64_t, d2) >> 32));
    if constexpr (KER == KER_LAPLACE) {
      u.add(w, d2, tab, kp);
      if (NOUT == 3) {
        const double qi = z ? 0.0 : w / d2;
        gx = __builtin_fma(qi, dx, gx);
        gy = __builtin_fma(qi, dy, gy);
      }
    } else {
      d2 = z ? 1.0 : d2;    // weight is 0 (or the result +inf) there
      if constexpr (NOUT == 1) {
        u.add(w, d2, tab, kp);
      } else {
        double k0, gk1;
        helm_k01(d2, rcp_nr(d2), tab, kp, k0, gk1);
        u.s = __builtin_fma(w, k0, u.s);
        const double qi = z ? 0.0 : -gk1 * w;
        gx = __builtin_fma(qi, dx, gx);
        gy = __builtin_fma(qi, dy, gy);
      }
    }
  }
  // gradient scale: Laplace -1/(2 pi) sum w (t-s)/d^2; Helmholtz 1/(2 pi) sum w (-k K1/r)(t-s)
  __device__ __forceinline__ double gscale() const {
    return KER == KER_LAPLACE ? -kInv2Pi : kInv2Pi;
  }
  __device__ __forceinline__ bool zhit() const { return dmin == 0u; }
};

// Items with few targets (a leaf holding at most kFewTargets of them, or the
// tail chunk of a larger one): instead of one lane per target, which leaves
// most lanes idle while the wave walks every near source, lane l works on
// target slot l mod TG (TG = the next power of two >= the item's targets) and
// takes every (64/TG)-th source and L2T term; the partial sums of a slot are
// then added by XOR butterflies over the lanes that share it, in a fixed
// order.  All results are linear in the partial sums.
constexpr int kFewTargets = 32;

struct FewLanes {
  int lg, ng, slot, sg;
  bool act;
};
__device__ __forceinline__ FewLanes few_lanes(int na) {
  FewLanes f;
  f.lg = na <= 1 ? 0 : 32 - __builtin_clz(na - 1);
  f.ng = 64 >> f.lg;
  const int lane = threadIdx.x & 63;
  f.slot = lane & ((1 << f.lg) - 1);
  f.sg = lane >> f.lg;
  f.act = f.slot < na;
  return f;
}
// L2T (chebeval2d1, bbfmm.py:27-32,185-199): sum_{i,j} c[i pc + j] T_i(x) T_j(y)
// over the rows i = i0, i0 + istep, ... (istep a power of two; lanes of one
// target slot split the rows).  Indices are unrolled to constants so the
// T values stay in registers.
__device__ __forceinline__ double cheb_l2t(double xr, double yr, double wadj,
                                           const double* __restrict__ c, int pc, int i0,
                                           int istep) {
  const double x = xr * wadj, y = yr * wadj;
  double tx[kMaxCheb], ty[kMaxCheb];
  tx[0] = 1.0;
  tx[1] = x;
  ty[0] = 1.0;
  ty[1] = y;
#pragma unroll
  for (int m = 2; m < kMaxCheb; ++m) {
    tx[m] = __builtin_fma(2.0 * x, tx[m - 1], -tx[m - 2]);
    ty[m] = __builtin_fma(2.0 * y, ty[m - 1], -ty[m - 2]);
  }
  double u = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxCheb; ++i) {
    if (i < pc && i >= i0 && ((i - i0) & (istep - 1)) == 0) {
      double r = 0.0;
#pragma unroll
      for (int j = 0; j < kMaxCheb; ++j)
        if (j < pc) r = __builtin_fma(c[i * pc + j], ty[j], r);
      u = __builtin_fma(tx[i], r, u);
    }
  }
  return u;
}

// The nine colleague source ranges of box g, loaded together: lane q < 9
// holds colleague q and its [bot, top) (an absent colleague an empty range),
// read back with readlane, so that an item pays one load latency for all
// colleagues instead of a dependent chain per colleague.  Call with the whole
// wave active.
struct CollLanes {
  int c, b, e;
};
__device__ __forceinline__ CollLanes coll_lanes(int g, const int* __restrict__ coll,
                                                const int2* __restrict__ crange) {
  const int q = threadIdx.x & 63;
  CollLanes L;
  L.c = q < 9 ? coll[(int64_t)g * 9 + q] : -1;
  const int2 r = q < 9 ? crange[(int64_t)g * 9 + q] : make_int2(0, 0);
  L.b = r.x;
  L.e = r.y;
  return L;
}

// sum over the lanes that share this lane's target slot
__device__ __forceinline__ double slot_sum(double v, int lg) {
  for (int off = 1 << lg; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ uint32_t slot_min(uint32_t v, int lg) {
  for (int off = 1 << lg; off < 64; off <<= 1)
    v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
  return v;
}

template <int PD, int KER, int NOUT, bool CHECK, bool DIP, bool CHEB>
__device__ __forceinline__ void eval_few(
    int g, int t0, int na, const int* __restrict__ tperm, const double* __restrict__ tsx,
    const double* __restrict__ tsy, const int* __restrict__ nlevel,
    const double* __restrict__ xmid, const double* __restrict__ ymid,
    const int* __restrict__ coll, const int2* __restrict__ crange, const int* __restrict__ bot, const int* __restrict__ top,
    const double4* __restrict__ src, const double2* __restrict__ dip,
    const double* __restrict__ le, const double* __restrict__ ops, int64_t opsz, int p,
    const double2* __restrict__ tab, double kp, int64_t nt, double* __restrict__ out,
    double4* __restrict__ stage, int scap) {
  const FewLanes F = few_lanes(na);
  const int t = t0 + F.slot;
  Target<KER, NOUT, CHECK> T;
  T.kp = helm_consts(kp);
  // idle slots repeat the item's first target (results not stored), so that
  // they take the same branch of the kernel evaluation as the active ones
  T.x = tsx[F.act ? t : t0];
  T.y = tsy[F.act ? t : t0];
  const int d = nlevel[g];
  double far = 0.0;   // Chebyshev L2T
  if (CHEB && d >= 2) {
    far = cheb_l2t(T.x - xmid[g], T.y - ymid[g], ops[d * opsz + ops_off_circles(PD)],
                   le + (int64_t)g * PD, p, F.sg, F.ng);
  } else if (d >= 2) {
    Target<KER, NOUT, false> L;
    L.kp = T.kp;
    L.x = T.x - xmid[g];
    L.y = T.y - ymid[g];
    const double* lx = ops + d * opsz + (int64_t)OP_NBLK * PD * PD;
    const double* ly = lx + PD;
    const double* lg = le + (int64_t)g * PD;
    for (int k = F.sg; k < p; k += F.ng) L.pair(lx[k], ly[k], lg[k], tab);
    T.u = L.u;
    T.gx = L.gx;
    T.gy = L.gy;
  }
  const CollLanes CL = coll_lanes(g, coll, crange);
  if (!DIP && scap > 0) {
    // The colleague sources, concatenated, pass through this wave's LDS area
    // scap at a time: every lane loads its share of a chunk at once (one
    // memory round trip per chunk instead of one per colleague), then the
    // source groups read them from LDS.
    int cb[9], cp[10];   // colleague starts and prefix offsets (wave-uniform)
    cp[0] = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      cb[q] = __builtin_amdgcn_readlane(CL.b, q);
      const int ce = __builtin_amdgcn_readlane(CL.e, q);
      cp[q + 1] = cp[q] + (ce > cb[q] ? ce - cb[q] : 0);
    }
    const int lane = threadIdx.x & 63;
    for (int base = 0; base < cp[9]; base += scap) {
      const int n = min(scap, cp[9] - base);
      for (int v = lane; v < n; v += 64) {
        const int vv = base + v;
        int j = vv + cb[0];
#pragma unroll
        for (int q = 1; q < 9; ++q) j = vv >= cp[q] ? vv - cp[q] + cb[q] : j;
        stage[v] = src[j];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (int v = F.sg; v < n; v += F.ng) {
        const double4 sj = stage[v];
        T.pair(sj.x, sj.y, sj.z, tab);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  // colleague 4 is the target's own box, the only one that can hold a
  // source at the target (a point and its box are found by the same
  // midpoint tests for sources and targets)
  auto colleague = [&](auto guard, int b, int e) {
    constexpr bool G = decltype(guard)::value;
    int j = b + F.sg;
    if constexpr (!DIP) {
      // four sources per lane in flight: the loads are per lane here (lanes
      // of different source groups read different sources)
      for (; j + 3 * F.ng < e; j += 4 * F.ng) {
        double4 s4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) s4[u] = src[j + u * F.ng];
#pragma unroll
        for (int u = 0; u < 4; ++u) T.template pair<G>(s4[u].x, s4[u].y, s4[u].z, tab);
      }
    }
    for (; j < e; j += F.ng) {
      const double4 sj = src[j];
      if constexpr (DIP) {
        const double2 q = dip[j];
        T.template pair_dip<G>(sj.x, sj.y, sj.z, q.x, q.y, tab);
      } else {
        T.template pair<G>(sj.x, sj.y, sj.z, tab);
      }
    }
  };
  for (int qn = 0; qn < 9 && (DIP || scap == 0); ++qn) {
    const int b = __builtin_amdgcn_readlane(CL.b, qn);
    const int e = __builtin_amdgcn_readlane(CL.e, qn);
    if (e <= b) continue;
    if (qn == 4) colleague(std::true_type{}, b, e);
    else colleague(std::false_type{}, b, e);
  }
  double u = slot_sum(DIP ? __builtin_fma(kInv2Pi, T.dip, T.u.result()) : T.u.result() + far,
                      F.lg);
  double gx = 0.0, gy = 0.0;
  if (NOUT == 3) {
    gx = slot_sum(T.gx, F.lg);
    gy = slot_sum(T.gy, F.lg);
  }
  const bool zh = CHECK ? false : slot_min(T.dmin, F.lg) == 0u;
  if (F.act && F.sg == 0) {
    const int64_t o = tperm[t];
    if (zh) u = DIP ? __builtin_nan("") : __builtin_inf();
    out[o] = u;
    if (NOUT == 3) {
      out[nt + o] = T.gscale() * gx;
      out[2 * nt + o] = T.gscale() * gy;
    }
  }
}

// The Helmholtz gradient instance is held to 3 waves per SIMD (168 VGPRs;
// it needs 169 unconstrained, which leaves 2).
template <int PD, int KER, int NOUT, bool CHECK, bool DIP, bool CHEB>
__global__ __launch_bounds__(64 * kEvalWaves)
__attribute__((amdgpu_waves_per_eu(KER == KER_MODHELM && NOUT == 3 ? 3 : 1))) void k_eval(
    const int4* __restrict__ items, const int* __restrict__ nitems,
    const int* __restrict__ evalnode, const int* __restrict__ toff,
    const int* __restrict__ tperm, const double* __restrict__ tsx, const double* __restrict__ tsy,
    const int* __restrict__ nlevel, const double* __restrict__ xmid,
    const double* __restrict__ ymid, const int* __restrict__ coll, const int2* __restrict__ crange, const int* __restrict__ bot,
    const int* __restrict__ top, const double4* __restrict__ src,
    const double2* __restrict__ dip,
    const double* __restrict__ le, const double* __restrict__ ops, int64_t opsz, int p,
    const double2* __restrict__ logtab, double kp, int64_t nt, double* __restrict__ out,
    int scap) {
  __shared__ double2 tab[kLogTableSize];
  extern __shared__ double4 stage_lds[];   // scap per wave when launched with it
  double4* stage = stage_lds + (int)(threadIdx.x >> 6) * scap;
  const int ni = *nitems;
  load_logtab(tab, logtab);
  __syncthreads();
  const ItemRange R = xcd_items(ni);
  for (int ii = R.first; ii < R.last; ii += R.step) {
    const int4 it = items[ii];
    // wave-uniform (ii is): in scalar registers, so loads indexed by them are scalar
    const int g = __builtin_amdgcn_readfirstlane(it.x);
    const int t0 = __builtin_amdgcn_readfirstlane(it.y), t1 = __builtin_amdgcn_readfirstlane(it.z);
    if (t1 - t0 <= kFewTargets) {
      eval_few<PD, KER, NOUT, CHECK, DIP, CHEB>(g, t0, t1 - t0, tperm, tsx, tsy, nlevel, xmid,
                                                ymid, coll, crange, bot, top, src, dip, le, ops, opsz, p,
                                                tab, kp, nt, out, stage, scap);
      continue;
    }
    const int t = t0 + (int)(threadIdx.x & 63);
    const bool act = t < t1;
    Target<KER, NOUT, CHECK> T;
    T.kp = helm_consts(kp);
    T.x = tsx[act ? t : t0];   // idle lanes: as in eval_few
    T.y = tsy[act ? t : t0];
    // L2T: helpers.py:219-235 (the equivalent density sits on the large circle);
    // Chebyshev expansions: bbfmm.py:185-199
    const int d = nlevel[g];
    double far = 0.0;
    if (CHEB && d >= 2) {
      far = cheb_l2t(T.x - xmid[g], T.y - ymid[g], ops[d * opsz + ops_off_circles(PD)],
                     le + (int64_t)g * PD, p, 0, 1);
    } else if (d >= 2) {
      const double* lx = ops + d * opsz + (int64_t)OP_NBLK * PD * PD;
      const double* ly = lx + PD;
      Target<KER, NOUT, false> L;
      L.kp = T.kp;
      L.x = T.x - xmid[g];
      L.y = T.y - ymid[g];
      const double* lg = le + (int64_t)g * PD;
      for (int k = 0; k < p; ++k) L.pair(lx[k], ly[k], lg[k], tab);
      T.u = L.u;
      T.gx = L.gx;
      T.gy = L.gy;
    }
    // P2P over the leaf's colleagues: helpers.py:180-217.  Only colleague 4,
    // the target's own box, can hold a source at the target (a point and its
    // box are found by the same midpoint tests for sources and targets), so
    // only its pairs test for coincidence.
    const CollLanes CL = coll_lanes(g, coll, crange);
    auto colleague = [&](auto guard, int b, int e) {
      constexpr bool G = decltype(guard)::value;
      int j = b;
      if constexpr (DIP) {
        for (; j < e; ++j) {
          const double4 sj = src[j];
          const double2 q = dip[j];
          T.template pair_dip<G>(sj.x, sj.y, sj.z, q.x, q.y, tab);
        }
        return;
      }
      for (; j + 4 <= e; j += 4) {
        double4 s4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) s4[u] = src[j + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) T.template pair<G>(s4[u].x, s4[u].y, s4[u].z, tab);
      }
      for (; j < e; ++j) {
        const double4 sj = src[j];
        T.template pair<G>(sj.x, sj.y, sj.z, tab);
      }
    };
    for (int qn = 0; qn < 9; ++qn) {
      const int b = __builtin_amdgcn_readlane(CL.b, qn);
      const int e = __builtin_amdgcn_readlane(CL.e, qn);
      if (e <= b) continue;
      if (qn == 4) colleague(std::true_type{}, b, e);
      else colleague(std::false_type{}, b, e);
    }
    if (act) {
      const int64_t o = tperm[t];
      double u = DIP ? __builtin_fma(kInv2Pi, T.dip, T.u.result()) : T.u.result() + far;
      // log(0): the reference's -log(0) = +inf; with a dipole also 0/0
      if (!CHECK && T.zhit()) u = DIP ? __builtin_nan("") : __builtin_inf();
      out[o] = u;
      if (NOUT == 3) {
        out[nt + o] = T.gscale() * T.gx;
        out[2 * nt + o] = T.gscale() * T.gy;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed precision near field.  Every source is stored in FP32 relative to the
// centre of its leaf in units of the leaf width, (x - c) * (1/w) computed in
// FP64 and rounded once; a target evaluated against a leaf colleague uses the
// same FP64 expression with the colleague's centre, so an exactly coincident
// pair gives exactly zero and is excluded as in FP64.  The log is the hardware
// log2 of d^2 / w^2 (O(1) values keep its absolute error small);
// sum_j t_j log d_j^2 = ln2 * sum_j t_j log2(d_j^2/w^2) + log(w^2) * sum_j t_j.
// Pairs closer than w/64 go to FP64: the coordinates' rounding (~6e-8 w)
// costs ~1e-7 w/d in log d.  The rarer such pairs, the more groups of four
// sources the loop runs without selects (w/16 -> w/64: 0.263 -> 0.223 ms on
// 1M points, error vs FP64 unchanged at ~1e-8).
constexpr float kNearF32 = 1.0f / 4096.0f;   // (w / 64)^2 in units of w^2

struct LevelConsts {
  double invw[kMaxLevels];
  double logw2[kMaxLevels];
};

__global__ __launch_bounds__(64) void k_prep_mixed(const int* __restrict__ evalnode,
                                                  const int* __restrict__ nlevel,
                                                  const int* __restrict__ bot,
                                                  const int* __restrict__ top,
                                                  const double* __restrict__ xmid,
                                                  const double* __restrict__ ymid,
                                                  const double4* __restrict__ src,
                                                  LevelConsts lc, f4* __restrict__ srcf,
                                                  double* __restrict__ tsum) {
  const int g = evalnode[blockIdx.x];
  const double iw = lc.invw[nlevel[g]];
  const double cx = xmid[g], cy = ymid[g];
  const int b = bot[g], e = top[g];
  double ts = 0.0;
  for (int j = b + threadIdx.x; j < e; j += 64) {
    const double4 sj = src[j];
    srcf[j] = f4{(float)((sj.x - cx) * iw), (float)((sj.y - cy) * iw), (float)sj.z, 0.f};
    ts += sj.z;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) ts += __shfl_xor(ts, off, 64);
  if (threadIdx.x == 0) tsum[g] = ts;
}

// Few-target items of k_eval_mixed (lane layout of eval_few).  The
// log(w^2) * sum(tau) term of a leaf colleague is added once per target, by
// the lanes of source group 0; every lane removes the charges it took in FP64.
template <int PD, bool CHECK, bool CHEB>
__device__ __forceinline__ void eval_few_mixed(
    int g, int t0, int na, const int* __restrict__ tperm, const double* __restrict__ tsx,
    const double* __restrict__ tsy, const int* __restrict__ nlevel, const int* __restrict__ flags,
    const double* __restrict__ xmid, const double* __restrict__ ymid,
    const int* __restrict__ coll, const int2* __restrict__ crange, const int* __restrict__ bot, const int* __restrict__ top,
    const double4* __restrict__ src, const f4* __restrict__ srcf, const double* __restrict__ tsum,
    const LevelConsts& lc, const double* __restrict__ le, const double* __restrict__ ops,
    int64_t opsz, int p, const double2* __restrict__ tab, double* __restrict__ out) {
  const FewLanes F = few_lanes(na);
  const int t = t0 + F.slot;
  Target<KER_LAPLACE, 1, CHECK> T;
  T.kp = HelmConsts{};
  // idle lanes repeat the item's first target (same branches as a real lane)
  T.x = tsx[F.act ? t : t0];
  T.y = tsy[F.act ? t : t0];
  const int d = nlevel[g];
  double far = 0.0;   // Chebyshev L2T
  if (CHEB && d >= 2) {
    far = cheb_l2t(T.x - xmid[g], T.y - ymid[g], ops[d * opsz + ops_off_circles(PD)],
                   le + (int64_t)g * PD, p, F.sg, F.ng);
  } else if (d >= 2) {
    Target<KER_LAPLACE, 1, false> L;
    L.kp = HelmConsts{};
    L.x = T.x - xmid[g];
    L.y = T.y - ymid[g];
    const double* lx = ops + d * opsz + ops_off_circles(PD);
    const double* ly = lx + PD;
    const double* lg = le + (int64_t)g * PD;
    for (int k = F.sg; k < p; k += F.ng) L.pair(lx[k], ly[k], lg[k], tab);
    T.u = L.u;
  }
  const double iw = lc.invw[d];
  const double lw2 = lc.logw2[d];
  double u32 = 0.0;
  const CollLanes CL = coll_lanes(g, coll, crange);
  for (int qn = 0; qn < 9; ++qn) {
    const int c = __builtin_amdgcn_readlane(CL.c, qn);
    const int b = __builtin_amdgcn_readlane(CL.b, qn);
    const int e = __builtin_amdgcn_readlane(CL.e, qn);
    if (e <= b) continue;
    if (flags[c] & NF_LEAF) {
      const float xr = (float)((T.x - xmid[c]) * iw);
      const float yr = (float)((T.y - ymid[c]) * iw);
      float acc = 0.f;
      double excl = 0.0;
      auto pair32 = [&](const f4& sj, int jj) {
        const float dx = xr - sj.x, dy = yr - sj.y;
        const float d2 = __builtin_fmaf(dy, dy, dx * dx);
        const bool nr = d2 < kNearF32;
        const float lg2 = __builtin_amdgcn_logf(nr ? 1.0f : d2);
        acc = __builtin_fmaf(nr ? 0.f : sj.z, lg2, acc);
        if (nr) {
          const double4 sd = src[jj];
          excl += sd.z;
          T.pair(sd.x, sd.y, sd.z, tab);
        }
      };
      int j = b + F.sg;
      for (; j + 3 * F.ng < e; j += 4 * F.ng) {   // four per-lane loads in flight
        f4 s4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) s4[u] = srcf[j + u * F.ng];
        float d2[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float dx = xr - s4[u].x, dy = yr - s4[u].y;
          d2[u] = __builtin_fmaf(dy, dy, dx * dx);
        }
        const float dm = fminf(fminf(d2[0], d2[1]), fminf(d2[2], d2[3]));
        if (__builtin_expect(__ballot(dm < kNearF32) == 0, 1)) {
#pragma unroll
          for (int u = 0; u < 4; ++u)
            acc = __builtin_fmaf(s4[u].z, __builtin_amdgcn_logf(d2[u]), acc);
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) pair32(s4[u], j + u * F.ng);
        }
      }
      for (; j < e; j += F.ng) pair32(srcf[j], j);
      u32 += (double)acc;
      T.u.a.at = __builtin_fma(lw2, (F.sg == 0 ? tsum[c] : 0.0) - excl, T.u.a.at);
    } else {
      int j = b + F.sg;
      for (; j + 3 * F.ng < e; j += 4 * F.ng) {
        double4 s4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) s4[u] = src[j + u * F.ng];
#pragma unroll
        for (int u = 0; u < 4; ++u) T.pair(s4[u].x, s4[u].y, s4[u].z, tab);
      }
      for (; j < e; j += F.ng) {
        const double4 sj = src[j];
        T.pair(sj.x, sj.y, sj.z, tab);
      }
    }
  }
  double u = slot_sum(__builtin_fma(-kInv4Pi, __builtin_fma(u32, kLn2, T.u.a.value()), far), F.lg);
  const bool zh = CHECK ? false : slot_min(T.dmin, F.lg) == 0u;
  if (F.act && F.sg == 0) out[tperm[t]] = zh ? __builtin_inf() : u;
}

template <int PD, bool CHECK, bool CHEB>
__global__ __launch_bounds__(64 * kEvalWaves) void k_eval_mixed(
    const int4* __restrict__ items, const int* __restrict__ nitems,
    const int* __restrict__ evalnode, const int* __restrict__ toff,
    const int* __restrict__ tperm, const double* __restrict__ tsx, const double* __restrict__ tsy,
    const int* __restrict__ nlevel, const int* __restrict__ flags,
    const double* __restrict__ xmid, const double* __restrict__ ymid,
    const int* __restrict__ coll, const int2* __restrict__ crange, const int* __restrict__ bot, const int* __restrict__ top,
    const double4* __restrict__ src, const f4* __restrict__ srcf,
    const double* __restrict__ tsum, LevelConsts lc, const double* __restrict__ le,
    const double* __restrict__ ops, int64_t opsz, int p, const double2* __restrict__ logtab,
    double* __restrict__ out) {
  __shared__ double2 tab[kLogTableSize];
  const int ni = *nitems;
  load_logtab(tab, logtab);
  __syncthreads();
  const ItemRange R = xcd_items(ni);
  for (int ii = R.first; ii < R.last; ii += R.step) {
    const int4 it = items[ii];
    // wave-uniform (ii is): in scalar registers, so loads indexed by them are scalar
    const int g = __builtin_amdgcn_readfirstlane(it.x);
    const int t0 = __builtin_amdgcn_readfirstlane(it.y), t1 = __builtin_amdgcn_readfirstlane(it.z);
    if (t1 - t0 <= kFewTargets) {
      eval_few_mixed<PD, CHECK, CHEB>(g, t0, t1 - t0, tperm, tsx, tsy, nlevel, flags, xmid,
                                      ymid, coll, crange, bot, top, src, srcf, tsum, lc, le, ops, opsz,
                                      p, tab, out);
      continue;
    }
    const int t = t0 + (int)(threadIdx.x & 63);
    const bool act = t < t1;
    Target<KER_LAPLACE, 1, CHECK> T;
    T.kp = HelmConsts{};
    // idle lanes repeat the item's first target: they then take the same
    // branches as a real lane (no near-pair path of their own)
    T.x = tsx[act ? t : t0];
    T.y = tsy[act ? t : t0];
    const int d = nlevel[g];
    double far = 0.0;   // Chebyshev L2T
    if (CHEB && d >= 2) {
      far = cheb_l2t(T.x - xmid[g], T.y - ymid[g], ops[d * opsz + ops_off_circles(PD)],
                     le + (int64_t)g * PD, p, 0, 1);
    } else if (d >= 2) {   // L2T in FP64 (the local densities carry large cancelling modes)
      const double* lx = ops + d * opsz + ops_off_circles(PD);
      const double* ly = lx + PD;
      Target<KER_LAPLACE, 1, false> L;
      L.kp = HelmConsts{};
      L.x = T.x - xmid[g];
      L.y = T.y - ymid[g];
      const double* lg = le + (int64_t)g * PD;
      for (int k = 0; k < p; ++k) L.pair(lx[k], ly[k], lg[k], tab);
      T.u = L.u;
    }
    const double iw = lc.invw[d];
    const double lw2 = lc.logw2[d];
    double u32 = 0.0;      // ln2-scaled part from the FP32 near field
    const CollLanes CL = coll_lanes(g, coll, crange);
    for (int qn = 0; qn < 9; ++qn) {
      const int c = __builtin_amdgcn_readlane(CL.c, qn);
      const int b = __builtin_amdgcn_readlane(CL.b, qn);
      const int e = __builtin_amdgcn_readlane(CL.e, qn);
      if (e <= b) continue;
      if (flags[c] & NF_LEAF) {
        const float xr = (float)((T.x - xmid[c]) * iw);
        const float yr = (float)((T.y - ymid[c]) * iw);
        float acc = 0.f;
        double excl = 0.0;   // charges of the pairs taken in FP64
        // pairs closer than w/64 (and coincident ones) go to FP64: the rounding
        // of the FP32 relative coordinates costs ~1e-7 w / d in log d
        auto near64 = [&](int jj) {
          const double4 sd = src[jj];
          excl += sd.z;
          T.pair(sd.x, sd.y, sd.z, tab);
        };
        int j = b;
        for (; j + 4 <= e; j += 4) {
          f4 s4[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) s4[u] = srcf[j + u];
          float d2[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const float dx = xr - s4[u].x, dy = yr - s4[u].y;
            d2[u] = __builtin_fmaf(dy, dy, dx * dx);
          }
          const float dm = fminf(fminf(d2[0], d2[1]), fminf(d2[2], d2[3]));
          if (__builtin_expect(__ballot(dm < kNearF32) == 0, 1)) {
            // no near pair in the wave's four sources: no selects
#pragma unroll
            for (int u = 0; u < 4; ++u)
              acc = __builtin_fmaf(s4[u].z, __builtin_amdgcn_logf(d2[u]), acc);
          } else {
            bool nr[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              nr[u] = d2[u] < kNearF32;
              const float lg2 = __builtin_amdgcn_logf(nr[u] ? 1.0f : d2[u]);
              acc = __builtin_fmaf(nr[u] ? 0.f : s4[u].z, lg2, acc);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (nr[u]) near64(j + u);
          }
        }
        for (; j < e; ++j) {
          const f4 sj = srcf[j];
          const float dx = xr - sj.x, dy = yr - sj.y;
          const float d2 = __builtin_fmaf(dy, dy, dx * dx);
          const bool nr = d2 < kNearF32;
          const float lg2 = __builtin_amdgcn_logf(nr ? 1.0f : d2);
          acc = __builtin_fmaf(nr ? 0.f : sj.z, lg2, acc);
          if (nr) near64(j);
        }
        u32 += (double)acc;
        T.u.a.at = __builtin_fma(lw2, tsum[c] - excl, T.u.a.at);
      } else {
        // colleague with children (adaptive trees): its sources are not stored
        // relative to it, so this range stays FP64
        int j = b;
        for (; j < e; ++j) {
          const double4 sj = src[j];
          T.pair(sj.x, sj.y, sj.z, tab);
        }
      }
    }
    if (act) {
      double u = __builtin_fma(-kInv4Pi, __builtin_fma(u32, kLn2, T.u.a.value()), far);
      if (!CHECK && T.zhit()) u = __builtin_inf();
      out[tperm[t]] = u;
    }
  }
}

// ---------------------------------------------------------------------------
// 2D Stokeslet (stokes_fmm.py:33-61, 102-184): for a force f at s and r = t - s,
//   u = (1/4 pi) [ -log|r| f + (r.f) r / |r|^2 ],   p = (1/2 pi) (r.f) / |r|^2.
// With L[g](t) = sum_s g_s log|t - s| and coordinates X = x - c0 relative to
// the root centre c0, since r_i r_j / |r|^2 f_j = (X_i(t) - X_i(s)) f_j d_j log|r|:
//   u_i = (1/4 pi) [ -L[f_i] + X_i(t) B - C_i ],  p = B / (2 pi),
//   B = d_x L[f_x] + d_y L[f_y],  C_i = d_x L[X_i f_x] + d_y L[X_i f_y].
// Double layer (stresslet, stokes_fmm.py:64-101,187-236): for a source with
// strength sigma and normal n,
//   u = (1/pi) (r.n)(r.sigma) r / |r|^4,
//   p = (1/pi) [-(n.sigma) / |r|^2 + 2 (r.n)(r.sigma) / |r|^4].
// With g = n.sigma, a = n_x s_x - n_y s_y, b = n_x s_y + n_y s_x and the
// quadrupole field Q[a, b](t) = sum (a d_xx + b d_xy) log|t - s| (harmonic:
// d_yy = -d_xx), since r_i r_j r_k / |r|^4 = (1/2)(delta_jk d_i log|r|
// - (X_i(t) - X_i(s)) d_j d_k log|r|):
//   u_i = (1/2 pi) [ d_i L[g] - X_i(t) Q[a, b] + Q[X_i a, X_i b] ],
//   p   = -(1/pi) Q[a, b].
// Both layers together therefore need only three fields, each harmonic away
// from the sources (a sum of log, d log and d d log terms):
//   D_i = -L[f_i] - C_i + 2 (d_i L[g] + Q[X_i a, X_i b])   (i = x, y),
//   D_2 = B - 2 Q[a, b],
//   u_i = (1/4 pi) [ D_i + X_i(t) D_2 ],   p = D_2 / (2 pi).
// They run through the scalar passes as a density dimension of three, with
// the fields themselves as check values (S2M) and read back by L2T as the
// potentials of their local equivalent densities; the near field evaluates
// the Stokeslet and the stresslet directly, with
// (r.n)(r.sigma) = (1/2) [g |r|^2 + a (r_x^2 - r_y^2) + 2 b r_x r_y].
constexpr int kStokesFields = 3;

// Sources in tree order as (x, y, f_x, f_y); tau holds f_x[n] then f_y[n].
__global__ void k_pack_stokes(const double* __restrict__ tau, const uint32_t* __restrict__ perm,
                              const double* __restrict__ sx, const double* __restrict__ sy,
                              int64_t n, double4* __restrict__ packed) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  packed[i] = make_double4(sx[i], sy[i], tau[j], tau[n + j]);
}

// Forces and double-layer densities: tau holds f_x, f_y, sigma_x, sigma_y
// ([n] each), dvec the normals n_x, n_y; (x, y, f_x, f_y) and (g, a, b, 0) in
// tree order.
__global__ void k_pack_stokes_dipole(const double* __restrict__ tau,
                                     const double* __restrict__ dvec,
                                     const uint32_t* __restrict__ perm,
                                     const double* __restrict__ sx, const double* __restrict__ sy,
                                     int64_t n, double4* __restrict__ packed,
                                     double4* __restrict__ sdip) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  packed[i] = make_double4(sx[i], sy[i], tau[j], tau[n + j]);
  const double gx = tau[2 * n + j], gy = tau[3 * n + j];
  const double nx = dvec[j], ny = dvec[n + j];
  sdip[i] = make_double4(nx * gx + ny * gy, nx * gx - ny * gy, nx * gy + ny * gx, 0.0);
}

// S2M of the three fields D_x, D_y, D_2 at the check points (r = check - s):
//   D_i += -f_i log|r| - X_i(s) (f.r)/|r|^2,  D_2 += (f.r)/|r|^2,
// DIP: D_i += 2 (g r_i/|r|^2 + X_i(s) q),  D_2 -= 2 q,
//   q = a d_xx log|r| + b d_xy log|r| = (a (r_y^2 - r_x^2) - 2 b r_x r_y)/|r|^4.
template <int PD, bool DIP>
__global__ __launch_bounds__(64 * kS2MWaves) void k_s2m_stokes(int nup,
                                                               const int* __restrict__ upnodes,
                                                   const int* __restrict__ nlevel,
                                                   const int* __restrict__ bot,
                                                   const int* __restrict__ top,
                                                   const int* __restrict__ row,
                                                   const double* __restrict__ xmid,
                                                   const double* __restrict__ ymid,
                                                   const double4* __restrict__ src,
                                                   const double* __restrict__ ops, int64_t opsz,
                                                   int p, const double2* __restrict__ logtab,
                                                   double c0x, double c0y, int64_t rstride,
                                                   const double4* __restrict__ sdip,
                                                   double* __restrict__ check) {
  __shared__ double2 tab[kLogTableSize];
  load_logtab(tab, logtab);
  __syncthreads();
  const int gi = blockIdx.x * kS2MWaves + (int)(threadIdx.x >> 6);
  if (gi >= nup) return;
  const int g = upnodes[gi];
  const int d = nlevel[g];
  const int k = threadIdx.x & 63;
  const int r = row[g];
  if (r < 0 || d < 2) return;   // wave-uniform
  const double* lx = ops + d * opsz + ops_off_circles(PD);
  const double* ly = lx + PD;
  const double ax = (k < p ? lx[k] : 1.0) + xmid[g];   // check point, absolute
  const double ay = (k < p ? ly[k] : 1.0) + ymid[g];
  const int b = __builtin_amdgcn_readfirstlane(bot[g]);
  const int t = __builtin_amdgcn_readfirstlane(top[g]);
  LogAcc lfx, lfy;                 // sum f_i log|r|^2
  double ex = 0.0, ey = 0.0, e2 = 0.0;
  for (int j = b; j < t; ++j) {
    const double4 sj = src[j];
    const double dx = ax - sj.x;
    const double dy = ay - sj.y;
    const double d2 = __builtin_fma(dy, dy, dx * dx);
    double le_, lt_;
    log_split(d2, tab, le_, lt_);
    lfx.ae = __builtin_fma(sj.z, le_, lfx.ae);
    lfx.at = __builtin_fma(sj.z, lt_, lfx.at);
    lfy.ae = __builtin_fma(sj.w, le_, lfy.ae);
    lfy.at = __builtin_fma(sj.w, lt_, lfy.at);
    const double id2 = rcp_nr(d2);
    const double X = sj.x - c0x, Y = sj.y - c0y;
    const double fr = __builtin_fma(sj.z, dx, sj.w * dy) * id2;
    ex = __builtin_fma(-X, fr, ex);
    ey = __builtin_fma(-Y, fr, ey);
    e2 += fr;
    if constexpr (DIP) {
      const double4 e = sdip[j];
      const double qv = __builtin_fma(e.y, (dy - dx) * (dy + dx), -2.0 * e.z * dx * dy) * id2 * id2;
      ex = __builtin_fma(2.0, __builtin_fma(e.x * dx, id2, X * qv), ex);
      ey = __builtin_fma(2.0, __builtin_fma(e.x * dy, id2, Y * qv), ey);
      e2 = __builtin_fma(-2.0, qv, e2);
    }
  }
  if (k < PD) {
    // -f_i log|r| = -(1/2) f_i log|r|^2
    check[r * PD + k] = k < p ? __builtin_fma(-0.5, lfx.value(), ex) : 0.0;
    check[(rstride + r) * PD + k] = k < p ? __builtin_fma(-0.5, lfy.value(), ey) : 0.0;
    check[(2 * rstride + r) * PD + k] = k < p ? e2 : 0.0;
  }
}

// One item of k_eval_stokes: L2T of the three local fields + Stokeslet P2P,
// outputs (u, v, p).  FEW selects the (target slot, source group) lane layout
// of eval_few; otherwise lane = target and the source loop index is uniform.
template <int PD, bool CHECK, bool FEW, bool DIP>
__device__ __forceinline__ void stokes_item(
    int g, int t0, int t1, const int* __restrict__ tperm, const double* __restrict__ tsx,
    const double* __restrict__ tsy, const int* __restrict__ nlevel,
    const double* __restrict__ xmid, const double* __restrict__ ymid,
    const int* __restrict__ coll, const int2* __restrict__ crange, const int* __restrict__ bot, const int* __restrict__ top,
    const double4* __restrict__ src, const double4* __restrict__ sdip,
    const double* __restrict__ le, int64_t nstride,
    const double* __restrict__ ops, int64_t opsz, int p, const double2* __restrict__ tab,
    double c0x, double c0y, int64_t nt, double* __restrict__ out) {
  FewLanes F;
  if constexpr (FEW) {
    F = few_lanes(t1 - t0);
  } else {
    F.lg = 6;
    F.ng = 1;
    F.slot = threadIdx.x & 63;
    F.sg = 0;
    F.act = t0 + F.slot < t1;
  }
  const int t = t0 + F.slot;
  const bool act = F.act;
  const double x = act ? tsx[t] : 0.0;
  const double y = act ? tsy[t] : 0.0;
  double u = 0.0, v = 0.0, pr = 0.0;
  const int d = nlevel[g];
  if (d >= 2) {
    const double* lx = ops + d * opsz + ops_off_circles(PD);
    const double* ly = lx + PD;
    const double* a = le + (int64_t)g * PD;
    const double xr = x - xmid[g], yr = y - ymid[g];
    // the three fields are the potentials -log(d^2)/(4 pi) of their local
    // equivalent densities: one log per equivalent point serves all three
    double l0 = 0.0, l1 = 0.0, l2 = 0.0;
    for (int k = F.sg; k < p; k += F.ng) {
      const double dx = xr - lx[k], dy = yr - ly[k];
      const double d2 = __builtin_fma(dy, dy, dx * dx);
      const double a0 = a[k], a1 = a[nstride * PD + k], a2 = a[2 * nstride * PD + k];
      double e, lt;
      log_split(d2, tab, e, lt);
      const double lg = __builtin_fma(e, kLn2, lt);
      l0 = __builtin_fma(a0, lg, l0);
      l1 = __builtin_fma(a1, lg, l1);
      l2 = __builtin_fma(a2, lg, l2);
    }
    const double Dx = -kInv4Pi * l0, Dy = -kInv4Pi * l1, D2 = -kInv4Pi * l2;
    const double X = x - c0x, Y = y - c0y;
    u = kInv4Pi * __builtin_fma(X, D2, Dx);
    v = kInv4Pi * __builtin_fma(Y, D2, Dy);
    pr = kInv2Pi * D2;
  }
  // near field: the Stokeslet itself (stokes_fmm.py:102-184)
  double lfx = 0.0, lfy = 0.0;   // sum f log(d^2)
  double su = 0.0, sv = 0.0, sp = 0.0;
  double du = 0.0, dv = 0.0, dp = 0.0;   // double layer
  double dmin = 1.0;
  const CollLanes CL = coll_lanes(g, coll, crange);
  for (int qn = 0; qn < 9; ++qn) {
    const int b = __builtin_amdgcn_readlane(CL.b, qn);
    const int e = __builtin_amdgcn_readlane(CL.e, qn);
    if (e <= b) continue;
    // A coincident pair (d2 = 0, so dx = dy = 0) continues with d2 = 1: every
    // term with a factor dx or dy then vanishes exactly, the stresslet
    // pressure -g/d2 + g/d2 cancels, and only the log term is zeroed (under
    // CHECK; otherwise the target is flagged through dmin).  Only colleague
    // 4, the target's own box, can hold a source at the target (a point and
    // its box are found by the same midpoint tests for sources and targets):
    // the other colleagues' pairs skip the test (guard false).  Forces and
    // stresslet densities stay in the wave-uniform registers they load into.
    auto pairS = [&](auto guard, const double4& sj, int jj) {
      constexpr bool G = decltype(guard)::value;
      const double dx = x - sj.x, dy = y - sj.y;
      const double dxx = dx * dx;
      double d2 = __builtin_fma(dy, dy, dxx);
      const bool z = G && d2 == 0.0;
      if (G && !CHECK) dmin = __builtin_fmin(dmin, d2);
      if (G) d2 = z ? 1.0 : d2;
      const double id2 = rcp_nr(d2);
      double le_, lt_;
      log_split(d2, tab, le_, lt_);
      double lg = __builtin_fma(le_, kLn2, lt_);
      if (G && CHECK) lg = z ? 0.0 : lg;
      lfx = __builtin_fma(sj.z, lg, lfx);
      lfy = __builtin_fma(sj.w, lg, lfy);
      const double rf = __builtin_fma(dx, sj.z, dy * sj.w) * id2;
      su = __builtin_fma(dx, rf, su);
      sv = __builtin_fma(dy, rf, sv);
      sp += rf;
      if constexpr (DIP) {   // the stresslet itself
        // w2 = 2 (r.n)(r.sigma)/d2^2, with 2 (r.n)(r.sigma) = g d2 +
        // a (r_x^2 - r_y^2) + 2 b r_x r_y; du, dv carry twice the sums
        const double4 e4 = sdip[jj];
        const double gg = e4.x, aa = e4.y, bb = e4.z;
        const double xy = dx * dy;
        const double t2 = __builtin_fma(-dy, dy, dxx);
        const double w2 =
            __builtin_fma(gg, d2, __builtin_fma(aa, t2, bb * (xy + xy))) * (id2 * id2);
        du = __builtin_fma(w2, dx, du);
        dv = __builtin_fma(w2, dy, dv);
        dp += __builtin_fma(-gg, id2, w2);
      }
    };
    auto colleague = [&](auto guard) {
      int j = b + F.sg;
      if constexpr (FEW) {   // four per-lane loads in flight
        for (; j + 3 * F.ng < e; j += 4 * F.ng) {
          double4 s4[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) s4[u] = src[j + u * F.ng];
#pragma unroll
          for (int u = 0; u < 4; ++u) pairS(guard, s4[u], j + u * F.ng);
        }
      }
      for (; j < e; j += F.ng) pairS(guard, src[j], j);
    };
    if (qn == 4) colleague(std::true_type{});
    else colleague(std::false_type{});
  }
  u += kInv4Pi * (-0.5 * lfx + su);
  v += kInv4Pi * (-0.5 * lfy + sv);
  pr += kInv2Pi * sp;
  if constexpr (DIP) {
    u += kInv2Pi * du;
    v += kInv2Pi * dv;
    pr += 2.0 * kInv2Pi * dp;
  }
  if constexpr (FEW) {
    u = slot_sum(u, F.lg);
    v = slot_sum(v, F.lg);
    pr = slot_sum(pr, F.lg);
    if (!CHECK)
      for (int off = 1 << F.lg; off < 64; off <<= 1)
        dmin = __builtin_fmin(dmin, __shfl_xor(dmin, off, 64));
  }
  if (act && F.sg == 0) {
    const int64_t o = tperm[t];
    if (!CHECK && dmin == 0.0) u = v = pr = __builtin_nan("");
    out[o] = u;
    out[nt + o] = v;
    out[2 * nt + o] = pr;
  }
}

// L2T of the three local fields + Stokeslet (and stresslet) P2P, outputs (u, v, p).
// The Stokeslet-only instance is held to 96 VGPRs (5 waves per SIMD instead of
// 4, no spills; 2% faster on 2M clustered points); with stresslets the same
// bound would spill, so that instance is left unconstrained.
template <int PD, bool CHECK, bool DIP>
__global__ __launch_bounds__(64 * kEvalWaves) __attribute__((amdgpu_waves_per_eu(DIP ? 1 : 5)))
void k_eval_stokes(
    const int4* __restrict__ items, const int* __restrict__ nitems,
    const int* __restrict__ evalnode, const int* __restrict__ toff,
    const int* __restrict__ tperm, const double* __restrict__ tsx, const double* __restrict__ tsy,
    const int* __restrict__ nlevel, const double* __restrict__ xmid,
    const double* __restrict__ ymid, const int* __restrict__ coll, const int2* __restrict__ crange, const int* __restrict__ bot,
    const int* __restrict__ top, const double4* __restrict__ src,
    const double4* __restrict__ sdip,
    const double* __restrict__ le, int64_t nstride, const double* __restrict__ ops, int64_t opsz,
    int p, const double2* __restrict__ logtab, double c0x, double c0y, int64_t nt,
    double* __restrict__ out) {
  __shared__ double2 tab[kLogTableSize];
  const int ni = *nitems;
  load_logtab(tab, logtab);
  __syncthreads();
  const ItemRange R = xcd_items(ni);
  for (int ii = R.first; ii < R.last; ii += R.step) {
    const int4 it = items[ii];
    // wave-uniform (ii is): in scalar registers, so loads indexed by them are scalar
    const int g = __builtin_amdgcn_readfirstlane(it.x);
    const int t0 = __builtin_amdgcn_readfirstlane(it.y), t1 = __builtin_amdgcn_readfirstlane(it.z);
    if (t1 - t0 <= kFewTargets)
      stokes_item<PD, CHECK, true, DIP>(g, t0, t1, tperm, tsx, tsy, nlevel, xmid, ymid, coll, crange,
                                        bot, top, src, sdip, le, nstride, ops, opsz, p, tab, c0x,
                                        c0y, nt, out);
    else
      stokes_item<PD, CHECK, false, DIP>(g, t0, t1, tperm, tsx, tsy, nlevel, xmid, ymid, coll, crange,
                                         bot, top, src, sdip, le, nstride, ops, opsz, p, tab,
                                         c0x, c0y, nt, out);
  }
}

inline int eval_grid(int64_t icap) {
  return (int)(ceil_div(ceil_div(icap, kNumXcd) + 1, kEvalWaves) * kNumXcd);
}

// Selects the k_eval instance of (PD, kernel, outputs, self-exclusion).
typedef decltype(&k_eval<32, KER_LAPLACE, 1, false, false, false>) EvalKernelFn;
template <int PD, int KER>
EvalKernelFn eval_kernel_k(int nout, bool check) {
  if (nout == 1)
    return check ? k_eval<PD, KER, 1, true, false, false> : k_eval<PD, KER, 1, false, false, false>;
  return check ? k_eval<PD, KER, 3, true, false, false> : k_eval<PD, KER, 3, false, false, false>;
}
// Laplace single + double layer sources
EvalKernelFn eval_kernel_dipole(int PD, bool check) {
  if (PD == 32)
    return check ? k_eval<32, KER_LAPLACE, 1, true, true, false>
                 : k_eval<32, KER_LAPLACE, 1, false, true, false>;
  return check ? k_eval<64, KER_LAPLACE, 1, true, true, false>
               : k_eval<64, KER_LAPLACE, 1, false, true, false>;
}
// Chebyshev (black-box) expansions: potential only
template <int PD>
EvalKernelFn eval_kernel_cheb_pd(int ker, bool check) {
  if (ker == KER_LAPLACE)
    return check ? k_eval<PD, KER_LAPLACE, 1, true, false, true>
                 : k_eval<PD, KER_LAPLACE, 1, false, false, true>;
  return check ? k_eval<PD, KER_MODHELM, 1, true, false, true>
               : k_eval<PD, KER_MODHELM, 1, false, false, true>;
}
EvalKernelFn eval_kernel_cheb(int PD, int ker, bool check) {
  return PD == 32 ? eval_kernel_cheb_pd<32>(ker, check) : eval_kernel_cheb_pd<64>(ker, check);
}
EvalKernelFn eval_kernel(int PD, int ker, int nout, bool check) {
  if (ker == KER_LAPLACE)
    return PD == 32 ? eval_kernel_k<32, KER_LAPLACE>(nout, check)
                    : eval_kernel_k<64, KER_LAPLACE>(nout, check);
  return PD == 32 ? eval_kernel_k<32, KER_MODHELM>(nout, check)
                  : eval_kernel_k<64, KER_MODHELM>(nout, check);
}

// Algorithmic work counts (for the roofline accounting).
__global__ __launch_bounds__(256) void k_eval_stats(const int* __restrict__ evalnode, int ne,
                                                   const int* __restrict__ tcnt,
                                                   const int* __restrict__ coll,
                                                   const int* __restrict__ bot,
                                                   const int* __restrict__ top,
                                                   const int* __restrict__ nlevel,
                                                   unsigned long long* __restrict__ out) {
  int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= ne) return;
  const int c = tcnt[b];
  if (c == 0) return;
  const int g = evalnode[b];
  long long ns = 0;
  for (int q = 0; q < 9; ++q) {
    int cc = coll[(int64_t)g * 9 + q];
    if (cc >= 0) ns += top[cc] - bot[cc];
  }
  atomicAdd(&out[0], (unsigned long long)((long long)c * ns));
  if (nlevel[g] >= 2) atomicAdd(&out[1], (unsigned long long)c);
  atomicAdd(&out[2], (unsigned long long)c);
}

__global__ __launch_bounds__(256) void k_tree_stats(int NN, const int* __restrict__ nlevel,
                                                   const int* __restrict__ flags,
                                                   const int* __restrict__ m2l,
                                                   const int* __restrict__ row,
                                                   const int* __restrict__ child,
                                                   const int* __restrict__ bot,
                                                   const int* __restrict__ top,
                                                   unsigned long long* __restrict__ out) {
  int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= NN) return;
  const int d = nlevel[g], f = flags[g];
  if (d < 2) return;
  if (!(f & NF_FAKE)) {
    int np = 0;
    for (int s = 0; s < kNumM2L; ++s) np += m2l[(int64_t)g * kNumM2L + s] >= 0;
    atomicAdd(&out[3], (unsigned long long)np);
    atomicAdd(&out[6], 1ull);
    if (d >= 3) atomicAdd(&out[8], 1ull);
  }
  if ((f & NF_UP) && row[g] >= 0) atomicAdd(&out[4], (unsigned long long)(top[g] - bot[g]));
  if ((f & NF_CHILDREN) && row[g] >= 0 && row[child[g]] >= 0) atomicAdd(&out[7], 1ull);
}

inline unsigned grid256(int64_t n) { return (unsigned)(n <= 0 ? 1 : ceil_div(n, 256)); }

}  // namespace

// ---------------------------------------------------------------------------
// Host side of the passes.
static void check_ops(const Plan& P) {
  const Tree& t = P.tree;
  FMM_REQUIRE(t.built, FMM_E_STATE, "no tree: call fmm2d_tree_build first");
  for (int d = 2; d < t.L; ++d)
    FMM_REQUIRE(d < P.ops_levels && P.ops_set[d], FMM_E_STATE,
                "operators for level %d not set (fmm2d_set_operators)", d);
}

// M2L launch tables of the current tree: all levels >= 2 and, on adaptive
// trees, the column order (k_m2l_order).
struct M2LTables {
  LevelTable lt;
  M2LOrderTable ot;
  int order_blocks;
  int perm_len;
  bool adaptive;
};
static M2LTables m2l_tables(const Plan& P) {
  const Tree& t = P.tree;
  M2LTables T{};
  LevelTable& lt = T.lt;
  for (int d = 2; d < t.L; ++d) {
    lt.d[lt.nlev] = d;
    lt.g0[lt.nlev] = (int)t.lev[d].node_off;
    lt.ngroups[lt.nlev] = (int)(t.lev[d].n_nodes / 4);
    T.adaptive = T.adaptive || t.lev[d].n_nodes != ((int64_t)1 << (2 * d));
    ++lt.nlev;
  }
  int nb = 0;
  for (int li = 0; li < lt.nlev; ++li) {
    lt.boff[li] = nb;
    nb += 4 * (int)ceil_div(lt.ngroups[li], kM2LCols);
  }
  lt.boff[lt.nlev] = nb;
  lt.perm = nullptr;
  if (T.adaptive) {
    M2LOrderTable& ot = T.ot;
    ot.nlev = lt.nlev;
    int ob = 0, po = 0;
    for (int li = 0; li < lt.nlev; ++li) {
      ot.g0[li] = lt.g0[li];
      ot.ngroups[li] = lt.ngroups[li];
      ot.nch[li] = (int)ceil_div(lt.ngroups[li], kM2LOrderChunk);
      ot.boff[li] = ob;
      ot.poff[li] = po;
      lt.poff[li] = po;
      ob += 4 * ot.nch[li];
      po += 4 * lt.ngroups[li];
    }
    ot.boff[lt.nlev] = ob;
    T.order_blocks = ob;
    T.perm_len = po;
  }
  return T;
}
// The column order, once per tree (on the stream that runs the M2L).
static void m2l_prepare(Plan& P, const M2LTables& T, hipStream_t s) {
  const Tree& t = P.tree;
  const int nd = P.nd;
  P.ple.ensure((t.n_nodes + 1) * P.PD * nd * sizeof(double));
  if (!T.adaptive || P.perm_tree == t.build_id) return;
  P.m2lperm.ensure((T.perm_len + 1) * sizeof(int));
  hipLaunchKernelGGL(k_m2l_order, dim3(T.order_blocks), dim3(kM2LOrderThreads), 0, s, T.ot,
                     t.flags.as<int>(),
                     t.m2l.as<int>(), P.m2lperm.as<int>());
  FMM_LAUNCH_CHECK();
  P.perm_tree = t.build_id;
}
// M2L of the table's levels [l0, l1).
// The mixed forms run the M2L on FP32 matrix cores, except for Chebyshev
// (black-box) plans: there the FP32 accumulation error of a slot's product
// alone reaches ~1e-7 of the potential (no dominant mode to split off, unlike
// the KIFMM's row means), so they keep the FP64 M2L.
// Black-box plans and Stokes plans with stresslets keep the FP64 M2L under the
// mixed forms: neither expansion has a dominant constant mode for the row-mean
// split to take out, and an FP32 product alone costs them 1e-7 to 4e-7.
static bool m2l_in_fp32(const Plan& P) {
  return P.precision != PREC_FP64 && !P.cheb && !(P.kernel_id == KER_STOKES && P.have_dip);
}

static void m2l_launch(Plan& P, const M2LTables& T, int l0, int l1, hipStream_t s) {
  const Tree& t = P.tree;
  const int PD = P.PD;
  const int64_t opsz = ops_per_level(PD);
  const int nd = P.nd;
  const int64_t rstride = t.n_rows + 1, nstride = t.n_nodes + 1;
  LevelTable lt{};
  lt.nlev = l1 - l0;
  if (lt.nlev <= 0) return;
  for (int i = 0; i < lt.nlev; ++i) {
    lt.d[i] = T.lt.d[l0 + i];
    lt.g0[i] = T.lt.g0[l0 + i];
    lt.ngroups[i] = T.lt.ngroups[l0 + i];
    lt.boff[i] = T.lt.boff[l0 + i] - T.lt.boff[l0];
    lt.poff[i] = T.lt.poff[l0 + i];
  }
  lt.boff[lt.nlev] = T.lt.boff[l1] - T.lt.boff[l0];
  lt.perm = T.adaptive ? P.m2lperm.as<int>() : nullptr;
  int nb = lt.boff[lt.nlev];
  if (nb == 0) return;
  if (m2l_in_fp32(P)) {
    // narrow tiles of 16 boxes, one block each
    LevelTable lw = lt;
    int nbw = 0;
    for (int li = 0; li < lw.nlev; ++li) {
      lw.boff[li] = nbw;
      nbw += 4 * (int)ceil_div(lw.ngroups[li], kM2LWCols);
    }
    lw.boff[lw.nlev] = nbw;
    lw.split = 1;
    const bool frame = P.kernel_id == KER_STOKES;
    constexpr int NW = kM2LF32Waves;
    auto kern = PD == 32
                    ? (frame ? k_m2l_waves_f32<32, true, NW> : k_m2l_waves_f32<32, false, NW>)
                    : (frame ? k_m2l_waves_f32<64, true, NW> : k_m2l_waves_f32<64, false, NW>);
    FMM_REQUIRE(frame || P.mult32.bytes >= (size_t)(rstride * PD * nd * sizeof(float)),
                FMM_E_INTERNAL, "FP32 multipole copy missing");
    hipLaunchKernelGGL(kern, dim3(nbw, nd), dim3(64 * NW), 0, s, lw, t.flags.as<int>(),
                       t.m2l.as<int>(), P.mult.as<double>(),
                       frame ? nullptr : P.mult32.as<float>(), P.qs.as<double>(),
                       P.ops.as<double>(), opsz, P.ple.as<double>(), rstride, nstride,
                       t.xmid.as<double>(), t.ymid.as<double>(), t.xmin + 0.5 * t.w0,
                       t.ymin + 0.5 * t.w0);
    FMM_LAUNCH_CHECK();
    return;
  }
  if (P.m2l_mode == FMM2D_M2L_DENSE && PD <= kM2LWavesF64MaxPD) {
    LevelTable lw = lt;
    int nbw = 0;
    for (int li = 0; li < lw.nlev; ++li) {
      lw.boff[li] = nbw;
      nbw += 4 * (int)ceil_div(lw.ngroups[li], kM2LWCols);
    }
    lw.boff[lw.nlev] = nbw;
    lw.split = 1;
    hipLaunchKernelGGL(k_m2l_waves_f64<32>, dim3(nbw, nd), dim3(64 * kM2LWWaves), 0, s, lw,
                       t.flags.as<int>(),
                       t.m2l.as<int>(), P.mult.as<double>(), P.ops.as<double>(), opsz,
                       P.ple.as<double>(), rstride, nstride);
    FMM_LAUNCH_CHECK();
    return;
  }
  // fewer than 3 tiles per CU: split the tiles' phases over several blocks
  lt.split = (int64_t)nb * nd < 3 * kNumCUs ? kM2LSplitF64 : 1;
  if (lt.split > 1) {
    // zero with a fill kernel (no memset node: the call may sit in a captured graph)
    fill_i32(P.ple.as<int>(), 2 * nstride * PD * nd, 0, s);
    nb *= lt.split;
  }
  constexpr int RS = kM2LRowSplit;
  auto kern = PD == 32 ? k_m2l_all<32, RS> : k_m2l_all<64, RS>;
  hipLaunchKernelGGL(kern, dim3(nb, nd), dim3(kM2LThreads * RS), 0, s, lt, t.flags.as<int>(),
                     t.m2l.as<int>(), P.mult.as<double>(), P.ops.as<double>(), opsz,
                     P.ple.as<double>(), rstride, nstride, P.m2l_mode == FMM2D_M2L_SVD ? 1 : 0);
  FMM_LAUNCH_CHECK();
}

// dvec: dipole directions (x [n] then y [n]) when the sources carry single
// and double layer densities (tau: charges [n] then dipole strengths [n]).
static void upward(Plan& P, const double* tau, const double* dvec, hipStream_t s) {
  Tree& t = P.tree;
  check_ops(P);
  const int PD = P.PD;
  const int64_t opsz = ops_per_level(PD);
  const bool stokes = P.kernel_id == KER_STOKES;
  const int nd = stokes ? kStokesFields : 1;
  P.nd = nd;
  P.tau_s.ensure(t.n * sizeof(double));
  P.src.ensure(t.n * sizeof(double4));
  P.have_srcf = false;
  if (stokes && dvec) {
    P.sdip.ensure((t.n + 1) * sizeof(double4));
    hipLaunchKernelGGL(k_pack_stokes_dipole, dim3(grid256(t.n)), dim3(256), 0, s, tau, dvec,
                       t.perm.as<uint32_t>(), t.sx.as<double>(), t.sy.as<double>(), t.n,
                       P.src.as<double4>(), P.sdip.as<double4>());
  } else if (stokes) {
    hipLaunchKernelGGL(k_pack_stokes, dim3(grid256(t.n)), dim3(256), 0, s, tau,
                       t.perm.as<uint32_t>(), t.sx.as<double>(), t.sy.as<double>(), t.n,
                       P.src.as<double4>());
  } else if (dvec) {
    P.dip.ensure((t.n + 1) * sizeof(double2));
    hipLaunchKernelGGL(k_pack_dipole, dim3(grid256(t.n)), dim3(256), 0, s, tau, dvec,
                       t.perm.as<uint32_t>(), t.sx.as<double>(), t.sy.as<double>(), t.n,
                       P.tau_s.as<double>(), P.src.as<double4>(), P.dip.as<double2>());
  } else {
    hipLaunchKernelGGL(k_pack_sources, dim3(grid256(t.n)), dim3(256), 0, s, tau,
                       t.perm.as<uint32_t>(), t.sx.as<double>(), t.sy.as<double>(), t.n,
                       P.tau_s.as<double>(), P.src.as<double4>());
  }
  FMM_LAUNCH_CHECK();
  P.have_dip = dvec != nullptr;
  P.up_done = true;
  const int64_t nr = t.n_rows;
  const int64_t rstride = nr + 1;   // rows per density
  P.check.ensure(rstride * PD * nd * sizeof(double));
  P.mult.ensure(rstride * PD * nd * sizeof(double));
  P.qs.ensure(rstride * nd * sizeof(double));
  // the FP32 M2L's operand copy of the multipoles (the Stokeslet's frame
  // fields form theirs from the FP64 rows)
  float* m32 = nullptr;
  if (m2l_in_fp32(P) && !stokes) {
    P.mult32.ensure(rstride * PD * nd * sizeof(float));
    m32 = P.mult32.as<float>();
  }
  if (t.L <= 2 || nr == 0) {
    P.have_exp = true;
    return;
  }
  P.mark(FMM2D_PHASE_S2M, false, s);
  if (t.n_up > 0 && stokes) {
    auto kern = dvec ? (PD == 32 ? k_s2m_stokes<32, true> : k_s2m_stokes<64, true>)
                     : (PD == 32 ? k_s2m_stokes<32, false> : k_s2m_stokes<64, false>);
    hipLaunchKernelGGL(kern, dim3(ceil_div(t.n_up, kS2MWaves)), dim3(64 * kS2MWaves), 0, s,
                       (int)t.n_up, t.upnodes.as<int>(), t.nlevel.as<int>(),
                       t.bot.as<int>(), t.top.as<int>(), t.row.as<int>(), t.xmid.as<double>(),
                       t.ymid.as<double>(), P.src.as<double4>(), P.ops.as<double>(), opsz, P.p,
                       P.logtab.as<double2>(), t.xmin + 0.5 * t.w0, t.ymin + 0.5 * t.w0,
                       rstride, dvec ? P.sdip.as<double4>() : nullptr, P.check.as<double>());
    FMM_LAUNCH_CHECK();
  } else if (t.n_up > 0 && P.cheb) {
    auto kern = PD == 32 ? k_s2m_cheb<32> : k_s2m_cheb<64>;
    hipLaunchKernelGGL(kern, dim3(t.n_up), dim3(64), 0, s, t.upnodes.as<int>(),
                       t.nlevel.as<int>(), t.bot.as<int>(), t.top.as<int>(), t.row.as<int>(),
                       t.xmid.as<double>(), t.ymid.as<double>(), P.src.as<double4>(),
                       P.ops.as<double>(), opsz, P.cheb, P.check.as<double>());
    FMM_LAUNCH_CHECK();
  } else if (t.n_up > 0) {
    auto kern = dvec ? (PD == 32 ? k_s2m<32, KER_LAPLACE, true> : k_s2m<64, KER_LAPLACE, true>)
                : P.kernel_id == KER_LAPLACE
                    ? (PD == 32 ? k_s2m<32, KER_LAPLACE, false> : k_s2m<64, KER_LAPLACE, false>)
                    : (PD == 32 ? k_s2m<32, KER_MODHELM, false> : k_s2m<64, KER_MODHELM, false>);
    hipLaunchKernelGGL(kern, dim3(ceil_div(t.n_up, kS2MWaves)), dim3(64 * kS2MWaves), 0, s,
                       (int)t.n_up, t.upnodes.as<int>(), t.nlevel.as<int>(),
                       t.bot.as<int>(), t.top.as<int>(), t.row.as<int>(), t.xmid.as<double>(),
                       t.ymid.as<double>(), P.src.as<double4>(), P.ops.as<double>(), opsz, P.p,
                       P.logtab.as<double2>(), P.kparam, dvec ? P.dip.as<double2>() : nullptr,
                       P.check.as<double>());
    FMM_LAUNCH_CHECK();
  }
  P.mark(FMM2D_PHASE_S2M, true, s);
  for (int d = t.L - 1; d >= 2; --d) {
    const int64_t r0 = t.row_off[d];
    const int nrows = (int)(t.row_off[d + 1] - r0);
    if (nrows == 0) continue;
    const double* opsL = P.ops.as<double>() + d * opsz;
    if (nrows <= kSmallLevelRows) {
      auto ks = PD == 32 ? k_mult_level_small<32> : k_mult_level_small<64>;
      hipLaunchKernelGGL(ks, dim3(ceil_div(nrows, 16), nd), dim3(64 * (PD / 16)), 0, s, r0,
                         nrows, d < t.L - 1 ? 1 : 0, t.rowinfo.as<int2>(), P.check.as<double>(), opsL,
                         P.mult.as<double>(), P.qs.as<double>(), rstride, m32);
      FMM_LAUNCH_CHECK();
      continue;
    }
    auto kern = PD == 32 ? k_mult_level<32> : k_mult_level<64>;
    hipLaunchKernelGGL(kern, dim3(ceil_div(nrows, 64), nd), dim3(256), 0, s, r0, nrows,
                       d < t.L - 1 ? 1 : 0, t.rowinfo.as<int2>(), P.check.as<double>(), opsL,
                       P.mult.as<double>(), P.qs.as<double>(), rstride, m32);
    FMM_LAUNCH_CHECK();
  }
  P.have_exp = false;
}

static void downward(Plan& P, hipStream_t s) {
  Tree& t = P.tree;
  check_ops(P);
  const int PD = P.PD;
  const int64_t opsz = ops_per_level(PD);
  const int nd = P.nd;
  const int64_t rstride = t.n_rows + 1, nstride = t.n_nodes + 1;
  P.le.ensure(nstride * PD * nd * sizeof(double));
  P.ple.ensure(nstride * PD * nd * sizeof(double));
  fill_i32(P.le.as<int>(), 2 * nstride * PD * nd, 0, s);   // 0.0 (no memset node)
  if (t.L <= 2) {
    P.have_exp = true;
    return;
  }
  const M2LTables T = m2l_tables(P);
  m2l_prepare(P, T, s);
  P.mark(FMM2D_PHASE_M2L, false, s);
  m2l_launch(P, T, 0, T.lt.nlev, s);
  P.mark(FMM2D_PHASE_M2L, true, s);
  FrameFix fx{nullptr, nullptr, 0.0, 0.0};
  if (P.kernel_id == KER_STOKES && m2l_in_fp32(P))   // the FP32 M2L ran in box frames
    fx = FrameFix{t.xmid.as<double>(), t.ymid.as<double>(), t.xmin + 0.5 * t.w0,
                  t.ymin + 0.5 * t.w0};
  for (int d = 2; d < t.L; ++d) {
    const int ngroups = (int)(t.lev[d].n_nodes / 4);
    if (ngroups == 0) continue;
    const double* opsL = P.ops.as<double>() + d * opsz;
    if (4 * (int64_t)ngroups <= kSmallLevelRows) {
      auto ks = PD == 32 ? k_local_level_small<32> : k_local_level_small<64>;
      hipLaunchKernelGGL(ks, dim3(4 * ceil_div(ngroups, 16), nd), dim3(64 * (PD / 16)), 0, s,
                         (int)t.lev[d].node_off, ngroups, d >= 3 ? 1 : 0, t.flags.as<int>(),
                         t.parent.as<int>(), P.ple.as<double>(), opsL, opsL - opsz,
                         P.le.as<double>(), nstride, fx);
      FMM_LAUNCH_CHECK();
      continue;
    }
    auto kern = PD == 32 ? k_local_level<32> : k_local_level<64>;
    hipLaunchKernelGGL(kern, dim3(4 * ceil_div(ngroups, 64), nd), dim3(256), 0, s,
                       (int)t.lev[d].node_off, ngroups, d >= 3 ? 1 : 0, t.flags.as<int>(),
                       t.parent.as<int>(), P.ple.as<double>(), opsL, opsL - opsz,
                       P.le.as<double>(), nstride, fx);
    FMM_LAUNCH_CHECK();
  }
  P.have_exp = true;
}

// Target-side scratch of one evaluation: target boxes, per-box counts and
// offsets, targets in leaf order and the work items.
struct EvalScratch {
  int *tnode, *tcnt, *toff, *tcur, *tperm, *icnt, *ioff, *ntot, *partials;
  double *tsx, *tsy;
  int4* items;
};

constexpr int64_t eval_item_cap(int64_t ne, int64_t nt) { return ne + nt / kTgtChunk + 1; }

// The plan's own scratch, grown to nt targets and ne evaluation boxes.
static EvalScratch plan_scratch(Plan& P, int64_t nt, int64_t ne) {
  P.tnode.ensure(nt * sizeof(int));
  P.tcnt.ensure((ne + 1) * sizeof(int));
  P.toff.ensure((ne + 1) * sizeof(int));
  P.tcur.ensure((nt + 1) * sizeof(int));   // per-target rank in its box
  P.tperm.ensure(nt * sizeof(int));
  P.tsx.ensure(nt * sizeof(double));
  P.tsy.ensure(nt * sizeof(double));
  P.icnt.ensure((ne + 1) * sizeof(int));
  P.ioff.ensure((ne + 1) * sizeof(int));
  P.items.ensure(eval_item_cap(ne, nt) * sizeof(int4));
  P.ntot.ensure(4 * sizeof(int));
  EvalScratch w;
  w.tnode = P.tnode.as<int>();
  w.tcnt = P.tcnt.as<int>();
  w.toff = P.toff.as<int>();
  w.tcur = P.tcur.as<int>();
  w.tperm = P.tperm.as<int>();
  w.icnt = P.icnt.as<int>();
  w.ioff = P.ioff.as<int>();
  w.ntot = P.ntot.as<int>();
  w.partials = nullptr;   // scan partials from the plan's PrimWork
  w.tsx = P.tsx.as<double>();
  w.tsy = P.tsy.as<double>();
  w.items = P.items.as<int4>();
  return w;
}

// Leaf-relative FP32 sources and per-leaf charge sums of the mixed-precision
// near field; formed once per set of expansions, by the downward pass, so
// that evaluations never allocate plan state.
static bool uses_mixed_near_field(const Plan& P) {
  return P.precision == PREC_MIXED && P.kernel_id == KER_LAPLACE;
}

static LevelConsts level_consts(const Tree& t) {
  LevelConsts lc{};
  for (int d = 0; d < t.L; ++d) {
    lc.invw[d] = 1.0 / t.lev[d].width;
    lc.logw2[d] = std::log(t.lev[d].width * t.lev[d].width);
  }
  return lc;
}

static void prep_mixed(Plan& P, hipStream_t s) {
  Tree& t = P.tree;
  const LevelConsts lc = level_consts(t);
  P.srcf.ensure((t.n + 1) * sizeof(f4));
  P.tsum.ensure((t.n_nodes + 1) * sizeof(double));
  if (t.n_eval > 0)
    hipLaunchKernelGGL(k_prep_mixed, dim3(t.n_eval), dim3(64), 0, s, t.evalnode.as<int>(),
                       t.nlevel.as<int>(), t.bot.as<int>(), t.top.as<int>(),
                       t.xmid.as<double>(), t.ymid.as<double>(), P.src.as<double4>(), lc,
                       P.srcf.as<f4>(), P.tsum.as<double>());
  FMM_LAUNCH_CHECK();
  P.have_srcf = true;
}

static void evaluate(Plan& P, const double* tx, const double* ty, int64_t nt, int ev, int check,
                     double* out, hipStream_t s, const EvalScratch* given) {
  Tree& t = P.tree;
  FMM_REQUIRE(P.have_exp, FMM_E_STATE, "no expansions: call fmm2d_upward and fmm2d_downward first");
  const bool stokes = P.kernel_id == KER_STOKES;
  FMM_REQUIRE(stokes ? (ev == EVAL_STOKES || ev == EVAL_STOKES_DIPOLE)
                     : (ev == EVAL_POT || ev == EVAL_GRAD ||
                        (ev == EVAL_DIPOLE && P.kernel_id == KER_LAPLACE)),
              FMM_E_ARG, "evaluator %d does not belong to kernel %d", ev, P.kernel_id);
  FMM_REQUIRE((ev != EVAL_DIPOLE && ev != EVAL_STOKES_DIPOLE) || P.have_dip, FMM_E_STATE,
              "the dipole evaluators need expansions from fmm2d_upward_dipole");
  FMM_REQUIRE(!P.cheb || ev == EVAL_POT, FMM_E_ARG,
              "a Chebyshev (black-box) plan evaluates the potential only");
  FMM_REQUIRE(!(stokes && ev == EVAL_STOKES && P.have_dip), FMM_E_STATE,
              "the expansions hold double-layer densities: evaluate with the Stokes dipole "
              "evaluator");
  FMM_REQUIRE(nt >= 0 && nt < ((int64_t)1 << 31), FMM_E_LIMIT, "too many targets");
  if (nt == 0) return;
  const int64_t ne = t.n_eval;
  const EvalScratch w = given ? *given : plan_scratch(P, nt, ne);
  P.last_tcnt = w.tcnt;
  const int64_t icap = eval_item_cap(ne, nt);
  fill_i32(w.tcnt, ne + 1, 0, s);
  LocateTable lt{};
  lt.L = t.L;
  lt.full_depth = t.complete ? t.L - 1 : -1;
  lt.xmin = t.xmin;
  lt.ymin = t.ymin;
  lt.w0 = t.w0;
  for (int d = 0; d <= t.L; ++d)
    lt.levoff[d] = (int)(d < t.L ? t.lev[d].node_off : t.n_nodes);
  for (int d = 0; d < kMaxLevels; ++d)
    lt.tab_off[d] = d < (int)t.nidx_off.size() ? t.nidx_off[d] : -1;
  lt.tab = t.nidx.as<int>();
  lt.code = t.code.as<uint64_t>();
  hipLaunchKernelGGL(k_locate_count, dim3(grid256(nt)), dim3(256), 0, s, tx, ty, nt, lt,
                     t.flags.as<int>(), t.child.as<int>(), t.xmid.as<double>(),
                     t.ymid.as<double>(), t.evalid.as<int>(), w.tnode,
                     w.tcur, w.tcnt);
  FMM_LAUNCH_CHECK();
  if (ne <= kBoxOffsetsMax) {
    hipLaunchKernelGGL(k_box_offsets, dim3(1), dim3(1024), 0, s, w.tcnt, (int)ne, w.toff, w.ioff,
                       w.ntot);
    FMM_LAUNCH_CHECK();
  } else {
    // toff has ne + 1 entries: scanning ne + 1 counts (the last is zero) gives toff[ne] = nt.
    exclusive_scan_i32(w.tcnt, w.toff, ne + 1, nullptr, P.pw, s, w.partials);
    hipLaunchKernelGGL(k_item_counts, dim3(grid256(ne)), dim3(256), 0, s, w.tcnt, (int)ne,
                       w.icnt);
    FMM_LAUNCH_CHECK();
    exclusive_scan_i32(w.icnt, w.ioff, ne, w.ntot, P.pw, s, w.partials);
  }
  hipLaunchKernelGGL(k_target_scatter, dim3(grid256(nt)), dim3(256), 0, s, tx, ty, nt,
                     w.tnode, w.tcur, w.toff, w.tperm,
                     w.tsx, w.tsy);
  FMM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_make_items, dim3(grid256(ne)), dim3(256), 0, s, w.ioff, (int)ne,
                     t.evalnode.as<int>(), w.toff, w.items);
  FMM_LAUNCH_CHECK();
  const int PD = P.PD;
  const int64_t opsz = ops_per_level(PD);
  if (stokes) {
    const bool dl = ev == EVAL_STOKES_DIPOLE;
    auto sk = dl ? (PD == 32 ? (check ? k_eval_stokes<32, true, true> : k_eval_stokes<32, false, true>)
                             : (check ? k_eval_stokes<64, true, true> : k_eval_stokes<64, false, true>))
                 : (PD == 32 ? (check ? k_eval_stokes<32, true, false> : k_eval_stokes<32, false, false>)
                             : (check ? k_eval_stokes<64, true, false> : k_eval_stokes<64, false, false>));
    P.mark(FMM2D_PHASE_EVAL_KERNEL, false, s);
    hipLaunchKernelGGL(sk, dim3(eval_grid(icap)), dim3(64 * kEvalWaves), 0, s, w.items, w.ntot,
                       t.evalnode.as<int>(), w.toff, w.tperm,
                       w.tsx, w.tsy, t.nlevel.as<int>(),
                       t.xmid.as<double>(), t.ymid.as<double>(), t.coll.as<int>(), t.crange.as<int2>(),
                       t.bot.as<int>(), t.top.as<int>(), P.src.as<double4>(),
                       dl ? P.sdip.as<double4>() : nullptr, P.le.as<double>(),
                       t.n_nodes + 1, P.ops.as<double>(), opsz, P.p, P.logtab.as<double2>(),
                       t.xmin + 0.5 * t.w0, t.ymin + 0.5 * t.w0, nt, out);
    FMM_LAUNCH_CHECK();
    P.mark(FMM2D_PHASE_EVAL_KERNEL, true, s);
    return;
  }
  // few targets per evaluation box on average: the FP64 near field with LDS
  // staging (below) is faster than the FP32 one, whose per-lane source loads
  // are latency-bound on short items, and it is the more accurate of the two
  const bool few_targets = ne > 0 && nt < (int64_t)kFewTargets * ne;
  if (P.precision == PREC_MIXED && ev == EVAL_POT && P.kernel_id == KER_LAPLACE && !few_targets) {
    const LevelConsts lc = level_consts(t);
    if (!P.have_srcf) {
      FMM_REQUIRE(!given, FMM_E_STATE,
                  "the mixed-precision near field is prepared by fmm2d_downward: call it "
                  "before an evaluation that allocates nothing");
      prep_mixed(P, s);
    }
    auto mk = P.cheb ? (check ? (PD == 32 ? k_eval_mixed<32, true, true> : k_eval_mixed<64, true, true>)
                              : (PD == 32 ? k_eval_mixed<32, false, true> : k_eval_mixed<64, false, true>))
                     : (check ? (PD == 32 ? k_eval_mixed<32, true, false> : k_eval_mixed<64, true, false>)
                              : (PD == 32 ? k_eval_mixed<32, false, false> : k_eval_mixed<64, false, false>));
    P.mark(FMM2D_PHASE_EVAL_KERNEL, false, s);
    hipLaunchKernelGGL(mk, dim3(eval_grid(icap)), dim3(64 * kEvalWaves), 0, s, w.items, w.ntot,
                       t.evalnode.as<int>(), w.toff, w.tperm,
                       w.tsx, w.tsy, t.nlevel.as<int>(),
                       t.flags.as<int>(), t.xmid.as<double>(), t.ymid.as<double>(),
                       t.coll.as<int>(), t.crange.as<int2>(), t.bot.as<int>(), t.top.as<int>(), P.src.as<double4>(),
                       P.srcf.as<f4>(), P.tsum.as<double>(), lc, P.le.as<double>(),
                       P.ops.as<double>(), opsz, P.cheb ? P.cheb : P.p, P.logtab.as<double2>(),
                       out);
    FMM_LAUNCH_CHECK();
    P.mark(FMM2D_PHASE_EVAL_KERNEL, true, s);
    return;
  }
  auto kern = P.cheb ? eval_kernel_cheb(PD, P.kernel_id, check)
              : ev == EVAL_DIPOLE ? eval_kernel_dipole(PD, check)
                                  : eval_kernel(PD, P.kernel_id, ev == EVAL_POT ? 1 : 3, check);
  P.mark(FMM2D_PHASE_EVAL_KERNEL, false, s);
  // few targets per evaluation box on average: the few-target items stage
  // their colleague sources through LDS (the extra LDS would only cost the
  // full-wave items occupancy)
  const int scap = (ev != EVAL_DIPOLE && few_targets) ? kStageSources : 0;
  const size_t dyn = (size_t)scap * kEvalWaves * sizeof(double4);
  hipLaunchKernelGGL(kern, dim3(eval_grid(icap)), dim3(64 * kEvalWaves), dyn, s, w.items, w.ntot,
                     t.evalnode.as<int>(), w.toff, w.tperm, w.tsx,
                     w.tsy, t.nlevel.as<int>(), t.xmid.as<double>(),
                     t.ymid.as<double>(), t.coll.as<int>(), t.crange.as<int2>(), t.bot.as<int>(), t.top.as<int>(),
                     P.src.as<double4>(), ev == EVAL_DIPOLE ? P.dip.as<double2>() : nullptr,
                     P.le.as<double>(), P.ops.as<double>(), opsz, P.cheb ? P.cheb : P.p,
                     P.logtab.as<double2>(), P.kparam, nt, out, scap);
  FMM_LAUNCH_CHECK();
  P.mark(FMM2D_PHASE_EVAL_KERNEL, true, s);
}

// ---------------------------------------------------------------------------
// Evaluation that allocates nothing: the target-side scratch lives in one
// file-scope device array, sized for kNoAllocMaxTargets targets and
// kNoAllocMaxBoxes evaluation boxes (each GPU the library runs on holds its
// own copy until the process ends).  Every call on a GPU shares that GPU's
// copy, so calls that overlap on one GPU race.
constexpr int64_t kNoAllocMaxTargets = (int64_t)1 << 20;
constexpr int64_t kNoAllocMaxBoxes = (int64_t)1 << 18;
constexpr int64_t kNoAllocMaxPartials = 1024;   // scan partials of kNoAllocMaxBoxes + 1 counts

constexpr int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// Byte offsets of the slices, in the order of EvalScratch's members.
struct NoAllocLayout {
  int64_t tnode, tcnt, toff, tcur, tperm, icnt, ioff, ntot, partials, tsx, tsy, items, total;
};
constexpr NoAllocLayout noalloc_layout() {
  NoAllocLayout l{};
  int64_t o = 0;
  l.tnode = o; o += align256(kNoAllocMaxTargets * 4);
  l.tcnt = o; o += align256((kNoAllocMaxBoxes + 1) * 4);
  l.toff = o; o += align256((kNoAllocMaxBoxes + 1) * 4);
  l.tcur = o; o += align256((kNoAllocMaxTargets + 1) * 4);
  l.tperm = o; o += align256(kNoAllocMaxTargets * 4);
  l.icnt = o; o += align256((kNoAllocMaxBoxes + 1) * 4);
  l.ioff = o; o += align256((kNoAllocMaxBoxes + 1) * 4);
  l.ntot = o; o += align256(4 * 4);
  l.partials = o; o += align256(kNoAllocMaxPartials * 4);
  l.tsx = o; o += align256(kNoAllocMaxTargets * 8);
  l.tsy = o; o += align256(kNoAllocMaxTargets * 8);
  l.items = o; o += align256(eval_item_cap(kNoAllocMaxBoxes, kNoAllocMaxTargets) * 16);
  l.total = o;
  return l;
}
constexpr NoAllocLayout kNoAlloc = noalloc_layout();

static __device__ __align__(256) unsigned char scratch[kNoAlloc.total];

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (the host-side name is only a stand-in).
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

static unsigned char* noalloc_scratch(int dev) {
  FMM_REQUIRE(dev >= 0 && dev < kMaxDevices, FMM_E_LIMIT, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    FMM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    FMM_REQUIRE(ptr != nullptr, FMM_E_INTERNAL, "no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

static EvalScratch noalloc_slices(unsigned char* base) {
  EvalScratch w;
  w.tnode = reinterpret_cast<int*>(base + kNoAlloc.tnode);
  w.tcnt = reinterpret_cast<int*>(base + kNoAlloc.tcnt);
  w.toff = reinterpret_cast<int*>(base + kNoAlloc.toff);
  w.tcur = reinterpret_cast<int*>(base + kNoAlloc.tcur);
  w.tperm = reinterpret_cast<int*>(base + kNoAlloc.tperm);
  w.icnt = reinterpret_cast<int*>(base + kNoAlloc.icnt);
  w.ioff = reinterpret_cast<int*>(base + kNoAlloc.ioff);
  w.ntot = reinterpret_cast<int*>(base + kNoAlloc.ntot);
  w.partials = reinterpret_cast<int*>(base + kNoAlloc.partials);
  w.tsx = reinterpret_cast<double*>(base + kNoAlloc.tsx);
  w.tsy = reinterpret_cast<double*>(base + kNoAlloc.tsy);
  w.items = reinterpret_cast<int4*>(base + kNoAlloc.items);
  return w;
}

}  // namespace fmm

// ===========================================================================
// C-ABI
using fmm::Plan;

#define ABI_BEGIN        \
  fmm::clear_error();    \
  try {
#define ABI_END                                                   \
  }                                                               \
  catch (const fmm::Failure& f) {                                 \
    return f.code;                                                \
  }                                                               \
  catch (const std::exception& e) {                               \
    fmm::set_error("internal error: %s", e.what());               \
    return fmm::FMM_E_INTERNAL;                                   \
  }                                                               \
  catch (...) {                                                   \
    fmm::set_error("internal error");                             \
    return fmm::FMM_E_INTERNAL;                                   \
  }                                                               \
  return fmm::FMM_OK;

#define NEED(ptr) FMM_REQUIRE((ptr) != nullptr, fmm::FMM_E_ARG, "%s is null", #ptr)

static Plan* as_plan(void* h) {
  FMM_REQUIRE(h != nullptr, fmm::FMM_E_ARG, "plan handle is null");
  return static_cast<Plan*>(h);
}

extern "C" {

int fmm2d_abi_version(void) { return FMM2D_ABI_VERSION; }

const char* fmm2d_last_error(void) { return fmm::last_error(); }

namespace {
// p equivalent points per box; cheb > 0: Chebyshev order of a black-box plan (p = cheb^2)
void create_plan(int kernel_id, const double* kparams, int nkparams, int p, int cheb,
                 int precision, void** out_plan) {
  NEED(out_plan);
  *out_plan = nullptr;
  FMM_REQUIRE(kernel_id == fmm::KER_LAPLACE || kernel_id == fmm::KER_MODHELM ||
                  kernel_id == fmm::KER_STOKES, fmm::FMM_E_ARG, "unknown kernel id %d", kernel_id);
  FMM_REQUIRE(kernel_id != fmm::KER_MODHELM || (nkparams == 1 && kparams && kparams[0] > 0.0),
              fmm::FMM_E_ARG, "the modified Helmholtz kernel needs one parameter k > 0");
  FMM_REQUIRE(p >= 4 && p <= 64, fmm::FMM_E_LIMIT, "Nequiv = %d outside [4, 64]", p);
  FMM_REQUIRE(precision == fmm::PREC_FP64 || precision == fmm::PREC_MIXED ||
                  precision == fmm::PREC_MIXED_M2L, fmm::FMM_E_ARG,
              "unknown precision %d", precision);
  FMM_REQUIRE(nkparams >= 0 && (nkparams == 0 || kparams), fmm::FMM_E_ARG, "bad kernel params");
  Plan* P = new Plan();
  try {
    P->kernel_id = kernel_id;
    P->kparam = nkparams > 0 ? kparams[0] : 0.0;
    P->p = p;
    P->cheb = cheb;
    P->PD = p <= 32 ? 32 : 64;
    P->precision = precision;
    FMM_HIP(hipGetDevice(&P->device));
    FMM_HIP(hipEventCreateWithFlags(&P->done_ev, hipEventDisableTiming));
    fmm::SlotTable st;
    for (int s = 0; s < fmm::kNumM2L; ++s) fmm::m2l_slot_offset(s, &st.ox[s], &st.oy[s]);
    FMM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fmm::c_slots), &st, sizeof st));
    fmm::ValidSlots vs{};
    for (int c = 0; c < 4; ++c) {
      const int cx = c >> 1, cy = c & 1;
      for (int s2 = 0; s2 < fmm::kNumM2L; ++s2)
        if (st.ox[s2] >= -2 - cx && st.ox[s2] <= 3 - cx && st.oy[s2] >= -2 - cy &&
            st.oy[s2] <= 3 - cy)
          vs.slot[c][vs.n[c]++] = s2;
    }
    FMM_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fmm::c_valid), &vs, sizeof vs));
    fmm::tree_init_device();
    double2 tab[fmm::kLogTableSize];
    fmm::LogTableHost::fill(tab);
    P->logtab.ensure(sizeof tab);
    FMM_HIP(hipMemcpy(P->logtab.ptr, tab, sizeof tab, hipMemcpyHostToDevice));
  } catch (...) {
    delete P;
    throw;
  }
  *out_plan = P;
}
}  // namespace

int fmm2d_plan_create(int kernel_id, const double* kparams, int nkparams, int p, int precision,
                      void** out_plan) {
  ABI_BEGIN
  create_plan(kernel_id, kparams, nkparams, p, 0, precision, out_plan);
  ABI_END
}

int fmm2d_plan_create_cheb(int kernel_id, const double* kparams, int nkparams, int order,
                           int precision, void** out_plan) {
  ABI_BEGIN
  NEED(out_plan);
  *out_plan = nullptr;
  FMM_REQUIRE(kernel_id == fmm::KER_LAPLACE || kernel_id == fmm::KER_MODHELM, fmm::FMM_E_ARG,
              "black-box plans take the Laplace or modified Helmholtz kernel (got %d)",
              kernel_id);
  FMM_REQUIRE(order >= 2 && order <= fmm::kMaxCheb, fmm::FMM_E_LIMIT,
              "Chebyshev order %d outside [2, %d]", order, fmm::kMaxCheb);
  FMM_REQUIRE(precision != fmm::PREC_MIXED || kernel_id == fmm::KER_LAPLACE, fmm::FMM_E_ARG,
              "the FP32 near field is Laplace only; use FP64 or FP32 M2L");
  create_plan(kernel_id, kparams, nkparams, order * order, order, precision, out_plan);
  ABI_END
}

int fmm2d_plan_destroy(void* plan) {
  ABI_BEGIN
  if (plan) {
    Plan* P = static_cast<Plan*>(plan);
    if (P->done_recorded) (void)hipEventSynchronize(P->done_ev);
    delete P;
  }
  ABI_END
}

int fmm2d_tree_build(void* plan, const double* x, const double* y, int64_t n, const double* bbox,
                     int64_t ncut, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  NEED(x);
  NEED(y);
  NEED(bbox);
  FMM_REQUIRE(ncut >= 1 && ncut <= (1 << 30), fmm::FMM_E_ARG, "Ncutoff %lld out of range",
              (long long)ncut);
  const double w0 = bbox[1] - bbox[0];   // tree.py:484 (the x extent sets the width)
  P->have_exp = false;
  P->up_done = false;
  P->mark(FMM2D_PHASE_TREE, false, (hipStream_t)stream);
  fmm::tree_build(P->tree, x, y, n, bbox[0], bbox[2], w0, (int)ncut, (hipStream_t)stream);
  P->mark(FMM2D_PHASE_TREE, true, (hipStream_t)stream);
  ABI_END
}

int fmm2d_tree_info(void* plan, int64_t* out, int64_t nout) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  NEED(out);
  const fmm::Tree& t = P->tree;
  FMM_REQUIRE(t.built, fmm::FMM_E_STATE, "no tree");
  const int L = t.L;
  FMM_REQUIRE(nout >= 9 + 3 * (int64_t)L, fmm::FMM_E_ARG, "tree_info needs %d slots",
              9 + 3 * L);
  out[0] = L;
  out[1] = t.n_nodes;
  out[2] = t.n_rows;
  out[3] = t.n_eval;
  out[4] = t.n_up;
  out[5] = t.n;
  out[6] = P->PD;
  out[7] = t.ncut;
  for (int d = 0; d < L; ++d) {
    out[8 + 3 * d] = t.lev[d].node_off;
    out[8 + 3 * d + 1] = t.lev[d].n_nodes;
    out[8 + 3 * d + 2] = t.row_off[d];
  }
  out[8 + 3 * L] = t.row_off[L];
  ABI_END
}

int fmm2d_tree_export(void* plan, int which, void* dst, int64_t bytes, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  NEED(dst);
  const fmm::Tree& t = P->tree;
  FMM_REQUIRE(t.built, fmm::FMM_E_STATE, "no tree");
  const void* src = nullptr;
  int64_t need = 0;
  const int64_t n = t.n, NN = t.n_nodes;
  switch (which) {
    case FMM2D_TREE_KEYS: src = t.keys.ptr; need = n * 8; break;
    case FMM2D_TREE_PERM: src = t.perm.ptr; need = n * 4; break;
    case FMM2D_TREE_SX: src = t.sx.ptr; need = n * 8; break;
    case FMM2D_TREE_SY: src = t.sy.ptr; need = n * 8; break;
    case FMM2D_TREE_CODE: src = t.code.ptr; need = NN * 8; break;
    case FMM2D_TREE_LEVEL: src = t.nlevel.ptr; need = NN * 4; break;
    case FMM2D_TREE_BOT: src = t.bot.ptr; need = NN * 4; break;
    case FMM2D_TREE_TOP: src = t.top.ptr; need = NN * 4; break;
    case FMM2D_TREE_PARENT: src = t.parent.ptr; need = NN * 4; break;
    case FMM2D_TREE_CHILD: src = t.child.ptr; need = NN * 4; break;
    case FMM2D_TREE_FLAGS: src = t.flags.ptr; need = NN * 4; break;
    case FMM2D_TREE_COLLEAGUES: src = t.coll.ptr; need = NN * 9 * 4; break;
    case FMM2D_TREE_XMID: src = t.xmid.ptr; need = NN * 8; break;
    case FMM2D_TREE_YMID: src = t.ymid.ptr; need = NN * 8; break;
    case FMM2D_TREE_ROW: src = t.row.ptr; need = NN * 4; break;
    case FMM2D_TREE_M2L: src = t.m2l.ptr; need = NN * fmm::kNumM2L * 4; break;
    case FMM2D_TREE_EVALID: src = t.evalid.ptr; need = NN * 4; break;
    default: FMM_REQUIRE(false, fmm::FMM_E_ARG, "unknown tree array %d", which);
  }
  FMM_REQUIRE(bytes >= need, fmm::FMM_E_ARG, "tree array %d needs %lld bytes", which,
              (long long)need);
  if (need > 0)
    FMM_HIP(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  ABI_END
}

int64_t fmm2d_operator_size(int p) {
  if (p < 4 || p > 64) return -1;
  return fmm::ops_per_level(p <= 32 ? 32 : 64);
}

int fmm2d_set_operators(void* plan, int level, const double* dev_ops, int64_t n, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  NEED(dev_ops);
  const int64_t opsz = fmm::ops_per_level(P->PD);
  FMM_REQUIRE(n == opsz, fmm::FMM_E_ARG, "operator block has %lld values, expected %lld",
              (long long)n, (long long)opsz);
  FMM_REQUIRE(level >= 0 && level < fmm::kMaxLevels, fmm::FMM_E_ARG, "bad level %d", level);
  hipStream_t s = (hipStream_t)stream;
  if (level >= P->ops_levels) {
    // grow to all levels at once, keeping what is set
    const int nl = fmm::kMaxLevels;
    fmm::DevBuf nb;
    nb.ensure(nl * opsz * sizeof(double));
    if (P->ops_levels > 0)
      FMM_HIP(hipMemcpyAsync(nb.ptr, P->ops.ptr, P->ops_levels * opsz * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
    FMM_HIP(hipStreamSynchronize(s));
    std::swap(nb.ptr, P->ops.ptr);
    std::swap(nb.bytes, P->ops.bytes);
    P->ops_levels = nl;
    P->ops_set.resize(nl, 0);
  }
  FMM_HIP(hipMemcpyAsync(P->ops.as<double>() + level * opsz, dev_ops, opsz * sizeof(double),
                         hipMemcpyDeviceToDevice, s));
  P->ops_set[level] = 1;
  ABI_END
}

int fmm2d_clear_operators(void* plan) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  std::fill(P->ops_set.begin(), P->ops_set.end(), 0);
  ABI_END
}

int fmm2d_upward(void* plan, const double* tau, int64_t n, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  NEED(tau);
  FMM_REQUIRE(P->tree.built, fmm::FMM_E_STATE, "no tree");
  FMM_REQUIRE(n == P->tree.n, fmm::FMM_E_ARG, "tau has %lld entries, tree has %lld sources",
              (long long)n, (long long)P->tree.n);
  P->mark(FMM2D_PHASE_UPWARD, false, (hipStream_t)stream);
  fmm::upward(*P, tau, nullptr, (hipStream_t)stream);
  P->mark(FMM2D_PHASE_UPWARD, true, (hipStream_t)stream);
  ABI_END
}

int fmm2d_upward_dipole(void* plan, const double* tau, const double* dvec, int64_t n,
                        void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  NEED(tau);
  NEED(dvec);
  FMM_REQUIRE(P->kernel_id == fmm::KER_LAPLACE || P->kernel_id == fmm::KER_STOKES,
              fmm::FMM_E_ARG, "dipole sources need the Laplace or Stokeslet kernel (plan has %d)",
              P->kernel_id);
  FMM_REQUIRE(!P->cheb, fmm::FMM_E_ARG, "black-box plans take single-layer charges only");
  FMM_REQUIRE(P->tree.built, fmm::FMM_E_STATE, "no tree");
  FMM_REQUIRE(n == P->tree.n, fmm::FMM_E_ARG, "n = %lld, tree has %lld sources", (long long)n,
              (long long)P->tree.n);
  P->mark(FMM2D_PHASE_UPWARD, false, (hipStream_t)stream);
  fmm::upward(*P, tau, dvec, (hipStream_t)stream);
  P->mark(FMM2D_PHASE_UPWARD, true, (hipStream_t)stream);
  ABI_END
}

int fmm2d_downward(void* plan, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  FMM_REQUIRE(P->tree.built && P->up_done, fmm::FMM_E_STATE, "call fmm2d_upward first");
  P->mark(FMM2D_PHASE_DOWNWARD, false, (hipStream_t)stream);
  fmm::downward(*P, (hipStream_t)stream);
  if (fmm::uses_mixed_near_field(*P)) fmm::prep_mixed(*P, (hipStream_t)stream);
  P->mark(FMM2D_PHASE_DOWNWARD, true, (hipStream_t)stream);
  ABI_END
}

int fmm2d_evaluate(void* plan, const double* tx, const double* ty, int64_t nt, int evaluator,
                   int check, double* out, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  if (nt != 0) {   // an empty target set may come with null pointers
    NEED(tx);
    NEED(ty);
    NEED(out);
  }
  P->mark(FMM2D_PHASE_EVALUATE, false, (hipStream_t)stream);
  fmm::evaluate(*P, tx, ty, nt, evaluator, check, out, (hipStream_t)stream, nullptr);
  P->mark(FMM2D_PHASE_EVALUATE, true, (hipStream_t)stream);
  ABI_END
}

int fmm2d_evaluate_noalloc(void* plan, const double* tx, const double* ty, int64_t nt,
                           int evaluator, int check, double* out, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream, false);   // no event node: the call may be captured
  if (nt != 0) {
    NEED(tx);
    NEED(ty);
    NEED(out);
  }
  FMM_REQUIRE(nt >= 0 && nt <= fmm::kNoAllocMaxTargets, fmm::FMM_E_LIMIT,
              "fmm2d_evaluate_noalloc takes at most %lld targets (got %lld); use fmm2d_evaluate",
              (long long)fmm::kNoAllocMaxTargets, (long long)nt);
  FMM_REQUIRE(P->tree.n_eval <= fmm::kNoAllocMaxBoxes, fmm::FMM_E_LIMIT,
              "fmm2d_evaluate_noalloc takes trees of at most %lld evaluation boxes (this tree "
              "has %lld); use fmm2d_evaluate",
              (long long)fmm::kNoAllocMaxBoxes, (long long)P->tree.n_eval);
  int dev = -1;
  FMM_HIP(hipGetDevice(&dev));
  FMM_REQUIRE(dev == P->device, fmm::FMM_E_ARG,
              "the plan belongs to device %d but device %d is current", P->device, dev);
  FMM_REQUIRE(fmm::scan_partials_i32(fmm::kNoAllocMaxBoxes + 1) <= fmm::kNoAllocMaxPartials,
              fmm::FMM_E_INTERNAL, "scan partials exceed the scratch array");
  const fmm::EvalScratch w = fmm::noalloc_slices(fmm::noalloc_scratch(dev));
  // not timed: phase timing creates its HIP events on first use
  struct NoTiming {
    Plan* P;
    bool was;
    ~NoTiming() { P->profile = was; }
  } nt_guard{P, P->profile};
  P->profile = false;
  fmm::evaluate(*P, tx, ty, nt, evaluator, check, out, (hipStream_t)stream, &w);
  ABI_END
}

int fmm2d_work_counts(void* plan, int64_t* out, int n) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  NEED(out);
  FMM_REQUIRE(n >= 9, fmm::FMM_E_ARG, "work_counts needs 9 slots");
  const fmm::Tree& t = P->tree;
  FMM_REQUIRE(t.built, fmm::FMM_E_STATE, "no tree");
  fmm::DevBuf acc;
  acc.ensure(16 * sizeof(unsigned long long));
  FMM_HIP(hipMemset(acc.ptr, 0, 16 * sizeof(unsigned long long)));
  if (P->last_tcnt && P->have_exp)
    hipLaunchKernelGGL(fmm::k_eval_stats, dim3(fmm::grid256(t.n_eval)), dim3(256), 0, 0,
                       t.evalnode.as<int>(), (int)t.n_eval, P->last_tcnt, t.coll.as<int>(),
                       t.bot.as<int>(), t.top.as<int>(), t.nlevel.as<int>(),
                       acc.as<unsigned long long>());
  hipLaunchKernelGGL(fmm::k_tree_stats, dim3(fmm::grid256(t.n_nodes)), dim3(256), 0, 0,
                     (int)t.n_nodes, t.nlevel.as<int>(), t.flags.as<int>(), t.m2l.as<int>(),
                     t.row.as<int>(), t.child.as<int>(), t.bot.as<int>(), t.top.as<int>(),
                     acc.as<unsigned long long>());
  FMM_LAUNCH_CHECK();
  unsigned long long h[16];
  FMM_HIP(hipMemcpy(h, acc.ptr, sizeof h, hipMemcpyDeviceToHost));
  for (int i = 0; i < 9; ++i) out[i] = (int64_t)h[i];
  out[5] = t.n_rows;
  ABI_END
}

int fmm2d_set_tree_mode(void* plan, int mode) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  FMM_REQUIRE(mode == FMM2D_TREE_AUTO || mode == FMM2D_TREE_GENERAL, fmm::FMM_E_ARG,
              "unknown tree mode %d", mode);
  P->tree.allow_full = mode == FMM2D_TREE_AUTO;
  ABI_END
}

int fmm2d_set_m2l_mode(void* plan, int mode) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  FMM_REQUIRE(mode == FMM2D_M2L_SVD || mode == FMM2D_M2L_DENSE, fmm::FMM_E_ARG,
              "unknown M2L mode %d", mode);
  P->m2l_mode = mode;
  ABI_END
}

int fmm2d_set_profiling(void* plan, int enable) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  P->profile = enable != 0;
  for (int i = 0; i < FMM2D_NUM_PHASES; ++i) {
    P->runs[i] = 0;
    P->open_run[i] = false;
  }
  ABI_END
}

int fmm2d_phase_times(void* plan, double* ms, int n) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  NEED(ms);
  FMM_REQUIRE(n >= FMM2D_NUM_PHASES, fmm::FMM_E_ARG, "phase_times needs %d slots",
              FMM2D_NUM_PHASES);
  for (int i = 0; i < FMM2D_NUM_PHASES; ++i) {
    ms[i] = -1.0;
    const int r = P->runs[i];
    if (r == 0) continue;
    double sum = 0.0;
    for (int k = 0; k < r; ++k) {
      FMM_HIP(hipEventSynchronize(P->ev[i][2 * k + 1]));
      float t = 0.f;
      FMM_HIP(hipEventElapsedTime(&t, P->ev[i][2 * k], P->ev[i][2 * k + 1]));
      sum += t;
    }
    ms[i] = sum / r;
    P->runs[i] = 0;
  }
  ABI_END
}

int fmm2d_export_expansions(void* plan, int which, double* dst, int64_t nvals, void* stream) {
  ABI_BEGIN
  Plan* P = as_plan(plan);
  fmm::PlanCall call(*P, stream);
  NEED(dst);
  const fmm::Tree& t = P->tree;
  FMM_REQUIRE(t.built, fmm::FMM_E_STATE, "no tree");
  const void* src = nullptr;
  int64_t need = 0;
  switch (which) {
    case FMM2D_EXP_CHECK: src = P->check.ptr; need = t.n_rows * P->PD; break;
    case FMM2D_EXP_MULTIPOLE: src = P->mult.ptr; need = t.n_rows * P->PD; break;
    case FMM2D_EXP_LOCAL: src = P->le.ptr; need = t.n_nodes * P->PD; break;
    case FMM2D_EXP_TAU_SORTED: src = P->tau_s.ptr; need = t.n; break;
    default: FMM_REQUIRE(false, fmm::FMM_E_ARG, "unknown expansion array %d", which);
  }
  FMM_REQUIRE(src != nullptr || need == 0, fmm::FMM_E_STATE, "expansion array %d not computed",
              which);
  FMM_REQUIRE(nvals >= need, fmm::FMM_E_ARG, "expansion array %d needs %lld values", which,
              (long long)need);
  if (need > 0)
    FMM_HIP(hipMemcpyAsync(dst, src, need * sizeof(double), hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
  ABI_END
}

}  // extern "C"
