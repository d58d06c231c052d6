// Shared host/device helpers for the flexmm_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace fmm {

// ---------------------------------------------------------------------------
// Error handling: every C-ABI entry point returns 0 on success, a negative
// FMM_E_* code for an argument error and a positive hipError_t for a runtime
// failure; the message is kept per thread and read by fmm2d_last_error().
enum {
  FMM_OK = 0,
  FMM_E_ARG = -1,       // invalid argument (null pointer, bad size ...)
  FMM_E_STATE = -2,     // call out of order (no tree, no expansions ...)
  FMM_E_LIMIT = -3,     // a documented limit was exceeded
  FMM_E_INTERNAL = -4,  // internal invariant violated
};

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();
const char* last_error();

struct Failure {
  int code;
};

#define FMM_HIP(expr)                                                             \
  do {                                                                            \
    hipError_t e__ = (expr);                                                      \
    if (e__ != hipSuccess) {                                                      \
      ::fmm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e__),    \
                       __FILE__, __LINE__);                                       \
      throw ::fmm::Failure{(int)e__};                                             \
    }                                                                             \
  } while (0)

#define FMM_REQUIRE(cond, code, ...)                                              \
  do {                                                                            \
    if (!(cond)) {                                                                \
      ::fmm::set_error(__VA_ARGS__);                                              \
      throw ::fmm::Failure{(code)};                                               \
    }                                                                             \
  } while (0)

#define FMM_LAUNCH_CHECK() FMM_HIP(hipGetLastError())

// Stream of the C-ABI call in progress on this thread (set by the entry
// points for the duration of the call; null: the default stream).
inline thread_local hipStream_t g_op_stream = nullptr;

// Device buffer that only grows; reused across calls so that steady-state
// calls allocate nothing.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  // Growing frees the old block after the stream of the current call has
  // drained, so that no kernel queued on it still uses the block (calls on a
  // plan are ordered on their streams; see the C-ABI header).
  void ensure(size_t nbytes) {
    if (nbytes <= bytes) return;
    if (ptr) {
      FMM_HIP(hipStreamSynchronize(g_op_stream));
      FMM_HIP(hipFree(ptr));
    }
    ptr = nullptr;
    bytes = 0;
    size_t want = nbytes < 256 ? 256 : nbytes;
    FMM_HIP(hipMalloc(&ptr, want));
    bytes = want;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
};

// Pinned host staging for the few scalars the host must read back.
struct HostBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void ensure(size_t nbytes) {
    if (nbytes <= bytes) return;
    if (ptr) FMM_HIP(hipHostFree(ptr));
    ptr = nullptr;
    FMM_HIP(hipHostMalloc(&ptr, nbytes, hipHostMallocDefault));
    bytes = nbytes;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
  ~HostBuf() {
    if (ptr) (void)hipHostFree(ptr);
  }
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
};

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Pinned, fine-grained host words that a kernel writes while its stream
// still runs: each word is (sequence << 32 | value), written by one 64-bit
// system-scope store, so the host can take a few results as soon as they are
// final instead of draining the stream first (the kernels queued after them
// then follow without a gap).  Every use writes all the words it reads under
// a new sequence number and the host takes each word once it carries that
// number; no order between words is assumed.
struct HostMailbox {
  static constexpr int kWords = 128;
  uint64_t* host = nullptr;   // host address of the words
  uint64_t* dev = nullptr;    // the same words as kernels address them
  uint32_t seq = 0;
  void ensure() {
    if (host) return;
    void* p = nullptr;
    FMM_HIP(hipHostMalloc(&p, kWords * sizeof(uint64_t), hipHostMallocCoherent));
    host = static_cast<uint64_t*>(p);
    for (int i = 0; i < kWords; ++i) host[i] = 0;   // sequence 0 is never used
    FMM_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev), host, 0));
  }
  uint32_t next() {
    seq = seq + 1 == 0 ? 1 : seq + 1;
    return seq;
  }
  // Words [0, n) of sequence sq into out: spins for up to 2 ms, then drains
  // the stream s (whose kernels write them), after which they must be there.
  void take(int n, uint32_t sq, hipStream_t s, int64_t* out) const {
    FMM_REQUIRE(n <= kWords, FMM_E_INTERNAL, "host mailbox: %d words requested", n);
    const auto t0 = std::chrono::steady_clock::now();
    for (int spin = 0;; ++spin) {
      int got = 0;
      for (; got < n; ++got) {
        const uint64_t v = __atomic_load_n(host + got, __ATOMIC_ACQUIRE);
        if ((uint32_t)(v >> 32) != sq) break;
        out[got] = (int32_t)(uint32_t)v;
      }
      if (got == n) return;
      if ((spin & 63) == 63 &&
          std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
        break;
#if defined(__x86_64__)
      __builtin_ia32_pause();
#endif
    }
    FMM_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i) {
      const uint64_t v = __atomic_load_n(host + i, __ATOMIC_ACQUIRE);
      FMM_REQUIRE((uint32_t)(v >> 32) == sq, FMM_E_INTERNAL,
                  "host mailbox: word %d not written by its kernel", i);
      out[i] = (int32_t)(uint32_t)v;
    }
  }
  ~HostMailbox() {
    if (host) (void)hipHostFree(host);
  }
  HostMailbox() = default;
  HostMailbox(const HostMailbox&) = delete;
  HostMailbox& operator=(const HostMailbox&) = delete;
};

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
// word i of a HostMailbox (device address mb) for sequence seq
__device__ __forceinline__ void mailbox_put(uint64_t* mb, int i, uint32_t seq, int value) {
  __hip_atomic_store(mb + i, ((uint64_t)seq << 32) | (uint32_t)value, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}
#endif


// ---------------------------------------------------------------------------
// Table-driven FP64 logarithm (P. T. P. Tang, "Table-driven implementation of
// the logarithm function in IEEE floating-point arithmetic", ACM TOMS 16(4),
// 1990), in an FMA variant split for accumulation:
//   x = m 2^e, m in [0.5, 1) (frexp);  j = top 9 fraction bits of m;
//   r = fma(m, T_j, -1) with T_j = 2/c_j, c_j = 1 + (j + 1/2)/512, |r| < 2^-10;
//   log(x) = e ln2 + t,  t = -log(T_j) + r p(r),
// with r p(r) = r (1 - r/2 + a3 r^2 + a4 r^3) the degree-4 polynomial closest
// to log1p(r) in the max norm on |r| <= 2^-10 (error below 2.4e-17; a3, a4
// fitted in extended precision), evaluated by Horner's rule with the table
// term as the last addend: four FMAs.  Callers accumulate sum(w e) and
// sum(w t) separately and multiply the first by ln 2 once.  The table
// {T_j, -log(T_j)} is filled on the host in extended precision.
constexpr int kLogTableBits = 9;
constexpr int kLogTableSize = 1 << kLogTableBits;
constexpr double kLn2 = 0.6931471805599453;
constexpr double kLog1pA3 = 0.3333334990879393;
constexpr double kLog1pA4 = -0.2500001566895226;

struct LogTableHost {
  static void fill(double2* out);  // out[kLogTableSize]
};

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
// log(x) = e * ln2 + t (x > 0).  For x == 0, e = 0 and t is finite garbage:
// callers handle coincident points themselves.
__device__ __forceinline__ void log_split(double x, const double2* __restrict__ tab, double& e,
                                          double& t) {
  const double m = __builtin_amdgcn_frexp_mant(x);
  e = (double)__builtin_amdgcn_frexp_exp(x);
  const int j = (__double2hiint(m) >> (20 - kLogTableBits)) & (kLogTableSize - 1);
  const double2 c = tab[j];
  const double r = __builtin_fma(m, c.x, -1.0);
  double q = __builtin_fma(kLog1pA4, r, kLog1pA3);
  q = __builtin_fma(q, r, -0.5);
  q = __builtin_fma(q, r, 1.0);
  t = __builtin_fma(q, r, c.y);
}

__device__ __forceinline__ double log_tab(double x, const double2* __restrict__ tab) {
  double e, t;
  log_split(x, tab, e, t);
  const double res = __builtin_fma(e, kLn2, t);
  return x == 0.0 ? -__builtin_inf() : res;
}
#endif

// ---------------------------------------------------------------------------
// XCD-aware block order.  MI355X dispatches consecutive workgroups round-robin
// to its 8 XCDs, each with its own L2; this bijection of [0, n) gives XCD x
// (blocks b with b % 8 == x) one contiguous range of logical blocks, so that
// neighbouring tiles, which share source data, meet in the same L2.
constexpr int kNumXcd = 8;
constexpr int kNumCUs = 256;   // MI355X compute units (launch-shape heuristics only)
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int xcd = b % kNumXcd, local = b / kNumXcd;
  const int per = n / kNumXcd, rem = n % kNumXcd;
  return xcd * per + min(xcd, rem) + local;
}

// ---------------------------------------------------------------------------
// Kernel identifiers shared with the Python registry.
enum KernelId : int {
  KER_LAPLACE = 0,   // -log(d^2)/(4 pi)
  KER_MODHELM = 1,   // K0(k r)/(2 pi)
  KER_STOKES = 2,    // 2D Stokeslet (velocity and pressure); far field as 3 harmonic fields
};
enum EvalId : int {
  EVAL_POT = 0,      // 1 output: u
  EVAL_GRAD = 1,     // 3 outputs: u, du/dtx, du/dty
  EVAL_STOKES = 2,   // 3 outputs: velocity u, v and pressure p
  EVAL_DIPOLE = 3,   // 1 output: u of Laplace single + double layer sources
  EVAL_STOKES_DIPOLE = 4,   // 3 outputs: u, v, p of Stokeslet + stresslet sources
};
enum Precision : int {
  PREC_FP64 = 0,
  PREC_MIXED = 1,       // FP32 M2L + FP32 leaf near field
  PREC_MIXED_M2L = 2,   // FP32 M2L only
};

}  // namespace fmm
