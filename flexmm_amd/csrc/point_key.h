// Morton keys of points in the quadtree's root box, for the device and the host.
//
// The key holds 31 levels of 2-bit digits, left aligned (level l at bits
// 64 - 2l, 65 - 2l), digit 2 (x > xmid) + (y > ymid) with the midpoints of the
// reference's FP64 recurrence (flexmm/tree.py:169-176, 301-304: xmid = xmin +
// w/2, children start at xmid - w/2 or xmid, class tree.py:32-34).
//
// point_key_fast gives the same key from one scaled coordinate per axis when
// the point is provably clear of every midpoint, and reports when it is not.
// The argument, in units u = w0 / 2^32 of the root width (every midpoint of
// levels 1..31 is an integer multiple of u from xmin0):
//   * the recurrence rounds twice per level (xmin + h, xm - h; h = w/2 is
//     exact), each time by at most half an ulp of a value of magnitude at most
//     M = max(|xmin0|, |xmin0 + w0|) up to rounding, so after l <= 31 levels
//     its midpoint is within 62 * 2^-53 * 2M = 62 * 2^-20 (M / w0) u of the
//     exact one, provided every earlier digit was the exact one;
//   * s = (p - xmin0) * (2^32 / w0) rounds three times, each by at most a
//     relative 2^-53, so |s - s_exact| <= 3 * 2^-53 * 2^32 < 2^-19 for s in
//     [0, 2^32).
// If the fraction of s is more than e = 2^-18 + 2^-13 (M / w0) from 0 and 1,
// the point lies further than both bounds from every midpoint, so (by
// induction over the levels) every digit of the recurrence is the exact
// classification: the level-l x digit is bit 32 - l of floor(s).
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define FMM_HD __host__ __device__
#define FMM_ROLLED _Pragma("unroll 1")
#else
#define FMM_HD
#define FMM_ROLLED
#endif

namespace fmm {

FMM_HD inline uint64_t spread_bits(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
FMM_HD inline uint32_t compact_bits(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return (uint32_t)x;
}
// Morton code of a cell: the digit of level l (from the top) is 2*xbit + ybit.
FMM_HD inline uint64_t morton(uint32_t ix, uint32_t iy) {
  return (spread_bits(ix) << 1) | spread_bits(iy);
}
FMM_HD inline void demorton(uint64_t c, uint32_t* ix, uint32_t* iy) {
  *ix = compact_bits(c >> 1);
  *iy = compact_bits(c);
}

constexpr int kKeyLevels = 31;

// The reference's recurrence, level by level.
FMM_HD inline uint64_t point_key_recurrence(double px, double py, double xmin, double ymin,
                                            double w) {
  uint64_t key = 0;
  FMM_ROLLED   // a rare path of the key kernel: kept small
  for (int l = 1; l <= kKeyLevels; ++l) {
    const double h = 0.5 * w;
    const double xm = xmin + h;
    const double ym = ymin + h;
    const int bx = px > xm;
    const int by = py > ym;
    key |= (uint64_t)(2 * bx + by) << (64 - 2 * l);
    xmin = bx ? xm : xm - h;
    ymin = by ? ym : ym - h;
    w = h;
  }
  return key;
}

// Per-axis constants of the fast form: scale = 2^32 / w0 and the margin e.
struct KeyScale {
  double scale, ex, ey;
};
inline KeyScale key_scale(double xmin0, double ymin0, double w0) {
  auto margin = [w0](double lo) {
    const double a = lo < 0 ? -lo : lo, b = lo + w0 < 0 ? -(lo + w0) : lo + w0;
    const double M = a > b ? a : b;
    return 0x1p-18 + 0x1p-13 * (M / w0);
  };
  return KeyScale{0x1p32 / w0, margin(xmin0), margin(ymin0)};
}

// floor of the scaled coordinate, and whether its fraction clears the margin
// (false as well for points outside [0, 2^32) u and for NaN)
FMM_HD inline bool key_axis(double p, double lo, double scale, double e, uint32_t* q) {
  const double s = (p - lo) * scale;
  if (!(s >= 0.0 && s < 0x1p32)) return false;
  const double f = s - __builtin_floor(s);
  *q = (uint32_t)s;
  return f > e && f < 1.0 - e;
}

// The key when both axes clear their margins (*ok true), else *ok false and
// the key must come from point_key_recurrence.
FMM_HD inline uint64_t point_key_fast(double px, double py, double xmin0, double ymin0,
                                      const KeyScale& k, bool* ok) {
  uint32_t ix = 0, iy = 0;
  const bool okx = key_axis(px, xmin0, k.scale, k.ex, &ix);
  const bool oky = key_axis(py, ymin0, k.scale, k.ey, &iy);
  *ok = okx && oky;
  return morton(ix >> 1, iy >> 1) << 2;
}

}  // namespace fmm
