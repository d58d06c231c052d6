"""The tree's point keys (flexmm_amd/csrc/point_key.h) on the host: the
scaled-coordinate form against the reference's FP64 midpoint recurrence
(flexmm/tree.py:169-176, 301-304), bit for bit, wherever the fast form
accepts a point.

The harness is the header compiled by g++.  Its inputs are uniform points and
the points where the recurrence is most fragile: its own rounded midpoints at
every level and the exact dyadic midpoints, each moved by -3 .. +3 ulps, in
boxes whose midpoints are exact (the unit box), inexact (offset, non-dyadic
widths) and badly scaled (a box far from the origin, where the recurrence's
rounding reaches the deep cells).  With the rounding margin removed the same
harness must find disagreements, so the margin is what makes the fast form
exact."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "flexmm_amd", "csrc")

HARNESS = r"""
#include "point_key.h"
#include <cmath>
#include <cstdio>
#include <cstring>
using namespace fmm;

static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {
  state ^= state << 13; state ^= state >> 7; state ^= state << 17;
  return state;
}
static double uni() { return (double)(next_u64() >> 11) * 0x1p-53; }

struct Count { long total = 0, fast = 0, bad = 0; };

static void check(double px, double py, double x0, double y0, double w, const KeyScale& k,
                  Count& c) {
  const uint64_t rec = point_key_recurrence(px, py, x0, y0, w);
  bool ok = false;
  const uint64_t f = point_key_fast(px, py, x0, y0, k, &ok);
  ++c.total;
  if (ok) {
    ++c.fast;
    if (f != rec) ++c.bad;
  }
}

int main(int argc, char** argv) {
  const bool nomargin = argc > 1 && !strcmp(argv[1], "nomargin");
  const double boxes[][3] = {{0.0, 0.0, 1.0},   {-0.37, 1.9, 2.9}, {0.1, 0.2, 0.3},
                             {-1.0, -1.0, 2.0}, {1e3, -1e3, 1e-3}, {12345.678, 0.5, 7.0},
                             {-3e-7, 2e-7, 1.3e-6}};
  for (const auto& b : boxes) {
    const double x0 = b[0], y0 = b[1], w = b[2];
    KeyScale k = key_scale(x0, y0, w);
    if (nomargin) k.ex = k.ey = 0.0;
    Count uniform, rounded, exact;
    for (int i = 0; i < 200000; ++i) check(x0 + uni() * w, y0 + uni() * w, x0, y0, w, k, uniform);
    for (int i = 0; i < 40000; ++i) {
      // a random path of the recurrence to level l; its rounded midpoint there
      const int l = 1 + (int)(next_u64() % 31);
      double xmin = x0, ymin = y0, ww = w, xm = 0, ym = 0;
      for (int j = 1; j <= l; ++j) {
        const double h = 0.5 * ww;
        xm = xmin + h;
        ym = ymin + h;
        const uint64_t r = next_u64();
        if (j < l) {
          xmin = (r & 1) ? xm : xm - h;
          ymin = (r & 2) ? ym : ym - h;
        }
        ww = h;
      }
      // the exact dyadic midpoint of a random cell of level l, rounded once
      const uint64_t cell = next_u64() & (((uint64_t)1 << (l - 1)) - 1);
      const double xe = x0 + std::ldexp((double)(2 * cell + 1), -l) * w;
      for (int d = -3; d <= 3; ++d) {
        double px = xm, pe = xe;
        for (int s = 0; s < (d < 0 ? -d : d); ++s) {
          px = std::nextafter(px, d < 0 ? -INFINITY : INFINITY);
          pe = std::nextafter(pe, d < 0 ? -INFINITY : INFINITY);
        }
        const double qy = (next_u64() & 1) ? ym : y0 + uni() * w;
        check(px, qy, x0, y0, w, k, rounded);
        check(qy - y0 + x0, px - x0 + y0, x0, y0, w, k, rounded);   // the same on the y axis
        check(pe, y0 + uni() * w, x0, y0, w, k, exact);
      }
    }
    // edges, outside points and non-finite coordinates take the recurrence
    Count edge;
    const double ex[] = {x0, x0 + w, x0 - w, x0 + 2 * w, NAN, INFINITY};
    for (double px : ex) check(px, y0 + 0.5 * w, x0, y0, w, k, edge);
    std::printf("%g %g %g uniform %ld %ld %ld rounded %ld %ld %ld exact %ld %ld %ld edge %ld %ld %ld\n",
                x0, y0, w, uniform.total, uniform.fast, uniform.bad, rounded.total, rounded.fast,
                rounded.bad, exact.total, exact.fast, exact.bad, edge.total, edge.fast, edge.bad);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("point_keys")
    src = d / "keys.cpp"
    src.write_text(HARNESS)
    exe = d / "keys"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, str(src), "-o",
                    str(exe)], check=True)
    return str(exe)


def _run(exe, *args):
    out = subprocess.run([exe, *args], check=True, capture_output=True, text=True).stdout
    rows = []
    for line in out.splitlines():
        f = line.split()
        box = tuple(float(v) for v in f[:3])
        vals = {f[i]: tuple(int(v) for v in f[i + 1:i + 4]) for i in range(3, len(f), 4)}
        rows.append((box, vals))
    return rows


def test_fast_keys_equal_the_recurrence(harness):
    rows = _run(harness)
    assert len(rows) == 7
    for box, vals in rows:
        print(box, vals)
        for kind, (total, fast, bad) in vals.items():
            assert bad == 0, (box, kind)
        # outside the box and non-finite: never the fast form
        assert vals["edge"][1] <= 2
        # most points take the fast form where the box is not badly scaled
        # (|corner| / width = M / w0 sets the margin)
        m_over_w = max(abs(box[0]), abs(box[0] + box[2]), abs(box[1]), abs(box[1] + box[2])) / box[2]
        if m_over_w <= 20:
            total, fast, _ = vals["uniform"]
            assert fast >= 0.99 * total, (box, fast, total)
    # the recurrence's own rounded midpoints (and points ulps away) are within
    # the margin, so they go to the recurrence
    unit = [v for b, v in rows if b == (0.0, 0.0, 1.0)][0]
    assert unit["rounded"][1] < unit["rounded"][0]


def test_without_the_margin_the_fast_form_disagrees(harness):
    """Plant the mistake of a zero margin: the harness must see it."""
    rows = _run(harness, "nomargin")
    assert sum(v["rounded"][2] + v["exact"][2] for _, v in rows) > 0
