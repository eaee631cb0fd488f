"""Host-side check of the ECM bucket filter (csrc/bucket_geom.h; compiled
with hipcc for the host, nothing runs on a GPU).

A bucket square leaves an ECM cell out of its list only when
hull_may_touch is false.  The device then never tests that cell for points
of the bucket, so the ray-cast containment test (quad_contains, the
reference's MathUtility::Contains) must be false, in float arithmetic, at
every point of the square whose y is not one of the cell's vertex y values
(those points go to the exact-y lists).  Random and degenerate quads
(repeated vertices, collinear, self-intersecting, slivers), random squares;
points sampled on a lattice over each rejected square, its border and its
corners."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "ecmgenerator_amd", "csrc")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r"""
#include "ecm_geom.h"
#include "bucket_geom.h"
#include <cstdio>
#include <cstdint>
using namespace ecm;
static uint64_t st = 0x2545f4914f6cdd1dull;
static uint32_t rnd() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 32); }
static float rf(float lo, float hi) { return lo + (hi - lo) * (rnd() / 4294967296.0f); }
int main() {
    long rejected = 0, kept = 0, points = 0, bad = 0;
    for (int it = 0; it < 60000; ++it) {
        float v[4][2];
        const float cx = rf(-800.0f, 800.0f), cy = rf(-800.0f, 800.0f), ext = it % 5 == 0 ? 0.5f : rf(1.0f, 60.0f);
        for (int k = 0; k < 4; ++k) { v[k][0] = cx + rf(-ext, ext); v[k][1] = cy + rf(-ext, ext); }
        if (it % 7 == 1) { v[2][0] = v[1][0]; v[2][1] = v[1][1]; }                  // triangle
        if (it % 7 == 2) { for (int k = 1; k < 4; ++k) { v[k][0] = v[0][0] + k * 3.0f; v[k][1] = v[0][1] + k * 2.0f; } }
        if (it % 7 == 3) { float t0 = v[1][0]; v[1][0] = v[2][0]; v[2][0] = t0; }    // may self-intersect
        if (it % 7 == 4) { v[3][0] = v[0][0]; v[3][1] = v[0][1] + 1e-3f; }          // sliver
        const double q[4][2] = {{v[0][0], v[0][1]}, {v[1][0], v[1][1]}, {v[2][0], v[2][1]}, {v[3][0], v[3][1]}};
        const float side = rf(0.5f, 8.0f);
        const float sx0 = cx + rf(-2.0f * ext - 10.0f, 2.0f * ext + 10.0f), sy0 = cy + rf(-2.0f * ext - 10.0f, 2.0f * ext + 10.0f);
        const float sx1 = sx0 + side, sy1 = sy0 + side;
        if (hull_may_touch(q, sx0, sy0, sx1, sy1, 0.05)) { ++kept; continue; }
        ++rejected;
        const int n = 24;
        for (int a = 0; a <= n; ++a)
            for (int b = 0; b <= n; ++b) {
                const float px = sx0 + (sx1 - sx0) * (float)a / n, py = sy0 + (sy1 - sy0) * (float)b / n;
                bool on_vertex_y = false;
                for (int k = 0; k < 4; ++k) on_vertex_y |= py == v[k][1];
                if (on_vertex_y) continue;
                ++points;
                const V2 p = mk(px, py);
                if (quad_contains(p, mk(v[0][0], v[0][1]), mk(v[1][0], v[1][1]), mk(v[2][0], v[2][1]), mk(v[3][0], v[3][1]))) ++bad;
            }
    }
    std::printf("%ld %ld %ld %ld\n", rejected, kept, points, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_rejected_buckets_hold_no_contained_point():
    d = tempfile.mkdtemp()
    try:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        open(src, "w").write(SRC)
        flags = ["-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                 "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
        subprocess.run([HIPCC, *flags, "-x", "hip", "-I", CSRC, src, "-o", exe], check=True,
                       capture_output=True, timeout=300)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        rejected, kept, points, bad = map(int, out.stdout.split())
        print(f"{rejected} squares rejected, {kept} kept; {points} points tested, {bad} contained", flush=True)
        assert rejected > 10000 and kept > 1000 and points > 5_000_000
        assert bad == 0
    finally:
        shutil.rmtree(d, ignore_errors=True)
