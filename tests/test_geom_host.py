"""Host-side checks of ecm_geom.h identities the kernels rely on (compiled
with hipcc for the host; nothing runs on a GPU).

rotate_pm(v, a) must equal (rotate(v, a), rotate(v, -a)) bit for bit: the
agent half-plane builds both VO legs from one sincos_f."""
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
#include <cstdio>
#include <cstring>
#include <cstdint>
using namespace ecm;
static uint64_t st = 0x9e3779b97f4a7c15ull;
static uint32_t rnd() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 32); }
static float rf(float lo, float hi) { return lo + (hi - lo) * (rnd() / 4294967296.0f); }
static bool same(V2 a, V2 b) { return std::memcmp(&a, &b, sizeof(V2)) == 0; }
int main() {
    long bad = 0, n = 0;
    const float special[] = {0.0f, -0.0f, 0.78539819f, -0.78539819f, 1.57079637f, -1.57079637f,
                             3.14159274f, 2.35619450f, 1e-30f, -1e-30f, 0.267949194f, 100.0f};
    for (int i = 0; i < 2000000; ++i) {
        float a;
        if (i < 12) a = special[i];
        else if (i % 3 == 0) a = rf(-1.6f, 1.6f);               // VO half-angles
        else if (i % 3 == 1) a = rf(-50.0f, 50.0f);
        else { uint32_t b = rnd(); std::memcpy(&a, &b, 4); if (!(a == a) || a > 1e6f || a < -1e6f) continue; }
        V2 v = mk(rf(-40.0f, 40.0f), rf(-40.0f, 40.0f));
        if (i % 7 == 0) v.x = 0.0f;
        V2 p, m;
        rotate_pm(v, a, p, m);
        ++n;
        if (!same(p, rotate(v, a)) || !same(m, rotate(v, -a))) ++bad;
    }
    std::printf("%ld %ld\n", n, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_rotate_pm_equals_two_rotations():
    d = tempfile.mkdtemp()
    try:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        open(src, "w").write(SRC)
        flags = ["-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                 "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
        subprocess.run([HIPCC, *flags, "-x", "hip", "-I", CSRC, src, "-o", exe], check=True,
                       capture_output=True, timeout=300)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        n, bad = map(int, out.stdout.split())
        print(f"{n} angles, {bad} mismatches", flush=True)
        assert n > 1_500_000 and bad == 0
    finally:
        shutil.rmtree(d, ignore_errors=True)
