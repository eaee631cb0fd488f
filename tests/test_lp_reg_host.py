"""Host-side check of the register-resident linear programs (csrc/lp_reg.h,
compiled with hipcc for the host; nothing runs on a GPU) against the
oracle's RandomizedLP / RandomizedLP3D (oracle_lp2d / oracle_lp3d, restating
ORCA.cpp:428-669) on a million random constraint lists: ORCA-like
half-planes near the speed circle, exactly parallel and antiparallel
copies, nearly parallel ones, NaN constraints, empty lists and obstacle
counts beyond what a list of agent constraints alone has.  The 2D LP's
failing index and velocity, and the 3D LP's velocity, must be equal bit for
bit (NaN equal to NaN)."""
import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CSRC = os.path.join(ROOT, "ecmgenerator_amd", "csrc")

SRC = r"""
#include "lp_reg.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdint>
extern "C" int oracle_lp2d(const float* c, int n, const float* opt, float maxSpeed, int useDir, float* outv);
extern "C" void oracle_lp3d(const float* c, int n, int nObst, float maxSpeed, int failed, float* outv);
using namespace ecm;
constexpr int OB = 4, K = 10;
static uint64_t st = 0x243f6a8885a308d3ull;
static uint32_t rnd() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 32); }
static float rf(float lo, float hi) { return lo + (hi - lo) * (rnd() / 4294967296.0f); }
static bool same(float a, float b) { return (a != a && b != b) || std::memcmp(&a, &b, 4) == 0; }
int main() {
    long n_case = 0, bad2 = 0, bad3 = 0, n_fail2 = 0, n_nan = 0, n_par = 0;
    float cons[4 * (OB + K)];
    for (int it = 0; it < 1000000; ++it) {
        const int nObst = (rnd() % 8 == 0) ? (int)(rnd() % (OB + 1)) : (int)(rnd() % 3);
        const int na = (int)(rnd() % (K + 1));
        const int n = nObst + na;
        const float maxSpeed = (rnd() & 1) ? 2.0f : rf(0.8f, 1.8f);
        const float spread = (rnd() & 1) ? 1.0f : 3.0f;
        bool has_nan = false, has_par = false;
        for (int c = 0; c < n; ++c) {
            float nx, ny;
            const uint32_t kind = rnd() % 16;
            if (c > 0 && kind < 2) {                        // exactly parallel or antiparallel copy
                const int o = (int)(rnd() % c);
                const float sgn = (kind == 0) ? 1.0f : -1.0f;
                nx = cons[4 * o] * sgn; ny = cons[4 * o + 1] * sgn; has_par = true;
            } else if (c > 0 && kind == 2) {                // nearly parallel
                const int o = (int)(rnd() % c);
                nx = cons[4 * o] + rf(-2e-5f, 2e-5f); ny = cons[4 * o + 1] + rf(-2e-5f, 2e-5f);
                const float l = std::sqrt(nx * nx + ny * ny); nx /= l; ny /= l;
            } else {
                const float a = rf(-3.14159265f, 3.14159265f);
                nx = std::cos(a); ny = std::sin(a);
            }
            const float off = rf(-spread, spread), tan = rf(-4.0f, 4.0f);
            float px = nx * off - ny * tan, py = ny * off + nx * tan;
            if (kind == 15 && (rnd() % 8) == 0) { nx = NAN; ny = NAN; has_nan = true; }   // a zero-w colliding pair
            cons[4 * c] = nx; cons[4 * c + 1] = ny; cons[4 * c + 2] = px; cons[4 * c + 3] = py;
        }
        const float opt[2] = {rf(-2.5f, 2.5f), rf(-2.5f, 2.5f)};
        // oracle
        float o2[2];
        const int r = oracle_lp2d(cons, n, opt, maxSpeed, 0, o2);
        // registers: obstacle c in slot OB - nObst + c, agent j in slot OB + j
        V2 on[OB], op[OB], an[K], ap[K];
        for (int s = 0; s < OB; ++s) {
            const int c = s - (OB - nObst);
            on[s] = c >= 0 ? mk(cons[4 * c], cons[4 * c + 1]) : mk(0.0f, 0.0f);
            op[s] = c >= 0 ? mk(cons[4 * c + 2], cons[4 * c + 3]) : mk(0.0f, 0.0f);
        }
        uint32_t aval = 0;
        for (int j = 0; j < K; ++j) {
            const int c = nObst + j;
            an[j] = j < na ? mk(cons[4 * c], cons[4 * c + 1]) : mk(0.0f, 0.0f);
            ap[j] = j < na ? mk(cons[4 * c + 2], cons[4 * c + 3]) : mk(0.0f, 0.0f);
            if (j < na) aval |= 1u << j;
        }
        V2 out;
        const int fs = lp2_reg<OB, K>(on, op, OB - nObst, an, ap, aval, mk(opt[0], opt[1]), maxSpeed, out);
        const int rr = fs == OB + K ? n : slot_to_index<OB>(fs, nObst);
        ++n_case; n_nan += has_nan; n_par += has_par;
        if (rr != r || !same(out.x, o2[0]) || !same(out.y, o2[1])) {
            if (bad2 < 5) std::printf("2D mismatch case %d: n %d nObst %d oracle %d (%g %g) reg %d (%g %g)\n", it, n, nObst,
                                      r, o2[0], o2[1], rr, out.x, out.y);
            ++bad2;
            continue;
        }
        if (r < n) {
            ++n_fail2;
            float o3[2] = {o2[0], o2[1]};
            oracle_lp3d(cons, n, nObst, maxSpeed, r, o3);
            V2 v3 = out;
            lp3d_reg<OB, K>(on, op, an, ap, nObst, na, r, maxSpeed,
                            [&](int c) { return make_float4(cons[4 * c], cons[4 * c + 1], cons[4 * c + 2], cons[4 * c + 3]); },
                            v3);
            if (!same(v3.x, o3[0]) || !same(v3.y, o3[1])) {
                if (bad3 < 5) std::printf("3D mismatch case %d: n %d nObst %d failed %d oracle (%g %g) reg (%g %g)\n", it, n,
                                          nObst, r, o3[0], o3[1], v3.x, v3.y);
                ++bad3;
            }
        }
    }
    std::printf("cases %ld 2D-failed %ld with-NaN %ld with-parallel %ld mismatches2d %ld mismatches3d %ld\n", n_case,
                n_fail2, n_nan, n_par, bad2, bad3);
    return (bad2 || bad3) ? 1 : 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_register_lps_equal_oracle():
    from oracle import oracle_py as O
    lib = O.build()
    libdir = os.path.dirname(lib)
    d = tempfile.mkdtemp()
    try:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        open(src, "w").write(SRC)
        flags = ["-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                 "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
        subprocess.run([HIPCC, *flags, "-x", "hip", "-I", CSRC, src, "-o", exe, "-L", libdir, "-lecm_oracle",
                        f"-Wl,-rpath,{libdir}"], check=True, capture_output=True, timeout=600)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        print(out.stdout, flush=True)
        last = out.stdout.strip().splitlines()[-1].split()
        vals = dict(zip(last[0::2], map(int, last[1::2])))
        assert out.returncode == 0, out.stdout
        assert vals["mismatches2d"] == 0 and vals["mismatches3d"] == 0
        # the generator reaches the branches that matter
        assert vals["2D-failed"] > 100_000 and vals["with-NaN"] > 10_000 and vals["with-parallel"] > 100_000
    finally:
        shutil.rmtree(d, ignore_errors=True)
