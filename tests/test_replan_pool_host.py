"""Host-side check of the replan worker pool (csrc/replan_pool.h, compiled
with g++; nothing runs on a GPU).  Thousands of batches, with the thread
count changing now and then (so that workers are made while a batch is
posted right after), must each run every item exactly once and return; an
item that throws makes run() report failure while the other items still
run.  The driver runs under a time limit, so a lost wake-up fails the test
instead of hanging it."""
import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT

GXX = shutil.which("g++")
CSRC = os.path.join(ROOT, "ecmgenerator_amd", "csrc")

SRC = r"""
#include "replan_pool.h"
#include <cstdio>
#include <stdexcept>
#include <vector>
int main() {
    ecm::ReplanPool pool;
    uint64_t st = 0x9e3779b97f4a7c15ull;
    auto rnd = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st >> 32); };
    std::vector<std::atomic<int>> hits(256);
    long bad = 0, resized = 0;
    size_t last = 0;
    for (int b = 0; b < 4000; ++b) {
        const int threads = (b % 37 == 0) ? 1 + (int)(rnd() % 16) : 8;
        const int n = (int)(rnd() % 200);
        for (int i = 0; i < n; ++i) hits[i].store(0);
        const bool ok = pool.run(threads, n, [&](int i) { hits[i].fetch_add(1); });
        for (int i = 0; i < n; ++i) bad += hits[i].load() != 1;
        bad += !ok;
        if (pool.workers() != last) { ++resized; last = pool.workers(); }
    }
    // a throwing item: run() reports it, every other item still runs
    for (int i = 0; i < 100; ++i) hits[i].store(0);
    const bool ok = pool.run(8, 100, [&](int i) {
        if (i == 17) throw std::runtime_error("item 17");
        hits[i].fetch_add(1);
    });
    long missing = 0;
    for (int i = 0; i < 100; ++i) missing += hits[i].load() != (i == 17 ? 0 : 1);
    printf("bad %ld resized %ld throw_reported %d missing %ld\n", bad, resized, ok ? 0 : 1, missing);
    return 0;
}
"""


@pytest.mark.skipif(GXX is None, reason="g++ not found")
def test_pool_runs_every_item_once_and_returns():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "pool.cpp")
        exe = os.path.join(d, "pool")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run([GXX, "-O2", "-std=c++17", "-pthread", "-I", CSRC, src, "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        words = r.stdout.split()
        got = dict(zip(words[0::2], (int(x) for x in words[1::2])))
        assert got["bad"] == 0
        assert got["resized"] > 10           # workers were made again and again
        assert got["throw_reported"] == 1
        assert got["missing"] == 0
