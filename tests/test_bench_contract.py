"""bench.py's contract pieces that need no GPU: the one-GPU refusal, and the
committed counter summary it reads for the roofline (HBM bytes per launch
and the VALU busy fraction of the dominant kernel)."""
import os
import subprocess
import sys

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_profiled_roofline_inputs():
    traffic, src = bench._profiled("c2", "knn_orca_main", "traffic")
    valu, src2 = bench._profiled("c2", "knn_orca_main", "valu_busy")
    assert src == src2 == os.path.join("profiles", "pmc_c2.md")
    # measured HBM traffic per launch is at least the whole step's algorithmic
    # bytes (92 B per agent-update over 1M agents) and below 4 KB per agent
    assert 92e6 <= traffic < 4e9
    assert 0.0 < valu <= 1.0
    assert bench._profiled("c2", "no_such_kernel", "traffic") == (None, None)
    assert bench._profiled("no_such_workload", "knn_orca_main", "traffic") == (None, None)
