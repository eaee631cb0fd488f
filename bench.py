"""Benchmark of the per-timestep crowd update on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line (rank 0): agent-updates/s over K timed steps with the
state resident in HBM, the roofline of the dominant kernel (HIP-event
timed inside the library on the handle's stream) and a CPU baseline: the
oracle (single-threaded restatement of the reference step with its KD-tree
query, linear obstacle scan and linear ECM cell scan) timed on a bounded
sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=None)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample-agents", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (reference for HBM counters)")
    return ap.parse_args()


def copy_calibration(mb: int) -> None:
    """One hipMemcpy device-to-device of `mb` MiB: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE give the counter-to-bytes factor."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0          # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is described in DESIGN.md "
              "and not built. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from ecmgenerator_amd import workloads as W
    from ecmgenerator_amd.simulator import Simulator

    name = args.workload or W.DEFAULT_BENCH
    t0 = time.time()
    wl = W.build(name, seed=args.seed)
    t_setup = time.time() - t0
    print(f"[bench] workload {name}: {wl.agents.n} agents, {wl.ecm.n_edge} ECM edges, "
          f"{wl.obstacles.n_rings} obstacles, setup {t_setup:.1f}s", file=sys.stderr, flush=True)
    sim = Simulator(wl.ecm, wl.obstacles, wl.box, wl.params, planner=wl.scene)
    ids = sim.SpawnAgents(wl.agents)
    n_spawned = int((ids >= 0).sum())
    print(f"[bench] spawned {n_spawned}", file=sys.stderr, flush=True)
    mean_path = float(wl.agents.path_cnt[ids >= 0].mean())

    for _ in range(args.warmup):
        sim.Update()
    sim.sync()
    st0 = sim.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.Update()
    sim.sync()
    t1 = time.perf_counter()
    st1 = sim.stats()
    # per-kernel HIP-event timing on the handle's stream, separate steps
    sim.set_timing(True)
    kt = np.zeros(3)
    n_kt = max(3, min(args.steps, 10))
    for _ in range(n_kt):
        sim.Update()
        kt += np.array(sim.last_timing())
    sim.set_timing(False)
    updates = st1["agent_updates"] - st0["agent_updates"]
    wall = t1 - t0
    ms_per_step = 1000.0 * wall / args.steps
    value = updates / wall
    kt_mean = kt / n_kt
    names = ["grid_build", "attraction_knn", "orca_lp_integrate"]
    dom = int(np.argmax(kt_mean))
    n_per_step = updates / args.steps
    # algorithmic HBM bytes per agent-update of each kernel
    bytes_per_agent = {
        "grid_build": 1 + 8 + 8 + 4 + 4 + 4 + 16 + 8 + 4,
        "attraction_knn": 16 + 8 + 8 + 8 * mean_path + 8 + 4 + 8 + 8 + 4 * wl.params.k + 4,
        "orca_lp_integrate": 16 + 8 + 1 + 4 + 8 + 4 + 4 * wl.params.k + 8 + 8,
    }
    achieved = bytes_per_agent[names[dom]] * n_per_step / (kt_mean[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None, "kernel": names[dom],
            "kernel_ms": {k: round(float(v), 4) for k, v in zip(names, kt_mean)}}

    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sim, wl, args.cpu_sample_agents)
    sim.close()
    line = {
        "metric": "agent-updates/s (plain GPU ORCA step)",
        "value": round(value, 1),
        "unit": "agent-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": name, **wl.describe(), "agents_spawned": n_spawned,
                   "mean_path_points": round(mean_path, 2)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stats": {k: st1[k] - st0[k] for k in st1},
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(sim, wl, n_sample):
    """Oracle (reference KD-tree query, linear obstacle + ECM scans), single
    thread: the KD-tree build over all agents + the per-agent update of a
    random sample of agents, extrapolated to the whole step."""
    from oracle import oracle_py as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_state
    st = oracle_state(sim)
    act = np.flatnonzero(st["active"])
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(act, size=min(n_sample, act.size), replace=False)).astype(np.int32)
    P = O.Problem(wl.params, st, wl.obstacles, wl.ecm, O.KNN_REFERENCE)
    t0 = time.perf_counter()
    O.lib().oracle_kdtree_build_only(P.st)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.step(wl.params, st, wl.obstacles, wl.ecm, subset=sample, knn_mode=O.KNN_REFERENCE)
    t_sample = time.perf_counter() - t0
    # the subset step includes one more tree build; remove it
    per_agent = max(t_sample - t_build, 1e-9) / sample.size
    step_s = t_build + per_agent * act.size
    return {"value": round(act.size / step_s, 2), "unit": "agent-updates/s", "cores": 1, "kind": "port",
            "sample": f"KD-tree build over all {act.size} agents ({t_build:.3f}s) + full update of "
                      f"{sample.size} random agents ({per_agent * 1e3:.3f} ms/agent), extrapolated to one step",
            "cpu": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
