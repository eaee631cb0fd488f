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
    ap.add_argument("--cpu-sample-agents", type=int, default=40000,
                    help="agents in the CPU baseline's sample (about 10 s of single-core work at 1M agents)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (reference for HBM counters)")
    ap.add_argument("--learned", default=None,
                    help="proposal-network weights (.npz) for the learned leg; default: the committed weights if present")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--theta", type=float, default=None, help="proposal logit threshold (default: from the weights file)")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels eagerly instead of replaying step graphs")
    ap.add_argument("--autotune", type=int, default=5,
                    help="steps per autotuner trial during warmup (0: default launch shape)")
    ap.add_argument("--launch-shape", default=None,
                    help="CELL_SCALE,KNN_BLOCK,ORCA_OBST_SLOTS to use instead of autotuning (with --autotune 0)")
    ap.add_argument("--prerun", type=int, default=None,
                    help="untimed steps before warmup (default: the workload's own, e.g. until the c2 jam forms)")
    return ap.parse_args()


def local_density(pos, active, cell=1.0):
    """Agents per m^2 on a grid of `cell` m squares: the fullest square,
    and the 99th percentile and mean of the density around each agent (the
    count of its own square)."""
    p = pos[active.astype(bool)].astype(np.float64)
    if p.shape[0] == 0:
        return {"peak": 0.0, "p99_agent": 0.0, "mean_agent": 0.0}
    ij = np.floor(p / cell).astype(np.int64)
    ij -= ij.min(axis=0)
    key = ij[:, 0] * (int(ij[:, 1].max()) + 1) + ij[:, 1]
    _, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    per_agent = cnt[inv] / (cell * cell)
    return {"peak": float(cnt.max() / (cell * cell)), "p99_agent": float(np.quantile(per_agent, 0.99)),
            "mean_agent": round(float(per_agent.mean()), 3),
            "frac_agents_ge3": round(float((per_agent >= 3.0).mean()), 4)}


def mean_path_points(sim):
    """Mean path length in points over the active agents (the P of the
    per-agent-update byte count 58 + 8 P)."""
    off, cnt, _ = sim.paths()
    act = sim.state()["active"].astype(bool)
    return float(cnt[act].mean()) if act.any() else 0.0


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
        print("bench.py: runs on one GPU only; sharding the step over several GPUs is not implemented. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from ecmgenerator_amd import workloads as W
    from ecmgenerator_amd.simulator import Simulator

    name = args.workload or W.DEFAULT_BENCH
    t0 = time.time()
    wl = W.build(name, seed=args.seed)
    t_setup = time.time() - t0
    batched = isinstance(wl, W.BatchWorkload)
    if batched:
        print(f"[bench] workload {name}: {len(wl.envs)} environments, {wl.n_agents} agents, setup {t_setup:.1f}s",
              file=sys.stderr, flush=True)
        sim, ids_per_env = wl.make_sim()
        n_spawned = int(sum((i >= 0).sum() for i in ids_per_env))
        cnts = np.concatenate([w.agents.path_cnt[i >= 0] for w, i in zip(wl.envs, ids_per_env)])
        mean_path = float(cnts.mean())
    else:
        print(f"[bench] workload {name}: {wl.agents.n} agents, {wl.ecm.n_edge} ECM edges, "
              f"{wl.obstacles.n_rings} obstacles, setup {t_setup:.1f}s", file=sys.stderr, flush=True)
        sim = Simulator(wl.ecm, wl.obstacles, wl.box, wl.params, planner=wl.scene)
        ids = sim.SpawnAgents(wl.agents)
        n_spawned = int((ids >= 0).sum())
        mean_path = float(wl.agents.path_cnt[ids >= 0].mean())
    print(f"[bench] spawned {n_spawned}", file=sys.stderr, flush=True)
    if not args.no_graph:
        sim.set_graph(True)
    prerun = args.prerun if args.prerun is not None else int(getattr(wl, "info", {}).get("prerun_steps", 0))
    if prerun:
        t0 = time.time()
        for i in range(prerun):
            sim.Update()
            if (i + 1) % 50 == 0:
                sim.sync()
                print(f"[bench] prerun step {i + 1}/{prerun} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
        sim.sync()

    import tempfile
    snap_dir = tempfile.mkdtemp(prefix="ecm_bench_")
    snap = os.path.join(snap_dir, "state.bin")
    sim.snapshot_save(snap)
    # the autotuner times its trials on the live workload (advancing it);
    # every timed step below starts again from the state after the prerun
    tuning = None
    if args.autotune > 0:
        tuning = sim.autotune(args.autotune)
        print(f"[bench] autotune: {tuning}", file=sys.stderr, flush=True)
    elif args.launch_shape:
        cs, kb, os_ = args.launch_shape.split(",")
        sim.set_tuning(cell_scale=float(cs), knn_block=int(kb), orca_obst_slots=int(os_))
        tuning = sim.get_tuning()
    # both legs time the same steps: the state just loaded, W untimed warmup
    # steps (the first re-creates buffers after the load and has no neighbour
    # seeds; the rest bring the GPU out of its idle state of the file load),
    # then K timed steps
    first_ms = _load_and_warm(sim, snap, args.warmup)
    print(f"[bench] plain: first step after load {first_ms:.2f} ms (untimed)", file=sys.stderr, flush=True)
    st0 = sim.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.Update()
    sim.sync()
    t1 = time.perf_counter()
    st1 = sim.stats()
    # the crowd at the end of the timed window: local density and path length
    window_state = sim.state()
    density = local_density(window_state["position"], window_state["active"])
    mean_path = mean_path_points(sim)
    # per-kernel HIP-event timing on the handle's stream, separate steps
    sim.set_timing(True)
    kt = np.zeros(6)
    n_kt = max(3, min(args.steps, 10))
    for _ in range(n_kt):
        sim.Update()
        kt += np.array(sim.last_timing())
    sim.set_timing(False)
    updates = st1["agent_updates"] - st0["agent_updates"]
    # beside (never as) the value: the rate when every step is followed by
    # GetPositionData, as a renderer drawing every step calls it; that read
    # waits for the step, brings the positions into slot order (k_unsort)
    # and copies every slot's position to the host over PCIe
    n_rb = max(3, min(args.steps, 10))
    pos_host = sim.pinned("position")         # a page-locked host buffer, kept from step to step
    sim.sync()
    t2 = time.perf_counter()
    for _ in range(n_rb):
        sim.Update()
        sim.GetPositionData(out=pos_host)
    t3 = time.perf_counter()
    readback = {"ms_per_step": round(1000.0 * (t3 - t2) / n_rb, 4),
                "value": round((updates / args.steps) * n_rb / (t3 - t2), 1),
                "bytes_to_host_per_step": int(pos_host.nbytes), "steps": n_rb,
                "what": "each step followed by GetPositionData into a page-locked buffer (slot-order pass + "
                        "device-to-host copy)"}
    wall = t1 - t0
    ms_per_step = 1000.0 * wall / args.steps
    value = updates / wall
    kt = kt / n_kt
    # phases (HIP events on the simulator's stream; no proposal phase in the
    # plain step), one kernel each except the grid build and the queue passes
    shape = sim.get_tuning()
    fused = wl.params.k in (5, 10) and shape["orca_obst_slots"] == 2
    phases = {"grid_build": kt[0], "attraction": kt[5]}
    if fused:           # k_knn_orca: the k-nearest pass and the main ORCA pass in one kernel
        phases["knn_orca_main"] = kt[1] - kt[5]
    else:
        phases["knn"] = kt[1] - kt[5]
        phases["orca_main"] = kt[4]
    phases["orca_queues"] = kt[3] - kt[4]
    # the roofline is quoted on the longest single kernel
    singles = [k for k in phases if k not in ("grid_build", "orca_queues")]
    dom_name = max(singles, key=lambda k: phases[k])
    dom_ms = phases[dom_name]
    n_per_step = updates / args.steps
    # algorithmic HBM bytes per agent-update: the state one update must read
    # and write once (pos 8, vel 8, radius 4, preferred speed 4, path offset
    # and count 8, active 1; the P path points the IRM scans, 8 P; writes of
    # pos 8, vel 8, attraction 8, active 1), 58 + 8 P.  Neighbour gathers,
    # the grid passes and verification outputs are not counted.
    b_agent = 58.0 + 8.0 * mean_path
    achieved = b_agent * n_per_step / (dom_ms * 1e-3) / 1e9
    # each kernel's own minimum state traffic per agent, for comparison
    k = wl.params.k
    own_bytes = {
        "grid_build": 1 + 8 + 8 + 4 + 4 + 4 + 16 + 8 + 4,
        # state 16 + 8, path offset and count 8, path points 8 P, attraction
        # point read and written 16, preferred velocity 8, destroyed flag 1
        "attraction": 16 + 8 + 8 + 8 * mean_path + 16 + 8 + 1,
        # state 16, seeds (previous list 4 k + count 4 + map 4), list and count written 4 k + 4
        "knn": 16 + 4 * k + 4 + 4 + 4 * k + 4,
        # state 16 + 8, destroyed flag, radius 4, preferred velocity 8, list
        # and count 4 k + 4, position and velocity written 8 + 8
        "orca_main": 16 + 8 + 1 + 4 + 8 + 4 + 4 * k + 8 + 8,
    }
    # the kernel running both reads the state, the list and its count once
    own_bytes["knn_orca_main"] = own_bytes["knn"] + own_bytes["orca_main"] - (16 + 4 * k + 4)
    traffic, tsrc = _profiled(name, dom_name, "traffic")
    valu, _ = _profiled(name, dom_name, "valu_busy")
    step_gbs = b_agent * n_per_step / (ms_per_step * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None if traffic is None else int(traffic),
            "traffic_source": tsrc, "valu_busy": valu,
            "kernel": dom_name, "kernel_ms": round(float(dom_ms), 4), "bytes_per_agent_update": round(b_agent, 2),
            "algorithmic_bytes_per_launch": int(b_agent * n_per_step),
            "whole_step": {"achieved": round(step_gbs, 3), "frac": round(step_gbs / HBM_PEAK_GBS, 6)},
            "kernel_own_bytes_per_agent": {k: round(v, 2) for k, v in own_bytes.items()},
            "phase_ms": {k: round(float(v), 4) for k, v in phases.items()}}

    graph_on, graph_why = sim.graph_status()
    learned = None
    wpath = args.learned or _default_weights()
    if not args.no_learned and wpath and os.path.exists(wpath):
        learned = learned_leg(sim, wl, snap, wpath, args)
    try:
        os.remove(snap)
        os.rmdir(snap_dir)
    except OSError:
        pass
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb)
    cpu = None
    if not args.no_cpu_baseline and not batched:
        cpu = cpu_baseline(sim, wl, args.cpu_sample_agents)
    sim.close()
    line = {
        "metric": "agent-updates/s at 1M agents, learned vs plain ORCA at 1e-4 m/s; % HBM peak",
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
                   "mean_path_points": round(mean_path, 2),
                   "timed_window_steps": [prerun + max(1, args.warmup) + 1, prerun + max(1, args.warmup) + args.steps]},
        "crowd": {"density_per_m2_1m_grid": density,
                  "lp3d_rate": round((st1["lp3d"] - st0["lp3d"]) / max(1, updates), 5),
                  "active_agents": int(window_state["active"].astype(bool).sum())},
        "launch": {"step_graphs": graph_on, "graph_fallback_reason": graph_why or None, "autotune": tuning},
        "roofline": roof,
        "with_position_readback": readback,
        "cpu_baseline": cpu,
        "learned": learned,
        "stats": {k: st1[k] - st0[k] for k in st1},
    }
    print(json.dumps(line), flush=True)


def _profiled(workload, kernel, what):
    """Value `what` ("traffic": HBM bytes per launch from FETCH_SIZE /
    WRITE_SIZE calibrated against a copy of known size; "valu_busy": the
    fraction of the chip's VALU issue slots the kernel kept busy) of `kernel`
    from the committed rocprofv3 --pmc summary of this workload, or
    (None, None) when there is none."""
    import re
    path = os.path.join(ROOT, "profiles", f"pmc_{workload}.md")
    if not os.path.exists(path):
        return None, None
    m = re.search(r"<!-- " + what + r": (\{.*?\}) -->", open(path).read())
    if not m:
        return None, None
    t = json.loads(m.group(1)).get(kernel)
    return (t, os.path.relpath(path, ROOT)) if t else (None, None)


def _load_and_warm(sim, snap, warmup):
    """Load the snapshot and run `warmup` (at least one) untimed steps;
    returns the first one's time in ms."""
    sim.snapshot_load(snap)
    sim.sync()
    t0 = time.perf_counter()
    sim.Update()
    sim.sync()
    first = 1000.0 * (time.perf_counter() - t0)
    for _ in range(max(0, warmup - 1)):
        sim.Update()
    sim.sync()
    return first


def _default_weights():
    from ecmgenerator_amd import learned as Lm
    return Lm.DEFAULT_WEIGHTS


def learned_leg(sim, wl, snap, wpath, args):
    """The learned-proposal step over the same K steps from the same state:
    throughput, acceptance, and the largest velocity difference to the plain
    step on the first step (implied ORCA-velocity difference, m/s)."""
    from ecmgenerator_amd.learned import ProposalNet
    with np.load(wpath, allow_pickle=False) as z:
        params = z["params"]
        meta = json.loads(str(z["meta"])) if "meta" in z else {}
    theta = args.theta if args.theta is not None else float(meta.get("theta", 0.0))
    net = ProposalNet(params=params)
    # first step from the snapshot, plain then learned, for the velocity check
    sim.snapshot_load(snap)
    act = sim.state()["active"].astype(bool)
    sim.Update()
    a = sim.state()
    sim.snapshot_load(snap)
    sim.set_learned(net, theta)
    sim.Update()
    b = sim.state()
    alive = act & a["active"].astype(bool) & b["active"].astype(bool)
    k_vel = wl.params.dt / wl.params.mass            # v_new - v_old = (v_orca - v_old) * dt / mass
    max_dv = float(np.abs(a["velocity"][alive] - b["velocity"][alive]).max() / k_vel) if alive.any() else 0.0
    max_dp = float(np.abs(a["position"][act] - b["position"][act]).max()) if act.any() else 0.0
    # timed: the same K steps as the plain leg
    first_ms = _load_and_warm(sim, snap, args.warmup)
    print(f"[bench] learned: first step after load {first_ms:.2f} ms (untimed)", file=sys.stderr, flush=True)
    st0 = sim.stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sim.Update()
    sim.sync()
    t1 = time.perf_counter()
    st1 = sim.stats()
    sim.set_timing(True)
    kt = np.zeros(6)
    n_kt = max(3, min(args.steps, 10))
    for _ in range(n_kt):
        sim.Update()
        kt += np.array(sim.last_timing())
    sim.set_timing(False)
    sim.set_learned(None)
    net.close()
    upd = st1["agent_updates"] - st0["agent_updates"]
    tried = st1["learned_tried"] - st0["learned_tried"]
    acc = st1["learned_accepted"] - st0["learned_accepted"]
    kt /= n_kt
    return {"value": round(upd / (t1 - t0), 1), "unit": "agent-updates/s",
            "ms_per_step": round(1000.0 * (t1 - t0) / args.steps, 4),
            "acceptance": round(acc / max(tried, 1), 4), "tried": tried, "accepted": acc,
            "rejected_full_2d_lp": tried - acc, "lp3d": st1["lp3d"] - st0["lp3d"],
            "max_dv_vs_plain": max_dv, "max_dp_vs_plain": max_dp, "tolerance_dv": 1e-4,
            "theta": theta, "weights": os.path.relpath(wpath, ROOT),
            "phase_ms": {"grid_build": round(float(kt[0]), 4), "attraction": round(float(kt[5]), 4),
                         "knn": round(float(kt[1] - kt[5]), 4),
                         "features_mlp": round(float(kt[2]), 4), "orca_verify_integrate": round(float(kt[3]), 4),
                         "orca_verify_main": round(float(kt[4]), 4)},
            # network FLOP rate over the whole proposal phase (features + MLP)
            "proposal_tflops": round(85248.0 * (upd / args.steps) / (kt[2] * 1e-3) / 1e12, 2) if kt[2] > 0 else None}


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
    # both timed legs run pinned to one of the cores this process may use
    allowed = os.sched_getaffinity(0)
    core = min(allowed)
    os.sched_setaffinity(0, {core})
    try:
        t0 = time.perf_counter()
        O.lib().oracle_kdtree_build_only(P.st)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.step(wl.params, st, wl.obstacles, wl.ecm, subset=sample, knn_mode=O.KNN_REFERENCE)
        t_sample = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, allowed)
    # the subset step includes one more tree build; remove it
    per_agent = max(t_sample - t_build, 1e-9) / sample.size
    step_s = t_build + per_agent * act.size
    return {"value": round(act.size / step_s, 2), "unit": "agent-updates/s", "cores": 1, "kind": "port",
            "sample": f"KD-tree build over all {act.size} agents ({t_build:.3f}s) + full update of "
                      f"{sample.size} random agents ({per_agent * 1e3:.3f} ms/agent), extrapolated to one step",
            "cpu": _cpu_model(), "pinned_core": core, "host_cpus": len(allowed)}


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
