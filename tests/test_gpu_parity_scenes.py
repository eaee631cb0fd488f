"""GPU step vs CPU oracle on the seeded scenarios with a built ECM and
planner paths: bit-exact per step from identical states.  c0 (5k agents) is
compared in full at several steps and on agent subsets at every step of its
600; c1 (100k) on random agent subsets (every agent reads only step-start
state, so a subset is exact) at steps 0, 1, 10, 100 and the last of 600;
c2 (1M agents in two opposing flows jammed in the 64 gaps of the barrier)
on stratified subsets -- agents in the gaps, in the compressed crowd in
front of the barrier, and uniform -- at steps 0, 1, 10, 100, 150 and the
last of 300, with the ORCA velocities compared too and the 3D linear
program (ORCA.cpp:592-669) shown to have decided many of the compared
agents."""
import numpy as np
import pytest

from ecmgenerator_amd import workloads as W
from ecmgenerator_amd.simulator import Params, Simulator
from helpers import compare_step, compare_step_close, oracle_state

pytestmark = pytest.mark.gpu


def random_subset(n):
    def choose(state, rng):
        act = np.flatnonzero(state["active"])
        return np.sort(rng.choice(act, size=min(n, act.size), replace=False))
    return choose


def c2_strata(n_gap=700, n_front=700, n_rand=400):
    """Agents of the c2 crowd to compare in one step: in and beside the gaps
    (|x| < 8 m; the barrier is x in [-2, 2]), in the crowd pressed against
    the barrier (8 m <= |x| < 40 m), and uniform over every active agent."""
    def choose(state, rng):
        act = np.flatnonzero(state["active"])
        x = np.abs(state["position"][act, 0])
        parts = []
        for pool, n in ((act[x < 8.0], n_gap), (act[(x >= 8.0) & (x < 40.0)], n_front), (act, n_rand)):
            if pool.size:
                parts.append(rng.choice(pool, size=min(n, pool.size), replace=False))
        return np.unique(np.concatenate(parts))
    return choose


def run(wl, steps, choose=None, keep=True, seed=0, every=1, learned=None, close=False, shape=None):
    """Step the workload, comparing the GPU with the oracle at `steps` on the
    agents `choose(state, rng)` picks (every agent when None), with the
    launch shape `shape` (cell scale, workgroup, ORCA obstacle slots) when
    given.  Returns the worst mismatch counts, the run's statistics and
    per-step tallies of the compared agents (how many the 3D LP decided,
    obstacle constraints)."""
    from oracle import oracle_py as O
    p = wl.params
    p.keep_orca_velocity = keep
    sim = Simulator(wl.ecm, wl.obstacles, wl.box, p, planner=wl.scene)
    ids = sim.SpawnAgents(wl.agents)
    assert (ids >= 0).all(), "every generated agent must spawn"
    if learned is not None:
        sim.set_learned(*learned)
    if shape is not None:
        sim.set_tuning(*shape)
        sim.set_graph(True)
    rng = np.random.default_rng(seed)
    worst, tally = {}, []
    for it in range(max(steps) + 1):
        if it in steps:
            before = oracle_state(sim)
            sub = None if choose is None else choose(before, rng)
            ora = O.step(p, before, wl.obstacles, wl.ecm, subset=sub)
            sim.Update()
            idx = np.flatnonzero(before["active"]) if sub is None else sub
            if close:
                res, err = compare_step_close(sim, before, ora, idx)
                print(f"{wl.name} step={it} max errors {err}", flush=True)
            else:
                res = compare_step(sim, before, ora, subset=sub)
            t = {"step": it, "compared": int(idx.size), "used_3d": int(ora["used_3d"][idx].sum()),
                 "max_obst_constraints": int(ora["diag"][idx, 1].max()) if idx.size else 0}
            tally.append(t)
            if it % every == 0 or any(res.values()):
                print(f"{wl.name} step={it} n={int(before['active'].sum())} {t} {res}", flush=True)
            for k, v in res.items():
                worst[k] = max(worst.get(k, 0), v)
        else:
            sim.Update()
    st = sim.stats()
    if learned is not None:
        sim.set_learned(None)
    sim.close()
    return worst, st, tally


def test_c0_full_parity():
    wl = W.build("c0", seed=1)
    worst, st, _ = run(wl, steps=(0, 1, 5, 20, 60))
    print(st)
    assert all(v == 0 for v in worst.values()), worst


def test_c0_every_step_subset_parity():
    """All 600 steps of the c0 run, each compared on a fresh random subset
    of 1,000 agents (about 120 comparisons per agent over the run)."""
    wl = W.build("c0", seed=1)
    worst, st, _ = run(wl, steps=range(600), choose=random_subset(1000), seed=3, every=100)
    assert all(v == 0 for v in worst.values()), worst


def test_c0_reference_literal_parity():
    wl = W.build("c0-ref", seed=2)
    worst, st, _ = run(wl, steps=(0, 3, 30))
    assert all(v == 0 for v in worst.values()), worst


def test_c1_subset_parity():
    wl = W.build("c1", seed=1)
    worst, st, _ = run(wl, steps=(0, 1, 10, 100, 599), choose=random_subset(1500))
    assert all(v == 0 for v in worst.values()), worst


def test_c2_stratified_parity_full_size():
    """Bit-exact positions, velocities, ORCA velocities, attraction points,
    neighbour lists and replans on ~1,800 agents per compared step, most of
    them in or in front of the gaps; the 3D LP decides hundreds of them.  Run
    with the launch shape the benchmark's autotuner picks at c2 (the main
    ORCA pass inside the k-nearest kernel) and step graphs."""
    wl = W.build("c2", seed=1)
    worst, st, tally = run(wl, steps=(0, 1, 10, 100, 150, 299), choose=c2_strata(), keep=True,
                           shape=(1.25, 256, 2))
    print(st, flush=True)
    assert all(v == 0 for v in worst.values()), worst
    late = [t for t in tally if t["step"] >= 100]
    assert all(t["used_3d"] >= 200 for t in late), tally
    assert max(t["max_obst_constraints"] for t in tally) >= 3, tally
    assert st["lp3d"] > 0.2 * st["agent_updates"] / 3, st


def test_c3_learned_step_against_oracle():
    """The learned-proposal step at 1M agents (the c2 crowd with the
    committed network and threshold) against the oracle's plain step on the
    stratified subsets: ORCA velocities and velocities within 1e-4 m/s,
    positions within max(1e-5 m, 1 ulp), discrete outputs exact.  Most
    compared agents are decided by an accepted proposal, the rest by the
    queued full solve."""
    import json
    from ecmgenerator_amd import learned as Lm
    from ecmgenerator_amd.learned import ProposalNet
    with np.load(Lm.DEFAULT_WEIGHTS, allow_pickle=False) as z:
        theta = float(json.loads(str(z["meta"]))["theta"])
    net = ProposalNet.load()
    wl = W.build("c2", seed=1)
    worst, st, tally = run(wl, steps=(0, 1, 10, 100, 150, 299), choose=c2_strata(), keep=True,
                           learned=(net, theta), close=True)
    net.close()
    print(st, flush=True)
    assert all(v == 0 for v in worst.values()), worst
    assert st["learned_tried"] > 0.3 * st["agent_updates"], st
    assert st["learned_accepted"] > 0.1 * st["learned_tried"], st
    assert sum(t["used_3d"] for t in tally) >= 600, tally


def test_c2_fp64_lp_resolve_error():
    """The GPU's ORCA velocities against an FP64 re-solve of the same LPs
    (same constraints, exact convex optimum) on a sample of c2 agents after
    the crowd has formed; reports the error distribution and the 3D-LP
    fallback rate.  Discrete-decision mismatches against the oracle are zero
    by the bit-exact tests above."""
    from oracle import oracle_py as O
    from oracle.lp_fp64 import qp_fp64
    wl = W.build("c2", seed=1)
    p = wl.params
    p.keep_orca_velocity = True
    sim = Simulator(wl.ecm, wl.obstacles, wl.box, p, planner=wl.scene)
    sim.SpawnAgents(wl.agents)
    for _ in range(20):
        sim.Update()
    before = oracle_state(sim)
    st0 = sim.stats()
    sim.Update()
    st1 = sim.stats()
    after = sim.state()
    nbr, cnt = sim.neighbors()
    rng = np.random.default_rng(3)
    act = np.flatnonzero(before["active"].astype(bool) & after["active"].astype(bool))
    sample = rng.choice(act, size=2000, replace=False)
    errs, n3d = [], 0
    for a in sample:
        cons, nob = O.constraints(p, before, wl.obstacles, int(a), nbr[a][: cnt[a]])
        vmax = p.max_speed
        r, _ = O.lp2d(cons, after["pref_velocity"][a], vmax)
        if r < len(cons):
            n3d += 1                    # 2D LP infeasible: the 3D LP decides
            continue
        opt = qp_fp64(cons, after["pref_velocity"][a], vmax)
        if opt is None:
            n3d += 1
            continue
        errs.append(float(np.linalg.norm(after["orca_velocity"][a].astype(np.float64) - opt)))
    e = np.array(errs)
    lp3d_rate = (st1["lp3d"] - st0["lp3d"]) / max(1, st1["agent_updates"] - st0["agent_updates"])
    print(f"c2 FP64 re-solve over {e.size} agents: max {e.max():.3e} m/s, p99.9 {np.quantile(e, 0.999):.3e}, "
          f"median {np.median(e):.3e}; 3D in sample {n3d}/{sample.size}; step 3D-LP rate {lp3d_rate:.4f}", flush=True)
    assert np.quantile(e, 0.999) <= 1e-4
    assert np.median(e) <= 1e-6
    sim.close()
