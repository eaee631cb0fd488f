"""GPU step vs CPU oracle on the seeded scenarios with a built ECM and
planner paths: bit-exact per step from identical states.  c0 (5k agents) is
compared in full at several steps and on agent subsets at every step of its
600; c1 (100k) and c2 (1M) on random agent subsets (every agent reads only
step-start state, so a subset is exact) at steps 0, 1, 10, 100 and the last
of each run (600 and 300 steps)."""
import numpy as np
import pytest

from ecmgenerator_amd import workloads as W
from ecmgenerator_amd.simulator import Params, Simulator
from helpers import compare_step, oracle_state

pytestmark = pytest.mark.gpu


def run(wl, steps, subset_n=None, keep=True, seed=0, every=1):
    from oracle import oracle_py as O
    p = wl.params
    p.keep_orca_velocity = keep
    sim = Simulator(wl.ecm, wl.obstacles, wl.box, p, planner=wl.scene)
    ids = sim.SpawnAgents(wl.agents)
    assert (ids >= 0).all(), "every generated agent must spawn"
    rng = np.random.default_rng(seed)
    worst = {}
    for it in range(max(steps) + 1):
        if it in steps:
            before = oracle_state(sim)
            sub = None
            if subset_n is not None:
                act = np.flatnonzero(before["active"])
                sub = np.sort(rng.choice(act, size=min(subset_n, act.size), replace=False))
            ora = O.step(p, before, wl.obstacles, wl.ecm, subset=sub)
            sim.Update()
            res = compare_step(sim, before, ora, subset=sub)
            if it % every == 0 or any(res.values()):
                print(f"{wl.name} step={it} n={int(before['active'].sum())} {res}", flush=True)
            for k, v in res.items():
                worst[k] = max(worst.get(k, 0), v)
        else:
            sim.Update()
    st = sim.stats()
    sim.close()
    return worst, st


def test_c0_full_parity():
    wl = W.build("c0", seed=1)
    worst, st = run(wl, steps=(0, 1, 5, 20, 60))
    print(st)
    assert all(v == 0 for v in worst.values()), worst


def test_c0_every_step_subset_parity():
    """All 600 steps of the c0 run, each compared on a fresh random subset
    of 1,000 agents (about 120 comparisons per agent over the run)."""
    wl = W.build("c0", seed=1)
    worst, st = run(wl, steps=range(600), subset_n=1000, seed=3, every=100)
    assert all(v == 0 for v in worst.values()), worst


def test_c0_reference_literal_parity():
    wl = W.build("c0-ref", seed=2)
    worst, st = run(wl, steps=(0, 3, 30))
    assert all(v == 0 for v in worst.values()), worst


def test_c1_subset_parity():
    wl = W.build("c1", seed=1)
    worst, st = run(wl, steps=(0, 1, 10, 100, 599), subset_n=1500)
    assert all(v == 0 for v in worst.values()), worst


def test_c2_subset_parity_full_size():
    wl = W.build("c2", seed=1)
    worst, st = run(wl, steps=(0, 1, 10, 100, 299), subset_n=600, keep=False)
    assert all(v == 0 for v in worst.values()), worst


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
