"""GPU step vs CPU oracle, bit for bit, from identical input states, on
small seeded scenes (one agent per lane path through the C ABI)."""
import numpy as np
import pytest

from ecmgenerator_amd.simulator import Params, Simulator
from helpers import compare_step, oracle_state, small_scene

pytestmark = pytest.mark.gpu


def run_parity(seed, params, n_agents=300, steps=(0, 1, 2, 5, 10, 20), n_rects=8, half=50.0):
    from oracle import oracle_py as O
    box, rects, obs, ecm, agents = small_scene(seed, half=half, n_rects=n_rects, n_agents=n_agents)
    sim = Simulator(ecm, obs, box, params)
    ids = sim.SpawnAgents(agents)
    assert (ids >= 0).sum() > 0.9 * agents.n
    worst = {}
    last = max(steps)
    for it in range(last + 1):
        if it in steps:
            before = oracle_state(sim)
            ora = O.step(params, before, obs, ecm)
            sim.Update()
            res = compare_step(sim, before, ora)
            print(f"seed={seed} step={it} {res}", flush=True)
            for k, v in res.items():
                worst[k] = max(worst.get(k, 0), v)
        else:
            sim.Update()
    sim.close()
    return worst


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_parity_default_params(seed):
    p = Params(max_agents=512, keep_orca_velocity=True)
    worst = run_parity(seed, p)
    assert all(v == 0 for v in worst.values()), worst


def test_parity_reference_literal_params():
    p = Params.reference(max_agents=512, keep_orca_velocity=True)
    worst = run_parity(7, p)
    assert all(v == 0 for v in worst.values()), worst


def test_parity_dense_crowd():
    p = Params(max_agents=2048, keep_orca_velocity=True)
    worst = run_parity(11, p, n_agents=1800, half=30.0, n_rects=4, steps=(0, 3, 8, 15))
    assert all(v == 0 for v in worst.values()), worst


def test_parity_pillar_forest_obstacle_overflow():
    """1 m pillars on a 3.5 m lattice: about 40% of the agents get several
    obstacle constraints (up to 14).  Those whose constraints outnumber
    their slots in the main ORCA tile take slots from their workgroup's
    spill area, and those with more obstacle constraints than the tile, or
    whose workgroup's spill area is used up, go to the overflow queue
    kernel; the result must still match the oracle bit for bit."""
    from oracle import oracle_py as O
    from ecmgenerator_amd import scenarios as S
    half = 20.0
    box = (-half, -half, half, half)
    rects = [(x, y, x + 1, y + 1) for x in np.arange(-15, 15, 3.5) for y in np.arange(-15, 15, 3.5)]
    obs = S.rects_to_obstacles(rects)
    ecm = S.empty_square_ecm(*box)
    rng = np.random.default_rng(5)
    agents = S.straight_agents(rng, box, 400, rects, margin=0.4)
    p = Params(max_agents=512, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    assert (sim.SpawnAgents(agents) >= 0).sum() > 300
    worst = {}
    used_3d = 0
    for it in range(6):
        before = oracle_state(sim)
        ora = O.step(p, before, obs, ecm)
        sim.Update()
        res = compare_step(sim, before, ora)
        used_3d += int(ora["used_3d"][before["active"].astype(bool)].sum())
        print(f"pillars step={it} {res} {sim.stats()}", flush=True)
        for k, v in res.items():
            worst[k] = max(worst.get(k, 0), v)
    st = sim.stats()
    sim.close()
    assert st["obst_overflow"] > 0, st
    # the 3D LP ran on the device for the agents the oracle sent there, some
    # of them with more obstacle constraints than its register tile holds
    print(f"3D LP: {st['lp3d']} agents, {st['lp3d_big']} of them from the LDS tile", flush=True)
    assert st["lp3d"] > 0 and st["lp3d"] == used_3d, (st, used_3d)
    assert st["lp3d_big"] > 0, st
    assert all(v == 0 for v in worst.values()), worst


def test_parity_pillar_forest_long_overflow_queue():
    """The same pillar lattice over 120 m with 6,000 agents: most of them
    have more than 4 obstacle constraints at the start, so the overflow
    queue holds thousands of agents, more than the overflow kernel's grid
    has wavefronts, and the kernel works one agent per lane instead of one
    agent per wavefront (the short queues of the test above).  Every agent
    must match the oracle bit for bit."""
    from oracle import oracle_py as O
    from ecmgenerator_amd import scenarios as S
    half = 60.0
    box = (-half, -half, half, half)
    rects = [(x, y, x + 1, y + 1) for x in np.arange(-half + 5, half - 5, 3.5)
             for y in np.arange(-half + 5, half - 5, 3.5)]
    obs = S.rects_to_obstacles(rects)
    ecm = S.empty_square_ecm(*box)
    rng = np.random.default_rng(5)
    agents = S.straight_agents(rng, box, 6000, rects, margin=0.4)
    p = Params(max_agents=6016, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    assert (sim.SpawnAgents(agents) >= 0).sum() > 5000    # overlapping starts are rejected
    worst = {}
    longest = 0
    for it in range(3):
        before = oracle_state(sim)
        ora = O.step(p, before, obs, ecm)
        ovf0 = sim.stats()["obst_overflow"]
        sim.Update()
        res = compare_step(sim, before, ora)
        longest = max(longest, sim.stats()["obst_overflow"] - ovf0)
        print(f"pillars-120m step={it} {res} overflow queue so far {longest}", flush=True)
        for k, v in res.items():
            worst[k] = max(worst.get(k, 0), v)
    sim.close()
    assert longest > 1024, longest          # the kernel's grid: 1,024 wavefronts
    assert all(v == 0 for v in worst.values()), worst


def test_parity_obstacle_buckets_longer_than_64_edges():
    """0.5 m squares on a 1.6 m lattice: up to 194 obstacle edges lie within
    the obstacle range of one point, more than the 64 that the overflow
    kernels' fixed tiles hold, so the scene takes the overflow passes with
    a tile sized for the longest bucket list (one agent per wavefront);
    agents have up to 56 obstacle constraints.  Every agent must match the
    oracle bit for bit, the 3D LP included."""
    from oracle import oracle_py as O
    from ecmgenerator_amd import scenarios as S
    half = 16.0
    box = (-half, -half, half, half)
    rects = [(x, y, x + 0.5, y + 0.5) for x in np.arange(-half + 2, half - 2, 1.6)
             for y in np.arange(-half + 2, half - 2, 1.6)]
    obs = S.rects_to_obstacles(rects)
    ecm = S.empty_square_ecm(*box)
    rng = np.random.default_rng(7)
    agents = S.straight_agents(rng, box, 300, rects, margin=0.4)
    p = Params(max_agents=320, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    assert (sim.SpawnAgents(agents) >= 0).sum() > 250
    worst = {}
    used_3d = 0
    most = 0
    for it in range(4):
        before = oracle_state(sim)
        ora = O.step(p, before, obs, ecm)
        sim.Update()
        res = compare_step(sim, before, ora)
        act = before["active"].astype(bool)
        used_3d += int(ora["used_3d"][act].sum())
        most = max(most, int(ora["diag"][act, 1].max()))
        print(f"wide buckets step={it} {res} most obstacle constraints {most}", flush=True)
        for k, v in res.items():
            worst[k] = max(worst.get(k, 0), v)
    st = sim.stats()
    sim.close()
    assert most > 16 and st["obst_overflow"] > 0, (most, st)
    assert st["lp3d"] == used_3d and used_3d > 0, (st, used_3d)
    assert all(v == 0 for v in worst.values()), worst
