"""Collision edge cases of the agent half-plane and the neighbour query, GPU
step vs CPU oracle bit for bit (NaN equal to NaN):

* a colliding pair whose relative velocity equals relPos / dt: w = 0, so
  the reference's `w / wLength` (ORCA.cpp:361-363) gives NaN; both agents
  end the step with NaN velocity and position.  One step only: nothing
  steps on from NaN positions.
* two agents at the same position: d2 = 0 <= 1e-4, so each is left out of
  the other's neighbour list (KDTree.cpp:112) and they pass through each
  other.
* an agent between two parallel walls: its two obstacle half-planes have
  opposite normals, so the 2D LP meets |det| <= 1e-4 (ORCA.cpp:526)."""
import numpy as np
import pytest

from ecmgenerator_amd import scenarios as S
from ecmgenerator_amd.simulator import Agents, Params, Simulator
from helpers import compare_step, oracle_state

pytestmark = pytest.mark.gpu

BOX = (-50.0, -50.0, 50.0, 50.0)


def make_sim(starts, goals, vel, rects=((-45.0, 38.0, -38.0, 45.0),)):
    obs = S.rects_to_obstacles(list(rects))
    ecm = S.empty_square_ecm(*BOX)
    n = len(starts)
    p = Params(max_agents=16)
    p.keep_orca_velocity = True
    sim = Simulator(ecm, obs, BOX, p)
    ag = Agents(np.array(starts, np.float32), np.full(n, 0.2, np.float32), np.full(n, 1.3, np.float32),
                np.arange(n, dtype=np.int32), np.ones(n, np.int32), np.array(goals, np.float32))
    ids = sim.SpawnAgents(ag, validate=False)
    assert (ids >= 0).all()
    v = sim.read("velocity")
    v[ids] = np.array(vel, np.float32)
    sim.write("velocity", v)
    return sim, obs, ecm, ids


def test_zero_w_colliding_pair_gives_the_reference_nan():
    from oracle import oracle_py as O
    # 0.25 / 0.1f rounds to 2.5f, so relVel - relPos / dt is exactly 0
    assert np.float32(0.25) / np.float32(0.1) == np.float32(2.5)
    sim, obs, ecm, ids = make_sim(
        starts=[(20.0, 10.0), (20.25, 10.0), (-20.0, -25.0), (-10.0, 30.0)],
        goals=[(40.0, 10.0), (0.0, 10.0), (-40.0, -25.0), (10.0, 30.0)],
        vel=[(2.5, 0.0), (0.0, 0.0), (0.3, 0.1), (-0.2, 0.0)])
    try:
        before = oracle_state(sim)
        ora = O.step(sim.params, before, obs, ecm)
        sim.Update()
        res = compare_step(sim, before, ora)
        after = sim.state()
        print(res, after["velocity"][ids].tolist(), flush=True)
        assert np.isnan(ora["velocity"][ids[:2]]).all(), "the pair must take the NaN branch"
        assert np.isfinite(ora["velocity"][ids[2:]]).all()
        assert all(v == 0 for v in res.values()), res
    finally:
        sim.close()


def test_coincident_pair_is_invisible_to_itself():
    from oracle import oracle_py as O
    sim, obs, ecm, ids = make_sim(
        starts=[(-20.0, -10.0), (-20.0, -10.0), (10.0, 14.0), (12.0, 14.0)],
        goals=[(-40.0, -10.0), (0.0, -10.0), (30.0, 14.0), (-10.0, 14.0)],
        vel=[(0.5, 0.0), (-0.5, 0.0), (0.0, 0.0), (0.0, 0.0)])
    try:
        for it in range(3):
            before = oracle_state(sim)
            ora = O.step(sim.params, before, obs, ecm)
            sim.Update()
            res = compare_step(sim, before, ora)
            print(f"step={it} {res}", flush=True)
            assert all(v == 0 for v in res.values()), res
            if it == 0:
                nbr, cnt = sim.neighbors()
                assert ids[1] not in nbr[ids[0], :cnt[ids[0]]] and ids[0] not in nbr[ids[1], :cnt[ids[1]]]
                assert cnt[ids[0]] == 2 and cnt[ids[1]] == 2
    finally:
        sim.close()


def test_corridor_parallel_walls():
    from oracle import oracle_py as O
    sim, obs, ecm, ids = make_sim(
        starts=[(0.0, 0.0), (3.0, 0.3), (-3.0, -0.4)],
        goals=[(9.0, 0.0), (-9.0, 0.3), (9.0, -0.4)],
        vel=[(1.0, 0.0), (-1.0, 0.0), (0.8, 0.0)],
        rects=((-10.0, 1.5, 10.0, 5.0), (-10.0, -5.0, 10.0, -1.5)))
    try:
        # premise, from the oracle's constraint builder (rows n.x, n.y, p.x, p.y):
        # the first agent's two wall half-planes are parallel
        st0 = oracle_state(sim)
        con, n_obst = O.constraints(sim.params, st0, obs, int(ids[0]), ids[1:])
        assert n_obst == 2
        (ax, ay), (bx, by) = con[0, :2], con[1, :2]
        assert abs(ay * bx - ax * by) <= 1e-4
        for it in range(5):
            before = oracle_state(sim)
            ora = O.step(sim.params, before, obs, ecm)
            sim.Update()
            res = compare_step(sim, before, ora)
            print(f"step={it} {res}", flush=True)
            assert all(v == 0 for v in res.values()), res
    finally:
        sim.close()
