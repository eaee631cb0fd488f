"""Slot pool, goal arrival, spawn validity, replanning and snapshots on the
GPU path (Simulator.cpp:97-124, 168-208, 295-311, 538-590)."""
import ctypes as C
import os

import numpy as np
import pytest

from ecmgenerator_amd import _native as N
from ecmgenerator_amd import scenarios as S
from ecmgenerator_amd.simulator import Agents, Params, Simulator

pytestmark = pytest.mark.gpu


def one(start, path, r=0.25, s=1.3):
    path = np.asarray(path, np.float32)
    return Agents(np.asarray([start], np.float32), np.asarray([r], np.float32), np.asarray([s], np.float32),
                  np.zeros(1, np.int32), np.asarray([len(path)], np.int32), path)


def empty_world(max_agents=8):
    ecm = S.empty_square_ecm(-50, -50, 50, 50)
    obs = S.rects_to_obstacles([])
    return Simulator(ecm, obs, (-50, -50, 50, 50), Params(max_agents=max_agents))


def test_lifo_pool_and_goal_destroy(native_gpu):
    sim = empty_world(4)
    a = sim.SpawnAgent((0.0, 10.0), [(0.0, 10.0), (0.0, 11.0)], 0.25, 1.3)   # within 2 m of its goal
    b = sim.SpawnAgent((20.0, 20.0), [(20.0, 20.0), (-30.0, -30.0)], 0.25, 1.3)
    assert (a, b) == (0, 1)                      # LIFO pool hands out 0, 1, ...
    assert sim.GetNumAgents() == 2 and sim.GetLastIndex() == 1
    sim.Update()
    act = sim.GetActiveFlags()
    assert not act[0] and act[1]                 # DestroyAgent at the goal
    assert sim.GetNumAgents() == 1
    c = sim.SpawnAgent((5.0, -5.0), [(5.0, -5.0), (40.0, 40.0)], 0.25, 1.3)
    assert c == 0                                # the freed slot is on top of the stack
    sim.DestroyAgent(1)
    sim.Update()
    assert sim.GetLastIndex() == 0               # UpdateMaxAgentIndex trims past slot 1
    sim.close()


def test_spawn_validity_and_full_pool(native_gpu):
    sim = empty_world(2)
    assert sim.SpawnAgent((1.0, 1.0), [(1.0, 1.0), (30.0, 30.0)], 0.3, 1.3) == 0
    # within its own clearance of an active agent -> rejected (Simulator.cpp:295-311)
    assert sim.SpawnAgent((1.1, 1.0), [(1.1, 1.0), (30.0, 30.0)], 0.3, 1.3) == -1
    assert sim.SpawnAgent((2.0, 1.0), [(2.0, 1.0), (30.0, 30.0)], 0.3, 1.3) == 1
    assert sim.SpawnAgent((9.0, 9.0), [(9.0, 9.0), (30.0, 30.0)], 0.3, 1.3) == -1   # pool empty
    with pytest.raises(N.NativeError):
        sim.SpawnAgents(Agents(np.zeros((1, 2), np.float32), np.full(1, 0.3, np.float32), np.full(1, 1.3, np.float32),
                               np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros((1, 2), np.float32)))
    sim.close()


def test_replan_uses_step_start_position(native_gpu):
    calls = []

    @N.REPLAN_FN
    def cb(user, aid, x, y, gx, gy, clr, out, cap, n):
        calls.append((aid, x, y, gx, gy, clr))
        pts = [(x, y), (gx, gy)]
        for i, (px, py) in enumerate(pts):
            out[2 * i] = px
            out[2 * i + 1] = py
        n[0] = 2
        return 0

    sim = empty_world(4)
    N.check(sim._L.ecm_sim_set_replan_callback(sim._h, cb, None))
    # an agent exactly on a cell vertex (the box centre) cannot be located
    # (Contains is false within EPSILON of a vertex), so IRM fails
    sim.SpawnAgent((0.0, 0.0), [(0.0, 0.0), (-40.0, 35.0)], 0.25, 1.3)
    v = np.zeros((4, 2), np.float32)
    v[0] = (0.7, -0.4)
    sim.write("velocity", v)
    sim.Update()
    assert sim.replans().tolist() == [0]
    assert len(calls) == 1
    aid, x, y, gx, gy, clr = calls[0]
    assert (aid, x, y, gx, gy) == (0, 0.0, 0.0, -40.0, 35.0) and abs(clr - 0.25) < 1e-7
    assert not np.allclose(sim.GetPositionData()[0], (0.0, 0.0))     # it did move this step
    off, cnt, pool = sim.paths()
    assert cnt[0] == 2 and pool[off[0]].tolist() == [0.0, 0.0]
    sim.close()


def test_snapshot_round_trip(native_gpu, tmp_path):
    from helpers import small_scene
    box, rects, obs, ecm, agents = small_scene(9, n_agents=150)
    p = Params(max_agents=200)
    a = Simulator(ecm, obs, box, p)
    a.SpawnAgents(agents)
    for _ in range(5):
        a.Update()
    f = str(tmp_path / "snap.bin")
    a.snapshot_save(f)
    b = Simulator(S.empty_square_ecm(-1, -1, 1, 1), S.rects_to_obstacles([]), box, p)
    b.snapshot_load(f)
    for _ in range(5):
        a.Update()
        b.Update()
    for k in ("position", "velocity", "attraction", "active"):
        assert np.array_equal(a.read(k), b.read(k)), k
    assert a.GetNumAgents() == b.GetNumAgents()
    a.close()
    b.close()
