"""The path pool and untrusted snapshots on the GPU path.

Paths are appended to one device pool by spawns, replans and UpdatePath
(Simulator.cpp:97-124); the points of replaced paths and of destroyed agents
are reclaimed by compacting the pool when they outnumber the live ones.  A
snapshot file is checked on the host before anything in it reaches the
device."""
import ctypes
import os
import struct
import tempfile

import numpy as np
import pytest

from ecmgenerator_amd import _native as N
from ecmgenerator_amd.simulator import Params, Simulator
from helpers import compare_step, oracle_state, small_scene

pytestmark = pytest.mark.gpu


def long_path(rng, start, goal, n):
    """n points from start to goal, the inner ones jittered off the line."""
    t = np.linspace(0.0, 1.0, n)[:, None]
    p = start[None, :] * (1 - t) + goal[None, :] * t
    p[1:-1] += rng.uniform(-1.0, 1.0, (n - 2, 2))
    return p.astype(np.float32)


def test_path_pool_compaction_keeps_live_paths(native_gpu):
    """UpdatePath 40 times over on 60 agents, with agents destroyed in
    between: the pool stays within a few times the live points (it would
    hold ~200k points without reclaiming), every active agent keeps exactly
    its latest path, and the next step equals the oracle's bit for bit."""
    from oracle import oracle_py as O
    box, rects, obs, ecm, agents = small_scene(8, half=40.0, n_rects=4, n_agents=80)
    p = Params(max_agents=128, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    ids = sim.SpawnAgents(agents)
    live = [int(i) for i in ids if i >= 0][:60]
    rng = np.random.default_rng(3)
    want = {}
    appended = 0
    for it in range(40):
        for k, i in enumerate(live):
            start = sim.state()["position"][i] if k == 0 else np.asarray(agents.start[0])
            goal = np.array([rng.uniform(-35, 35), rng.uniform(-35, 35)], np.float32)
            path = long_path(rng, np.asarray(start, np.float32), goal, 80 + (it + k) % 40)
            sim.SetPath(i, path)
            want[i] = path
            appended += path.shape[0]
        if it % 10 == 9:
            sim.DestroyAgent(live.pop())
        sim.Update()
    off, cnt, pool = sim.paths()
    act = sim.state()["active"].astype(bool)
    live_pts = int(cnt[act].sum())
    print(f"appended {appended} points, pool {pool.shape[0]}, live {live_pts}", flush=True)
    # compaction runs when the pool must grow while its dead points
    # outnumber the live ones, so it never holds much more than 4x the live
    assert pool.shape[0] < appended // 4
    assert pool.shape[0] <= 4 * live_pts + 1000
    for i in live:
        if act[i]:
            assert np.array_equal(pool[off[i]: off[i] + cnt[i]], want[i]), i
    before = oracle_state(sim)
    ora = O.step(p, before, obs, ecm)
    sim.Update()
    res = compare_step(sim, before, ora)
    assert all(v == 0 for v in res.values()), res
    sim.close()


def _layout(n_slots, pu):
    """Byte offsets of the snapshot fields (ecm_sim_snapshot_save)."""
    o = {"pu": 16 + ctypes.sizeof(N.SimParams) + 12}
    o["fs"] = o["pu"] + 8
    o["pos"] = o["fs"] + 8
    o["off"] = o["pos"] + n_slots * (8 * 4 + 4 * 2 + 1)
    o["cnt"] = o["off"] + 4 * n_slots
    o["pool"] = o["cnt"] + 4 * n_slots
    o["stack"] = o["pool"] + 8 * pu
    return o


def test_corrupt_snapshots_are_rejected(native_gpu):
    """A snapshot whose path count, path offsets, free-slot stack or length
    is wrong is refused with an error before it changes the handle, and a
    good snapshot loads afterwards as before."""
    from oracle import oracle_py as O
    box, rects, obs, ecm, agents = small_scene(9, half=30.0, n_rects=3, n_agents=50)
    p = Params(max_agents=64, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    sim.SpawnAgents(agents)
    for _ in range(3):
        sim.Update()
    with tempfile.TemporaryDirectory() as d:
        good = os.path.join(d, "good.bin")
        sim.snapshot_save(good)
        raw = bytearray(open(good, "rb").read())
        pu = struct.unpack_from("<Q", raw, _layout(64, 0)["pu"])[0]
        lay = _layout(64, pu)
        fs = struct.unpack_from("<Q", raw, lay["fs"])[0]
        act = sim.state()["active"].astype(bool)
        a = int(np.flatnonzero(act)[0])
        cases = {}
        b = bytearray(raw); struct.pack_into("<Q", b, lay["pu"], 1 << 40); cases["huge path count"] = b
        b = bytearray(raw); struct.pack_into("<i", b, lay["off"] + 4 * a, pu + 5); cases["path offset"] = b
        b = bytearray(raw); struct.pack_into("<i", b, lay["cnt"] + 4 * a, 0); cases["empty active path"] = b
        if fs:
            b = bytearray(raw); struct.pack_into("<i", b, lay["stack"], -7); cases["stack entry"] = b
            b = bytearray(raw); struct.pack_into("<i", b, lay["stack"], a); cases["active slot on the stack"] = b
        cases["truncated"] = raw[: len(raw) // 2]
        state0 = sim.state()
        for name, blob in cases.items():
            bad = os.path.join(d, "bad.bin")
            open(bad, "wb").write(bytes(blob))
            with pytest.raises(N.NativeError):
                sim.snapshot_load(bad)
            print(f"{name}: rejected ({N.lib().ecm_last_error().decode()})", flush=True)
            for k in ("position", "velocity", "active"):
                assert np.array_equal(sim.state()[k], state0[k]), (name, k)
        sim.snapshot_load(good)
    before = oracle_state(sim)
    ora = O.step(p, before, obs, ecm)
    sim.Update()
    res = compare_step(sim, before, ora)
    assert all(v == 0 for v in res.values()), res
    sim.close()


def test_repacked_pool_keeps_each_agents_path(native_gpu):
    """The pool is repacked in Morton order every 64 steps: over 200 steps
    (three repacks, agents reaching their goals and being destroyed, a spawn
    batch in between) every active agent keeps exactly the points it was
    given, the pool shrinks to the live points, and a step after the last
    repack equals the oracle's bit for bit.  No planner is attached, so no
    path changes on its own."""
    from oracle import oracle_py as O
    box, rects, obs, ecm, agents = small_scene(12, half=40.0, n_rects=4, n_agents=400)
    p = Params(max_agents=512, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    ids = [int(i) for i in sim.SpawnAgents(agents) if i >= 0]

    def snapshot_paths():
        off, cnt, pool = sim.paths()
        act = sim.state()["active"].astype(bool)
        return {i: pool[off[i]: off[i] + cnt[i]].copy() for i in np.flatnonzero(act)}, pool.shape[0], int(cnt[act].sum())

    want, _, _ = snapshot_paths()
    n_destroyed = 0
    live_before = None
    for step in range(200):
        sim.Update()
        if step == 100:
            # spawn the agents that reached their goals again, with new ids' paths recorded
            _, _, _, _, more = small_scene(13, half=40.0, n_rects=4, n_agents=60)
            for i in sim.SpawnAgents(more):
                if i >= 0:
                    off, cnt, pool = sim.paths()
                    want[int(i)] = pool[off[i]: off[i] + cnt[i]].copy()
        if step in (63, 64, 127, 128, 191, 192, 199):
            got, pool_n, live = snapshot_paths()
            for i, pts in got.items():
                assert np.array_equal(pts, want[i]), (step, i)
            n_destroyed = len(want) - len(got)
            print(f"step {step}: {len(got)} active, pool {pool_n} points, live {live}", flush=True)
            # the repack before steps 64, 128 and 192 keeps the paths of the
            # agents active after the step before
            if step in (64, 128, 192):
                assert pool_n == live_before, (step, pool_n, live_before)
            live_before = live
    assert n_destroyed > 0
    before = oracle_state(sim)
    ora = O.step(p, before, obs, ecm)
    sim.Update()
    res = compare_step(sim, before, ora)
    assert all(v == 0 for v in res.values()), res
    sim.close()
