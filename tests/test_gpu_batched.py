"""Batched multi-scenario stepping: several independent environments in
one handle, advanced by the same launches.  The environments' world boxes
overlap on purpose (all centred on the origin): agents of different
environments at the same place must not see each other."""
import os
import tempfile

import numpy as np
import pytest

from ecmgenerator_amd.simulator import Params, Simulator
from helpers import compare_step, oracle_state, small_scene

pytestmark = pytest.mark.gpu

SCENES = [(1, 50.0, 8, 300), (2, 30.0, 4, 500), (3, 80.0, 12, 200)]


def build(p):
    scenes = [small_scene(seed, half=half, n_rects=nr, n_agents=na) for seed, half, nr, na in SCENES]
    box, _, obs, ecm, agents = scenes[0]
    sim = Simulator(ecm, obs, box, p)
    assert (sim.SpawnAgents(agents) >= 0).sum() > 0.9 * agents.n
    for box, _, obs, ecm, agents in scenes[1:]:
        e = sim.AddEnvironment(ecm, obs, box)
        assert (sim.SpawnAgents(agents, env=e) >= 0).sum() > 0.9 * agents.n
    assert sim.num_environments() == len(SCENES)
    return sim, scenes


def test_environments_match_oracle_per_environment():
    from oracle import oracle_py as O
    p = Params(max_agents=2048, keep_orca_velocity=True)
    sim, scenes = build(p)
    env = sim.agent_environments()
    assert sorted(set(env[env >= 0].tolist())) == [0, 1, 2]
    for it in range(8):
        before = oracle_state(sim)
        env = sim.agent_environments()
        oras = []
        for e, (box, _, obs, ecm, agents) in enumerate(scenes):
            st = dict(before)
            st["active"] = (before["active"].astype(bool) & (env == e)).astype(np.uint8)
            oras.append((st, O.step(p, st, obs, ecm)))
        sim.Update()
        for e, (st, ora) in enumerate(oras):
            res = compare_step(sim, st, ora, subset=np.flatnonzero(env == e))
            print(f"step {it} env {e}: {res}", flush=True)
            assert all(v == 0 for v in res.values()), (it, e, res)
    sim.close()


def test_multi_environment_snapshot_round_trip():
    p = Params(max_agents=2048)
    sim, _ = build(p)
    for _ in range(3):
        sim.Update()
    with tempfile.TemporaryDirectory() as d:
        snap = os.path.join(d, "s.bin")
        sim.snapshot_save(snap)
        for _ in range(4):
            sim.Update()
        a = sim.state()
        sim.snapshot_load(snap)
        for _ in range(4):
            sim.Update()
        b = sim.state()
    for k in ("position", "velocity", "attraction", "active"):
        assert np.array_equal(a[k], b[k]), k
    sim.close()


def _eq_rows(a, b):
    """Rows of a and b equal bit for bit (NaN equal to NaN)."""
    e = (a == b) | (np.isnan(a) & np.isnan(b))
    return e.all(axis=1) if e.ndim == 2 else e


@pytest.mark.timeout(1200)
def test_c4_all_environments_against_oracle():
    """The c4 workload (256 environments of 100-1000 m with their own
    obstacles, ECM and planner; ~2.2M agents) stepped by one graph per step:
    every environment compared bit for bit with the oracle stepping it
    alone, on 24 agents per environment at steps 0, 1 and the last of 300."""
    from ecmgenerator_amd import workloads as W
    from oracle import oracle_py as O
    wl = W.build("c4", seed=1)
    p = wl.params
    p.keep_orca_velocity = True
    sim, ids = wl.make_sim()
    sim.set_graph(True)
    rng = np.random.default_rng(5)
    keys = ("position", "velocity", "attraction", "active")
    compared = 0
    for it in range(300):
        if it not in (0, 1, 299):
            sim.Update()
            continue
        before = oracle_state(sim)
        env = sim.agent_environments()
        act = before["active"].astype(bool)
        rows = []
        for e, w in enumerate(wl.envs):
            members = np.flatnonzero(act & (env == e))
            if members.size == 0:
                continue
            sub = np.sort(rng.choice(members, size=min(24, members.size), replace=False))
            st = dict(before)
            st["active"] = (act & (env == e)).astype(np.uint8)
            ora = O.step(p, st, w.obstacles, w.ecm, subset=sub)
            rows.append((e, sub, {k: ora[k][sub].copy() for k in keys + ("pref_velocity", "v_orca", "nbr",
                                                                         "nbr_cnt", "replan")}))
            del ora
        sim.Update()
        after = sim.state()
        nbr, cnt = sim.neighbors()
        rp = np.zeros(sim.max_agents, np.uint8)
        rp[sim.replans()] = 1
        bad = {}
        for e, sub, o in rows:
            alive = o["active"].astype(bool)
            res = {k: int(np.sum(~_eq_rows(after[k][sub], o[k]))) for k in keys}
            res["pref_velocity"] = int(np.sum(~_eq_rows(after["pref_velocity"][sub][alive], o["pref_velocity"][alive])))
            res["v_orca"] = int(np.sum(~_eq_rows(after["orca_velocity"][sub][alive], o["v_orca"][alive])))
            res["nbr"] = int(np.sum((nbr[sub][alive] != o["nbr"][alive]).any(axis=1)))
            res["nbr_cnt"] = int(np.sum(cnt[sub][alive] != o["nbr_cnt"][alive]))
            res["replan"] = int(np.sum(rp[sub] != o["replan"]))
            if any(res.values()):
                bad[e] = res
            compared += sub.size
        print(f"c4 step {it}: {len(rows)} environments, {sum(s.size for _, s, _ in rows)} agents compared, "
              f"mismatching environments {bad}", flush=True)
        assert not bad, (it, bad)
    assert compared > 3 * 256 * 20
    sim.close()
