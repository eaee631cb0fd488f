"""Named benchmark / test workloads.

Each workload builds (box, obstacles, ECM, agents, params).  c0 / c1 / c2
come from the seeded scenario generator (ECM built from the obstacles,
paths from the ECM planner).  `proxy-*` workloads use the closed-form ECM
of the empty box and straight two-point paths.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import scenarios as S
from .simulator import Agents, Ecm, Obstacles, Params


@dataclass
class Workload:
    name: str
    box: tuple
    obstacles: Obstacles
    ecm: Ecm
    agents: Agents
    params: Params
    scene: Any = None
    info: dict = field(default_factory=dict)

    def describe(self) -> dict:
        d = {"world_m": [self.box[2] - self.box[0], self.box[3] - self.box[1]],
             "obstacles": self.obstacles.n_rings, "ecm_edges": self.ecm.n_edge,
             "agents": self.agents.n, "k": self.params.k, "time_horizon": self.params.time_horizon,
             "time_horizon_obst": self.params.time_horizon_obst, "max_speed": self.params.max_speed,
             "dt": self.params.dt}
        d.update(self.info)
        return d


def _proxy(name, half, n_rects, side, n_agents, seed):
    rng = np.random.default_rng(seed)
    box = (-half, -half, half, half)
    rects = S.random_rects(rng, box, n_rects, side[0], side[1])
    obs = S.rects_to_obstacles(rects)
    ecm = S.empty_square_ecm(*box)
    agents = S.straight_agents(rng, box, n_agents, rects)
    params = Params(max_agents=n_agents)
    return Workload(name, box, obs, ecm, agents, params, info={"ecm": "empty-box closed form", "paths": "straight"})


def _scenario(name, config, seed, n_agents=0, params=None):
    from . import scene as SC
    sc, agents, goals = SC.generate(config, seed=seed, n_agents=n_agents)
    ecm = sc.ecm()
    obs = sc.obstacles()
    p = params or Params()
    p.max_agents = agents.n
    return Workload(name, sc.box, obs, ecm, agents, p, scene=sc,
                    info={"ecm": "built (segment Voronoi of the rectangles)", "paths": "ECM planner",
                          "seed": seed})


@dataclass
class BatchWorkload:
    """Independent environments advanced together (batched multi-scenario
    stepping): one Workload per environment."""
    name: str
    envs: list
    params: Params
    info: dict = field(default_factory=dict)

    @property
    def n_agents(self) -> int:
        return sum(w.agents.n for w in self.envs)

    def describe(self) -> dict:
        n = [w.agents.n for w in self.envs]
        sides = [w.box[2] - w.box[0] for w in self.envs]
        d = {"environments": len(self.envs), "agents": int(sum(n)), "agents_per_env_min": int(min(n)),
             "agents_per_env_max": int(max(n)), "side_m_min": float(min(sides)), "side_m_max": float(max(sides)),
             "obstacles": int(sum(w.obstacles.n_rings for w in self.envs)), "k": self.params.k,
             "time_horizon": self.params.time_horizon, "time_horizon_obst": self.params.time_horizon_obst,
             "max_speed": self.params.max_speed, "dt": self.params.dt}
        d.update(self.info)
        return d

    def make_sim(self):
        """A Simulator holding every environment, agents spawned; returns
        (sim, spawned ids per environment)."""
        from .simulator import Simulator
        p = self.params
        p.max_agents = self.n_agents
        w0 = self.envs[0]
        sim = Simulator(w0.ecm, w0.obstacles, w0.box, p, planner=w0.scene)
        ids = [sim.SpawnAgents(w0.agents)]
        for w in self.envs[1:]:
            e = sim.AddEnvironment(w.ecm, w.obstacles, w.box, planner=w.scene)
            ids.append(sim.SpawnAgents(w.agents, env=e))
        return sim, ids


def _batch(name, seed, n_env=256):
    """c4: n_env squares with sides U[100, 1000] m, U{10..200} rectangles
    with sides in [2, 15] m (capped at one per 250 m^2 so that they fit),
    agent counts log-uniform in [500, 50,000] (capped at 0.2 per m^2)."""
    from . import scene as SC
    rng = np.random.default_rng(seed)
    envs = []
    for e in range(n_env):
        side = float(rng.uniform(100.0, 1000.0))
        half = float(np.round(side / 2.0))
        n_rects = min(int(rng.integers(10, 201)), int(4 * half * half / 250.0))
        n_agents = int(round(float(np.exp(rng.uniform(np.log(500.0), np.log(50000.0))))))
        n_agents = min(n_agents, int(0.2 * 4 * half * half))
        sc, agents, _ = SC.generate_box(half, n_rects, 2, 15, n_agents, seed=int(rng.integers(1, 1 << 62)))
        envs.append(Workload(f"{name}/{e}", sc.box, sc.obstacles(), sc.ecm(), agents, Params(max_agents=agents.n),
                             scene=sc, info={"half": half, "n_rects": n_rects}))
    return BatchWorkload(name, envs, Params(), info={"ecm": "built (segment Voronoi of the rectangles)",
                                                     "paths": "ECM planner", "seed": seed})


def build(name: str, seed: int = 1) -> Workload:
    """c0: 200 m, 40 obstacles, 5k agents; c1: 1000 m, 400 obstacles, 100k
    agents; c2: 2000 m, 1,500 obstacles (barrier with 64 gaps of 8 m), two
    opposing flows of 500k agents starting at 2 agents/m^2 in 125 m bands
    against the barrier, goals in the band across; c0-ref: c0 with the reference's literal
    parameters (k = 5, T = 10 s, maxSpeed = preferred speed); c4: 256
    independent environments stepped together (c4-small: 8 of them)."""
    if name == "c0":
        return _scenario(name, 0, seed)
    if name == "c0-ref":
        return _scenario(name, 0, seed, params=Params.reference())
    if name == "c1":
        return _scenario(name, 1, seed)
    if name == "c2":
        wl = _scenario(name, 2, seed)
        # the flows meet in the gaps within seconds; 150 steps (15 s) puts a
        # timed window in the middle of the 300-step run, inside the jam
        wl.info["prerun_steps"] = 150
        return wl
    if name == "c4":
        return _batch(name, seed)
    if name == "c4-small":
        return _batch(name, seed, n_env=8)
    if name == "proxy-c0":
        return _proxy(name, 100.0, 40, (2, 15), 5000, seed)
    if name == "proxy-c1":
        return _proxy(name, 500.0, 400, (2, 25), 100000, seed)
    if name == "proxy-c2":
        return _proxy(name, 1000.0, 1500, (2, 25), 1000000, seed)
    raise KeyError(f"unknown workload {name}")


DEFAULT_BENCH = "c2"
