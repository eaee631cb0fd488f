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


def build(name: str, seed: int = 1) -> Workload:
    """c0: 200 m, 40 obstacles, 5k agents; c1: 1000 m, 400 obstacles, 100k
    agents; c2: 2000 m, 1,500 obstacles (barrier with 64 gaps of 8 m), two
    opposing flows of 500k agents; c0-ref: c0 with the reference's literal
    parameters (k = 5, T = 10 s, maxSpeed = preferred speed)."""
    if name == "c0":
        return _scenario(name, 0, seed)
    if name == "c0-ref":
        return _scenario(name, 0, seed, params=Params.reference())
    if name == "c1":
        return _scenario(name, 1, seed)
    if name == "c2":
        return _scenario(name, 2, seed)
    if name == "proxy-c0":
        return _proxy(name, 100.0, 40, (2, 15), 5000, seed)
    if name == "proxy-c1":
        return _proxy(name, 500.0, 400, (2, 25), 100000, seed)
    if name == "proxy-c2":
        return _proxy(name, 1000.0, 1500, (2, 25), 1000000, seed)
    raise KeyError(f"unknown workload {name}")


DEFAULT_BENCH = "c2"
