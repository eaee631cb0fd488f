"""Named benchmark / test workloads.

Each workload builds (box, obstacles, ECM, agents, params).  `proxy-*`
workloads use the closed-form ECM of the empty box and straight two-point
paths; they exercise the full step at the configured sizes but with a
trivial ECM.
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


def build(name: str, seed: int = 1) -> Workload:
    if name == "proxy-c0":
        return _proxy(name, 100.0, 40, (2, 15), 5000, seed)
    if name == "proxy-c1":
        return _proxy(name, 500.0, 400, (2, 25), 100000, seed)
    if name == "proxy-c2":
        return _proxy(name, 1000.0, 1500, (2, 25), 1000000, seed)
    raise KeyError(f"unknown workload {name}")


DEFAULT_BENCH = "proxy-c2"
