"""An agent with a finite bound from its previous neighbours finds its k
nearest by walking the rectangle of grid cells the bound reaches: the
non-empty cells become ranges of sorted positions (consecutive ones
joined), a few at a time, and one flat loop visits their candidates.  The
scenes below stress that walk against the oracle, bit for bit, over several
steps (the first step has no seeds and searches by rings; the later ones
walk):

* agents on an exact lattice, so many candidates tie in squared distance
  and ids decide;
* grid cells much smaller than the neighbour distances (long walks, many
  batches of ranges) and much larger (a few cells hold everything);
* lattices that reach the world border, so walks are clipped there;
* a sparse crowd whose k-th neighbours lie tens of metres away."""
import numpy as np
import pytest

from ecmgenerator_amd import scenarios as S
from ecmgenerator_amd.simulator import Agents, Params, Simulator
from helpers import compare_step, oracle_state

pytestmark = pytest.mark.gpu


def lattice_agents(half: float, spacing: float, speed: float) -> Agents:
    g = np.arange(-half + 0.5, half - 0.5 + 1e-3, spacing, dtype=np.float64)
    xx, yy = np.meshgrid(g, g)
    start = np.stack([xx.ravel(), yy.ravel()], 1).astype(np.float32)
    goal = (-start[:, ::-1]).astype(np.float32)     # across the centre
    n = start.shape[0]
    path = np.stack([start, goal], 1).reshape(-1, 2)
    return Agents(start, np.full(n, 0.25, np.float32), np.full(n, speed, np.float32),
                  np.arange(n, dtype=np.int32) * 2, np.full(n, 2, np.int32), path)


def run(agents: Agents, half: float, params: Params, steps: int, tag: str):
    from oracle import oracle_py as O
    box = (-half, -half, half, half)
    obs = S.rects_to_obstacles([])
    ecm = S.empty_square_ecm(*box)
    sim = Simulator(ecm, obs, box, params)
    ids = sim.SpawnAgents(agents)
    assert (ids >= 0).all(), int((ids < 0).sum())
    for it in range(steps):
        before = oracle_state(sim)
        ora = O.step(params, before, obs, ecm)
        sim.Update()
        res = compare_step(sim, before, ora)
        print(f"{tag} step {it}: {res}", flush=True)
        assert all(v == 0 for v in res.values()), (tag, it, res)
    sim.close()


@pytest.mark.parametrize("cell", [0.5, 1.5, 40.0])
def test_lattice_ties(cell):
    """Agents that barely move keep the lattice's exact distance ties into
    the seeded steps."""
    half = 20.0
    agents = lattice_agents(half, 1.5, 1e-6)
    p = Params(max_agents=1024, keep_orca_velocity=True, grid_cell=cell)
    run(agents, half, p, 4, f"lattice cell {cell}")


def test_lattice_moving_to_border():
    """Agents cross the world towards the opposite side, up to the border."""
    half = 15.0
    agents = lattice_agents(half, 1.25, 1.3)
    p = Params(max_agents=1024, keep_orca_velocity=True, grid_cell=0.75)
    run(agents, half, p, 5, "moving lattice")


def test_sparse_far_neighbours():
    """40 agents in a 120 m square on 0.5 m cells: each walk covers
    thousands of cells and tens of range batches."""
    rng = np.random.default_rng(5)
    half = 60.0
    agents = S.straight_agents(rng, (-half, -half, half, half), 40)
    p = Params(max_agents=64, keep_orca_velocity=True, grid_cell=0.5)
    run(agents, half, p, 4, "sparse")
