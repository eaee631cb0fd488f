"""Step graphs and the autotuner.  A handle replaying captured step graphs
must produce the eager handle's results bit for bit, step after step, with
host work (state reads, spawns, destroys, replans, a switch to the learned
proposal) between the replays; and no launch shape the autotuner tries may
change a result."""
import numpy as np
import pytest

from ecmgenerator_amd.scene import generate
from ecmgenerator_amd.simulator import Agents, Params, Simulator

pytestmark = pytest.mark.gpu

KEYS = ("position", "velocity", "attraction", "pref_velocity", "active", "orca_velocity")


def make(sc, agents, graph, extra=64):
    p = Params(max_agents=agents.n + extra, keep_orca_velocity=True)
    sim = Simulator(sc.ecm(), sc.obstacles(), sc.box, p, planner=sc)
    sim.SpawnAgents(agents)
    sim.set_graph(graph)
    return sim


def same(a, b):
    sa, sb = a.state(), b.state()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(a.replans(), b.replans())


def subset(agents, idx):
    pts = [agents.path_xy[agents.path_off[i]: agents.path_off[i] + agents.path_cnt[i]] for i in idx]
    cnt = np.array([len(p) for p in pts], np.int32)
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int32)
    return Agents(agents.start[idx], agents.radius[idx], agents.pref_speed[idx], off, cnt, np.concatenate(pts))


def test_step_graphs_replay_like_eager():
    from ecmgenerator_amd.learned import ProposalNet
    sc, agents, _ = generate(3, seed=5)
    first = np.arange(agents.n - 40)
    late = np.arange(agents.n - 40, agents.n)
    a = make(sc, subset(agents, first), False)
    b = make(sc, subset(agents, first), True)
    net = ProposalNet(seed=4)
    for it in range(36):
        a.Update()
        b.Update()
        same(a, b)
        if it == 8:            # more agents: new grid layout, graphs recaptured
            ia = a.SpawnAgents(subset(agents, late))
            ib = b.SpawnAgents(subset(agents, late))
            assert np.array_equal(ia, ib) and (ia >= 0).sum() > 20
        if it == 15:
            live = np.flatnonzero(a.state()["active"])
            a.DestroyAgent(int(live[3]))
            b.DestroyAgent(int(live[3]))
        if it == 22:           # learned proposal: a different graph
            a.set_learned(net, 0.0)
            b.set_learned(net, 0.0)
    on, why = b.graph_status()
    assert on, why
    st = b.stats()
    print(f"graph run: {st}", flush=True)
    assert st["learned_tried"] > 0
    for s in (a, b):
        s.set_learned(None)
        s.close()
    net.close()


def test_autotune_changes_speed_not_results():
    sc, agents, _ = generate(3, seed=6)
    a = make(sc, agents, False)
    b = make(sc, agents, True)
    t = b.autotune(steps=2)
    print(f"autotune picked {t}", flush=True)
    assert t["knn_block"] in (16, 64, 128, 256) and t["orca_obst_slots"] in (1, 2, 4) and t["best_ms"] > 0
    n = b.stats()["steps"]
    a.Update(steps=n)
    same(a, b)
    for _ in range(5):
        a.Update()
        b.Update()
        same(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("k", [5, 10])
def test_orca_pass_shapes_agree(k):
    """The main ORCA pass as its own kernel with 4 obstacle slots (shape 1),
    inside the k-nearest kernel (shape 2, k 5 and 10; also with 16 agents per
    wavefront, knn_block 16) and with 8 obstacle slots (shape 4) gives the
    same steps bit for bit, with step graphs and eager launches, replans
    included."""
    sc, agents, _ = generate(0, seed=7)
    sims = []
    for oslots, graph, block in ((1, False, 256), (2, True, 256), (2, False, 256), (4, True, 256), (2, True, 16),
                                 (1, True, 16)):
        p = Params(max_agents=agents.n + 64, keep_orca_velocity=True, k=k)
        sim = Simulator(sc.ecm(), sc.obstacles(), sc.box, p, planner=sc)
        sim.SpawnAgents(agents)
        sim.set_tuning(knn_block=block, orca_obst_slots=oslots)
        sim.set_graph(graph)
        sims.append(sim)
    for step in range(40):
        for sim in sims:
            sim.Update()
        if step % 8 == 7:
            for sim in sims[1:]:
                same(sims[0], sim)
    print(f"k {k}: replans {sims[0].stats()['replans']}, 3D LP {sims[0].stats()['lp3d']}", flush=True)
    for sim in sims:
        sim.close()
