"""Learned ORCA proposal on the GPU: the MFMA network kernels against the
CPU oracle (forward, loss, gradient), the logged training samples against
the C oracle, the verified learned step against the plain step, and what a
trained network gains over an untrained one."""
import os
import tempfile

import numpy as np
import pytest

from ecmgenerator_amd.learned import ProposalNet
from ecmgenerator_amd.simulator import Params, Simulator
from helpers import oracle_state, small_scene
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def rand_inputs(seed, n):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 64), np.float32)
    x[:, :61] = rng.normal(0, 1.5, (n, 61)).astype(np.float32)
    lab = (rng.integers(0, 1 << 14, n) | (rng.integers(0, 1 << 14, n) << 16)).astype(np.uint32)
    v = rng.normal(0, 1, (n, 2)).astype(np.float32)
    return x, lab, v


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_mfma_forward_matches_oracle():
    net = ProposalNet(seed=3)
    p = net.get_params()
    x, _, _ = rand_inputs(4, 200)          # 200 = 6 full 32-agent tiles + a partial one
    y = net.forward(x)
    yo, _ = M.forward(p, x, bf16=True)
    e = rel(y, yo)
    print(f"forward rel err {e:.2e} max abs {np.abs(y - yo).max():.2e}", flush=True)
    # observed ~8e-8: the kernel and the oracle round the same BF16
    # activations and differ only in FP32 summation order
    assert e < 1e-6
    assert np.abs(y - yo).max() < 1e-5 * np.abs(yo).max()
    net.close()


def test_loss_and_gradient_match_oracle():
    net = ProposalNet(seed=5)
    p = net.get_params()
    x, lab, v = rand_inputs(6, 300)
    loss, g = net.loss_grad(x, lab, v, pos_weight=2.0, vel_weight=0.7)
    lo, go = M.loss_grad(p, x, lab, v, pos_weight=2.0, vel_weight=0.7, bf16=True)
    print(f"loss {loss:.6f} oracle {lo:.6f}", flush=True)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    o = 0
    for name, shp in M.SHAPES:
        n = int(np.prod(shp))
        e = rel(g[o:o + n], go[o:o + n])
        print(f"grad {name} rel err {e:.2e}", flush=True)
        assert e < 6e-4, name          # observed <= 6e-5
        o += n
    net.close()


def dense_scene(seed=11, n_agents=1800, half=30.0, n_rects=4):
    box, rects, obs, ecm, agents = small_scene(seed, half=half, n_rects=n_rects, n_agents=n_agents)
    return box, obs, ecm, agents


def test_logged_samples_match_oracle():
    """Features (after BF16 rounding), labels and LP velocities of one
    logged step equal the C oracle's for every agent whose 2D LP succeeds
    in the main ORCA pass (compared as multisets: the log order is the
    order of the device's appends)."""
    from oracle import oracle_py as O
    box, obs, ecm, agents = dense_scene(n_agents=600)
    p = Params(max_agents=1024, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    sim.SpawnAgents(agents)
    for _ in range(5):
        sim.Update()
    before = oracle_state(sim)
    ora = O.step(p, before, obs, ecm)
    net = ProposalNet(seed=0)
    net.reserve(4096)
    n = sim.log_step(net)
    x, lab, v = net.read(0, n)
    nbr, cnt = sim.neighbors()
    want = []
    for a in np.flatnonzero(before["active"].astype(bool) & ora["active"].astype(bool)):
        if ora["used_3d"][a]:
            continue
        nb = nbr[a][: cnt[a]]
        cons, nob = O.constraints(p, before, obs, a, nb)
        if nob > 6:          # overflow agents go to the plain overflow kernel
            continue
        f, l = O.proposal_sample(p, before, obs, a, nb, ora["pref_velocity"][a], ora["v_orca"][a])
        want.append((M.bf16_round(f).astype(np.float32), l, ora["v_orca"][a]))
    print(f"logged {n}, oracle {len(want)}", flush=True)
    assert n == len(want) and n > 400

    def key(f, l, vv):
        return tuple(f.tolist()) + (l,) + tuple(np.asarray(vv, np.float32).tolist())
    got = sorted(key(x[i], int(lab[i]), v[i]) for i in range(n))
    exp = sorted(key(*w) for w in want)
    bad = sum(1 for a, b in zip(got, exp) if a != b)
    assert bad == 0, f"{bad} of {n} samples differ"
    act = np.array([bin(int(l) & 0x3FFF).count("1") for l in lab])
    print(f"active slots per sample: mean {act.mean():.2f}, histogram {np.bincount(act).tolist()}", flush=True)
    net.close()
    sim.close()


def one_step_pair(net, theta, seed=11, warm=8):
    """Plain and learned step from the same state: (max |v_orca diff|,
    max |position diff|, tried, accepted, constraints proposed)."""
    box, obs, ecm, agents = dense_scene(seed=seed)
    p = Params(max_agents=2048, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    sim.SpawnAgents(agents)
    for _ in range(warm):
        sim.Update()
    with tempfile.TemporaryDirectory() as d:
        snap = os.path.join(d, "s.bin")
        sim.snapshot_save(snap)
        act = sim.state()["active"].astype(bool)
        sim.Update()
        a = sim.state()
        sim.snapshot_load(snap)
        sim.set_learned(net, theta)
        st0 = sim.stats()
        sim.Update()
        st1 = sim.stats()
        b = sim.state()
    sim.set_learned(None)
    sim.close()
    alive = act & a["active"].astype(bool)
    dv = float(np.abs(a["orca_velocity"][alive] - b["orca_velocity"][alive]).max())
    dp = float(np.abs(a["position"][act] - b["position"][act]).max())
    tried = st1["learned_tried"] - st0["learned_tried"]
    acc = st1["learned_accepted"] - st0["learned_accepted"]
    prop = st1["learned_proposed"] - st0["learned_proposed"]
    return dv, dp, tried, acc, prop


@pytest.mark.parametrize("theta", [-1e9, 0.0, 1e9])
def test_learned_step_matches_plain_step(theta):
    """Whatever the network proposes (every slot, an untrained guess, no
    slot), the verified step equals the plain step within 1e-4 m/s."""
    net = ProposalNet(seed=1)
    dv, dp, tried, acc, _ = one_step_pair(net, theta)
    print(f"theta {theta}: max dv {dv:.2e} max dp {dp:.2e} accepted {acc}/{tried}", flush=True)
    assert tried > 1500
    assert dv <= 1e-4 and dp <= 1e-5
    if theta < 0:
        assert acc > 0.5 * tried          # every slot proposed: a superset of the active set
    net.close()


def test_trained_network_gains_over_untrained():
    """Train on samples logged from one dense scene, then compare the
    acceptance of the trained and the untrained network on a different
    scene's step (both exact: verified against the plain step)."""
    box, obs, ecm, agents = dense_scene(seed=21)
    p = Params(max_agents=2048, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    sim.SpawnAgents(agents)
    net = ProposalNet(seed=2)
    net.reserve(200_000)
    for i in range(60):
        if i % 2 == 0:
            sim.log_step(net)
        else:
            sim.Update()
    sim.close()
    n = net.dataset_size()
    losses = net.train(epochs=12, batch=1024, lr=3e-3, seed=0)
    print(f"samples {n}, losses {np.round(losses, 4).tolist()}", flush=True)
    assert losses[-1] < 0.9 * losses[0]
    untrained = ProposalNet(seed=2)
    dv0, _, tried0, acc0, prop0 = one_step_pair(untrained, 0.0, seed=11)
    dv1, _, tried1, acc1, prop1 = one_step_pair(net, 0.0, seed=11)
    _, _, tried2, acc2, prop2 = one_step_pair(untrained, -1e9, seed=11)
    print(f"theta 0: untrained accepted {acc0}/{tried0} = {acc0 / tried0:.3f} ({prop0 / tried0:.2f} constraints "
          f"proposed per agent); trained {acc1}/{tried1} = {acc1 / tried1:.3f} ({prop1 / tried1:.2f}); every slot "
          f"{acc2}/{tried2} = {acc2 / tried2:.3f} ({prop2 / tried2:.2f})", flush=True)
    assert dv0 <= 1e-4 and dv1 <= 1e-4
    # the trained proposal is accepted more often than the untrained one
    # while proposing a fraction of the constraints
    assert acc1 / tried1 > acc0 / tried0 + 0.1
    assert prop1 / tried1 < 0.5 * prop0 / tried0
    assert prop1 / tried1 < 0.5 * prop2 / tried2
    net.close()
    untrained.close()


def test_committed_weights_proposal():
    """The committed network (trained on ~19M samples from the 100k- and
    1M-agent workloads) on an unseen dense scene: exact, and most agents
    accept a proposal far smaller than every slot."""
    from ecmgenerator_amd import learned as Lm
    import json
    with np.load(Lm.DEFAULT_WEIGHTS, allow_pickle=False) as z:
        theta = float(json.loads(str(z["meta"]))["theta"])
    net = ProposalNet.load()
    dv, dp, tried, acc, prop = one_step_pair(net, theta, seed=31)
    print(f"committed weights theta {theta}: accepted {acc}/{tried} = {acc / tried:.3f}, "
          f"{prop / tried:.2f} constraints proposed per agent, max dv {dv:.2e}", flush=True)
    assert dv <= 1e-4 and dp <= 1e-5
    assert acc > 0.8 * tried
    net.close()
