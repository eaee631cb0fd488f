"""Between steps the agent state lives in the last step's sorted outputs and
the slot-indexed arrays are brought up to date only when the host asks.
These tests change agents through the slot arrays between steps and check
that the next steps (rebuilt from the slot arrays, then from sorted state
again) still equal the oracle bit for bit, and that reads see the state of
the last step."""
import numpy as np
import pytest

from ecmgenerator_amd.simulator import Params, Simulator
from helpers import compare_step, oracle_state, small_scene

pytestmark = pytest.mark.gpu


def check_step(sim, p, obs, ecm, tag):
    from oracle import oracle_py as O
    before = oracle_state(sim)
    ora = O.step(p, before, obs, ecm)
    sim.Update()
    res = compare_step(sim, before, ora)
    print(f"{tag}: {res}", flush=True)
    assert all(v == 0 for v in res.values()), (tag, res)
    return ora


def test_writes_between_steps():
    box, _, obs, ecm, agents = small_scene(61, half=30.0, n_rects=4, n_agents=1200)
    p = Params(max_agents=2048, keep_orca_velocity=True)
    sim = Simulator(ecm, obs, box, p)
    sim.SpawnAgents(agents)
    for _ in range(4):
        sim.Update()
    rng = np.random.default_rng(0)
    act = sim.read("active").astype(bool)
    pos = sim.read("position")
    pos[act] += rng.normal(0, 0.05, (int(act.sum()), 2)).astype(np.float32)
    sim.write("position", pos)
    vel = sim.read("velocity")
    vel[act] *= np.float32(0.5)
    sim.write("velocity", vel)
    att = sim.read("attraction")
    sim.write("attraction", att)
    for it in range(3):                    # slot entry, then sorted entry twice
        check_step(sim, p, obs, ecm, f"after writes step {it}")
    sim.close()


def test_reads_see_the_last_step():
    """Positions read after a step equal the oracle's; reading twice, or
    reading between two steps, changes nothing in the following step."""
    box, _, obs, ecm, agents = small_scene(62, half=30.0, n_rects=4, n_agents=900)
    p = Params(max_agents=1024, keep_orca_velocity=True)
    a = Simulator(ecm, obs, box, p)
    b = Simulator(ecm, obs, box, p)
    a.SpawnAgents(agents)
    b.SpawnAgents(agents)
    for it in range(6):
        a.Update()
        b.Update()
        if it % 2 == 0:                    # b is read between steps, a is not
            s1 = b.state()
            s2 = b.state()
            for k in ("position", "velocity", "attraction", "pref_velocity", "orca_velocity"):
                assert np.array_equal(s1[k], s2[k]), k
    sa, sb = a.state(), b.state()
    for k in ("position", "velocity", "attraction", "pref_velocity", "orca_velocity", "active"):
        assert np.array_equal(sa[k], sb[k]), k
    ora = check_step(a, p, obs, ecm, "after reads")
    a.close()
    b.close()


def test_position_reads_into_kept_and_page_locked_buffers():
    """A renderer's read after every step: positions alone (the
    positions-only slot-order pass), into a kept pageable buffer and into a
    page-locked one from Simulator.pinned, equal a fresh read and, after the
    whole-state pass of a later read of another component, still equal it;
    a buffer of the wrong shape or type is refused."""
    box, _, obs, ecm, agents = small_scene(63, half=30.0, n_rects=4, n_agents=900)
    p = Params(max_agents=1024)
    sim = Simulator(ecm, obs, box, p)
    sim.SpawnAgents(agents)
    kept = sim.GetPositionData()
    pinned = sim.pinned("position")
    for _ in range(4):
        sim.Update()
        sim.GetPositionData(out=pinned)              # positions only
        sim.GetPositionData(out=kept)
        vel = sim.read("velocity")                   # the whole state
        fresh = sim.GetPositionData()
        assert np.array_equal(pinned, fresh)
        assert np.array_equal(kept, fresh)
        assert vel.shape == (1024, 2)
    with pytest.raises(ValueError):
        sim.GetPositionData(out=np.zeros((1024, 3), np.float32))
    with pytest.raises(ValueError):
        sim.GetPositionData(out=np.zeros((1024, 2), np.float64))
    sim.close()
