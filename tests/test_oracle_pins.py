"""Pins of the CPU oracle (oracle/ecm_oracle.c) against what the mathematics
fixes: closed forms in FP64 (tests/golden/closed_forms.json, written by
tests/golden/make_golden.py), brute force, FP64 re-solves of the LPs, libm,
and hand-traced known answers of the reference KD-tree (KDTree.cpp:85-202)."""
import json
import math
import os

import numpy as np

from oracle.lp_fp64 import qp_fp64
import pytest

from ecmgenerator_amd import scenarios as S
from ecmgenerator_amd.simulator import Params

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "closed_forms.json")))


def f32ulp(x):
    return np.spacing(np.float32(abs(x)))


# ---------------------------------------------------------------- trig
def test_trig_against_libm(oracle):
    xs = np.concatenate([np.linspace(-4, 4, 20001), np.linspace(-0.01, 0.01, 2001), [0.0, 1.0, -1.0, 0.2679492]])
    for x in xs.astype(np.float32):
        for f, ref in ((oracle.atanf, math.atan), (oracle.sinf, math.sin), (oracle.cosf, math.cos)):
            got, want = f(float(x)), ref(float(x))
            assert abs(got - want) <= 2.5 * f32ulp(want) + 1e-12, (f, x, got, want)


# ----------------------------------------------------- ECM / IRM (closed form)
@pytest.mark.parametrize("case", GOLD["irm"])
def test_irm_closed_form(oracle, case):
    L = case["L"]
    ecm = S.empty_square_ecm(-L, -L, L, L)
    x, y = case["pos"]
    ok, r, edge = oracle.retract(ecm, x, y)
    assert ok
    assert np.allclose(r, case["retract"], atol=2e-5)
    ok, a = oracle.attraction(ecm, x, y, np.array(case["path"], np.float32))
    assert ok
    assert np.allclose(a, case["attraction"], atol=5e-4 * max(1.0, L / 10))


def test_locate_first_match_and_vertex_rule(oracle):
    ecm = S.empty_square_ecm(-50, -50, 50, 50)
    # cells: edge e centre->corner e, left cell 2e, right cell 2e+1
    assert oracle.locate(ecm, 5.0, 30.0) == 0        # top region, right half: left cell of edge 0
    assert oracle.locate(ecm, 30.0, 5.0) == 1        # right region, upper half: right cell of edge 0
    assert oracle.locate(ecm, -5.0, 30.0) == 3       # top region, left half: right cell of edge 1
    # within EPSILON of a cell vertex -> not contained (UtilityFunctions.cpp:65-66)
    assert oracle.locate(ecm, 0.00005, 0.00003) == -1
    assert oracle.locate(ecm, 60.0, 0.0) == -1


# ------------------------------------------------------------ k-nearest
def _state(pos, k=10, n=None):
    n = n or len(pos)
    st = dict(active=np.ones(n, np.uint8), position=np.asarray(pos, np.float32), velocity=np.zeros((n, 2), np.float32),
              clearance=np.full(n, 0.25, np.float32), pref_speed=np.full(n, 1.3, np.float32),
              attraction=np.zeros((n, 2), np.float32), path_off=np.zeros(n, np.int32), path_cnt=np.ones(n, np.int32),
              path_xy=np.array([[1e4, 1e4]], np.float32))
    return st


def test_exact_knn_matches_brute_force(oracle):
    rng = np.random.default_rng(0)
    pos = rng.uniform(-10, 10, (400, 2)).astype(np.float32)
    pos[5] = pos[7]                       # coincident pair: invisible to each other (d2 <= 1e-4)
    pos[9] = pos[3] + np.float32(0.004)   # d2 = 3.2e-5 <= 1e-4
    pos[10] = pos[11] + np.array([1.0, 0.0], np.float32)   # a tie in distance for some agents
    p = Params(k=10)
    st = _state(pos)
    out = oracle.step(p, st, None, S.empty_square_ecm(-1e4, -1e4, 1e4, 1e4))
    for i in range(len(pos)):
        d = pos - pos[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32)
        cand = [(float(d2[j]), j) for j in range(len(pos)) if d2[j] > np.float32(1e-4)]
        cand.sort()
        want = [j for _, j in cand[:10]]
        assert out["nbr"][i][: out["nbr_cnt"][i]].tolist() == want
    assert 7 not in out["nbr"][5] and 5 not in out["nbr"][7]


def test_reference_kdtree_known_answers(oracle):
    """Hand-traced KDTree.cpp:22-202 with k = 2.
    Case A, agents 0..3 on the x axis at 0, 1, 2, 3: tree [1, 0, 2, -, -, -, 3].
    Agent 0 (first query of the step) inserts agent 1, then meets itself
    (d2 = 0) before k are found, which forces kFound = k (:141-143); slot 1
    keeps the zero-initialised cache value, i.e. agent 0 itself.
    Case B, agents at (0,0), (1,0), (5,0), (5,3): tree [1, 0, 2, -, -, -, 3].
    Agent 3 inserts agent 1 (d2 25), then agent 2 (d2 9) fills the list; the
    fill-slot swap (:118-139) writes agent 2 into both slots and drops agent 1;
    the left subtree is pruned against slot k-1 (:185)."""
    from oracle.oracle_py import KNN_REFERENCE
    p = Params(k=2)
    ecm = S.empty_square_ecm(-1e4, -1e4, 1e4, 1e4)
    out = oracle.step(p, _state([(0, 0), (1, 0), (2, 0), (3, 0)]), None, ecm, knn_mode=KNN_REFERENCE)
    assert out["nbr_cnt"][0] == 2 and out["nbr"][0].tolist() == [1, 0]
    out = oracle.step(p, _state([(0, 0), (1, 0), (5, 0), (5, 3)]), None, ecm, knn_mode=KNN_REFERENCE)
    assert out["nbr_cnt"][3] == 2 and out["nbr"][3].tolist() == [2, 2]
    # the exact mode returns the true 2-nearest sets for both
    out = oracle.step(p, _state([(0, 0), (1, 0), (5, 0), (5, 3)]), None, ecm)
    assert out["nbr"][3].tolist() == [2, 1]


# --------------------------------------------------------------- ORCA
@pytest.mark.parametrize("case", GOLD["agent"])
def test_agent_halfplane_closed_form(oracle, case):
    p = Params(k=1, time_horizon=case["T"], time_horizon_obst=case["T"], dt=case.get("dt", 0.1))
    st = _state([case["pa"], case["pb"]])
    st["velocity"] = np.array([case["va"], case["vb"]], np.float32)
    st["clearance"] = np.array([case["ra"], case["rb"]], np.float32)
    cs, nob = oracle.constraints(p, st, None, 0, [1])
    assert nob == 0 and cs.shape[0] == 1
    assert np.allclose(cs[0, :2], case["normal"], atol=2e-6)
    assert np.allclose(cs[0, 2:], case["point"], atol=2e-6)


@pytest.mark.parametrize("case", GOLD["obstacle"])
def test_obstacle_halfplane_closed_form(oracle, case):
    obs = S.rects_to_obstacles([(-1, -1, 1, 1)])
    p = Params(k=1, time_horizon=case["T"], time_horizon_obst=case["T"])
    st = _state([case["pos"]])
    st["clearance"] = np.array([case["r"]], np.float32)
    cs, nob = oracle.constraints(p, st, obs, 0, [])
    assert nob == 1
    assert np.allclose(cs[0, :2], case["normal"], atol=1e-6)
    assert np.allclose(cs[0, 2:], case["point"], atol=1e-6)


def _shape_obstacles(ring):
    from ecmgenerator_amd.simulator import Obstacles
    xy = np.asarray(ring, np.float32)
    return Obstacles(xy, np.array([0, xy.shape[0]], np.int32))


@pytest.mark.parametrize("case", GOLD["obstacle_branches"],
                         ids=lambda c: f"{c['shape']}{c['edge']}-{'+'.join(c['branches'])}")
def test_obstacle_branch_geometry(oracle, case):
    """Every branch of the obstacle half-plane (ORCA.cpp:91-332): the branch
    the oracle takes and the half-plane it emits ({u : n . u >= n . p})
    against the FP64 geometry of the truncated velocity obstacle
    (tests/golden/make_golden.py, obstacle_model)."""
    obs = _shape_obstacles(GOLD["shapes"][case["shape"]])
    p = Params(k=1, time_horizon=case["T"], time_horizon_obst=case["T"])
    st = _state([case["pos"]])
    st["velocity"] = np.array([case["vel"]], np.float32)
    st["clearance"] = np.array([case["r"]], np.float32)
    n, c, branches = oracle.obstacle_halfplane(p, st, obs, 0, case["edge"])
    assert branches == set(case["branches"]), (branches, case["branches"])
    assert n == (1 if case["emitted"] else 0)
    if case["emitted"]:
        nrm = c[:2].astype(np.float64)
        assert np.allclose(nrm, case["normal"], atol=1e-5), (nrm, case["normal"])
        assert abs(float(nrm @ c[2:].astype(np.float64)) - case["offset"]) < 1e-5


def test_obstacle_branch_cases_cover_every_branch():
    hit = set().union(*(set(c["branches"]) for c in GOLD["obstacle_branches"]))
    from oracle.oracle_py import OB_BITS
    assert hit == set(OB_BITS)


def test_obstacle_filter_brute_force(oracle):
    """FindNearestObstacles (Simulator.cpp:259-292) against an FP64 brute
    force: an edge is kept when the agent is strictly right of it (outside a
    CCW ring) and the segment comes closer than the range; kept in scan order.
    Random rings of squares, L shapes and triangles; points within 1e-3 of
    either threshold are skipped (float rounding may decide those)."""
    rng = np.random.default_rng(4)
    rings = []
    for i in range(30):
        shape = list(GOLD["shapes"].values())[i % 3]
        sc, ang = rng.uniform(0.5, 4.0), rng.uniform(0, 2 * np.pi)
        off = rng.uniform(-20, 20, 2)
        R = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
        rings.append((np.asarray(shape) * sc) @ R.T + off)
    from ecmgenerator_amd.simulator import Obstacles
    xy = np.concatenate(rings).astype(np.float32)
    off = np.cumsum([0] + [r.shape[0] for r in rings]).astype(np.int32)
    obs = Obstacles(xy, off)
    nxt = np.empty(xy.shape[0], np.int64)
    for a, b in zip(off[:-1], off[1:]):
        nxt[a:b] = np.roll(np.arange(a, b), -1)
    A = xy.astype(np.float64)
    B = A[nxt]
    checked = kept_total = 0
    for _ in range(600):
        q = rng.uniform(-25, 25, 2)
        rg = rng.uniform(2.0, 6.0)
        sld = (A[:, 0] - q[0]) * (B[:, 1] - A[:, 1]) - (A[:, 1] - q[1]) * (B[:, 0] - A[:, 0])
        ab = B - A
        t = np.clip(np.einsum("ij,ij->i", q - A, ab) / np.einsum("ij,ij->i", ab, ab), 0, 1)
        d = np.linalg.norm(q - (A + ab * t[:, None]), axis=1)
        if np.any(np.abs(sld) < 1e-3) or np.any(np.abs(d - rg) < 1e-3):
            continue
        want = np.flatnonzero((sld < 0) & (d < rg))
        got = oracle.nearest_obstacles(obs, np.float32(q[0]), np.float32(q[1]), np.float32(rg))
        assert got.tolist() == want.tolist()
        checked += 1
        kept_total += want.size
    assert checked > 400 and kept_total > 300


def test_lp2d_matches_fp64_qp(oracle):
    rng = np.random.default_rng(1)
    checked = 0
    for trial in range(400):
        m = rng.integers(1, 12)
        ang = rng.uniform(0, 2 * np.pi, m)
        n = np.stack([np.cos(ang), np.sin(ang)], 1)
        p = rng.uniform(-0.8, 0.8, (m, 2))
        cons = np.concatenate([n, p], 1).astype(np.float32)
        pref = rng.uniform(-2.5, 2.5, 2).astype(np.float32)
        r, v = oracle.lp2d(cons, pref, 2.0)
        opt = qp_fp64(cons, pref, 2.0)
        if r < m:
            # failure is reported only when the FP64 problem is infeasible (or
            # within rounding of it)
            assert opt is None or min(np.einsum("ij,ij->i", cons[:, :2].astype(float), opt - cons[:, 2:])) < 1e-4
            continue
        assert opt is not None
        assert np.linalg.norm(v - opt) < 1e-4, (trial, v, opt)
        checked += 1
    assert checked > 100


def test_lp3d_minimises_max_penetration(oracle):
    """Infeasible agent constraints: the 3D LP returns the velocity that
    minimises the largest violation (RVO2-style linearProgram3); compared
    with an FP64 minimax solve on a fine polygon of the speed disk."""
    from scipy.optimize import linprog
    rng = np.random.default_rng(2)
    done = 0
    for trial in range(200):
        m = int(rng.integers(3, 8))
        ang = rng.uniform(0, 2 * np.pi, m)
        n = np.stack([np.cos(ang), np.sin(ang)], 1)
        p = n * rng.uniform(0.5, 2.5, (m, 1))          # lines pushed outward: often infeasible
        cons = np.concatenate([n, p], 1).astype(np.float32)
        r, v0 = oracle.lp2d(cons, np.zeros(2, np.float32), 2.0)
        if r >= m:
            continue
        v = oracle.lp3d(cons, 0, 2.0, r, v0)
        pen = lambda x: max(float(c[0] * (c[2] - x[0]) + c[1] * (c[3] - x[1])) for c in cons.astype(float))
        K = 256
        th = np.linspace(0, 2 * np.pi, K, endpoint=False)
        # variables (vx, vy, t): minimise t, n.(p - v) <= t, polygon |v| <= 2
        A = [[-c[0], -c[1], -1.0] for c in cons.astype(float)] + [[math.cos(a), math.sin(a), 0.0] for a in th]
        b = [-(c[0] * c[2] + c[1] * c[3]) for c in cons.astype(float)] + [2.0 * math.cos(math.pi / K)] * K
        res = linprog([0, 0, 1], A_ub=A, b_ub=b, bounds=[(None, None)] * 3, method="highs")
        assert res.success
        assert np.linalg.norm(v) <= 2.0 + 1e-4
        assert pen(v) <= res.x[2] + 2e-3, (trial, pen(v), res.x[2])
        done += 1
    assert done > 20


def test_integration_closed_form(oracle):
    """One agent, no neighbours or obstacles, far from the goal on an empty
    square: v_orca = preferred velocity (speed <= maxSpeed), then
    v' = v + (v_orca - v) * (1/0.8) * dt and p' = p + v' dt (Simulator.cpp:592-635)."""
    ecm = S.empty_square_ecm(-50, -50, 50, 50)
    st = _state([(1.0, 30.0)])
    st["velocity"] = np.array([[0.3, -0.2]], np.float32)
    st["path_xy"] = np.array([[1.0, 30.0], [1.0, -45.0]], np.float32)
    st["path_cnt"] = np.array([2], np.int32)
    out = oracle.step(Params(k=10), st, None, ecm)
    pv = out["pref_velocity"][0].astype(np.float64)
    assert np.allclose(np.linalg.norm(pv), 1.3, atol=1e-6)
    assert np.allclose(out["v_orca"][0], pv, atol=1e-7)
    v = np.array([0.3, -0.2])
    v1 = v + (pv - v) * 1.25 * 0.1
    assert np.allclose(out["velocity"][0], v1, atol=1e-6)
    assert np.allclose(out["position"][0], np.array([1.0, 30.0]) + v1 * 0.1, atol=1e-5)


def test_goal_arrival_and_destroy(oracle):
    """d2 to the goal < 20^2: attraction = goal; < 2^2: the agent is destroyed
    and keeps its state (Simulator.cpp:554-566)."""
    ecm = S.empty_square_ecm(-50, -50, 50, 50)
    st = _state([(0.0, 10.0), (5.0, 5.0)])
    st["path_xy"] = np.array([[0.0, 10.0], [0.0, 0.0], [5.0, 5.0], [5.0, 6.0]], np.float32)
    st["path_off"] = np.array([0, 2], np.int32)
    st["path_cnt"] = np.array([2, 2], np.int32)
    st["velocity"] = np.array([[0.1, 0.0], [0.0, 0.2]], np.float32)
    out = oracle.step(Params(k=10), st, None, ecm)
    assert out["attraction"][0].tolist() == [0.0, 0.0] and out["active"][0] == 1
    assert out["active"][1] == 0 and out["position"][1].tolist() == [5.0, 5.0]
    assert out["velocity"][1].tolist() == [0.0, np.float32(0.2)]
