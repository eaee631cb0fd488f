"""The proposal network's CPU oracle (oracle/mlp_oracle.py) and the feature
/ label restatement in the C oracle, pinned without a GPU: finite
differences for the backward pass, BF16 rounding known answers, the MFMA
input permutation, and closed-form feature / label cases."""
import numpy as np

from ecmgenerator_amd import scenarios as S
from ecmgenerator_amd.simulator import Params
from oracle import mlp_oracle as M
from oracle import oracle_py as O


def rand_problem(seed, n=12):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.N_PARAMS)
    x = np.zeros((n, 64))
    x[:, :61] = rng.normal(0, 1, (n, 61))
    lab = rng.integers(0, 1 << 14, n).astype(np.uint64) | (rng.integers(0, 1 << 14, n).astype(np.uint64) << 16)
    v = rng.normal(0, 1, (n, 2))
    return p, x, lab.astype(np.uint32), v


def test_backward_matches_central_differences():
    p, x, lab, v = rand_problem(1)
    _, g = M.loss_grad(p, x, lab, v, pos_weight=2.0, vel_weight=0.7, bf16=False)
    rng = np.random.default_rng(2)
    idx = rng.choice(M.N_PARAMS, 300, replace=False)
    eps = 1e-5
    worst = 0.0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = M.loss_grad(pp, x, lab, v, 2.0, 0.7, bf16=False)
        lm, _ = M.loss_grad(pm, x, lab, v, 2.0, 0.7, bf16=False)
        fd = (lp - lm) / (2 * eps)
        worst = max(worst, abs(fd - g[i]) / max(1e-3, abs(fd), abs(g[i])))
    assert worst < 1e-5, worst


def test_loss_known_answer_zero_network():
    # all-zero parameters: logits 0 -> softplus(0) = ln 2 per valid slot
    # (both classes, pos_weight 1); velocity head 0 -> 0.5 |v/2|^2
    p = np.zeros(M.N_PARAMS)
    x = np.zeros((2, 64))
    lab = np.array([(0b101) | (0b111 << 16), (0b1) | (0b1 << 16)], np.uint32)
    v = np.array([[2.0, 0.0], [0.0, -4.0]])
    loss, g = M.loss_grad(p, x, lab, v, pos_weight=1.0, vel_weight=1.0, bf16=False)
    expect = (4 * np.log(2) + 0.5 * (1.0 + 4.0)) / 2
    assert abs(loss - expect) < 1e-12
    # db4 = mean over the batch of (sigmoid(0) - t) on valid slots, -(t_v) on the velocity head
    db4 = g[-16:]
    assert abs(db4[0] - (0.5 - 1 + 0.5 - 1) / 2) < 1e-12
    assert abs(db4[1] - 0.5 / 2) < 1e-12 and abs(db4[3]) < 1e-12
    assert abs(db4[14] - (-1.0) / 2) < 1e-12 and abs(db4[15] - 2.0 / 2) < 1e-12


def test_bf16_rounding_known_answers():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    r = M.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even mantissa
    assert r[2] == 1 + 2 ** -6              # tie -> even mantissa (up)
    assert r[3] == -2.5
    assert np.isfinite(r[4]) and abs(r[4] - 3.0e38) / 3.0e38 < 2 ** -8
    assert r[5] == 2.0 ** -133              # subnormal: nearest multiple of the BF16 spacing 2^-133


def test_hidden_permutation_maps_accumulator_rows():
    perm = [M.hidden_perm(k) for k in range(128)]
    assert sorted(perm) == list(range(128))
    # chunk kc = 2 mi + h of the next layer only holds rows of m-tile mi,
    # and lane half g = k_local // 8 holds accumulator rows 8 j + 4 g + r
    for k in range(128):
        kc, kl = divmod(k, 16)
        mi, h = divmod(kc, 2)
        row = perm[k] - 32 * mi
        assert 0 <= row < 32
        g, e = divmod(kl, 8)
        reg = 8 * h + e                      # accumulator register index
        assert row == 8 * (reg // 4) + 4 * g + reg % 4


def _state(pos, vel, rad, speed):
    n = len(pos)
    return dict(position=np.array(pos, np.float32), velocity=np.array(vel, np.float32),
                clearance=np.full(n, rad, np.float32), pref_speed=np.full(n, speed, np.float32),
                attraction=np.zeros((n, 2), np.float32), active=np.ones(n, np.uint8),
                path_off=np.zeros(n, np.int32), path_cnt=np.ones(n, np.int32), path_xy=np.zeros((1, 2), np.float32))


def test_features_and_label_closed_form_head_on():
    # agent 0 at the origin moving +x at 1 m/s, agent 1 at (1.5, 0) moving
    # -x: the agent constraint binds the preferred velocity (1, 0), so slot 0
    # is active at the LP velocity and exists
    st = _state([(0, 0), (1.5, 0)], [(1, 0), (-1, 0)], 0.25, 1.3)
    p = Params()
    obs = S.rects_to_obstacles([])
    cons, nob = O.constraints(p, st, obs, 0, np.array([1], np.int32))
    assert nob == 0 and len(cons) == 1
    v, _ = O.lp2d(cons, np.array([1.0, 0.0], np.float32), 2.0)
    feat, lab = O.proposal_sample(p, st, obs, 0, [1], [1.0, 0.0], v)
    assert feat[0] == np.float32(1.5 * 0.25) and feat[1] == 0.0
    assert feat[2] == np.float32(-2.0 * 0.5) and feat[3] == 0.0
    assert np.all(feat[4:56] == 0.0)
    assert feat[56] == 0.5 and feat[57] == 0.0 and feat[58] == 0.5 and feat[59] == 0.0
    assert feat[60] == 1.0 and np.all(feat[61:] == 0.0)
    assert lab & 1 and (lab >> 16) & 1 and (lab >> 30) & 1
    assert (lab & 0x3FFE) == 0 and ((lab >> 16) & 0x3FFE) == 0


def test_features_nearest_obstacle_edges_order():
    # agent 1 m right of a 2x2 square centred at the origin: the square's
    # right edge (x = 1) is nearest, the top and bottom edges are not kept
    # (the agent is on their inner side), the left edge is out of sight
    st = _state([(2.0, 0.0)], [(0, 0)], 0.25, 1.3)
    p = Params()
    obs = S.rects_to_obstacles([(-1, -1, 1, 1)])
    feat, lab = O.proposal_sample(p, st, obs, 0, [], [0.0, 0.0], [0.0, 0.0])
    # CCW ring (1,1) (-1,1) (-1,-1) (1,-1): the edge (1,-1) -> (1,1) faces +x
    assert np.allclose(feat[40:44], np.array([1 - 2.0, -1, 1 - 2.0, 1], np.float32) * 0.25)
    assert np.all(feat[44:56] == 0.0)
