"""Shared helpers for the parity tests: build a scene, mirror the GPU state
for the oracle, and compare step outputs."""
from __future__ import annotations

import numpy as np

from ecmgenerator_amd import scenarios as S
from ecmgenerator_amd.simulator import Params, Simulator


def small_scene(seed: int, half: float = 50.0, n_rects: int = 8, n_agents: int = 300, side=(2, 10)):
    rng = np.random.default_rng(seed)
    box = (-half, -half, half, half)
    rects = S.random_rects(rng, box, n_rects, side[0], side[1])
    obs = S.rects_to_obstacles(rects)
    ecm = S.empty_square_ecm(*box)
    agents = S.straight_agents(rng, box, n_agents, rects)
    return box, rects, obs, ecm, agents


def oracle_state(sim: Simulator) -> dict:
    st = sim.state()
    off, cnt, pool = sim.paths()
    st = dict(st)
    st["path_off"] = off
    st["path_cnt"] = np.maximum(cnt, 1)
    st["path_xy"] = pool if pool.shape[0] else np.zeros((1, 2), np.float32)
    return st


def compare_step(sim: Simulator, before: dict, ora: dict, check_orca=True, subset=None, detail=False):
    """Bitwise comparison of one GPU step against the oracle step from the
    same input state (restricted to `subset` when given).  Returns a dict of
    mismatch counts."""
    after = sim.state()
    act0 = before["active"].astype(bool)
    if subset is not None:
        m = np.zeros_like(act0)
        m[subset] = True
        act0 = act0 & m
    res = {}
    for key in ("position", "velocity", "attraction"):
        a, b = after[key][act0], ora[key][act0]
        res[key] = int(np.sum(~((a == b) | (np.isnan(a) & np.isnan(b))).all(axis=1)))
    alive = act0 & ora["active"].astype(bool)
    a, b = after["pref_velocity"][alive], ora["pref_velocity"][alive]
    res["pref_velocity"] = int(np.sum(~((a == b) | (np.isnan(a) & np.isnan(b))).all(axis=1)))
    res["active"] = int(np.sum(after["active"][act0] != ora["active"][act0]))
    nbr, cnt = sim.neighbors()
    res["nbr_cnt"] = int(np.sum(cnt[alive] != ora["nbr_cnt"][alive]))
    res["nbr"] = int(np.sum((nbr[alive] != ora["nbr"][alive]).any(axis=1)))
    if check_orca and sim.params.keep_orca_velocity:
        a, b = after["orca_velocity"][alive], ora["v_orca"][alive]
        res["v_orca"] = int(np.sum(~((a == b) | (np.isnan(a) & np.isnan(b))).all(axis=1)))
    rp = np.zeros(sim.max_agents, np.uint8)
    rp[sim.replans()] = 1
    res["replan"] = int(np.sum(rp[act0] != ora["replan"][act0]))
    if detail:
        bad = np.flatnonzero(act0 & ~((after["velocity"] == ora["velocity"]) | (np.isnan(after["velocity"]) & np.isnan(ora["velocity"]))).all(axis=1))
        res["bad_ids"] = bad[:20].tolist()
    return res


def ulp32(x):
    """Spacing of float32 numbers at |x|."""
    a = np.abs(np.asarray(x, np.float32))
    return np.spacing(a).astype(np.float64)


def compare_step_close(sim: Simulator, before: dict, ora: dict, subset, tol_v=1e-4, tol_p=1e-5):
    """Tolerance comparison of one GPU step against the oracle on `subset`:
    ORCA velocities and velocities within tol_v m/s, positions within
    max(tol_p m, 1 ulp of the position), discrete outputs (active flags,
    neighbour lists, replans) exact.  Returns (mismatch counts, max errors)."""
    after = sim.state()
    act0 = np.zeros_like(before["active"].astype(bool))
    act0[subset] = True
    act0 &= before["active"].astype(bool)
    alive = act0 & ora["active"].astype(bool)
    res, err = {}, {}
    dv = np.abs(after["orca_velocity"][alive].astype(np.float64) - ora["v_orca"][alive]).max(axis=1)
    res["v_orca"] = int(np.sum(~(dv <= tol_v)))
    err["v_orca"] = float(np.nanmax(dv)) if dv.size else 0.0
    dv = np.abs(after["velocity"][act0].astype(np.float64) - ora["velocity"][act0]).max(axis=1)
    res["velocity"] = int(np.sum(~(dv <= tol_v)))
    err["velocity"] = float(np.nanmax(dv)) if dv.size else 0.0
    dp = np.abs(after["position"][act0].astype(np.float64) - ora["position"][act0])
    gate = np.maximum(tol_p, ulp32(ora["position"][act0]))
    res["position"] = int(np.sum(~(dp <= gate).all(axis=1)))
    err["position"] = float(np.nanmax(dp)) if dp.size else 0.0
    res["active"] = int(np.sum(after["active"][act0] != ora["active"][act0]))
    nbr, cnt = sim.neighbors()
    res["nbr"] = int(np.sum((nbr[alive] != ora["nbr"][alive]).any(axis=1)))
    res["nbr_cnt"] = int(np.sum(cnt[alive] != ora["nbr_cnt"][alive]))
    rp = np.zeros(sim.max_agents, np.uint8)
    rp[sim.replans()] = 1
    res["replan"] = int(np.sum(rp[act0] != ora["replan"][act0]))
    return res, err
