"""ctypes binding of the CPU oracle (oracle/_build/libecm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libecm_oracle.so")

KNN_EXACT = 0
KNN_REFERENCE = 1


class OParams(C.Structure):
    _fields_ = [("k", C.c_int), ("time_horizon", C.c_float), ("time_horizon_obst", C.c_float),
                ("max_speed", C.c_float), ("max_speed_from_pref", C.c_int), ("dt", C.c_float),
                ("mass", C.c_float), ("arrival_radius", C.c_float), ("delete_radius", C.c_float),
                ("knn_mode", C.c_int)]


class OState(C.Structure):
    _fields_ = [("n", C.c_int), ("active", C.c_void_p), ("pos", C.c_void_p), ("vel", C.c_void_p),
                ("radius", C.c_void_p), ("pref_speed", C.c_void_p), ("attraction", C.c_void_p),
                ("path_off", C.c_void_p), ("path_cnt", C.c_void_p), ("path_pts", C.c_void_p)]


class OObst(C.Structure):
    _fields_ = [("n_rings", C.c_int), ("ring_off", C.c_void_p), ("xy", C.c_void_p)]


class OEcm(C.Structure):
    _fields_ = [("n_vert", C.c_int), ("vert_xy", C.c_void_p), ("n_edge", C.c_int),
                ("he_target", C.c_void_p), ("he_pts", C.c_void_p)]


class OOut(C.Structure):
    _fields_ = [("pos", C.c_void_p), ("vel", C.c_void_p), ("attraction", C.c_void_p), ("pref_vel", C.c_void_p),
                ("active", C.c_void_p), ("replan", C.c_void_p), ("v_orca", C.c_void_p), ("nbr", C.c_void_p),
                ("nbr_cnt", C.c_void_p), ("diag", C.c_void_p), ("used_3d", C.c_void_p)]


_lib = None


def build() -> str:
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_step.restype = C.c_int
        L.oracle_step.argtypes = [C.POINTER(OParams), C.POINTER(OState), C.POINTER(OObst), C.POINTER(OEcm),
                                  C.c_void_p, C.c_int, C.POINTER(OOut)]
        L.oracle_locate_cell.restype = C.c_int
        L.oracle_locate_cell.argtypes = [C.POINTER(OEcm), C.c_float, C.c_float]
        L.oracle_retract.restype = C.c_int
        L.oracle_retract.argtypes = [C.POINTER(OEcm), C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        L.oracle_attraction.restype = C.c_int
        L.oracle_attraction.argtypes = [C.POINTER(OEcm), C.c_float, C.c_float, C.c_void_p, C.c_int, C.c_void_p]
        L.oracle_lp2d.restype = C.c_int
        L.oracle_lp2d.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_float, C.c_int, C.c_void_p]
        L.oracle_lp3d.restype = None
        L.oracle_lp3d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_int, C.c_void_p]
        L.oracle_constraints.restype = C.c_int
        L.oracle_constraints.argtypes = [C.POINTER(OParams), C.POINTER(OState), C.POINTER(OObst), C.c_int,
                                         C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        for f in ("oracle_atanf", "oracle_sinf", "oracle_cosf"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [C.c_float]
        L.oracle_proposal_sample.restype = C.c_int
        L.oracle_proposal_sample.argtypes = [C.POINTER(OParams), C.POINTER(OState), C.POINTER(OObst), C.c_int,
                                             C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_obstacle_halfplane.restype = C.c_int
        L.oracle_obstacle_halfplane.argtypes = [C.POINTER(OParams), C.POINTER(OState), C.POINTER(OObst), C.c_int,
                                                C.c_int, C.c_void_p, C.POINTER(C.c_uint)]
        L.oracle_nearest_obstacles.restype = C.c_int
        L.oracle_nearest_obstacles.argtypes = [C.POINTER(OObst), C.c_float, C.c_float, C.c_float, C.c_void_p,
                                               C.c_int]
        L.oracle_kdtree_build_only.restype = C.c_int
        L.oracle_kdtree_build_only.argtypes = [C.POINTER(OState)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Problem:
    """Holds numpy arrays alive for one oracle call."""

    def __init__(self, params, state: dict, obstacles, ecm, knn_mode=KNN_EXACT):
        self.keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep.append(a)
            return a

        n = int(state["active"].shape[0])
        self.n = n
        self.prm = OParams(int(params.k), float(params.time_horizon), float(params.time_horizon_obst),
                           float(params.max_speed), int(bool(params.max_speed_from_pref)), float(params.dt),
                           float(params.mass), float(params.arrival_radius), float(params.delete_radius),
                           int(knn_mode))
        self.st = OState(n, _p(arr(state["active"], np.uint8)), _p(arr(state["position"], np.float32)),
                         _p(arr(state["velocity"], np.float32)), _p(arr(state["clearance"], np.float32)),
                         _p(arr(state["pref_speed"], np.float32)), _p(arr(state["attraction"], np.float32)),
                         _p(arr(state["path_off"], np.int32)), _p(arr(state["path_cnt"], np.int32)),
                         _p(arr(state["path_xy"], np.float32)))
        if obstacles is not None and obstacles.n_rings > 0:
            self.obs = OObst(obstacles.n_rings, _p(arr(obstacles.ring_off, np.int32)), _p(arr(obstacles.xy, np.float32)))
        else:
            self.obs = OObst(0, None, None)
        self.ecm = OEcm(ecm.n_vert, _p(arr(ecm.vert_xy, np.float32)), ecm.n_edge,
                        _p(arr(ecm.he_target, np.int32)), _p(arr(ecm.he_pts, np.float32)))
        self.state = state


def step(params, state: dict, obstacles, ecm, subset=None, knn_mode=KNN_EXACT):
    """One oracle step from `state`; returns a dict of outputs (copies of the
    inputs where an agent was not processed)."""
    P = Problem(params, state, obstacles, ecm, knn_mode)
    n, k = P.n, int(params.k)
    out = {
        "position": np.array(state["position"], np.float32, copy=True),
        "velocity": np.array(state["velocity"], np.float32, copy=True),
        "attraction": np.array(state["attraction"], np.float32, copy=True),
        "pref_velocity": np.array(state.get("pref_velocity", np.zeros((n, 2))), np.float32, copy=True),
        "active": np.array(state["active"], np.uint8, copy=True),
        "replan": np.zeros(n, np.uint8),
        "v_orca": np.full((n, 2), np.nan, np.float32),
        "nbr": np.full((n, k), -1, np.int32),
        "nbr_cnt": np.zeros(n, np.int32),
        "diag": np.full((n, 4), -9, np.int32),
        "used_3d": np.zeros(n, np.uint8),
    }
    o = OOut(*[_p(out[key]) for key in ("position", "velocity", "attraction", "pref_velocity", "active", "replan",
                                        "v_orca", "nbr", "nbr_cnt", "diag", "used_3d")])
    sub = None if subset is None else np.ascontiguousarray(subset, dtype=np.int32)
    rc = lib().oracle_step(C.byref(P.prm), C.byref(P.st), C.byref(P.obs), C.byref(P.ecm), _p(sub),
                           0 if sub is None else int(sub.shape[0]), C.byref(o))
    if rc != 0:
        raise RuntimeError(f"oracle_step failed: {rc}")
    return out


def ecm_struct(ecm):
    keep = [np.ascontiguousarray(ecm.vert_xy, np.float32), np.ascontiguousarray(ecm.he_target, np.int32),
            np.ascontiguousarray(ecm.he_pts, np.float32)]
    s = OEcm(ecm.n_vert, _p(keep[0]), ecm.n_edge, _p(keep[1]), _p(keep[2]))
    return s, keep


def locate(ecm, x, y) -> int:
    s, keep = ecm_struct(ecm)
    return int(lib().oracle_locate_cell(C.byref(s), float(x), float(y)))


def retract(ecm, x, y):
    s, keep = ecm_struct(ecm)
    out = np.zeros(2, np.float32)
    e = np.zeros(1, np.int32)
    ok = lib().oracle_retract(C.byref(s), float(x), float(y), _p(out), _p(e))
    return bool(ok), out, int(e[0])


def attraction(ecm, x, y, path):
    s, keep = ecm_struct(ecm)
    p = np.ascontiguousarray(path, np.float32)
    out = np.zeros(2, np.float32)
    ok = lib().oracle_attraction(C.byref(s), float(x), float(y), _p(p), int(p.shape[0]), _p(out))
    return bool(ok), out


def lp2d(constraints, opt, max_speed, use_dir=False):
    c = np.ascontiguousarray(constraints, np.float32).reshape(-1, 4)
    o = np.ascontiguousarray(opt, np.float32)
    out = np.zeros(2, np.float32)
    r = lib().oracle_lp2d(_p(c), int(c.shape[0]), _p(o), float(max_speed), int(use_dir), _p(out))
    return int(r), out


def lp3d(constraints, n_obst, max_speed, failed, vel):
    c = np.ascontiguousarray(constraints, np.float32).reshape(-1, 4)
    v = np.array(vel, np.float32, copy=True)
    lib().oracle_lp3d(_p(c), int(c.shape[0]), int(n_obst), float(max_speed), int(failed), _p(v))
    return v


def proposal_sample(params, state, obstacles, agent, nbrs, pref_vel, v):
    """(features[64] float32 before BF16 rounding, label word) of the
    proposal network for one agent and LP velocity v."""
    P = Problem(params, state, obstacles, _EmptyEcm())
    nb = np.ascontiguousarray(nbrs, np.int32)
    pv = np.ascontiguousarray(pref_vel, np.float32)
    vv = np.ascontiguousarray(v, np.float32)
    feat = np.zeros(64, np.float32)
    lab = np.zeros(1, np.uint32)
    lib().oracle_proposal_sample(C.byref(P.prm), C.byref(P.st), C.byref(P.obs), int(agent), _p(nb), int(nb.shape[0]),
                                 _p(pv), _p(vv), _p(feat), _p(lab))
    return feat, int(lab[0])


def constraints(params, state, obstacles, agent, nbrs):
    P = Problem(params, state, obstacles, _EmptyEcm())
    nb = np.ascontiguousarray(nbrs, np.int32)
    out = np.zeros((256, 4), np.float32)
    nob = np.zeros(1, np.int32)
    n = lib().oracle_constraints(C.byref(P.prm), C.byref(P.st), C.byref(P.obs), int(agent), _p(nb),
                                 int(nb.shape[0]), _p(out), 256, _p(nob))
    return out[:n].copy(), int(nob[0])


# branch bits of obstacle_halfplane (ORACLE_OB_* in ecm_oracle.h)
OB_BITS = ["case1", "case1_skip", "case2", "case2_skip", "case3", "vertex_l", "vertex_l_skip", "vertex_r",
           "vertex_r_skip", "concave_l", "concave_r", "foreign_l", "foreign_r", "circle_l", "circle_r",
           "cutoff_line", "leg_l", "leg_l_skip", "leg_r", "leg_r_skip"]


def obstacle_halfplane(params, state, obstacles, agent, edge):
    """(constraints emitted (0/1), (nx, ny, px, py), set of branch names) of
    obstacle edge `edge` for `agent` (ORCA.cpp:70-333)."""
    P = Problem(params, state, obstacles, _EmptyEcm())
    out = np.zeros(4, np.float32)
    b = C.c_uint(0)
    n = lib().oracle_obstacle_halfplane(C.byref(P.prm), C.byref(P.st), C.byref(P.obs), int(agent), int(edge), _p(out),
                                        C.byref(b))
    if n < 0:
        raise RuntimeError("oracle_obstacle_halfplane failed")
    return n, out, {name for i, name in enumerate(OB_BITS) if (b.value >> i) & 1}


def nearest_obstacles(obstacles, x, y, range_):
    """FindNearestObstacles (Simulator.cpp:259-292): kept edges in scan order."""
    keep = [np.ascontiguousarray(obstacles.ring_off, np.int32), np.ascontiguousarray(obstacles.xy, np.float32)]
    o = OObst(obstacles.n_rings, _p(keep[0]), _p(keep[1]))
    out = np.zeros(4096, np.int32)
    n = lib().oracle_nearest_obstacles(C.byref(o), float(x), float(y), float(range_), _p(out), out.shape[0])
    if n < 0:
        raise RuntimeError("oracle_nearest_obstacles failed")
    return out[:min(n, out.shape[0])].copy()


class _EmptyEcm:
    n_vert = 0
    n_edge = 0
    vert_xy = np.zeros((0, 2), np.float32)
    he_target = np.zeros((0, 2), np.int32)
    he_pts = np.zeros((0, 8), np.float32)


def atanf(x): return float(lib().oracle_atanf(float(x)))
def sinf(x): return float(lib().oracle_sinf(float(x)))
def cosf(x): return float(lib().oracle_cosf(float(x)))
