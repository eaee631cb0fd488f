"""Python mirror of ECM::Simulation::Simulator (ECMAgentSimulator/Simulator.h:59-188)
over the C ABI of libecmcrowd.so.

Method names follow the reference (SpawnAgent, DestroyAgent, Initialize,
Update, Reset, Get*Data); the component getters return numpy copies of the
device state instead of raw pointers.  The step itself runs on the GPU; there
is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as N

COMPONENTS = {
    "position": (0, np.float32, 2),
    "velocity": (1, np.float32, 2),
    "pref_velocity": (2, np.float32, 2),
    "attraction": (3, np.float32, 2),
    "clearance": (4, np.float32, 1),
    "pref_speed": (5, np.float32, 1),
    "active": (6, np.uint8, 1),
    "orca_velocity": (7, np.float32, 2),
}


@dataclass
class Params:
    """ecm_sim_params; defaults are the benchmark values (k=10, T=T_obst=2 s,
    maxSpeed 2 m/s, dt 0.1 s).  Params.reference() gives the reference's
    literals (Simulator.cpp:55, ORCA.h:102-103, Simulator.cpp:671-674)."""
    max_agents: int = 10000
    k: int = 10
    time_horizon: float = 2.0
    time_horizon_obst: float = 2.0
    max_speed: float = 2.0
    max_speed_from_pref: bool = False
    dt: float = 0.1
    mass: float = 0.8
    arrival_radius: float = 20.0
    delete_radius: float = 2.0
    grid_cell: float = 0.0
    keep_orca_velocity: bool = False

    @classmethod
    def reference(cls, **kw) -> "Params":
        p = cls(k=5, time_horizon=10.0, time_horizon_obst=10.0, max_speed_from_pref=True)
        for a, b in kw.items():
            setattr(p, a, b)
        return p

    def to_c(self) -> N.SimParams:
        c = N.SimParams()
        for f in N.SimParams._fields_:
            v = getattr(self, f[0])
            setattr(c, f[0], int(v) if f[1] is C.c_int else float(v))
        return c


@dataclass
class Ecm:
    """ECM graph, query side (ECM.h:20-50)."""
    vert_xy: np.ndarray          # [nv, 2] float32
    he_target: np.ndarray        # [ne, 2] int32: he0.v_target, he1.v_target
    he_pts: np.ndarray           # [ne, 8] float32: he0.cl, he0.cr, he1.cl, he1.cr
    vert_clearance: Optional[np.ndarray] = None
    he_next: Optional[np.ndarray] = None

    @property
    def n_vert(self) -> int:
        return int(self.vert_xy.shape[0])

    @property
    def n_edge(self) -> int:
        return int(self.he_target.shape[0])


@dataclass
class Obstacles:
    """CCW rings in (obstacle, vertex) order (Environment.cpp:285-315)."""
    xy: np.ndarray               # [nv, 2] float32
    ring_off: np.ndarray         # [n_rings + 1] int32

    @property
    def n_rings(self) -> int:
        return int(self.ring_off.shape[0] - 1)


@dataclass
class Agents:
    start: np.ndarray            # [n, 2]
    radius: np.ndarray           # [n]
    pref_speed: np.ndarray       # [n]
    path_off: np.ndarray         # [n]
    path_cnt: np.ndarray         # [n]
    path_xy: np.ndarray          # [P, 2]

    @property
    def n(self) -> int:
        return int(self.start.shape[0])


class Simulator:
    """Simulator(ecm, planner, environment, maxAgents, simStepTime) (Simulator.h:63-74)."""

    def __init__(self, ecm: Ecm, obstacles: Obstacles, world, params: Params, planner=None):
        self._L = N.lib()
        self.params = params
        self.world = tuple(float(v) for v in world)
        self._h = C.c_void_p()
        N.check(self._L.ecm_sim_create(C.byref(params.to_c()), *self.world, C.byref(self._h)))
        self.max_agents = params.max_agents
        self.SetObstacles(obstacles)
        self.SetEcm(ecm)
        self._planner = planner
        self._cb = None
        self._pinned = []
        if planner is not None:
            self.AttachPlanner(planner)

    # ---- lifecycle ------------------------------------------------------
    def close(self) -> None:
        if self._h:
            self._L.ecm_sim_destroy(self._h)
            self._h = C.c_void_p()
        for p in getattr(self, "_pinned", []):
            self._L.ecm_host_free(p)
        self._pinned = []

    def pinned(self, name: str) -> np.ndarray:
        """A page-locked host buffer shaped for component `name`, for
        read(name, out=...) every step: the copy then runs at full PCIe
        speed.  The buffer belongs to the simulator and is freed by close();
        do not use it afterwards."""
        cid, dt, w = COMPONENTS[name]
        shape = (self.max_agents, w) if w > 1 else (self.max_agents,)
        nbytes = int(np.prod(shape)) * np.dtype(dt).itemsize
        p = C.c_void_p()
        N.check(self._L.ecm_host_alloc(nbytes, C.byref(p)))
        self._pinned.append(p)
        raw = (C.c_ubyte * nbytes).from_address(p.value)
        return np.frombuffer(raw, dtype=dt).reshape(shape)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Initialize(self) -> None:        # Simulator.cpp:21-60 (buffers exist after create)
        return None

    def SetObstacles(self, obs: Obstacles, env: int = 0) -> None:
        xy = np.ascontiguousarray(obs.xy, dtype=np.float32)
        ro = np.ascontiguousarray(obs.ring_off, dtype=np.int32)
        N.check(self._L.ecm_sim_env_set_obstacles(self._h, int(env), N.fptr(xy), N.iptr(ro), obs.n_rings))

    def SetEcm(self, ecm: Ecm, env: int = 0) -> None:
        v = np.ascontiguousarray(ecm.vert_xy, dtype=np.float32)
        t = np.ascontiguousarray(ecm.he_target, dtype=np.int32)
        p = np.ascontiguousarray(ecm.he_pts, dtype=np.float32)
        N.check(self._L.ecm_sim_env_set_ecm(self._h, int(env), ecm.n_vert, N.fptr(v), ecm.n_edge, N.iptr(t),
                                            N.fptr(p)))

    def AttachPlanner(self, scene, env: int = 0) -> None:
        """UpdatePath on IRM failure (Simulator.cpp:579-587) through a scene's planner."""
        fn = self._L.ecm_scene_replan_callback()
        N.check(self._L.ecm_sim_env_set_replan_callback(self._h, int(env), fn, scene.handle))
        # the scene planner is thread-safe (read-only ECM and index)
        N.check(self._L.ecm_sim_set_replan_threads(self._h, max(1, min(16, os.cpu_count() or 1))))
        self._planners = getattr(self, "_planners", {})
        self._planners[int(env)] = scene
        if env == 0:
            self._planner = scene

    # ---- environments (batched multi-scenario stepping) -------------------
    def AddEnvironment(self, ecm: Ecm, obstacles: Obstacles, world, planner=None) -> int:
        """A further independent scene advanced by the same step; returns its id."""
        e = C.c_int()
        N.check(self._L.ecm_sim_add_env(self._h, *[float(v) for v in world], C.byref(e)))
        self.SetObstacles(obstacles, e.value)
        self.SetEcm(ecm, e.value)
        if planner is not None:
            self.AttachPlanner(planner, e.value)
        return e.value

    def num_environments(self) -> int:
        n = C.c_int()
        N.check(self._L.ecm_sim_num_envs(self._h, C.byref(n)))
        return n.value

    def agent_environments(self) -> np.ndarray:
        out = np.empty(self.max_agents, np.int32)
        N.check(self._L.ecm_sim_agent_envs(self._h, N.iptr(out)))
        return out

    # ---- agents ---------------------------------------------------------
    def SpawnAgents(self, agents: Agents, validate: bool = True, env: int = 0) -> np.ndarray:
        """SpawnAgent (Simulator.cpp:168-200) for a batch into environment
        `env`; -1 where rejected."""
        n = agents.n
        ids = np.full(n, -1, dtype=np.int32)
        st = np.ascontiguousarray(agents.start, dtype=np.float32)
        r = np.ascontiguousarray(agents.radius, dtype=np.float32)
        s = np.ascontiguousarray(agents.pref_speed, dtype=np.float32)
        po = np.ascontiguousarray(agents.path_off, dtype=np.int32)
        pc = np.ascontiguousarray(agents.path_cnt, dtype=np.int32)
        px = np.ascontiguousarray(agents.path_xy, dtype=np.float32)
        N.check(self._L.ecm_sim_env_spawn(self._h, int(env), n, N.fptr(st), N.fptr(r), N.fptr(s), N.iptr(po),
                                          N.iptr(pc), N.fptr(px), int(validate), N.iptr(ids)))
        return ids

    def SpawnAgent(self, start, goal_path: Sequence, clearance: float, preferred_speed: float) -> int:
        path = np.asarray(goal_path, dtype=np.float32).reshape(-1, 2)
        a = Agents(np.asarray([start], np.float32), np.asarray([clearance], np.float32),
                   np.asarray([preferred_speed], np.float32), np.zeros(1, np.int32),
                   np.asarray([path.shape[0]], np.int32), path)
        return int(self.SpawnAgents(a)[0])

    def DestroyAgent(self, idx: int) -> None:
        N.check(self._L.ecm_sim_destroy_agent(self._h, int(idx)))

    def Reset(self) -> None:
        N.check(self._L.ecm_sim_reset(self._h))

    def SetPath(self, idx: int, path) -> None:
        p = np.ascontiguousarray(np.asarray(path, np.float32).reshape(-1, 2))
        N.check(self._L.ecm_sim_set_path(self._h, int(idx), N.fptr(p), p.shape[0]))

    # ---- stepping -------------------------------------------------------
    def Update(self, dt: float = 0.0, steps: int = 1) -> None:
        """Simulator::Update (Simulator.cpp:314-323); dt is ignored as in the reference."""
        N.check(self._L.ecm_sim_step(self._h, int(steps)))

    def sync(self) -> None:
        N.check(self._L.ecm_sim_sync(self._h))

    # ---- getters (Simulator.h:114-126) ---------------------------------
    def read(self, name: str, out: np.ndarray = None) -> np.ndarray:
        """Component `name` of every slot; into `out` when given (a buffer
        kept from step to step, as a renderer keeps its own)."""
        cid, dt, w = COMPONENTS[name]
        shape = (self.max_agents, w) if w > 1 else (self.max_agents,)
        if out is None:
            a = np.empty(shape, dtype=dt)
        else:
            if out.shape != shape or out.dtype != np.dtype(dt) or not out.flags.c_contiguous:
                raise ValueError(f"{name} needs a C-contiguous {np.dtype(dt)} array of shape {shape}")
            a = out
        N.check(self._L.ecm_sim_read(self._h, cid, N.vptr(a), a.nbytes))
        return a

    def write(self, name: str, a: np.ndarray) -> None:
        cid, dt, w = COMPONENTS[name]
        a = np.ascontiguousarray(a, dtype=dt)
        N.check(self._L.ecm_sim_write(self._h, cid, N.vptr(a), a.nbytes))

    def GetPositionData(self, out=None): return self.read("position", out)
    def GetVelocityData(self): return self.read("velocity")
    def GetPreferredVelocityData(self): return self.read("pref_velocity")
    def GetAttractionPointData(self): return self.read("attraction")
    def GetClearanceData(self): return self.read("clearance")
    def GetActiveFlags(self): return self.read("active").astype(bool)
    def GetSimulationStepTime(self): return self.params.dt

    def GetNumAgents(self) -> int:
        v = C.c_int()
        N.check(self._L.ecm_sim_num_agents(self._h, C.byref(v)))
        return v.value

    def GetLastIndex(self) -> int:
        v = C.c_int()
        N.check(self._L.ecm_sim_last_index(self._h, C.byref(v)))
        return v.value

    def neighbors(self):
        k = self.params.k
        nbr = np.empty((self.max_agents, k), np.int32)
        cnt = np.empty(self.max_agents, np.int32)
        N.check(self._L.ecm_sim_get_neighbors(self._h, N.iptr(nbr), N.iptr(cnt)))
        return nbr, cnt

    def paths(self):
        """(path_off, path_cnt, path_xy) of every slot, as stored on the device."""
        used = C.c_size_t()
        N.check(self._L.ecm_sim_read_paths(self._h, None, None, None, 0, C.byref(used)))
        off = np.empty(self.max_agents, np.int32)
        cnt = np.empty(self.max_agents, np.int32)
        pool = np.empty((max(1, used.value), 2), np.float32)
        N.check(self._L.ecm_sim_read_paths(self._h, N.iptr(off), N.iptr(cnt), N.fptr(pool), pool.shape[0],
                                           C.byref(used)))
        return off, cnt, pool[: used.value]

    def replans(self) -> np.ndarray:
        n = C.c_int()
        buf = np.empty(self.max_agents, np.int32)
        N.check(self._L.ecm_sim_get_replans(self._h, N.iptr(buf), self.max_agents, C.byref(n)))
        return buf[: n.value].copy()

    def stats(self) -> dict:
        keys = ["steps", "agent_updates", "lp3d", "locate_fallback", "replans", "destroyed",
                "obst_overflow", "learned_tried", "learned_accepted", "learned_proposed", "lp3d_big"]
        a = (C.c_uint64 * len(keys))()
        N.check(self._L.ecm_sim_stats(self._h, a, len(keys)))
        return {k: int(v) for k, v in zip(keys, a)}

    def set_timing(self, on: bool) -> None:
        N.check(self._L.ecm_sim_set_timing(self._h, int(on)))

    def last_timing(self):
        """[grid build, attraction + k-nearest, learned proposal, ORCA + LP +
        integration, the main ORCA pass alone, the attraction pass alone]
        device ms of the last step (with set_timing(True))."""
        a = (C.c_float * 6)()
        N.check(self._L.ecm_sim_last_timing(self._h, a, 6))
        return list(a)

    def set_graph(self, on: bool = True) -> None:
        """Replay the device part of each step as captured hipGraphs."""
        N.check(self._L.ecm_sim_set_graph(self._h, int(on)))

    def graph_status(self):
        """(graphs in use, reason they are not)."""
        a = C.c_int()
        why = C.create_string_buffer(512)
        N.check(self._L.ecm_sim_graph_status(self._h, C.byref(a), why, 512))
        return bool(a.value), why.value.decode()

    def tuning(self) -> dict:
        t = N.SimTuning()
        N.check(self._L.ecm_sim_get_tuning(self._h, C.byref(t)))
        return {"cell_scale": t.cell_scale, "knn_block": t.knn_block, "orca_obst_slots": t.orca_obst_slots,
                "best_ms": t.best_ms}

    def set_tuning(self, cell_scale=1.25, knn_block=256, orca_obst_slots=2) -> None:
        t = N.SimTuning(float(cell_scale), int(knn_block), int(orca_obst_slots), 0.0)
        N.check(self._L.ecm_sim_set_tuning(self._h, C.byref(t)))

    def get_tuning(self) -> dict:
        """The launch shape in use."""
        t = N.SimTuning()
        N.check(self._L.ecm_sim_get_tuning(self._h, C.byref(t)))
        return {"cell_scale": round(t.cell_scale, 3), "knn_block": t.knn_block, "orca_obst_slots": t.orca_obst_slots,
                "best_ms": round(t.best_ms, 4)}

    def autotune(self, steps: int = 3) -> dict:
        """Pick the launch shape by timing trials on this workload (advances
        the simulation); returns the chosen shape and its step time."""
        t = N.SimTuning()
        N.check(self._L.ecm_sim_autotune(self._h, int(steps), C.byref(t)))
        return {"cell_scale": round(t.cell_scale, 3), "knn_block": t.knn_block, "orca_obst_slots": t.orca_obst_slots,
                "best_ms": round(t.best_ms, 4)}

    def set_learned(self, net, theta: float = 0.0) -> None:
        """Use the learned ORCA proposal of `net` (a learned.ProposalNet), or
        the plain LP with net=None."""
        N.check(self._L.ecm_sim_set_learned(self._h, net.handle if net is not None else None, float(theta)))
        self._learned = net

    def log_step(self, net) -> int:
        """One plain step that appends training samples to `net`'s dataset."""
        n = C.c_int()
        N.check(self._L.ecm_sim_log_step(self._h, net.handle, C.byref(n)))
        return n.value

    def snapshot_save(self, path: str) -> None:
        N.check(self._L.ecm_sim_snapshot_save(self._h, path.encode()))

    def snapshot_load(self, path: str) -> None:
        N.check(self._L.ecm_sim_snapshot_load(self._h, path.encode()))

    def state(self) -> dict:
        """Copy of every per-slot component (used for parity snapshots)."""
        return {k: self.read(k) for k in COMPONENTS if k != "orca_velocity" or self.params.keep_orca_velocity}
