"""Python mirror of the host scene (include/ecm_scene.h): Environment with a
walkable box and obstacles (ECMGenerator/Environment.h:18-85), its ECM
(ECMGenerator::GenerateECM) and the path planner (ECMPathPlanner::FindPath),
plus the seeded benchmark scenarios."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .simulator import Agents, Ecm, Obstacles


class Scene:
    def __init__(self, box=None, handle=None):
        self._L = N.lib()
        if handle is not None:
            self.handle = handle
        else:
            self.handle = C.c_void_p()
            N.check(self._L.ecm_scene_create(*[float(v) for v in box], C.byref(self.handle)))
        self.box = box

    def close(self):
        if self.handle:
            self._L.ecm_scene_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def AddObstacle(self, ring) -> None:
        r = np.ascontiguousarray(np.asarray(ring, np.float32).reshape(-1, 2))
        N.check(self._L.ecm_scene_add_obstacle(self.handle, N.fptr(r), r.shape[0]))

    def ComputeECM(self, resolution: float = 0.5, threads: int = 0) -> None:
        N.check(self._L.ecm_scene_build_ecm(self.handle, float(resolution), int(threads)))

    def ecm(self) -> Ecm:
        nv, ne = C.c_int(), C.c_int()
        N.check(self._L.ecm_scene_ecm_counts(self.handle, C.byref(nv), C.byref(ne)))
        v = np.zeros((nv.value, 2), np.float32)
        clr = np.zeros(nv.value, np.float32)
        t = np.zeros((ne.value, 2), np.int32)
        nx = np.zeros((ne.value, 2), np.int32)
        p = np.zeros((ne.value, 8), np.float32)
        N.check(self._L.ecm_scene_ecm_export(self.handle, N.fptr(v), N.fptr(clr), N.iptr(t), N.iptr(nx), N.fptr(p)))
        return Ecm(v, t, p, clr, nx)

    def obstacles(self) -> Obstacles:
        nr, nv = C.c_int(), C.c_int()
        N.check(self._L.ecm_scene_obstacle_counts(self.handle, C.byref(nr), C.byref(nv)))
        xy = np.zeros((max(1, nv.value), 2), np.float32)
        ro = np.zeros(nr.value + 1, np.int32)
        N.check(self._L.ecm_scene_obstacle_export(self.handle, N.fptr(xy), N.iptr(ro)))
        return Obstacles(xy[: nv.value], ro)

    def FindCell(self, x, y) -> int:
        return int(self._L.ecm_scene_locate(self.handle, float(x), float(y)))

    def InsideObstacle(self, x, y) -> bool:
        return bool(self._L.ecm_scene_inside_obstacle(self.handle, float(x), float(y)))

    def FindPath(self, start, goal, clearance):
        buf = np.zeros((4096, 2), np.float32)
        n = C.c_int()
        rc = self._L.ecm_scene_plan_path(self.handle, float(start[0]), float(start[1]), float(goal[0]),
                                         float(goal[1]), float(clearance), N.fptr(buf), 4096, C.byref(n))
        if rc == 1:
            return None
        N.check(rc)
        return buf[: n.value].copy()


def generate(config: int, seed: int = 1, n_agents: int = 0):
    """Seeded scenario (ecm_scenario_generate): returns (Scene, Agents, goals)."""
    L = N.lib()
    sh, ah = C.c_void_p(), C.c_void_p()
    N.check(L.ecm_scenario_generate(int(config), C.c_uint64(seed), int(n_agents), C.byref(sh), C.byref(ah)))
    half = {0: 100.0, 1: 500.0, 2: 1000.0}.get(config, 30.0)
    return _export(L, sh, ah, half)


def generate_box(half: float, n_rects: int, side_lo: int, side_hi: int, n_agents: int, seed: int):
    """A square scenario of the given size (ecm_scenario_generate_box)."""
    L = N.lib()
    sh, ah = C.c_void_p(), C.c_void_p()
    N.check(L.ecm_scenario_generate_box(float(half), int(n_rects), int(side_lo), int(side_hi), int(n_agents),
                                        C.c_uint64(seed), C.byref(sh), C.byref(ah)))
    return _export(L, sh, ah, float(half))


def _export(L, sh, ah, half):
    n, npts = C.c_int(), C.c_longlong()
    N.check(L.ecm_scenario_counts(ah, C.byref(n), C.byref(npts)))
    st = np.zeros((n.value, 2), np.float32)
    go = np.zeros((n.value, 2), np.float32)
    r = np.zeros(n.value, np.float32)
    s = np.zeros(n.value, np.float32)
    off = np.zeros(n.value, np.int32)
    cnt = np.zeros(n.value, np.int32)
    path = np.zeros((max(1, npts.value), 2), np.float32)
    N.check(L.ecm_scenario_export(ah, N.fptr(st), N.fptr(go), N.fptr(r), N.fptr(s), N.iptr(off), N.iptr(cnt),
                                  N.fptr(path)))
    L.ecm_scenario_destroy(ah)
    scene = Scene(box=(-half, -half, half, half), handle=sh)
    return scene, Agents(st, r, s, off, cnt, path[: npts.value]), go
