"""Synthetic scenes for the crowd step.

Obstacles are axis-aligned rectangles with CCW vertices, in the order the
reference's editor produces them (Simulator.cpp:404-407), corners snapped to
integer metres (the reference's ECM input is integral, ECMGenerator.cpp:37-43).
Agent radius ~ U[0.2, 0.3] m, preferred speed ~ N(1.3, 0.2) clipped to
[0.8, 1.8] m/s; all draws come from numpy's PCG64 seeded generator.
"""
from __future__ import annotations

import numpy as np

from .simulator import Agents, Ecm, Obstacles


def empty_square_ecm(x0: float, y0: float, x1: float, y1: float) -> Ecm:
    """ECM of an empty box: its medial axis, i.e. the four half-diagonals
    from the corners to the centre (exact for a square).  Vertex 0 is the
    centre, 1..4 the corners; edge e runs centre -> corner, its left feature
    is the boundary side met turning counter-clockwise."""
    cx, cy = (x0 + x1) / 2.0, (y0 + y1) / 2.0
    corners = [(x1, y1), (x0, y1), (x0, y0), (x1, y0)]
    verts = np.array([(cx, cy)] + corners, np.float32)
    tgt = []
    pts = []
    for i, (kx, ky) in enumerate(corners):
        # sides meeting at corner i: left (CCW) and right (CW) of centre->corner
        if i == 0:    # top-right: left = top (y=y1), right = right side (x=x1)
            cl0, cr0 = (cx, y1), (x1, cy)
        elif i == 1:  # top-left: left = left side (x=x0), right = top
            cl0, cr0 = (x0, cy), (cx, y1)
        elif i == 2:  # bottom-left: left = bottom, right = left side
            cl0, cr0 = (cx, y0), (x0, cy)
        else:         # bottom-right: left = right side, right = bottom
            cl0, cr0 = (x1, cy), (cx, y0)
        tgt.append((i + 1, 0))
        # he0: centre -> corner, closest points at the centre
        # he1: corner -> centre, closest points at the corner (both = corner)
        pts.append((cl0[0], cl0[1], cr0[0], cr0[1], kx, ky, kx, ky))
    return Ecm(verts, np.array(tgt, np.int32), np.array(pts, np.float32))


def rect_ring(cx: float, cy: float, hw: float, hh: float) -> np.ndarray:
    """CCW ring of an axis-aligned rectangle (Simulator.cpp:404-407)."""
    return np.array([(cx + hw, cy + hh), (cx - hw, cy + hh), (cx - hw, cy - hh), (cx + hw, cy - hh)], np.float32)


def random_rects(rng: np.random.Generator, box, n: int, side_lo: float, side_hi: float, gap: float = 2.0,
                 max_tries: int = 200000) -> list:
    """Non-overlapping integer-cornered rectangles inside `box`, at least
    `gap` apart from each other and from the boundary; (x0, y0, x1, y1) each."""
    bx0, by0, bx1, by1 = box
    out = []
    tries = 0
    while len(out) < n and tries < max_tries:
        tries += 1
        w = int(np.round(rng.uniform(side_lo, side_hi)))
        h = int(np.round(rng.uniform(side_lo, side_hi)))
        x0 = int(np.floor(rng.uniform(bx0 + gap, bx1 - gap - w)))
        y0 = int(np.floor(rng.uniform(by0 + gap, by1 - gap - h)))
        r = (x0, y0, x0 + w, y0 + h)
        if r[2] > bx1 - gap or r[3] > by1 - gap:
            continue
        if any(not (r[2] + gap <= q[0] or q[2] + gap <= r[0] or r[3] + gap <= q[1] or q[3] + gap <= r[1]) for q in out):
            continue
        out.append(r)
    return out


def rects_to_obstacles(rects) -> Obstacles:
    rings = [rect_ring((a + c) / 2.0, (b + d) / 2.0, (c - a) / 2.0, (d - b) / 2.0) for a, b, c, d in rects]
    if not rings:
        return Obstacles(np.zeros((0, 2), np.float32), np.zeros(1, np.int32))
    xy = np.concatenate(rings).astype(np.float32)
    off = np.arange(len(rings) + 1, dtype=np.int32) * 4
    return Obstacles(xy, off)


def inside_any(rects, x, y, margin=0.0):
    x = np.asarray(x)
    y = np.asarray(y)
    m = np.zeros(x.shape, bool)
    for a, b, c, d in rects:
        m |= (x > a - margin) & (x < c + margin) & (y > b - margin) & (y < d + margin)
    return m


def agent_draws(rng: np.random.Generator, n: int):
    r = rng.uniform(0.2, 0.3, n).astype(np.float32)
    s = np.clip(rng.normal(1.3, 0.2, n), 0.8, 1.8).astype(np.float32)
    return r, s


def straight_agents(rng: np.random.Generator, box, n: int, rects=(), margin: float = 1.0) -> Agents:
    """Agents with uniform start/goal in free space and straight two-point
    paths [start, goal] (what ECMPathPlanner::FindPath returns when start and
    goal retract onto the same ECM edge, ECMPathPlanner.cpp:144-150)."""
    bx0, by0, bx1, by1 = box
    pts = []
    while sum(len(p) for p in pts) < 2 * n:
        m = 4 * n
        x = rng.uniform(bx0 + margin, bx1 - margin, m)
        y = rng.uniform(by0 + margin, by1 - margin, m)
        ok = ~inside_any(rects, x, y, margin)
        pts.append(np.stack([x[ok], y[ok]], 1))
    allp = np.concatenate(pts)[: 2 * n].astype(np.float32)
    start, goal = allp[:n], allp[n:]
    r, s = agent_draws(rng, n)
    path = np.stack([start, goal], 1).reshape(-1, 2)
    return Agents(start, r, s, np.arange(n, dtype=np.int32) * 2, np.full(n, 2, np.int32), path)
