"""Writes tests/golden/*.json: FP64 closed-form expected values for the
oracle's known-answer tests.  The reference itself cannot run here (it needs
Boost.Polygon, which is absent), so every expected value below is derived
from the geometry that the reference's functions compute, in double
precision, independently of the oracle's float code:

* retraction / clearance / IRM on the ECM of the empty square [-L, L]^2, whose
  medial axis is the two diagonals (ECM.cpp:20-96, IRMPathFollower.cpp:14-115);
* ORCA agent half-planes for configurations whose answer has a closed form
  (ORCA.cpp:339-423);
* obstacle half-planes for an agent facing one side of a CCW square
  (ORCA.cpp:70-333, the "cutoff line" and collision cases);
* every branch of the obstacle half-plane (ORCA.cpp:91-332) from the
  geometry of the truncated velocity obstacle (obstacle_model), on a convex
  square, an L-shaped hexagon with a concave corner and a triangle with an
  obtuse corner.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def irm_cases():
    """Agent in the top region (y > |x|) of the empty square [-L, L]^2: the
    retraction ray is perpendicular to the top side, straight down onto the
    diagonal, so the retracted point is (x, |x|); the clearance (distance to
    the edge's left feature) is L - |x| by symmetry.  The attraction point is
    the intersection of the clearance circle with the path furthest along it
    (goal outside the circle)."""
    L = 50.0
    out = []
    for (x, y, path) in [
        (3.0, 20.0, [(-45.0, -40.0), (48.0, 46.0)]),
        (-7.5, 30.0, [(-45.0, 45.0), (-10.0, -48.0)]),
        (10.0, 40.0, [(-48.0, 20.0), (0.0, 20.0), (48.0, -48.0)]),
    ]:
        rx, ry = x, abs(x)
        R = L - ry
        gx, gy = path[-1]
        assert (gx - rx) ** 2 + (gy - ry) ** 2 >= R * R
        best = None
        for (ax, ay), (bx, by) in zip(path[:-1], path[1:]):
            dx, dy = bx - ax, by - ay
            fx, fy = ax - rx, ay - ry
            a = dx * dx + dy * dy
            b = 2 * (fx * dx + fy * dy)
            c = fx * fx + fy * fy - R * R
            disc = b * b - 4 * a * c
            if disc < 0:
                continue
            for t in ((-b - math.sqrt(disc)) / (2 * a), (-b + math.sqrt(disc)) / (2 * a)):
                if 0 <= t <= 1:
                    best = (ax + t * dx, ay + t * dy)   # later segments win
        out.append({"L": L, "pos": [x, y], "retract": [rx, ry], "clearance": R, "path": path,
                    "attraction": list(best)})
    return out


def agent_halfplane(pa, va, pb, vb, ra, rb, T, dt):
    """ORCA.cpp:339-423 in FP64: branch chosen by the reference's rules
    (colliding; IsLeftOfVector(leftLegTip - VOPos, relVel - VOPos) for the
    base circle; Dot(Left(VOPos)/|VOPos|, relVel) >= 0 for the left leg)."""
    vop = ((pb[0] - pa[0]) / T, (pb[1] - pa[1]) / T)
    R = (ra + rb) / T
    rv = (va[0] - vb[0], va[1] - vb[1])
    rp = (pb[0] - pa[0], pb[1] - pa[1])
    if math.hypot(*rp) < ra + rb:
        w = (rv[0] - rp[0] / dt, rv[1] - rp[1] / dt)
        lw = math.hypot(*w)
        uw = (w[0] / lw, w[1] / lw)
        U = (uw[0] * ((ra + rb) / dt - lw), uw[1] * ((ra + rb) / dt - lw))
        return uw, (va[0] + U[0] * 0.5, va[1] + U[1] * 0.5), "colliding"
    lv = math.hypot(*vop)
    th = math.atan(R / lv)
    rot = lambda v, a: (v[0] * math.cos(a) - v[1] * math.sin(a), v[0] * math.sin(a) + v[1] * math.cos(a))
    lt, rt = rot(vop, th), rot(vop, -th)
    b = (lt[0] - vop[0], lt[1] - vop[1])
    w = (rv[0] - vop[0], rv[1] - vop[1])
    if b[0] * w[1] - b[1] * w[0] > 0:
        lw = math.hypot(*w)
        n = (w[0] / lw, w[1] / lw)
        return n, (va[0] + n[0] * (R - lw) * 0.5, va[1] + n[1] * (R - lw) * 0.5), "base-circle"
    left_side = (-vop[1] / lv) * rv[0] + (vop[0] / lv) * rv[1] >= 0
    leg = lt if left_side else rt
    ll = math.hypot(*leg)
    d = (leg[0] / ll, leg[1] / ll)
    s = rv[0] * d[0] + rv[1] * d[1]
    U = (d[0] * s - rv[0], d[1] * s - rv[1])
    n = (-d[1], d[0]) if left_side else (d[1], -d[0])
    return n, (va[0] + U[0] * 0.5, va[1] + U[1] * 0.5), "left-leg" if left_side else "right-leg"


def agent_cases():
    out = []
    for (pa, va, pb, vb, ra, rb, T, dt) in [
        ((0.0, 0.0), (1.0, 0.0), (5.0, 0.0), (-1.0, 0.0), 0.5, 0.5, 2.0, 0.1),
        ((0.0, 0.0), (1.2, 0.0), (4.0, 0.0), (-1.0, 0.0), 0.3, 0.25, 2.0, 0.1),
        ((0.0, 0.0), (1.0, 0.8), (6.0, 0.0), (0.0, 0.0), 0.3, 0.3, 2.0, 0.1),
        ((0.0, 0.0), (3.2, 0.5), (6.0, 0.0), (0.0, 0.0), 0.3, 0.3, 2.0, 0.1),
        ((0.0, 0.0), (4.0, -0.3), (6.0, 0.0), (0.0, 0.0), 0.3, 0.3, 2.0, 0.1),
        ((0.0, 0.0), (0.5, 0.0), (0.4, 0.0), (-0.5, 0.0), 0.3, 0.3, 2.0, 0.1),
    ]:
        n, p, br = agent_halfplane(pa, va, pb, vb, ra, rb, T, dt)
        out.append({"pa": pa, "va": va, "pb": pb, "vb": vb, "ra": ra, "rb": rb, "T": T, "dt": dt,
                    "normal": n, "point": p, "branch": br})
    return out


def obstacle_cases():
    out = []
    # CCW square [-1, 1]^2 (vertices (1,1), (-1,1), (-1,-1), (1,-1)); an agent
    # at rest below the bottom side (edge (-1,-1) -> (1,-1)), far enough that
    # no collision case applies.  Only that edge faces the agent; the
    # velocity projects onto the cutoff line, so the normal is
    # Left(-seg) = Left((-1, 0)) = (0, -1) and the point is
    # leftCutoff + normal * r / T with leftCutoff = (L - pos) / T: the agent
    # may approach at most (distance - r) / T.
    for (pos, r, T) in [((0.0, -3.0), 0.25, 2.0), ((0.2, -2.5), 0.3, 2.0)]:
        Lx, Ly = -1.0, -1.0
        lc = ((Lx - pos[0]) / T, (Ly - pos[1]) / T)
        n = (0.0, -1.0)
        p = (lc[0] + n[0] / T * r, lc[1] + n[1] / T * r)
        out.append({"pos": pos, "vel": (0.0, 0.0), "r": r, "T": T, "normal": n, "point": p, "case": "cutoff-line"})
    # collision case 3: agent overlapping the bottom side's interior:
    # constraint through the origin with normal Right(seg) = (0, -1)
    out.append({"pos": (0.3, -1.1), "vel": (0.0, 0.0), "r": 0.25, "T": 2.0, "normal": (0.0, -1.0),
                "point": (0.0, 0.0), "case": "collision-3"})
    return out


# ------------------------------------------------- obstacle branch geometry
# Polygons (CCW) for the obstacle branch cases: a convex square, an L-shaped
# hexagon with one concave corner at (1, 1), and a triangle with an obtuse
# convex corner at (2, 1) and acute ones (collision case 2 emits only at an
# acute corner: past the end vertex on the outer side, the agent must lie on
# the far side of the next edge's direction).
SHAPES = {
    "square": [(1.0, 1.0), (-1.0, 1.0), (-1.0, -1.0), (1.0, -1.0)],
    "ell": [(0.0, 0.0), (2.0, 0.0), (2.0, 1.0), (1.0, 1.0), (1.0, 2.0), (0.0, 2.0)],
    "obtuse": [(0.0, 0.0), (4.0, 0.0), (2.0, 1.0)],
}


def _sub(a, b): return (a[0] - b[0], a[1] - b[1])
def _add(a, b): return (a[0] + b[0], a[1] + b[1])
def _mul(a, k): return (a[0] * k, a[1] * k)
def _dot(a, b): return a[0] * b[0] + a[1] * b[1]
def _cross(a, b): return a[0] * b[1] - a[1] * b[0]
def _norm(a): return math.hypot(a[0], a[1])
def _unit(a): return _mul(a, 1.0 / _norm(a))
def _left(a): return (-a[1], a[0])
def _right(a): return (a[1], -a[0])
def _rot(a, t): return (a[0] * math.cos(t) - a[1] * math.sin(t), a[0] * math.sin(t) + a[1] * math.cos(t))


def _seg_dist(q, a, b):
    ab = _sub(b, a)
    t = max(0.0, min(1.0, _dot(_sub(q, a), ab) / _dot(ab, ab)))
    return _norm(_sub(q, _add(a, _mul(ab, t))))


def _ray_dist(q, a, d):
    t = max(0.0, _dot(_sub(q, a), d))
    return _norm(_sub(q, _add(a, _mul(d, t))))


def obstacle_model(ring, e, p, v, r, T):
    """The obstacle half-plane of edge e (ring[e] -> ring[e + 1]) for an agent
    at p with velocity v, radius r, horizon T, from the geometry of the
    truncated velocity obstacle, in FP64:
    * collision: the agent's disc meets the segment's start vertex (before
      the segment), end vertex (past it) or interior; the half-plane then
      passes through the origin with its normal pointing from the obstacle
      to the agent (from the vertex, or the edge's outward normal);
    * otherwise the velocity obstacle is bounded by legs from the origin
      tangent to the discs of radius r at the vertices it sees (one vertex
      when the agent's disc reaches behind the edge's line beyond an end),
      a leg lying along the edge where the vertex is concave, a leg turned
      onto a neighbouring edge that it would cut into (then that edge is
      responsible and this one emits nothing there), and, at horizon T,
      the discs of radius r / T around (vertex - p) / T joined by the
      segment's offset copy (the cutoff line).  That set is everything
      within r / T of a skeleton (the cutoff centres, the segment between
      them, the leg rays from them), and v is projected onto the boundary
      part around the skeleton point nearest to v: a cutoff circle when that
      point is a centre, else the cutoff line or a leg.
    Returns the branch names (as the oracle reports them), whether a
    half-plane is emitted, its unit normal n and offset n . point (the
    half-plane is {u : n . u >= offset}), and the distance margins that make
    the case unambiguous."""
    m = len(ring)
    def vert(i): return ring[i % m]
    def convex(i): return _cross(_sub(vert(i), vert(i - 1)), _sub(vert(i + 1), vert(i))) >= 0.0
    L, R = vert(e), vert(e + 1)
    seg = _sub(R, L)
    su = _unit(seg)
    s = _dot(_sub(p, L), seg) / _dot(seg, seg)
    dline = abs(_cross(su, _sub(p, L)))
    d1, d2 = _norm(_sub(p, L)), _norm(_sub(p, R))
    side = _cross(seg, _sub(p, L))                 # < 0: right of L -> R, the side the filter keeps
    margins = [abs(s), abs(s - 1.0), abs(dline - r), abs(d1 - r), abs(d2 - r)]
    out = {"margins": margins, "right_of_edge": side < 0.0}
    if s < 0 and d1 <= r:
        if convex(e):
            return dict(out, branches={"case1"}, emitted=True, normal=_unit(_sub(p, L)), offset=0.0)
        return dict(out, branches={"case1_skip"}, emitted=False)
    if s > 1 and d2 <= r:
        nd = _unit(_sub(vert(e + 2), R))
        margins.append(abs(_cross(_sub(R, p), nd)))
        if convex(e + 1) and _cross(_sub(R, p), nd) >= 0.0:
            return dict(out, branches={"case2"}, emitted=True, normal=_unit(_sub(p, R)), offset=0.0)
        return dict(out, branches={"case2_skip"}, emitted=False)
    if 0 <= s < 1 and dline <= r:
        return dict(out, branches={"case3"}, emitted=True, normal=_right(su), offset=0.0)
    br = set()
    if s < 0 and dline <= r:                      # the start vertex alone
        if not convex(e):
            return dict(out, branches={"vertex_l_skip"}, emitted=False)
        br.add("vertex_l")
        A, prevA, ca = L, vert(e - 1), True
        B, nextB, cb = L, R, True
    elif s > 1 and dline <= r:                    # the end vertex alone
        if not convex(e + 1):
            return dict(out, branches={"vertex_r_skip"}, emitted=False)
        br.add("vertex_r")
        A, prevA, ca = R, L, True
        B, nextB, cb = R, vert(e + 2), True
    else:
        A, prevA, ca = L, vert(e - 1), convex(e)
        B, nextB, cb = R, vert(e + 2), convex(e + 1)
    single = A == B

    def tangent(V, sign):
        rel = _sub(V, p)
        return _rot(_unit(rel), sign * math.asin(r / _norm(rel)))
    if ca:
        legL = tangent(A, 1.0)
    else:
        legL = _mul(su, -1.0)
        br.add("concave_l")
    if cb:
        legR = tangent(B, -1.0)
    else:
        legR = su
        br.add("concave_r")
    forL = forR = False
    u = _unit(_sub(A, prevA))
    if ca:
        margins.append(abs(_cross(legL, _mul(u, -1.0))))
        if _cross(legL, _mul(u, -1.0)) >= 0.0:
            legL, forL = _mul(u, -1.0), True
            br.add("foreign_l")
    w = _unit(_sub(nextB, B))
    if cb:
        margins.append(abs(_cross(legR, w)))
        if _cross(legR, w) <= 0.0:
            legR, forR = w, True
            br.add("foreign_r")
    rho = r / T
    cA, cB = _mul(_sub(A, p), 1.0 / T), _mul(_sub(B, p), 1.0 / T)
    nseg = _right(su)
    # The truncated velocity obstacle is the set of points within rho of its
    # skeleton: the cutoff centres cA, cB, the segment between them and the
    # leg rays from cA along legL and from cB along legR.  v is projected onto
    # the boundary part around the skeleton point nearest to v.
    def proj_seg(a, b):
        ab = _sub(b, a)
        return _dot(_sub(v, a), ab) / _dot(ab, ab)
    def proj_ray(a, d):
        return _dot(_sub(v, a), d)
    pieces = []                                   # (name, distance, raw parameter, clamped-at-start)
    tL = proj_ray(cA, legL)
    pieces.append(("leg_l", _norm(_sub(v, _add(cA, _mul(legL, max(0.0, tL))))), tL))
    tR = proj_ray(cB, legR)
    pieces.append(("leg_r", _norm(_sub(v, _add(cB, _mul(legR, max(0.0, tR))))), tR))
    if not single:
        ts = proj_seg(cA, cB)
        q = _add(cA, _mul(_sub(cB, cA), min(1.0, max(0.0, ts))))
        pieces.append(("cutoff_line", _norm(_sub(v, q)), ts))
    D = min(d for _, d, _ in pieces)
    # where the nearest skeleton point lies
    interior = [n for n, d, t in pieces
                if d <= D + 1e-12 and (t > 0.0 if n != "cutoff_line" else 0.0 < t < 1.0)]
    if interior:
        best = interior[0]
        par = dict((n, t) for n, _, t in pieces)[best]
        margins.append(par if best != "cutoff_line" else min(par, 1.0 - par))
        others = [d for n, d, _ in pieces if n != best]
        out["feature_gap"] = min(others) - D if others else 1.0
    else:
        # an end point: cA when the projections onto everything around it fall
        # before it, else cB
        at_a = _norm(_sub(v, cA)) <= D + 1e-12
        best = "circle_l" if at_a else "circle_r"
        margins.extend(abs(t) if n != "cutoff_line" else (abs(t) if at_a else abs(t - 1.0)) for n, _, t in pieces)
        out["feature_gap"] = 1.0
    if best == "leg_l" and forL:
        return dict(out, branches=br | {"leg_l_skip"}, emitted=False)
    if best == "leg_r" and forR:
        return dict(out, branches=br | {"leg_r_skip"}, emitted=False)
    if best == "circle_l" or best == "circle_r":
        c = cA if best == "circle_l" else cB
        n = _unit(_sub(v, c))
        off = _dot(n, c) + rho
    elif best == "cutoff_line":
        n, off = nseg, _dot(nseg, cA) + rho
    elif best == "leg_l":
        n = _left(legL)
        off = _dot(n, cA) + rho
    else:
        n = _right(legR)
        off = _dot(n, cB) + rho
    return dict(out, branches=br | {best}, emitted=True, normal=n, offset=off)


# (shape, edge, required branch names) for the cases; every branch of the
# obstacle half-plane is required by at least one of them
BRANCH_TARGETS = [
    ("square", 2, {"case1"}), ("ell", 3, {"case1_skip"}), ("obtuse", 0, {"case2"}), ("square", 2, {"case2_skip"}),
    ("square", 2, {"case3"}),
    ("square", 2, {"vertex_l", "circle_l"}), ("square", 2, {"vertex_l", "leg_r"}),
    ("square", 2, {"vertex_r", "leg_l"}), ("obtuse", 0, {"vertex_r", "circle_l"}),
    ("ell", 3, {"vertex_l_skip"}), ("ell", 2, {"vertex_r_skip"}),
    ("ell", 3, {"concave_l", "cutoff_line"}), ("ell", 3, {"concave_l", "leg_r"}),
    ("ell", 2, {"concave_r", "cutoff_line"}), ("ell", 2, {"concave_r", "leg_l"}),
    ("square", 2, {"foreign_l", "leg_l_skip"}), ("square", 2, {"foreign_r", "leg_r_skip"}),
    ("square", 2, {"circle_l"}), ("square", 2, {"circle_r"}), ("square", 2, {"cutoff_line"}),
    ("obtuse", 0, {"leg_l"}), ("obtuse", 0, {"leg_r"}), ("square", 2, {"foreign_l", "cutoff_line"}),
]


def obstacle_branch_cases(per_target=3):
    """Seeded search for unambiguous configurations of every target: agent
    positions within 3 m of the edge and on its right (outer) side,
    velocities up to 2 m/s, radius in [0.2, 0.3], T 2 s.  A case is kept only when every decision is at least
    1e-3 (and 2e-3 m/s between boundary features) from its threshold, so
    float rounding cannot flip it."""
    import random
    rnd = random.Random(7)
    cases = []
    for shape, e, want in BRANCH_TARGETS:
        ring = SHAPES[shape]
        L, R = ring[e], ring[(e + 1) % len(ring)]
        cx, cy = (L[0] + R[0]) / 2, (L[1] + R[1]) / 2
        found = 0
        for _ in range(2_000_000):
            near = rnd.random() < 0.5
            a, b = (L, R) if rnd.random() < 0.5 else (R, L)
            if near:                                     # close to a vertex
                p = (a[0] + rnd.uniform(-0.7, 0.7), a[1] + rnd.uniform(-0.7, 0.7))
            else:
                p = (cx + rnd.uniform(-3.0, 3.0), cy + rnd.uniform(-3.0, 3.0))
            v = (rnd.uniform(-2.0, 2.0), rnd.uniform(-2.0, 2.0))
            r = rnd.uniform(0.2, 0.3)
            p = (round(p[0], 3), round(p[1], 3))
            v = (round(v[0], 3), round(v[1], 3))
            r = round(r, 3)
            try:
                res = obstacle_model(ring, e, p, v, r, 2.0)
            except (ValueError, ZeroDivisionError):
                continue
            # the obstacle filter (Simulator.cpp:259-292) passes only edges
            # that have the agent strictly on their right
            if not res["right_of_edge"] or not want <= res["branches"]:
                continue
            if min(res["margins"]) < 1e-3 or res.get("feature_gap", 1.0) < 2e-3:
                continue
            case = {"shape": shape, "edge": e, "pos": p, "vel": v, "r": r, "T": 2.0,
                    "branches": sorted(res["branches"]), "emitted": res["emitted"]}
            if res["emitted"]:
                case["normal"] = list(res["normal"])
                case["offset"] = res["offset"]
            cases.append(case)
            found += 1
            if found == per_target:
                break
        assert found == per_target, (shape, e, want)
    return cases


def main():
    data = {"irm": irm_cases(), "agent": agent_cases(), "obstacle": obstacle_cases(),
            "shapes": SHAPES, "obstacle_branches": obstacle_branch_cases()}
    # every agent branch of ORCA.cpp:356-422 is represented
    assert {c["branch"] for c in data["agent"]} == {"colliding", "base-circle", "left-leg", "right-leg"}
    # ... and every obstacle branch of ORCA.cpp:91-332
    hit = set().union(*(set(c["branches"]) for c in data["obstacle_branches"]))
    assert hit == {"case1", "case1_skip", "case2", "case2_skip", "case3", "vertex_l", "vertex_l_skip", "vertex_r",
                   "vertex_r_skip", "concave_l", "concave_r", "foreign_l", "foreign_r", "circle_l", "circle_r",
                   "cutoff_line", "leg_l", "leg_l_skip", "leg_r", "leg_r_skip"}, hit
    with open(os.path.join(HERE, "closed_forms.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(HERE, "closed_forms.json"))


if __name__ == "__main__":
    main()
