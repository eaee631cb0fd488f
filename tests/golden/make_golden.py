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
  (ORCA.cpp:70-333, the "cutoff line" and collision cases).

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


def main():
    data = {"irm": irm_cases(), "agent": agent_cases(), "obstacle": obstacle_cases()}
    # every agent branch of ORCA.cpp:356-422 is represented
    assert {c["branch"] for c in data["agent"]} == {"colliding", "base-circle", "left-leg", "right-leg"}
    with open(os.path.join(HERE, "closed_forms.json"), "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", os.path.join(HERE, "closed_forms.json"))


if __name__ == "__main__":
    main()
