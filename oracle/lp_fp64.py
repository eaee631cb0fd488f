"""FP64 re-solve of the 2D ORCA linear program (the convex problem the
incremental LP of ORCA.cpp:428-587 solves): min |v - pref|^2 subject to
every half-plane n.(v - p) >= 0 and |v| <= maxSpeed.

TEST INFRASTRUCTURE ONLY: the checker of the LP results in tests/.
"""
import math

import numpy as np


def qp_fp64(cons, pref, vmax):
    """min |v - pref|^2 s.t. n.(v - p) >= 0, |v| <= vmax, by enumerating the
    active sets of size <= 2 (lines and the speed circle)."""
    cons = np.asarray(cons, np.float64).reshape(-1, 4)
    N, P = cons[:, :2], cons[:, 2:]
    pref = np.asarray(pref, np.float64)

    def feas(v, tol=1e-7):
        return np.linalg.norm(v) <= vmax + tol and np.all(np.einsum("ij,ij->i", N, v - P) >= -tol)

    cands = []
    q = pref if np.linalg.norm(pref) <= vmax else pref / np.linalg.norm(pref) * vmax
    cands.append(q)
    for i in range(len(cons)):
        n, p = N[i], P[i]
        d = np.array([n[1], -n[0]])
        t = d @ (pref - p)
        cands.append(p + d * t)
        b = d @ p
        disc = b * b + vmax * vmax - p @ p
        if disc >= 0:
            for s in (-b - math.sqrt(disc), -b + math.sqrt(disc)):
                cands.append(p + d * s)
        for j in range(i):
            A = np.array([N[i], N[j]])
            if abs(np.linalg.det(A)) < 1e-12:
                continue
            cands.append(np.linalg.solve(A, np.array([N[i] @ P[i], N[j] @ P[j]])))
    feasible = [c for c in cands if feas(c)]
    if not feasible:
        return None
    return min(feasible, key=lambda c: np.linalg.norm(c - pref))
