// bucket_geom.h -- host-side geometry of the ECM cell buckets (no device
// code): whether a bucket square must list an ECM cell.
#pragma once

namespace ecm {

// Whether a point of the square [sx0, sx1] x [sy0, sy1] can lie within m of
// the convex hull of the quad q: false only when a supporting line of the
// hull (through two of its vertices, every vertex on one side) leaves the
// square grown by m strictly on the other side.  The ray-cast containment
// test (MathUtility::Contains, UtilityFunctions.cpp:54-86) finds a point
// farther than m from the hull outside the quad whatever its rounding: the
// crossings it counts lie on the quad's edges, at least m away in x, and
// the rounding of their abscissae is far below m.  Points on a vertex's
// exact y, where the test's open y intervals skip crossings, are covered by
// the exact-y lists instead.
inline bool hull_may_touch(const double (&q)[4][2], double sx0, double sy0, double sx1, double sy1, double m) {
    const double X[4] = {sx0 - m, sx1 + m, sx1 + m, sx0 - m};
    const double Y[4] = {sy0 - m, sy0 - m, sy1 + m, sy1 + m};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const double dx = q[j][0] - q[i][0], dy = q[j][1] - q[i][1];
            if (i == j || (dx == 0.0 && dy == 0.0)) continue;
            bool support = true;                   // every vertex left of or on i -> j
            for (int k = 0; k < 4; ++k)
                if (dx * (q[k][1] - q[i][1]) - dy * (q[k][0] - q[i][0]) < 0.0) support = false;
            if (!support) continue;
            bool apart = true;                     // every grown corner strictly right
            for (int c = 0; c < 4; ++c)
                if (dx * (Y[c] - q[i][1]) - dy * (X[c] - q[i][0]) >= 0.0) apart = false;
            if (apart) return false;
        }
    return true;
}

}  // namespace ecm
