// ecm_geom.h -- float geometry primitives for the crowd step, usable on the
// host and on gfx950.  Each routine keeps the reference's IEEE binary32
// expression order (compiled with -ffp-contract=off and correctly rounded
// division / square root), so host and device evaluate identically.
//
// Reference: ECMGenerator/ECMDataTypes.h:13-104 (Vec2, Point),
// ECMGenerator/UtilityFunctions.cpp:15-349 (MathUtility),
// ECMGenerator/Configuration.h:11-15 (EPSILON, MAX_FLOAT).
#pragma once

#include <hip/hip_runtime.h>

#define ECM_HD __host__ __device__ __forceinline__

namespace ecm {

constexpr float kEps = 0.0001f;                 // Configuration.h:14
constexpr float kMaxFloat = 3.40282347e+38f;    // Configuration.h:11

struct V2 {
    float x, y;
};

ECM_HD V2 mk(float x, float y) { V2 r; r.x = x; r.y = y; return r; }
ECM_HD V2 operator+(V2 a, V2 b) { return mk(a.x + b.x, a.y + b.y); }
ECM_HD V2 operator-(V2 a, V2 b) { return mk(a.x - b.x, a.y - b.y); }
ECM_HD V2 operator*(V2 a, float s) { return mk(a.x * s, a.y * s); }
ECM_HD V2 operator/(V2 a, float s) { return mk(a.x / s, a.y / s); }
ECM_HD V2 neg(V2 a) { return mk(a.x * -1.0f, a.y * -1.0f); }   // v * -1.0f

// MathUtility::Dot / Determinant / Right / Left (UtilityFunctions.cpp:172-231)
ECM_HD float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
ECM_HD float det(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
ECM_HD V2 right(V2 v) { return mk(v.y, -v.x); }
ECM_HD V2 left(V2 v) { return mk(-v.y, v.x); }

// std::min / std::max as the reference calls them on floats: (b < a) ? b : a
// and (a < b) ? b : a, so a NaN second argument leaves the first unchanged
ECM_HD float std_min(float a, float b) { return (b < a) ? b : a; }
ECM_HD float std_max(float a, float b) { return (a < b) ? b : a; }

// Vec2::Length / LengthSquared (ECMDataTypes.h:35-43)
ECM_HD float len(V2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
ECM_HD float len2(V2 v) { return v.x * v.x + v.y * v.y; }

// Vec2::Normalize: a zero vector is left unchanged (ECMDataTypes.h:45-52)
ECM_HD V2 normalize(V2 v) {
    float l = len(v);
    if (l == 0.0f) return v;
    return mk(v.x / l, v.y / l);
}
// Vec2::Normalized (ECMDataTypes.h:54-60); a zero vector gives (0, 0) here
// (the reference returns an uninitialised Vec2).
ECM_HD V2 normalized(V2 v) {
    float l = len(v);
    if (l == 0.0f) return mk(0.0f, 0.0f);
    return mk(v.x / l, v.y / l);
}

// Point::Approximate (ECMDataTypes.cpp:97-100)
ECM_HD bool approx(V2 p, V2 a) {
    return p.x < (a.x + kEps) && p.x > (a.x - kEps) && p.y < (a.y + kEps) && p.y > (a.y - kEps);
}

// MathUtility::SquareDistance(p1, p2), difference taken as p2 - p1
// (UtilityFunctions.cpp:32-39)
ECM_HD float sq_dist(V2 p1, V2 p2) {
    float dx = p2.x - p1.x, dy = p2.y - p1.y;
    return dx * dx + dy * dy;
}

// MathUtility::GetClosestPointOnSegment (UtilityFunctions.cpp:287-306)
ECM_HD V2 closest_on_segment(V2 p, V2 a, V2 b) {
    if (approx(a, b)) return a;
    V2 s = b - a;
    V2 q = p - a;
    float t2 = sq_dist(a, b);
    float u = (q.x * s.x + q.y * s.y) / t2;
    if (u > 1.0f) u = 1.0f;
    if (u < 0.0f) u = 0.0f;
    return mk(a.x + u * s.x, a.y + u * s.y);
}

// MathUtility::IsLeftOfSegment (UtilityFunctions.cpp:193-196)
ECM_HD bool left_of_segment(V2 s0, V2 s1, V2 p) {
    return (s1.x - s0.x) * (p.y - s0.y) - (s1.y - s0.y) * (p.x - s0.x) > 0;
}

// MathUtility::Contains for a 4-vertex ring q0->q1->q2->q3->q0, ray cast with
// the reference's vertex and vertical-edge rules (UtilityFunctions.cpp:54-86).
ECM_HD bool ring_edge_toggles(V2 p, V2 a, V2 b) {
    if (p.y > fminf(a.y, b.y) && p.y < fmaxf(a.y, b.y) && p.x < fmaxf(a.x, b.x)) {
        float xi = (p.y - a.y) * (b.x - a.x) / (b.y - a.y) + a.x;
        return a.x == b.x || p.x < xi;
    }
    return false;
}
ECM_HD bool quad_contains(V2 p, V2 q0, V2 q1, V2 q2, V2 q3) {
    if (approx(p, q0) || approx(p, q1) || approx(p, q2) || approx(p, q3)) return false;
    bool in = false;
    in ^= ring_edge_toggles(p, q0, q1);
    in ^= ring_edge_toggles(p, q1, q2);
    in ^= ring_edge_toggles(p, q2, q3);
    in ^= ring_edge_toggles(p, q3, q0);
    return in;
}

// MathUtility::GetRayToLineSegmentIntersection (UtilityFunctions.cpp:323-349).
// `abs(dot) < 0.000001` is the float abs (MSVC); for a float f,
// (double)f < 1e-6 holds exactly when f <= 1e-6f.
ECM_HD bool ray_segment(V2 o, V2 d, V2 p1, V2 p2, V2& out) {
    V2 v1 = o - p1;
    V2 v2 = p2 - p1;
    V2 v3 = mk(-d.y, d.x);
    float dd = dot(v2, v3);
    if (fabsf(dd) <= 1e-6f) return false;
    float t1 = det(v2, v1) / dd;
    float t2 = dot(v1, v3) / dd;
    if (t1 >= 0.0f && t2 >= 0.0f && t2 <= 1.0f) {
        out = mk(o.x + d.x * t1, o.y + d.y * t1);
        return true;
    }
    return false;
}

// ---------------------------------------------------------------- trig --
// atan after Cody & Waite (1980, ch. 11): |x| > 1 -> pi/2 - atan(1/x);
// |x| > 2 - sqrt(3) -> pi/6 + atan(((x(sqrt3-1) - 1/2) - 1/2 + x)/(sqrt3 + x));
// core: odd Taylor series to degree 15.  sin/cos: two-constant pi/2
// reduction, Taylor cores on [-pi/4, pi/4].  A fixed sequence of binary32
// operations, so host and device agree bit for bit.
ECM_HD float atan_f(float x) {
    float a = fabsf(x);
    bool inv = false, red = false;
    if (a > 1.0f) { a = 1.0f / a; inv = true; }
    if (a > 0.267949194f) {
        a = (((a * 0.732050808f - 0.5f) - 0.5f) + a) / (a + 1.73205081f);
        red = true;
    }
    float z = a * a;
    float p = 6.66666701e-02f;
    p = p * z - 7.69230798e-02f;
    p = p * z + 9.09090936e-02f;
    p = p * z - 1.11111112e-01f;
    p = p * z + 1.42857149e-01f;
    p = p * z - 2.00000003e-01f;
    p = p * z + 3.33333343e-01f;
    float r = a - a * (z * p);
    if (red) r = r + 5.23598790e-01f;
    if (inv) r = 1.57079637f - r;
    return x < 0.0f ? -r : r;
}

ECM_HD float sin_poly(float r) {
    float z = r * r;
    float p = -2.50521079e-08f;
    p = p * z + 2.75573188e-06f;
    p = p * z - 1.98412701e-04f;
    p = p * z + 8.33333377e-03f;
    p = p * z - 1.66666672e-01f;
    return r + r * (z * p);
}
ECM_HD float cos_poly(float r) {
    float z = r * r;
    float p = 2.08767559e-09f;
    p = p * z - 2.75573188e-07f;
    p = p * z + 2.48015876e-05f;
    p = p * z - 1.38888892e-03f;
    p = p * z + 4.16666679e-02f;
    p = p * z - 5.00000000e-01f;
    return 1.0f + z * p;
}
// returns (sin x, cos x)
ECM_HD V2 sincos_f(float x) {
    float k = rintf(x * 6.36619747e-01f);
    int q = (int)k;
    float r = (x - k * 1.57079637f) - k * (-4.37113883e-08f);
    float s = sin_poly(r), c = cos_poly(r);
    switch (q & 3) {
        case 0: return mk(s, c);
        case 1: return mk(c, -s);
        case 2: return mk(-s, -c);
        default: return mk(-c, s);
    }
}

// Uniform-grid cell of a coordinate: floor((v - v0) * inv) clamped to
// [0, side); NaN maps to 0.
ECM_HD int grid_cell(float v, float v0, float inv, int side) {
    float f = (v - v0) * inv;
    f = fmaxf(f, 0.0f);
    f = fminf(f, (float)(side - 1));
    return (int)f;
}

// MathUtility::RotateVector (UtilityFunctions.cpp:233-242)
ECM_HD V2 rotate(V2 v, float rad) {
    V2 sc = sincos_f(rad);
    float cs = sc.y, sn = sc.x;
    return mk(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
}

// rotate(v, rad) and rotate(v, -rad) from one sincos_f: sincos_f is odd /
// even in x bit for bit (rintf, the two-constant reduction and both
// polynomials are symmetric, and the quadrant swap maps q to -q), and
// v.y * -sn = -(v.y * sn), a - (-b) = a + b, (-a) + b = b - a exactly, so
// both results equal the two separate rotations bit for bit.
ECM_HD void rotate_pm(V2 v, float rad, V2& plus, V2& minus) {
    V2 sc = sincos_f(rad);
    float cs = sc.y, sn = sc.x;
    plus = mk(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
    minus = mk(v.x * cs + v.y * sn, v.y * cs - v.x * sn);
}

}  // namespace ecm
