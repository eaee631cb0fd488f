/*
 * ecm_oracle.c -- CPU oracle (test infrastructure only; see ecm_oracle.h).
 *
 * Plain C11 restatement of the reference per-step crowd update.  Every
 * function names the reference file:line it follows.  Expression order is
 * kept as written in the reference so that an IEEE-754 binary32 evaluation
 * without contraction (-ffp-contract=off) is reproducible bit for bit.
 *
 * Deliberate, documented choices where the reference is undefined or
 * platform dependent:
 *   - abs(dot) in GetRayToLineSegmentIntersection (UtilityFunctions.cpp:330)
 *     is the float abs, as MSVC (the reference's compiler) resolves it:
 *     (double)fabsf(dot) < 1e-6  <=>  fabsf(dot) <= 1e-6f.
 *   - std::powf(x, 2.0f) (Simulator.cpp:277) is x * x.
 *   - Vec2::Normalized() of a zero vector returns an uninitialised Vec2
 *     (ECMDataTypes.h:54-60); here it returns (0, 0).
 *   - atan/cos/sin: the oracle's own routines (below), pinned against libm.
 *   - KD-tree std::sort ties (KDTree.cpp:67,71) are ordered by slot id.
 */
#include "ecm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS 0.0001f                     /* Configuration.h:14 */
#define MAXF 3.40282347e+38f            /* Configuration.h:11 */

typedef struct { float x, y; } v2;

static inline v2 V(float x, float y) { v2 r; r.x = x; r.y = y; return r; }
/* Vec2/Point operators, ECMDataTypes.h:20-33, 84-98 */
static inline v2 vsub(v2 a, v2 b) { return V(a.x - b.x, a.y - b.y); }
static inline v2 vadd(v2 a, v2 b) { return V(a.x + b.x, a.y + b.y); }
static inline v2 vmul(v2 a, float s) { return V(a.x * s, a.y * s); }
static inline v2 vdiv(v2 a, float s) { return V(a.x / s, a.y / s); }
/* MathUtility::Dot / Determinant / Right / Left, UtilityFunctions.cpp:172-231 */
static inline float dot(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
static inline float det(v2 a, v2 b) { return a.x * b.y - a.y * b.x; }
static inline v2 right(v2 v) { return V(v.y, -v.x); }
static inline v2 left(v2 v) { return V(-v.y, v.x); }
/* Vec2::Length / LengthSquared (ECMDataTypes.h:35-43); sqrt of a float sum,
 * correctly rounded either way. */
static inline float vlen(v2 v) { return sqrtf(v.x * v.x + v.y * v.y); }
static inline float vlen2(v2 v) { return v.x * v.x + v.y * v.y; }
/* Vec2::Normalize (ECMDataTypes.h:45-52): zero vector left unchanged */
static inline v2 vnormalize(v2 v) {
    float l = vlen(v);
    if (l == 0.0f) return v;
    v.x /= l; v.y /= l;
    return v;
}
/* Vec2::Normalized (ECMDataTypes.h:54-60) */
static inline v2 vnormalized(v2 v) {
    float l = vlen(v);
    if (l == 0.0f) return V(0.0f, 0.0f);
    return V(v.x / l, v.y / l);
}
/* Point::Approximate (ECMDataTypes.cpp:97-100) */
static inline int approx(v2 p, v2 a) {
    return p.x < (a.x + EPS) && p.x > (a.x - EPS) && p.y < (a.y + EPS) && p.y > (a.y - EPS);
}
/* MathUtility::SquareDistance(p1, p2): diff = p2 - p1 (UtilityFunctions.cpp:32-39) */
static inline float sqdist(v2 p1, v2 p2) {
    v2 d = V(p2.x - p1.x, p2.y - p1.y);
    return d.x * d.x + d.y * d.y;
}

/* ------------------------------------------------------------------ trig */
/* atan: Cody & Waite, "Software Manual for the Elementary Functions" (1980),
 * ch. 11: |x| > 1 -> pi/2 - atan(1/x); |x| > 2 - sqrt(3) ->
 * pi/6 + atan((sqrt(3) x - 1) / (sqrt(3) + x)), with the numerator written
 * ((x (sqrt3 - 1) - 1/2) - 1/2) + x to limit cancellation; core: the odd
 * Taylor series of atan to degree 15 on |z| <= 0.268. */
float oracle_atanf(float x) {
    float a = fabsf(x);
    int inv = 0, red = 0;
    if (a > 1.0f) { a = 1.0f / a; inv = 1; }
    if (a > 0.267949194f) {
        a = (((a * 0.732050808f - 0.5f) - 0.5f) + a) / (a + 1.73205081f);
        red = 1;
    }
    float z = a * a;
    float p = 6.66666701e-02f;          /* 1/15  */
    p = p * z - 7.69230798e-02f;        /* 1/13  */
    p = p * z + 9.09090936e-02f;        /* 1/11  */
    p = p * z - 1.11111112e-01f;        /* 1/9   */
    p = p * z + 1.42857149e-01f;        /* 1/7   */
    p = p * z - 2.00000003e-01f;        /* 1/5   */
    p = p * z + 3.33333343e-01f;        /* 1/3   */
    float r = a - a * (z * p);
    if (red) r = r + 5.23598790e-01f;   /* pi/6  */
    if (inv) r = 1.57079637f - r;       /* pi/2  */
    return x < 0.0f ? -r : r;
}

/* sin/cos: Cody-Waite two-constant reduction by pi/2, Taylor cores on
 * [-pi/4, pi/4] (sin to degree 11, cos to degree 12). */
static float sin_core(float r) {
    float z = r * r;
    float p = -2.50521079e-08f;         /* -1/11! */
    p = p * z + 2.75573188e-06f;        /*  1/9!  */
    p = p * z - 1.98412701e-04f;        /* -1/7!  */
    p = p * z + 8.33333377e-03f;        /*  1/5!  */
    p = p * z - 1.66666672e-01f;        /* -1/3!  */
    return r + r * (z * p);
}
static float cos_core(float r) {
    float z = r * r;
    float p = 2.08767559e-09f;          /*  1/12! */
    p = p * z - 2.75573188e-07f;        /* -1/10! */
    p = p * z + 2.48015876e-05f;        /*  1/8!  */
    p = p * z - 1.38888892e-03f;        /* -1/6!  */
    p = p * z + 4.16666679e-02f;        /*  1/4!  */
    p = p * z - 5.00000000e-01f;        /* -1/2!  */
    return 1.0f + z * p;
}
static float reduce_pio2(float x, int* q) {
    float k = rintf(x * 6.36619747e-01f);           /* 2/pi */
    *q = (int)k;
    return (x - k * 1.57079637f) - k * (-4.37113883e-08f);
}
float oracle_sinf(float x) {
    int q;
    float r = reduce_pio2(x, &q);
    switch (q & 3) {
        case 0: return sin_core(r);
        case 1: return cos_core(r);
        case 2: return -sin_core(r);
        default: return -cos_core(r);
    }
}
float oracle_cosf(float x) {
    int q;
    float r = reduce_pio2(x, &q);
    switch (q & 3) {
        case 0: return cos_core(r);
        case 1: return -sin_core(r);
        case 2: return -cos_core(r);
        default: return sin_core(r);
    }
}

/* MathUtility::RotateVector (UtilityFunctions.cpp:233-242) */
static inline v2 rotate(v2 v, float rad) {
    float cs = oracle_cosf(rad);
    float sn = oracle_sinf(rad);
    float px = v.x * cs - v.y * sn;
    float py = v.x * sn + v.y * cs;
    return V(px, py);
}

/* MathUtility::GetClosestPointOnSegment (UtilityFunctions.cpp:287-306) */
static v2 closest_on_segment(v2 p, v2 s1, v2 s2) {
    if (approx(s1, s2)) return s1;
    v2 seg = vsub(s2, s1);
    v2 ps = vsub(p, s1);
    float t2 = sqdist(s1, s2);
    float d = (ps.x * seg.x + ps.y * seg.y) / t2;
    if (d > 1.0f) d = 1.0f;
    if (d < 0.0f) d = 0.0f;
    return V(s1.x + d * seg.x, s1.y + d * seg.y);
}

/* MathUtility::Contains(p, polygon) (UtilityFunctions.cpp:54-86), for the
 * closed 4-vertex ring q[0] -> q[1] -> q[2] -> q[3] -> q[0]. */
static int contains4(v2 p, const v2* q) {
    int inside = 0;
    for (int i = 0; i < 4; ++i) {
        v2 a = q[i], b = q[(i + 1) & 3];
        if (approx(p, a)) return 0;
        if (approx(p, b)) return 0;
        if (p.y > fminf(a.y, b.y)) {
            if (p.y < fmaxf(a.y, b.y)) {
                if (p.x < fmaxf(a.x, b.x)) {
                    float xi = (p.y - a.y) * (b.x - a.x) / (b.y - a.y) + a.x;
                    if (a.x == b.x || p.x < xi) inside = !inside;
                }
            }
        }
    }
    return inside;
}

/* MathUtility::IsLeftOfSegment (UtilityFunctions.cpp:193-196) */
static inline int is_left_of_segment(v2 s0, v2 s1, v2 p) {
    return (s1.x - s0.x) * (p.y - s0.y) - (s1.y - s0.y) * (p.x - s0.x) > 0;
}

/* MathUtility::GetRayToLineSegmentIntersection (UtilityFunctions.cpp:323-349) */
static int ray_segment(v2 o, v2 d, v2 p1, v2 p2, v2* out) {
    v2 v1 = vsub(o, p1);
    v2 vv = vsub(p2, p1);
    v2 v3 = V(-d.y, d.x);
    float dt = dot(vv, v3);
    if (fabsf(dt) <= 1e-6f) return 0;     /* float abs < 0.000001 (double) */
    float t1 = det(vv, v1) / dt;
    float t2 = dot(v1, v3) / dt;
    if (t1 >= 0.0f && (t2 >= 0.0f && t2 <= 1.0f)) {
        out->x = o.x + d.x * t1;
        out->y = o.y + d.y * t1;
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------- ECM */
static inline v2 evert(const oracle_ecm* e, int i) { return V(e->vert_xy[2 * i], e->vert_xy[2 * i + 1]); }
static inline v2 hept(const oracle_ecm* e, int edge, int which) {
    /* which: 0 he0.cl, 1 he0.cr, 2 he1.cl, 3 he1.cr */
    return V(e->he_pts[8 * edge + 2 * which], e->he_pts[8 * edge + 2 * which + 1]);
}

/* ECMCellCollection::Construct (ECMCellCollection.cpp:18-46) builds two cells
 * per edge: left (he0.cl, he1.cr) and right (he0.cr, he1.cl).
 * PointLocationQueryLinear (:57-90) tests the ring
 * (v(he1.target), boundary.p0, boundary.p1, v(he0.target)). */
int oracle_locate_cell(const oracle_ecm* e, float x, float y) {
    v2 p = V(x, y);
    int ncell = 2 * e->n_edge;
    for (int c = 0; c < ncell; ++c) {
        int edge = c >> 1;
        v2 q[4];
        q[0] = evert(e, e->he_target[2 * edge + 1]);
        q[3] = evert(e, e->he_target[2 * edge + 0]);
        if ((c & 1) == 0) { q[1] = hept(e, edge, 0); q[2] = hept(e, edge, 3); }
        else              { q[1] = hept(e, edge, 1); q[2] = hept(e, edge, 2); }
        if (contains4(p, q)) return c;
    }
    return -1;
}

/* ECM::RetractPoint (ECM.cpp:20-96) */
static int retract_cell(const oracle_ecm* e, v2 loc, int cell, v2* out) {
    int edge = cell >> 1;
    v2 p1 = evert(e, e->he_target[2 * edge + 1]);
    v2 p2 = evert(e, e->he_target[2 * edge + 0]);
    /* halfEdgeToP2Idx is always 0: p2 is vertex(he0.target) (ECM.cpp:36) */
    v2 o1, o2, ray;
    if (is_left_of_segment(p1, p2, loc)) {
        o1 = hept(e, edge, 0);            /* he0.closest_left  */
        o2 = hept(e, edge, 3);            /* he1.closest_right */
        if (approx(o1, o2)) {
            v2 a = vsub(p1, o1), b = vsub(p2, o1);
            ray = vadd(a, b);
        } else {
            v2 v = vsub(o2, o1);
            ray = V(v.y, -v.x);
        }
        ray = vnormalize(ray);
    } else {
        o1 = hept(e, edge, 1);            /* he0.closest_right */
        o2 = hept(e, edge, 2);            /* he1.closest_left  */
        if (approx(o1, o2)) {
            v2 a = vsub(p1, o1), b = vsub(p2, o1);
            ray = vadd(a, b);
        } else {
            v2 v = vsub(o2, o1);
            ray = V(-v.y, v.x);
        }
        ray = vnormalize(ray);
    }
    return ray_segment(loc, ray, p1, p2, out);
}

int oracle_retract(const oracle_ecm* e, float x, float y, float* out_xy, int* out_edge) {
    int cell = oracle_locate_cell(e, x, y);
    if (out_edge) *out_edge = cell < 0 ? -1 : cell >> 1;
    if (cell < 0) return 0;
    v2 r = V(0.0f, 0.0f);
    int ok = retract_cell(e, V(x, y), cell, &r);
    if (ok) { out_xy[0] = r.x; out_xy[1] = r.y; }
    return ok;
}

/* IRMPathFollower::FindAttractionPoint (IRMPathFollower.cpp:14-115).
 * `out` keeps its input value unless overwritten (Point() at the call site). */
static int find_attraction(const oracle_ecm* e, v2 pos, const float* path, int np,
                           v2* out, int* cell_out) {
    int cell = oracle_locate_cell(e, pos.x, pos.y);
    if (cell_out) *cell_out = cell;
    if (cell < 0) return 0;
    v2 rl = V(0.0f, 0.0f);
    if (!retract_cell(e, pos, cell, &rl)) return 0;
    int edge = cell >> 1;
    v2 oa = hept(e, edge, 0);            /* he0.closest_left  */
    v2 ob = hept(e, edge, 3);            /* he1.closest_right */
    v2 co = closest_on_segment(rl, oa, ob);
    float clearance = vlen(vsub(rl, co));
    float clr2 = clearance * clearance;
    v2 goal = V(path[2 * (np - 1)], path[2 * (np - 1) + 1]);
    v2 pg = vsub(goal, rl);
    if (vlen2(pg) < clr2) { *out = goal; return 1; }
    int success = 0;
    for (int i = 0; i < np - 1; ++i) {
        v2 p1 = vsub(V(path[2 * i], path[2 * i + 1]), rl);
        v2 p2 = vsub(V(path[2 * i + 2], path[2 * i + 3]), rl);
        v2 ed = vsub(p2, p1);
        float el2 = vlen2(ed);
        float dtm = det(p1, p2);
        float disc = clr2 * el2 - dtm * dtm;
        if (disc < EPS) continue;
        success = 1;
        float sgn = ed.y < 0.0f ? -1.0f : 1.0f;
        float sq = sqrtf(disc);
        v2 i1 = V((dtm * ed.y + sgn * ed.x * sq) / el2, (-dtm * ed.x + fabsf(ed.y) * sq) / el2);
        v2 i2 = V((dtm * ed.y - sgn * ed.x * sq) / el2, (-dtm * ed.x - fabsf(ed.y) * sq) / el2);
        v2 g1 = vadd(p1, rl), g2 = vadd(p2, rl);
        v2 gi1 = vadd(i1, rl), gi2 = vadd(i2, rl);
        v2 edg = vsub(g2, g1);
        float t1 = dot(vsub(gi1, g1), edg) / el2;
        float t2 = dot(vsub(gi2, g1), edg) / el2;
        float maxT = -1.0f;
        if (t1 >= 0.0f && t1 <= 1.0f) { *out = gi1; maxT = t1; }
        if (t2 >= 0.0f && t2 <= 1.0f) { if (t2 > maxT) *out = gi2; }
    }
    return success;
}

int oracle_attraction(const oracle_ecm* e, float x, float y, const float* path, int np, float* out_xy) {
    v2 o = V(0.0f, 0.0f);
    int ok = find_attraction(e, V(x, y), path, np, &o, NULL);
    out_xy[0] = o.x; out_xy[1] = o.y;
    return ok;
}

/* ------------------------------------------------------------ obstacles */
typedef struct {
    int nv;
    v2* p;
    int* next;
    int* prev;
    unsigned char* convex;
} obst_soa;

/* Obstacle::Initialize(vertexPositions) (ECMDataTypes.cpp:23-61) */
static int obst_build(const oracle_obstacles* o, obst_soa* s) {
    int nv = o->n_rings > 0 ? o->ring_off[o->n_rings] : 0;
    s->nv = nv;
    s->p = (v2*)malloc(sizeof(v2) * (nv + 1));
    s->next = (int*)malloc(sizeof(int) * (nv + 1));
    s->prev = (int*)malloc(sizeof(int) * (nv + 1));
    s->convex = (unsigned char*)malloc(nv + 1);
    if (!s->p || !s->next || !s->prev || !s->convex) return -1;
    for (int i = 0; i < nv; ++i) s->p[i] = V(o->xy[2 * i], o->xy[2 * i + 1]);
    for (int r = 0; r < o->n_rings; ++r) {
        int b = o->ring_off[r], m = o->ring_off[r + 1] - b;
        if (m < 2) return -2;   /* Environment.cpp:288-292 rejects < 2 */
        for (int i = 0; i < m; ++i) {
            int v = b + i;
            if (i == 0) { s->prev[v] = b + m - 1; s->next[v] = b + 1; }
            else { s->prev[v] = b + i - 1; s->next[v] = b + (i + 1) % m; }
        }
        for (int i = 0; i < m; ++i) {
            int v = b + i;
            if (m == 2) s->convex[v] = 1;
            else s->convex[v] = det(vsub(s->p[s->prev[v]], s->p[s->next[v]]), vsub(s->p[v], s->p[s->prev[v]])) >= 0.0f;
        }
    }
    return 0;
}
static void obst_free(obst_soa* s) { free(s->p); free(s->next); free(s->prev); free(s->convex); }

/* Simulator::FindNearestObstacles (Simulator.cpp:259-292) */
static int nearest_obstacles(const obst_soa* s, v2 pos, float range2, int* out, int cap) {
    int n = 0;
    for (int v = 0; v < s->nv; ++v) {
        v2 a = s->p[v], b = s->p[s->next[v]];
        float sld = det(vsub(a, pos), vsub(b, a));          /* LineLeftDistance */
        float sq = sld * sld / sqdist(a, b);
        if (sq < range2) {
            if (sld < 0.0f) {
                v2 c = closest_on_segment(pos, a, b);
                if (sqdist(c, pos) < range2) {
                    if (n < cap) out[n] = v;
                    ++n;
                }
            }
        }
    }
    return n;
}

/* ------------------------------------------------------------ constraints */
typedef struct { v2 n; v2 p; } con;     /* Constraint (ORCA.h:26-80): normal, point */

/* Constraint::Contains, methodB (ORCA.h:49-60) */
static inline int con_contains(con c, v2 q) { return det(right(c.n), vsub(c.p, q)) <= 0.0f; }

static inline void push(con* cs, int* n, int cap, v2 p, v2 nrm) {
    if (*n < cap) { cs[*n].n = nrm; cs[*n].p = p; }
    ++*n;
}

/* Branches of the obstacle half-plane (ORCA.cpp:70-333) taken for the edges
 * processed while g_obst_branches points somewhere (oracle_obstacle_halfplane);
 * bits as ORACLE_OB_* in ecm_oracle.h. */
static unsigned* g_obst_branches = NULL;
#define OB_HIT(bit) do { if (g_obst_branches) *g_obst_branches |= (bit); } while (0)

/* ORCA::GenerateConstraints (ORCA.cpp:60-424) */
static int gen_constraints_ex(const oracle_params* prm, const oracle_state* st, const obst_soa* s,
                              int e, const int* obst, int n_obst_nb, const int* nbr, int n_nbr,
                              con* cs, int cap, int* n_obst_out, int* con_start);
static int gen_constraints(const oracle_params* prm, const oracle_state* st, const obst_soa* s,
                           int e, const int* obst, int n_obst_nb, const int* nbr, int n_nbr,
                           con* cs, int cap, int* n_obst_out) {
    return gen_constraints_ex(prm, st, s, e, obst, n_obst_nb, nbr, n_nbr, cs, cap, n_obst_out, NULL);
}

/* con_start (optional, n_obst_nb + 1 entries): constraint count before each
 * obstacle edge is processed, and after the last, so that edge i emitted
 * constraint con_start[i] iff con_start[i + 1] > con_start[i]. */
static int gen_constraints_ex(const oracle_params* prm, const oracle_state* st, const obst_soa* s,
                              int e, const int* obst, int n_obst_nb, const int* nbr, int n_nbr,
                              con* cs, int cap, int* n_obst_out, int* con_start) {
    int nc = 0;
    v2 position = V(st->pos[2 * e], st->pos[2 * e + 1]);
    v2 velocity = V(st->vel[2 * e], st->vel[2 * e + 1]);
    float rad = st->radius[e];
    const float inf = INFINITY;

    for (int i = 0; i < n_obst_nb; ++i) {                   /* :70-333 */
        if (con_start) con_start[i] = nc;
        int L = obst[i];
        int R = s->next[L];
        v2 rp1 = vsub(s->p[L], position);
        v2 rp2 = vsub(s->p[R], position);
        v2 seg = vsub(s->p[R], s->p[L]);
        const float sp = dot(vmul(rp1, -1.0f), seg) / vlen2(seg);
        const float distSqLine = vlen2(vsub(vmul(rp1, -1.0f), vmul(seg, sp)));
        const float distSq1 = vlen2(rp1);
        const float distSq2 = vlen2(rp2);
        seg = vnormalize(seg);
        const float radiusSq = rad * rad;

        if (sp < 0.0f && distSq1 <= radiusSq) {              /* case 1 :92-105 */
            OB_HIT(s->convex[L] ? ORACLE_OB_CASE1 : ORACLE_OB_CASE1_SKIP);
            if (s->convex[L]) push(cs, &nc, cap, V(0.0f, 0.0f), vnormalized(vmul(rp1, -1.0f)));
            continue;
        } else if (sp > 1.0f && distSq2 <= radiusSq) {       /* case 2 :108-121 */
            v2 rnd = vnormalize(vsub(s->p[s->next[R]], s->p[R]));
            const int emit = s->convex[R] && det(rp2, rnd) >= 0.0f;
            OB_HIT(emit ? ORACLE_OB_CASE2 : ORACLE_OB_CASE2_SKIP);
            if (emit) push(cs, &nc, cap, V(0.0f, 0.0f), vnormalized(vmul(rp2, -1.0f)));
            continue;
        } else if (sp >= 0.0f && sp < 1.0f && distSqLine <= radiusSq) {  /* case 3 :124-135 */
            OB_HIT(ORACLE_OB_CASE3);
            push(cs, &nc, cap, V(0.0f, 0.0f), right(seg));
            continue;
        }

        v2 leftLeg, rightLeg;
        if (sp < 0.0f && distSqLine <= radiusSq) {           /* :146-169 */
            OB_HIT(s->convex[L] ? ORACLE_OB_VERTEX_L : ORACLE_OB_VERTEX_L_SKIP);
            if (!s->convex[L]) continue;
            R = L;
            const float leg1 = sqrtf(distSq1 - radiusSq);
            leftLeg = vdiv(V(rp1.x * leg1 - rp1.y * rad, rp1.x * rad + rp1.y * leg1), distSq1);
            rightLeg = vdiv(V(rp1.x * leg1 + rp1.y * rad, -rp1.x * rad + rp1.y * leg1), distSq1);
        } else if (sp > 1.0f && distSqLine <= radiusSq) {    /* :171-183 */
            OB_HIT(s->convex[R] ? ORACLE_OB_VERTEX_R : ORACLE_OB_VERTEX_R_SKIP);
            if (!s->convex[R]) continue;
            L = R;
            const float leg2 = sqrtf(distSq2 - radiusSq);
            leftLeg = vdiv(V(rp2.x * leg2 - rp2.y * rad, rp2.x * rad + rp2.y * leg2), distSq2);
            rightLeg = vdiv(V(rp2.x * leg2 + rp2.y * rad, -rp2.x * rad + rp2.y * leg2), distSq2);
        } else {                                             /* :186-212 */
            if (s->convex[L]) {
                const float leg1 = sqrtf(distSq1 - radiusSq);
                leftLeg = vdiv(V(rp1.x * leg1 - rp1.y * rad, rp1.x * rad + rp1.y * leg1), distSq1);
            } else {
                OB_HIT(ORACLE_OB_CONCAVE_L);
                leftLeg = vmul(seg, -1.0f);
            }
            if (s->convex[R]) {
                const float leg2 = sqrtf(distSq2 - radiusSq);
                rightLeg = vdiv(V(rp2.x * leg2 + rp2.y * rad, -rp2.x * rad + rp2.y * leg2), distSq2);
            } else {
                OB_HIT(ORACLE_OB_CONCAVE_R);
                rightLeg = seg;
            }
        }

        /* foreign legs :218-239 */
        int leftForeign = 0, rightForeign = 0;
        v2 lnd = vnormalize(vsub(s->p[L], s->p[s->prev[L]]));
        if (s->convex[L] && det(leftLeg, vmul(lnd, -1.0f)) >= 0.0f) {
            OB_HIT(ORACLE_OB_FOREIGN_L);
            leftLeg = vmul(lnd, -1.0f);
            leftForeign = 1;
        }
        v2 rnd = vnormalize(vsub(s->p[s->next[R]], s->p[R]));
        if (s->convex[R] && det(rightLeg, rnd) <= 0.0f) {
            OB_HIT(ORACLE_OB_FOREIGN_R);
            rightLeg = rnd;
            rightForeign = 1;
        }

        float invT = 1.0f / prm->time_horizon_obst;          /* :241 */
        const v2 leftCutoff = vmul(vsub(s->p[L], position), invT);
        const v2 rightCutoff = vmul(vsub(s->p[R], position), invT);
        const v2 cutoffVec = vsub(rightCutoff, leftCutoff);
        const int same = (L == R);
        const float t = same ? 0.5f : dot(vsub(velocity, leftCutoff), cutoffVec) / vlen2(cutoffVec);
        const float tLeft = dot(vsub(velocity, leftCutoff), leftLeg);
        const float tRight = dot(vsub(velocity, rightCutoff), rightLeg);

        if ((t < 0.0f && tLeft < 0.0f) || (same && tLeft < 0.0f && tRight < 0.0f)) {   /* :259-268 */
            OB_HIT(ORACLE_OB_CIRCLE_L);
            v2 uw = vnormalized(vsub(velocity, leftCutoff));
            v2 p = vadd(leftCutoff, vmul(vmul(uw, invT), rad));
            push(cs, &nc, cap, p, uw);
            continue;
        } else if (t > 1.0f && tRight < 0.0f) {              /* :270-280 */
            OB_HIT(ORACLE_OB_CIRCLE_R);
            v2 uw = vnormalized(vsub(velocity, rightCutoff));
            v2 p = vadd(rightCutoff, vmul(vmul(uw, invT), rad));
            push(cs, &nc, cap, p, uw);
            continue;
        }

        const float dCut = (t < 0.0f || t > 1.0f || same) ? inf : vlen2(vsub(velocity, vadd(leftCutoff, vmul(cutoffVec, t))));
        const float dLeft = (tLeft < 0.0f) ? inf : vlen2(vsub(velocity, vadd(leftCutoff, vmul(leftLeg, tLeft))));
        const float dRight = (tRight < 0.0f) ? inf : vlen2(vsub(velocity, vadd(rightCutoff, vmul(rightLeg, tRight))));

        if (dCut <= dLeft && dCut <= dRight) {               /* :289-301 */
            OB_HIT(ORACLE_OB_CUTOFF_LINE);
            v2 nrm = left(vmul(seg, -1.0f));
            v2 p = vadd(leftCutoff, vmul(vmul(nrm, invT), rad));
            push(cs, &nc, cap, p, nrm);
            continue;
        } else if (dLeft <= dRight) {                        /* :303-317 */
            OB_HIT(leftForeign ? ORACLE_OB_LEG_L_SKIP : ORACLE_OB_LEG_L);
            if (leftForeign) continue;
            v2 nrm = left(leftLeg);
            v2 p = vadd(leftCutoff, vmul(vmul(nrm, rad), invT));
            push(cs, &nc, cap, p, nrm);
            continue;
        } else {                                             /* :319-332 */
            OB_HIT(rightForeign ? ORACLE_OB_LEG_R_SKIP : ORACLE_OB_LEG_R);
            if (rightForeign) continue;
            v2 nrm = right(rightLeg);
            v2 p = vadd(rightCutoff, vmul(vmul(nrm, rad), invT));
            push(cs, &nc, cap, p, nrm);
            continue;
        }
    }
    *n_obst_out = nc;                                        /* :335 */
    if (con_start) con_start[n_obst_nb] = nc;

    const float T = prm->time_horizon;
    const float dt = prm->dt;
    for (int i = 0; i < n_nbr; ++i) {                        /* :339-423 */
        int nb = nbr[i];
        v2 npos = V(st->pos[2 * nb], st->pos[2 * nb + 1]);
        v2 nvel = V(st->vel[2 * nb], st->vel[2 * nb + 1]);
        float nrad = st->radius[nb];
        v2 VOPos = V((npos.x - position.x) / T, (npos.y - position.y) / T);
        float VOPosLength = vlen(VOPos);
        float combined = nrad + rad;
        float VORadius = combined / T;
        v2 relVel = V(velocity.x - nvel.x, velocity.y - nvel.y);
        v2 relPos = V(npos.x - position.x, npos.y - position.y);
        float relPosLength = vlen(relPos);
        if (relPosLength < combined) {                       /* :356-372 */
            const v2 w = vsub(relVel, vdiv(relPos, dt));
            const float wl = vlen(w);
            const v2 uw = vdiv(w, wl);
            const v2 U = vmul(uw, combined / dt - wl);
            push(cs, &nc, cap, vadd(velocity, vmul(U, 0.5f)), uw);
            continue;
        }
        float tanAngleFactor = VORadius / VOPosLength;       /* :375-378 */
        float half = oracle_atanf(tanAngleFactor);
        v2 VOLeftLeg = rotate(VOPos, half);
        v2 VORightLeg = rotate(VOPos, -half);
        float sqFromCentre = sqdist(VOPos, relVel);
        int below = det(vsub(VOLeftLeg, VOPos), vsub(relVel, VOPos)) > 0;   /* IsLeftOfVector */
        if (below) {                                         /* :385-397 */
            float distToEdge = VORadius - sqrtf(sqFromCentre);
            v2 ln = vnormalize(vsub(relVel, VOPos));
            v2 p = vadd(velocity, vmul(vmul(ln, distToEdge), 0.5f));
            push(cs, &nc, cap, p, ln);
        } else {                                             /* :398-422 */
            const v2 lperp = vdiv(left(VOPos), VOPosLength);
            int closerLeft = dot(lperp, relVel) >= 0;
            if (closerLeft) {
                const v2 ll = vnormalized(VOLeftLeg);
                float lp = dot(relVel, ll);                  /* GetClosestPointOnLineThroughOrigin */
                const v2 U = vsub(vmul(ll, lp), relVel);
                push(cs, &nc, cap, vadd(velocity, vmul(U, 0.5f)), left(ll));
            } else {
                const v2 rl = vnormalized(VORightLeg);
                float lp = dot(relVel, rl);
                const v2 U = vsub(vmul(rl, lp), relVel);
                push(cs, &nc, cap, vadd(velocity, vmul(U, 0.5f)), right(rl));
            }
        }
    }
    return nc;
}

/* ------------------------------------------------------------------- LPs */
/* std::min(a, b) = (b < a) ? b : a ; std::max(a, b) = (a < b) ? b : a */
static inline float std_min(float a, float b) { return (b < a) ? b : a; }
static inline float std_max(float a, float b) { return (a < b) ? b : a; }

/* ORCA::RandomizedLP (ORCA.cpp:428-587) */
static int lp2d(const con* cs, int n, v2 opt, float maxSpeed, int useDir, v2* out) {
    if (useDir) *out = vmul(opt, maxSpeed);
    else if (vlen(opt) > maxSpeed) *out = vmul(vnormalized(opt), maxSpeed);
    else *out = opt;
    if (n == 0) return n;
    for (int i = 0; i < n; ++i) {
        con h = cs[i];
        if (con_contains(h, *out)) continue;
        v2 dir = right(h.n);
        float dpd = dot(dir, h.p);
        float disc = dpd * dpd + maxSpeed * maxSpeed - dot(h.p, h.p);
        if (disc <= 0.0f) return i;
        float sq = sqrtf(disc);
        float lft = -dpd - sq;
        float rgt = -dpd + sq;
        for (int j = 0; j < i; ++j) {
            con hj = cs[j];
            float den = det(right(h.n), right(hj.n));
            float num = det(right(hj.n), vsub(h.p, hj.p));
            if (fabsf(den) <= EPS) {
                if (num < 0.0f) return i;
                else continue;
            }
            const float t = num / den;
            if (den >= 0.0f) rgt = std_min(rgt, t);
            else lft = std_max(lft, t);
            if (lft > rgt) return i;
        }
        if (useDir) {
            if (dot(opt, right(h.n)) > 0) *out = vadd(h.p, vmul(right(h.n), rgt));
            else *out = vadd(h.p, vmul(right(h.n), lft));
        } else {
            float t = dot(right(h.n), vsub(opt, h.p));
            if (t < lft) *out = vadd(h.p, vmul(right(h.n), lft));
            else if (t > rgt) *out = vadd(h.p, vmul(right(h.n), rgt));
            else *out = vadd(h.p, vmul(right(h.n), t));
        }
    }
    return n;
}

/* ORCA::RandomizedLP3D (ORCA.cpp:592-669) */
static void lp3d(const con* cs, int n, int nObst, float maxSpeed, int failed, v2* out, con* scratch) {
    float maxPen = 0.0f;
    for (int i = failed; i < n; ++i) {
        const con ci = cs[i];
        v2 dir = right(ci.n);
        if (det(dir, vsub(ci.p, *out)) <= maxPen) continue;
        int np = 0;
        for (int j = 0; j < nObst; ++j) scratch[np++] = cs[j];
        for (int j = nObst; j < i; ++j) {
            const con cj = cs[j];
            float d = det(right(ci.n), right(cj.n));
            v2 newP;
            if (fabsf(d) <= EPS) {
                if (dot(ci.n, cj.n) > 0) continue;
                newP = vmul(vadd(ci.p, cj.p), 0.5f);
            } else {
                float t = det(right(cj.n), vsub(ci.p, cj.p)) / d;
                newP = vadd(ci.p, vmul(right(ci.n), t));
            }
            scratch[np].p = newP;
            scratch[np].n = vnormalized(vsub(cj.n, ci.n));
            ++np;
        }
        const v2 temp = *out;
        if (lp2d(scratch, np, ci.n, maxSpeed, 1, out) < np) *out = temp;
        maxPen = det(dir, vsub(ci.p, *out));
    }
}

int oracle_lp2d(const float* c, int n, const float* opt, float maxSpeed, int useDir, float* outv) {
    con* cs = (con*)malloc(sizeof(con) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) { cs[i].n = V(c[4 * i], c[4 * i + 1]); cs[i].p = V(c[4 * i + 2], c[4 * i + 3]); }
    v2 o;
    int r = lp2d(cs, n, V(opt[0], opt[1]), maxSpeed, useDir, &o);
    outv[0] = o.x; outv[1] = o.y;
    free(cs);
    return r;
}

void oracle_lp3d(const float* c, int n, int nObst, float maxSpeed, int failed, float* outv) {
    con* cs = (con*)malloc(sizeof(con) * (n > 0 ? n : 1));
    con* sc = (con*)malloc(sizeof(con) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) { cs[i].n = V(c[4 * i], c[4 * i + 1]); cs[i].p = V(c[4 * i + 2], c[4 * i + 3]); }
    v2 o = V(outv[0], outv[1]);
    lp3d(cs, n, nObst, maxSpeed, failed, &o, sc);
    outv[0] = o.x; outv[1] = o.y;
    free(cs); free(sc);
}

/* -------------------------------------------------------------- KD-tree */
typedef struct {
    int* tree;
    int size;       /* tree array size */
    int maxDepth;
    const float* pos;
} kdtree;

static const float* g_sort_pos;
static int g_sort_axis;
static int kd_cmp(const void* a, const void* b) {
    int ia = *(const int*)a, ib = *(const int*)b;
    float va = g_sort_pos[2 * ia + g_sort_axis], vb = g_sort_pos[2 * ib + g_sort_axis];
    if (va < vb) return -1;
    if (vb < va) return 1;
    return (ia > ib) - (ia < ib);     /* ties by slot id */
}

/* KDTree::ConstructRecursive (KDTree.cpp:59-83) */
static void kd_build_r(kdtree* t, int index, int lb, int rb, int depth, int* idx) {
    if (lb >= rb) return;
    g_sort_axis = (depth % 2 == 0) ? 0 : 1;
    qsort(idx + lb, (size_t)(rb - lb), sizeof(int), kd_cmp);
    int mid = lb + (rb - 1 - lb) / 2;
    t->tree[index] = idx[mid];
    kd_build_r(t, index * 2 + 1, lb, mid, depth + 1, idx);
    kd_build_r(t, index * 2 + 2, mid + 1, rb, depth + 1, idx);
}

/* KDTree::Construct (KDTree.cpp:22-57) */
static int kd_build(kdtree* t, const oracle_state* st) {
    int size = 0;
    for (int i = 0; i < st->n; ++i) size += st->active[i] ? 1 : 0;
    t->pos = st->pos;
    t->tree = NULL; t->size = 0; t->maxDepth = -1;
    if (size == 0) return 0;
    int* idx = (int*)malloc(sizeof(int) * size);
    int c = 0;
    for (int i = 0; i < st->n; ++i) if (st->active[i]) idx[c++] = i;
    t->maxDepth = (int)ceil(log2((double)size + 1.0) - 1.0);
    t->size = (int)pow(2.0, (double)(t->maxDepth + 1)) - 1;
    t->tree = (int*)malloc(sizeof(int) * t->size);
    for (int i = 0; i < t->size; ++i) t->tree[i] = -1;
    g_sort_pos = st->pos;
    kd_build_r(t, 0, 0, size, 0, idx);
    free(idx);
    return size;
}

/* KDTree::KNearestAgents_R (KDTree.cpp:98-202), defects included. */
static void kd_query_r(const kdtree* t, v2 target, int cur, int k, int* kFound, int depth,
                       int* nidx, float* nd2) {
    if (depth > t->maxDepth) return;
    if (t->tree[cur] == -1) return;
    int node = t->tree[cur];
    v2 cp = V(t->pos[2 * node], t->pos[2 * node + 1]);
    v2 diff = V(cp.x - target.x, cp.y - target.y);
    float d2 = diff.x * diff.x + diff.y * diff.y;
    if (*kFound < k && d2 > EPS) {
        nidx[*kFound] = node;
        nd2[*kFound] = d2;
        (*kFound)++;
        if (*kFound == k) {
            float largest = d2;
            int li = k - 1;
            for (int i = 0; i < k - 1; ++i)
                if (nd2[i] > largest) { largest = nd2[i]; li = i; }
            nd2[0] = largest;
            nidx[0] = nidx[li];
            nd2[li] = d2;
            nidx[li] = node;
        }
    } else {
        *kFound = k;
        if (d2 < nd2[0] && d2 > EPS) {
            nd2[0] = d2;
            nidx[0] = node;
            float largest = d2;
            int li = 0;
            for (int i = 1; i < k; ++i)
                if (nd2[i] > largest) { largest = nd2[i]; li = i; }
            nd2[0] = largest;
            nidx[0] = nidx[li];
            nd2[li] = d2;
            nidx[li] = node;
        }
    }
    float cv = depth % 2 == 0 ? cp.x : cp.y;
    float tv = depth % 2 == 0 ? target.x : target.y;
    if (tv < cv) {
        kd_query_r(t, target, cur * 2 + 1, k, kFound, depth + 1, nidx, nd2);
        float dd = tv - cv;
        float sq = dd * dd;
        if (sq < nd2[k - 1]) kd_query_r(t, target, cur * 2 + 2, k, kFound, depth + 1, nidx, nd2);
    } else {
        kd_query_r(t, target, cur * 2 + 2, k, kFound, depth + 1, nidx, nd2);
        float dd = tv - cv;
        float sq = dd * dd;
        if (sq < nd2[k - 1]) kd_query_r(t, target, cur * 2 + 1, k, kFound, depth + 1, nidx, nd2);
    }
}

int oracle_kdtree_build_only(const oracle_state* st) {
    kdtree t;
    kd_build(&t, st);
    int s = t.size;
    free(t.tree);
    return s;
}

/* Exact k-nearest: the k smallest d2 > EPS over step-start active agents,
 * ordered by (d2, slot id); d2 computed as in KDTree.cpp:106-108. */
static int knn_exact(const oracle_state* st, int e, int k, int* out, float* outd) {
    int cnt = 0;
    float tx = st->pos[2 * e], ty = st->pos[2 * e + 1];
    for (int j = 0; j < st->n; ++j) {
        if (!st->active[j]) continue;
        float dx = st->pos[2 * j] - tx, dy = st->pos[2 * j + 1] - ty;
        float d2 = dx * dx + dy * dy;
        if (!(d2 > EPS)) continue;
        if (cnt == k && !(d2 < outd[k - 1] || (d2 == outd[k - 1] && j < out[k - 1]))) continue;
        int p = cnt < k ? cnt : k - 1;
        if (cnt < k) ++cnt;
        while (p > 0 && (outd[p - 1] > d2 || (outd[p - 1] == d2 && out[p - 1] > j))) {
            outd[p] = outd[p - 1]; out[p] = out[p - 1]; --p;
        }
        outd[p] = d2; out[p] = j;
    }
    return cnt;
}

/* ------------------------------------------------------------------ step */
#define MAXC 256

int oracle_step(const oracle_params* prm, const oracle_state* st, const oracle_obstacles* obs,
                const oracle_ecm* ecm, const int* subset, int n_subset, oracle_out* out) {
    if (!prm || !st || !out || prm->k <= 0 || prm->k > 64) return -1;
    obst_soa s;
    memset(&s, 0, sizeof(s));
    if (obs && obs->n_rings > 0) { if (obst_build(obs, &s) != 0) { obst_free(&s); return -2; } }
    const int k = prm->k;
    const float massRecip = 1.0f / prm->mass;           /* Simulator.cpp:623 */
    const float dt = prm->dt;
    const float arr2 = prm->arrival_radius * prm->arrival_radius;
    const float del2 = prm->delete_radius * prm->delete_radius;

    kdtree tree;
    memset(&tree, 0, sizeof(tree));
    if (prm->knn_mode == ORACLE_KNN_REFERENCE) kd_build(&tree, st);
    int* cache = (int*)calloc((size_t)k, sizeof(int));   /* ORCA::m_NeighborCache, zero-initialised */
    float* cd = (float*)malloc(sizeof(float) * k);
    int* nb = (int*)malloc(sizeof(int) * k);
    int* obl = (int*)malloc(sizeof(int) * MAXC);
    con* cs = (con*)malloc(sizeof(con) * MAXC);
    con* scr = (con*)malloc(sizeof(con) * MAXC);

    int count = subset ? n_subset : st->n;
    for (int it = 0; it < count; ++it) {
        int e = subset ? subset[it] : it;
        if (e < 0 || e >= st->n || !st->active[e]) continue;
        v2 pos = V(st->pos[2 * e], st->pos[2 * e + 1]);
        v2 vel = V(st->vel[2 * e], st->vel[2 * e + 1]);
        v2 attr = V(st->attraction[2 * e], st->attraction[2 * e + 1]);
        int np = st->path_cnt[e];
        const float* path = st->path_pts + 2 * (size_t)st->path_off[e];
        int cell = -1;
        unsigned char alive = 1, replan = 0;

        /* UpdateAttractionPointSystem (Simulator.cpp:538-590) */
        v2 goal = V(path[2 * (np - 1)], path[2 * (np - 1) + 1]);
        float dGoal = (pos.x - goal.x) * (pos.x - goal.x) + (pos.y - goal.y) * (pos.y - goal.y);
        if (dGoal < arr2) {
            attr = goal;
            if (dGoal < del2) alive = 0;                     /* DestroyAgent */
        } else {
            v2 ap = V(0.0f, 0.0f);                           /* Point attractionPoint; */
            if (ecm && find_attraction(ecm, pos, path, np, &ap, &cell)) attr = ap;
            else replan = 1;                                 /* UpdatePath (host) */
        }
        out->attraction[2 * e] = attr.x; out->attraction[2 * e + 1] = attr.y;
        out->replan[e] = replan;
        out->active[e] = alive;
        if (out->diag) out->diag[4 * e] = cell;
        if (!alive) {
            /* a destroyed agent is skipped by every later system this step */
            out->pos[2 * e] = pos.x; out->pos[2 * e + 1] = pos.y;
            out->vel[2 * e] = vel.x; out->vel[2 * e + 1] = vel.y;
            if (out->nbr_cnt) out->nbr_cnt[e] = 0;
            continue;
        }

        /* ApplySteeringForce (Simulator.cpp:638-657) */
        float speed = st->pref_speed[e];
        v2 des = vnormalize(vsub(attr, pos));
        des = vmul(des, speed);
        out->pref_vel[2 * e] = des.x; out->pref_vel[2 * e + 1] = des.y;

        /* ApplyObstacleAvoidanceForce -> ORCA::GetVelocity (Simulator.cpp:659-686, ORCA.cpp:14-57) */
        float maxSpeed = prm->max_speed_from_pref ? speed : prm->max_speed;
        int nn;
        if (prm->knn_mode == ORACLE_KNN_REFERENCE) {
            for (int i = 0; i < k; ++i) cd[i] = MAXF;
            nn = 0;
            if (tree.tree) kd_query_r(&tree, pos, 0, k, &nn, 0, cache, cd);
            memcpy(nb, cache, sizeof(int) * k);
        } else {
            nn = knn_exact(st, e, k, nb, cd);
        }
        if (out->nbr) for (int i = 0; i < k; ++i) out->nbr[(size_t)e * k + i] = i < nn ? nb[i] : -1;
        if (out->nbr_cnt) out->nbr_cnt[e] = nn;

        float range = prm->time_horizon_obst * maxSpeed + st->radius[e];   /* ORCA.cpp:27 */
        int nob = s.nv > 0 ? nearest_obstacles(&s, pos, range * range, obl, MAXC) : 0;
        if (nob > MAXC - 64) nob = MAXC - 64;
        int nObst = 0;
        int nc = gen_constraints(prm, st, &s, e, obl, nob, nb, nn, cs, MAXC, &nObst);
        v2 o;
        int fail = lp2d(cs, nc, des, maxSpeed, 0, &o);
        int used3d = 0;
        if (fail < nc) { lp3d(cs, nc, nObst, maxSpeed, fail, &o, scr); used3d = 1; }
        if (out->v_orca) { out->v_orca[2 * e] = o.x; out->v_orca[2 * e + 1] = o.y; }
        if (out->used_3d) out->used_3d[e] = (unsigned char)used3d;
        if (out->diag) { out->diag[4 * e + 1] = nObst; out->diag[4 * e + 2] = nc; out->diag[4 * e + 3] = fail; }

        /* force (Simulator.cpp:676-677), UpdateVelocitySystem (:619-635),
         * UpdatePositionSystem (:592-606) */
        v2 F = V(o.x - vel.x, o.y - vel.y);
        v2 nv = V(vel.x + F.x * massRecip * dt, vel.y + F.y * massRecip * dt);
        v2 npn = V(pos.x + (nv.x * dt), pos.y + (nv.y * dt));
        out->vel[2 * e] = nv.x; out->vel[2 * e + 1] = nv.y;
        out->pos[2 * e] = npn.x; out->pos[2 * e + 1] = npn.y;
    }

    free(tree.tree); free(cache); free(cd); free(nb); free(obl); free(cs); free(scr);
    obst_free(&s);
    return 0;
}

int oracle_constraints(const oracle_params* prm, const oracle_state* st, const oracle_obstacles* obs,
                       int agent, const int* nbrs, int n_nbrs, float* out, int cap, int* n_obst) {
    obst_soa s;
    memset(&s, 0, sizeof(s));
    if (obs && obs->n_rings > 0) { if (obst_build(obs, &s) != 0) { obst_free(&s); return -2; } }
    float maxSpeed = prm->max_speed_from_pref ? st->pref_speed[agent] : prm->max_speed;
    float range = prm->time_horizon_obst * maxSpeed + st->radius[agent];
    int obl[MAXC];
    v2 pos = V(st->pos[2 * agent], st->pos[2 * agent + 1]);
    int nob = s.nv > 0 ? nearest_obstacles(&s, pos, range * range, obl, MAXC) : 0;
    if (nob > MAXC - 64) nob = MAXC - 64;
    con cs[MAXC];
    int nc = gen_constraints(prm, st, &s, agent, obl, nob, nbrs, n_nbrs, cs, MAXC, n_obst);
    for (int i = 0; i < nc && i < cap; ++i) {
        out[4 * i] = cs[i].n.x; out[4 * i + 1] = cs[i].n.y;
        out[4 * i + 2] = cs[i].p.x; out[4 * i + 3] = cs[i].p.y;
    }
    obst_free(&s);
    return nc;
}

int oracle_obstacle_halfplane(const oracle_params* prm, const oracle_state* st, const oracle_obstacles* obs, int agent,
                              int edge, float* out4, unsigned* branches) {
    obst_soa s;
    if (obst_build(obs, &s) != 0) { obst_free(&s); return -1; }
    if (edge < 0 || edge >= s.nv) { obst_free(&s); return -1; }
    con c[2];
    int nob = 0;
    unsigned b = 0;
    g_obst_branches = &b;
    gen_constraints(prm, st, &s, agent, &edge, 1, NULL, 0, c, 2, &nob);
    g_obst_branches = NULL;
    if (nob > 0) { out4[0] = c[0].n.x; out4[1] = c[0].n.y; out4[2] = c[0].p.x; out4[3] = c[0].p.y; }
    if (branches) *branches = b;
    obst_free(&s);
    return nob;
}

/* Simulator::FindNearestObstacles (Simulator.cpp:259-292): the edges kept for
 * an agent at (x, y) with the given range, in scan order. */
int oracle_nearest_obstacles(const oracle_obstacles* obs, float x, float y, float range, int* out, int cap) {
    obst_soa s;
    if (obst_build(obs, &s) != 0) { obst_free(&s); return -1; }
    int n = nearest_obstacles(&s, V(x, y), range * range, out, cap);
    obst_free(&s);
    return n;
}

/* ------------------------------------------- learned-proposal samples */
/* Inputs and label of the proposal network for one agent (the layout the
 * library's step logs; no reference counterpart):
 *   feat[64]: neighbour j < 10 relative position * 0.25 and velocity * 0.5
 *   (j over nbrs in the order given), endpoints of the 4 nearest obstacle
 *   edges kept by FindNearestObstacles (by squared distance to the closest
 *   point, then edge index) relative to the agent * 0.25, preferred
 *   velocity * 0.5, velocity * 0.5, radius * 4, each clamped to +-8;
 *   *label: bit s (s < 14) slot s active at the LP velocity v (within
 *   1e-3), bit 16 + s slot s holds a constraint, bit 30 every active
 *   constraint is behind a slot. */
int oracle_proposal_sample(const oracle_params* prm, const oracle_state* st, const oracle_obstacles* obs,
                           int agent, const int* nbrs, int n_nbrs, const float* pref_vel, const float* v,
                           float* feat, unsigned* label) {
    obst_soa s;
    memset(&s, 0, sizeof(s));
    if (obs && obs->n_rings > 0) { if (obst_build(obs, &s) != 0) { obst_free(&s); return -2; } }
    const float maxSpeed = prm->max_speed_from_pref ? st->pref_speed[agent] : prm->max_speed;
    const float range = prm->time_horizon_obst * maxSpeed + st->radius[agent];
    const v2 pos = V(st->pos[2 * agent], st->pos[2 * agent + 1]);
    const v2 vel = V(st->vel[2 * agent], st->vel[2 * agent + 1]);
    int obl[MAXC] = {0};
    int nob = s.nv > 0 ? nearest_obstacles(&s, pos, range * range, obl, MAXC) : 0;
    if (nob > MAXC - 64) nob = MAXC - 64;
    con cs[MAXC];
    int start[MAXC + 1];
    int nObst = 0;
    const int nc = gen_constraints_ex(prm, st, &s, agent, obl, nob, nbrs, n_nbrs, cs, MAXC, &nObst, start);
    /* 4 nearest kept edges by (dd, edge) */
    float nd[4];
    int ne[4], nci[4];
    for (int i = 0; i < 4; ++i) { nd[i] = INFINITY; ne[i] = 0x7fffffff; nci[i] = -1; }
    for (int i = 0; i < nob; ++i) {
        const int e = obl[i];
        const v2 c = closest_on_segment(pos, s.p[e], s.p[s.next[e]]);
        const float dd = sqdist(c, pos);
        const int ci = start[i + 1] > start[i] ? start[i] : -1;
        int at = 4;
        for (int j = 3; j >= 0; --j)
            if (dd < nd[j] || (dd == nd[j] && e < ne[j])) at = j;
        if (at < 4) {
            for (int j = 3; j > at; --j) { nd[j] = nd[j - 1]; ne[j] = ne[j - 1]; nci[j] = nci[j - 1]; }
            nd[at] = dd; ne[at] = e; nci[at] = ci;
        }
    }
#define FCLAMP(x) fminf(fmaxf((x), -8.0f), 8.0f)
    for (int i = 0; i < 64; ++i) feat[i] = 0.0f;
    for (int j = 0; j < n_nbrs && j < 10; ++j) {
        const int nb = nbrs[j];
        feat[4 * j + 0] = FCLAMP((st->pos[2 * nb] - pos.x) * 0.25f);
        feat[4 * j + 1] = FCLAMP((st->pos[2 * nb + 1] - pos.y) * 0.25f);
        feat[4 * j + 2] = FCLAMP((st->vel[2 * nb] - vel.x) * 0.5f);
        feat[4 * j + 3] = FCLAMP((st->vel[2 * nb + 1] - vel.y) * 0.5f);
    }
    for (int i = 0; i < 4; ++i) {
        if (ne[i] == 0x7fffffff) continue;
        const v2 a = s.p[ne[i]], b = s.p[s.next[ne[i]]];
        feat[40 + 4 * i + 0] = FCLAMP((a.x - pos.x) * 0.25f);
        feat[40 + 4 * i + 1] = FCLAMP((a.y - pos.y) * 0.25f);
        feat[40 + 4 * i + 2] = FCLAMP((b.x - pos.x) * 0.25f);
        feat[40 + 4 * i + 3] = FCLAMP((b.y - pos.y) * 0.25f);
    }
    feat[56] = FCLAMP(pref_vel[0] * 0.5f);
    feat[57] = FCLAMP(pref_vel[1] * 0.5f);
    feat[58] = FCLAMP(vel.x * 0.5f);
    feat[59] = FCLAMP(vel.y * 0.5f);
    feat[60] = FCLAMP(st->radius[agent] * 4.0f);
#undef FCLAMP
    const v2 o = V(v[0], v[1]);
    unsigned act = 0, valid = 0;
    int repr = 1;
    for (int c = 0; c < nc && c < MAXC; ++c) {
        const int a = det(right(cs[c].n), vsub(cs[c].p, o)) >= -1e-3f;
        int slot = -1;
        if (c >= nObst) {
            if (c - nObst < 10) slot = c - nObst;
        } else {
            for (int i = 0; i < 4; ++i)
                if (nci[i] == c) slot = 10 + i;
        }
        if (slot >= 0) {
            valid |= 1u << slot;
            if (a) act |= 1u << slot;
        } else if (a) {
            repr = 0;
        }
    }
    *label = act | (valid << 16) | (repr ? (1u << 30) : 0u);
    obst_free(&s);
    return nc;
}
