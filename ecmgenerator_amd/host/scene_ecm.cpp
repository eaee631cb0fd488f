// scene_ecm.cpp -- construction of the Explicit Corridor Map for a box with
// axis-aligned rectangular obstacles, without Boost.Polygon.
//
// The reference builds a segment Voronoi diagram with
// boost::polygon::construct_voronoi (ECMGenerator.cpp:235-256) and keeps
// its finite primary edges outside obstacles (ECMGenerator.cpp:60-232).
// Here the same diagram is computed directly for this input class:
//   sites ("features") are the open sides of every rectangle (facing out),
//   its four corners, and the four sides of the walkable box (facing in);
//   1. every grid cell proposes the triples of features that can be nearest
//      somewhere inside it (a 1-Lipschitz bound, so no vertex is missed);
//   2. each triple is solved exactly in double precision for the points
//      equidistant to its three features (lines, parabolas, circles);
//   3. a candidate is a Voronoi vertex when no feature is closer (checked
//      against every rectangle through a bucket grid) -- degenerate
//      vertices with four or more features are merged;
//   4. for every feature pair, the vertices on their bisector are sorted
//      along it and consecutive ones are joined when the bisector point
//      between them is itself equidistant and closest to the pair;
//   5. half-edges get the closest points on their left/right features at
//      their source vertex (ECMGenerator.cpp:172-195), cells follow as
//      ECMCellCollection::Construct (ECMCellCollection.cpp:18-46).
// Edges between a segment and its own endpoint (Boost's secondary edges)
// are not ECM edges; their vertices become degree-2 ECM vertices.
#include "scene.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <unordered_map>

#include "../csrc/ecm_geom.h"

namespace ecm_host {

namespace {

enum Kind { HSEG = 0, VSEG = 1, POINT = 2 };

struct Feature {
    int kind;
    double c, lo, hi;   // segment: line coordinate and extent
    int s;              // segment: outward normal sign along its axis
    double px, py;      // point
    int rect;           // owner rectangle, -1 for the walkable box
    int end0, end1;     // segment: feature ids of its endpoints (-1 if none)
    double nx() const { return kind == VSEG ? (double)s : 0.0; }
    double ny() const { return kind == HSEG ? (double)s : 0.0; }
    // signed distance for segments (valid flag = foot inside the extent)
    double dist(double x, double y, bool& valid, double tol) const {
        switch (kind) {
            case HSEG: valid = x >= lo - tol && x <= hi + tol; return s * (y - c);
            case VSEG: valid = y >= lo - tol && y <= hi + tol; return s * (x - c);
            default: valid = true; return std::hypot(x - px, y - py);
        }
    }
    void closest(double x, double y, double& qx, double& qy) const {
        switch (kind) {
            case HSEG: qx = std::min(std::max(x, lo), hi); qy = c; return;
            case VSEG: qx = c; qy = std::min(std::max(y, lo), hi); return;
            default: qx = px; qy = py; return;
        }
    }
};

struct Builder {
    const Scene& sc;
    std::vector<Feature> F;
    RectGrid rg;
    double tol_d = 1e-6;        // distance equality / emptiness tolerance (m)
    double tol_ext = 1e-9;      // segment extent tolerance

    explicit Builder(const Scene& s) : sc(s) {}

    void make_features() {
        // walkable box sides, facing inward: bottom, right, top, left
        auto seg = [&](int kind, double c, double lo, double hi, int s, int rect) {
            Feature f;
            f.kind = kind; f.c = c; f.lo = lo; f.hi = hi; f.s = s; f.px = f.py = 0; f.rect = rect;
            f.end0 = f.end1 = -1;
            F.push_back(f);
            return (int)F.size() - 1;
        };
        auto pt = [&](double x, double y, int rect) {
            Feature f;
            f.kind = POINT; f.c = f.lo = f.hi = 0; f.s = 0; f.px = x; f.py = y; f.rect = rect;
            f.end0 = f.end1 = -1;
            F.push_back(f);
            return (int)F.size() - 1;
        };
        seg(HSEG, sc.Y0, sc.X0, sc.X1, +1, -1);
        seg(VSEG, sc.X1, sc.Y0, sc.Y1, -1, -1);
        seg(HSEG, sc.Y1, sc.X0, sc.X1, -1, -1);
        seg(VSEG, sc.X0, sc.Y0, sc.Y1, +1, -1);
        for (int r = 0; r < (int)sc.rects.size(); ++r) {
            const RectD& q = sc.rects[r];
            int b = seg(HSEG, q.y0, q.x0, q.x1, -1, r);     // bottom
            int rt = seg(VSEG, q.x1, q.y0, q.y1, +1, r);    // right
            int t = seg(HSEG, q.y1, q.x0, q.x1, +1, r);     // top
            int l = seg(VSEG, q.x0, q.y0, q.y1, -1, r);     // left
            int c00 = pt(q.x0, q.y0, r), c10 = pt(q.x1, q.y0, r), c11 = pt(q.x1, q.y1, r), c01 = pt(q.x0, q.y1, r);
            F[b].end0 = c00; F[b].end1 = c10;
            F[rt].end0 = c10; F[rt].end1 = c11;
            F[t].end0 = c01; F[t].end1 = c11;
            F[l].end0 = c00; F[l].end1 = c01;
        }
    }

    // distance from (x,y) to the nearest obstacle or box side
    double clearance(double x, double y) const {
        double bx = std::min(std::min(x - sc.X0, sc.X1 - x), std::min(y - sc.Y0, sc.Y1 - y));
        double best;
        rg.nearest(x, y, bx, best);
        return std::min(best, bx);
    }

    // distance to the closed segment / point (1-Lipschitz everywhere)
    double closed_dist(int f, double x, double y) const {
        const Feature& q = F[f];
        double qx, qy;
        q.closest(x, y, qx, qy);
        return std::hypot(x - qx, y - qy);
    }

    bool is_endpoint(int s, int p) const {
        return F[s].kind != POINT && F[p].kind == POINT && (F[s].end0 == p || F[s].end1 == p);
    }

    // linear form of a segment's distance: D = a x + b y + e
    void lin(int f, double& a, double& b, double& e) const {
        a = F[f].nx(); b = F[f].ny();
        e = -(a * (F[f].kind == VSEG ? F[f].c : 0.0) + b * (F[f].kind == HSEG ? F[f].c : 0.0));
    }

    // points equidistant to the three features (before validation)
    void solve(int a, int b, int c, std::vector<std::pair<double, double>>& out) const {
        out.clear();
        int id[3] = {a, b, c};
        // segment + own endpoint: the point lies on the perpendicular at it
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                if (i == j || !is_endpoint(id[i], id[j])) continue;
                int g = id[3 - i - j];
                if (is_endpoint(g, id[j]) || is_endpoint(id[i], g)) return;
                const Feature& S = F[id[i]];
                const Feature& P = F[id[j]];
                double ux = S.nx(), uy = S.ny();         // P + t u, t >= 0
                double t;
                if (F[g].kind == POINT) {
                    double wx = P.px - F[g].px, wy = P.py - F[g].py;
                    double den = 2.0 * (ux * wx + uy * wy);
                    if (std::fabs(den) < 1e-15) return;
                    t = -(wx * wx + wy * wy) / den;
                } else {
                    double ga, gb, ge;
                    lin(g, ga, gb, ge);
                    double dg = ga * P.px + gb * P.py + ge;
                    double den = 1.0 - (ga * ux + gb * uy);
                    if (std::fabs(den) < 1e-15) return;
                    t = dg / den;
                }
                if (t > 0) out.push_back({P.px + t * ux, P.py + t * uy});
                return;
            }
        int pts[3], lns[3], np = 0, nl = 0;
        for (int q : id) (F[q].kind == POINT ? pts[np++] : lns[nl++]) = q;
        if (nl == 3) {
            double a1, b1, e1, a2, b2, e2, a3, b3, e3;
            lin(lns[0], a1, b1, e1); lin(lns[1], a2, b2, e2); lin(lns[2], a3, b3, e3);
            // (a1-a2) x + (b1-b2) y = e2 - e1 ; (a1-a3) x + (b1-b3) y = e3 - e1
            double A = a1 - a2, B = b1 - b2, C = e2 - e1, D = a1 - a3, E = b1 - b3, G = e3 - e1;
            double det = A * E - B * D;
            if (std::fabs(det) < 1e-12) return;
            out.push_back({(C * E - B * G) / det, (A * G - C * D) / det});
            return;
        }
        if (np == 3) {
            const Feature &P = F[pts[0]], &Q = F[pts[1]], &R = F[pts[2]];
            double A = 2 * (Q.px - P.px), B = 2 * (Q.py - P.py), C = Q.px * Q.px + Q.py * Q.py - P.px * P.px - P.py * P.py;
            double D = 2 * (R.px - P.px), E = 2 * (R.py - P.py), G = R.px * R.px + R.py * R.py - P.px * P.px - P.py * P.py;
            double det = A * E - B * D;
            if (std::fabs(det) < 1e-12) return;
            out.push_back({(C * E - B * G) / det, (A * G - C * D) / det});
            return;
        }
        // one quadratic: a line L (p0 + t d) intersected with |p - P| = D_s(p)
        double p0x, p0y, dx, dy;
        const Feature* P;
        int sline;
        if (nl == 2) {
            double a1, b1, e1, a2, b2, e2;
            lin(lns[0], a1, b1, e1); lin(lns[1], a2, b2, e2);
            double A = a1 - a2, B = b1 - b2, C = e2 - e1;   // A x + B y = C
            double nn = A * A + B * B;
            if (nn < 1e-18) return;
            p0x = A * C / nn; p0y = B * C / nn;
            double l = std::sqrt(nn);
            dx = -B / l; dy = A / l;
            P = &F[pts[0]];
            sline = lns[0];
        } else {
            const Feature &Pa = F[pts[0]], &Pb = F[pts[1]];
            double mx = 0.5 * (Pa.px + Pb.px), my = 0.5 * (Pa.py + Pb.py);
            double vx = Pb.px - Pa.px, vy = Pb.py - Pa.py, l = std::hypot(vx, vy);
            if (l < 1e-12) return;
            p0x = mx; p0y = my; dx = -vy / l; dy = vx / l;
            P = &Pa;
            sline = lns[0];
        }
        double la, lb, le;
        lin(sline, la, lb, le);
        double wx = p0x - P->px, wy = p0y - P->py;
        double E0 = la * p0x + lb * p0y + le, f = la * dx + lb * dy;
        double qa = 1.0 - f * f, qb = 2.0 * (wx * dx + wy * dy - E0 * f), qc = wx * wx + wy * wy - E0 * E0;
        if (std::fabs(qa) < 1e-12) {
            if (std::fabs(qb) < 1e-15) return;
            double t = -qc / qb;
            out.push_back({p0x + t * dx, p0y + t * dy});
            return;
        }
        double disc = qb * qb - 4 * qa * qc;
        if (disc < 0) {
            if (disc > -1e-9 * (qb * qb + 1)) disc = 0;
            else return;
        }
        double sq = std::sqrt(disc);
        double t1 = (-qb - sq) / (2 * qa), t2 = (-qb + sq) / (2 * qa);
        out.push_back({p0x + t1 * dx, p0y + t1 * dy});
        if (sq > 0) out.push_back({p0x + t2 * dx, p0y + t2 * dy});
    }

    bool in_free_space(double x, double y) const {
        if (x < sc.X0 - tol_d || x > sc.X1 + tol_d || y < sc.Y0 - tol_d || y > sc.Y1 + tol_d) return false;
        return true;
    }

    // feature set of a validated vertex: all features at distance r
    void features_at(double x, double y, double r, std::vector<int>& out) const {
        out.clear();
        for (int f = 0; f < 4; ++f) {
            bool v;
            double d = F[f].dist(x, y, v, tol_ext);
            if (v && std::fabs(d - r) <= tol_d) out.push_back(f);
        }
        static thread_local std::vector<int> near;
        rg.within(x, y, r + tol_d, near);
        for (int ri : near)
            for (int k = 0; k < 8; ++k) {
                int f = 4 + 8 * ri + k;
                bool v;
                double d = F[f].dist(x, y, v, tol_ext);
                if (v && std::fabs(d - r) <= tol_d) out.push_back(f);
            }
    }

    // ---- bisector of a feature pair: point(u), tangent(u), param of a point
    struct Bis {
        int kind = -1;        // 0 line, 1 parabola, 2 degenerate ray
        double ox, oy, dx, dy; // line: origin/dir; parabola: focus/tangent of directrix
        double nx, ny, h;      // parabola: directrix normal (toward focus side) and focus height
    };
    bool bisector(int a, int b, Bis& B) const {
        const Feature &fa = F[a], &fb = F[b];
        if (fa.kind != POINT && fb.kind != POINT) {
            double a1, b1, e1, a2, b2, e2;
            lin(a, a1, b1, e1); lin(b, a2, b2, e2);
            double A = a1 - a2, Bc = b1 - b2, C = e2 - e1;
            double nn = A * A + Bc * Bc;
            if (nn < 1e-18) return false;
            double l = std::sqrt(nn);
            B.kind = 0; B.ox = A * C / nn; B.oy = Bc * C / nn; B.dx = -Bc / l; B.dy = A / l;
            return true;
        }
        if (fa.kind == POINT && fb.kind == POINT) {
            double vx = fb.px - fa.px, vy = fb.py - fa.py, l = std::hypot(vx, vy);
            if (l < 1e-12) return false;
            B.kind = 0; B.ox = 0.5 * (fa.px + fb.px); B.oy = 0.5 * (fa.py + fb.py); B.dx = -vy / l; B.dy = vx / l;
            return true;
        }
        const Feature& P = fa.kind == POINT ? fa : fb;
        int s = fa.kind == POINT ? b : a;
        double la, lb, le;
        lin(s, la, lb, le);
        double h = la * P.px + lb * P.py + le;
        if (h < -1e-9) return false;
        B.ox = P.px; B.oy = P.py; B.nx = la; B.ny = lb; B.dx = -lb; B.dy = la; B.h = h;
        B.kind = h < 1e-9 ? 2 : 1;
        return true;
    }
    static void bpoint(const Bis& B, double u, double& x, double& y) {
        if (B.kind == 0) { x = B.ox + u * B.dx; y = B.oy + u * B.dy; return; }
        if (B.kind == 2) { x = B.ox + u * B.nx; y = B.oy + u * B.ny; return; }
        double v = (u * u - B.h * B.h) / (2 * B.h);
        x = B.ox + u * B.dx + v * B.nx;
        y = B.oy + u * B.dy + v * B.ny;
    }
    static void btangent(const Bis& B, double u, double& tx, double& ty) {
        if (B.kind == 0) { tx = B.dx; ty = B.dy; return; }
        if (B.kind == 2) { tx = B.nx; ty = B.ny; return; }
        tx = B.dx + (u / B.h) * B.nx;
        ty = B.dy + (u / B.h) * B.ny;
    }
    static double bparam(const Bis& B, double x, double y) {
        double wx = x - B.ox, wy = y - B.oy;
        if (B.kind == 2) return wx * B.nx + wy * B.ny;
        return wx * B.dx + wy * B.dy;
    }
};

inline uint64_t key3(int a, int b, int c) {
    int v[3] = {a, b, c};
    std::sort(v, v + 3);
    return ((uint64_t)v[0] << 42) | ((uint64_t)v[1] << 21) | (uint64_t)v[2];
}
inline uint64_t key2(int a, int b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)a << 32) | (uint32_t)b;
}

}  // namespace

void RectGrid::build(const std::vector<RectD>& r, double X0, double Y0, double X1, double Y1) {
    rects = &r;
    double ext = std::max(X1 - X0, Y1 - Y0);
    cell = std::max(4.0, ext / 512.0);
    inv = 1.0 / cell;
    x0 = X0; y0 = Y0;
    side = std::max(1, (int)std::ceil(ext / cell));
    std::vector<std::vector<int>> l((size_t)side * side);
    for (int i = 0; i < (int)r.size(); ++i) {
        int a = cx(r[i].x0), b = cx(r[i].x1), c = cy(r[i].y0), d = cy(r[i].y1);
        for (int y = c; y <= d; ++y)
            for (int x = a; x <= b; ++x) l[(size_t)y * side + x].push_back(i);
    }
    start.assign(l.size() + 1, 0);
    for (size_t i = 0; i < l.size(); ++i) start[i + 1] = start[i] + (int)l[i].size();
    ids.clear();
    for (auto& v : l) ids.insert(ids.end(), v.begin(), v.end());
}
int RectGrid::cx(double x) const { return std::min(side - 1, std::max(0, (int)std::floor((x - x0) * inv))); }
int RectGrid::cy(double y) const { return std::min(side - 1, std::max(0, (int)std::floor((y - y0) * inv))); }
double RectGrid::box_dist(const RectD& q, double x, double y) {
    double dx = std::max(std::max(q.x0 - x, 0.0), x - q.x1);
    double dy = std::max(std::max(q.y0 - y, 0.0), y - q.y1);
    return std::hypot(dx, dy);
}
int RectGrid::nearest(double x, double y, double limit, double& best) const {
    best = limit;
    int arg = -1;
    if (side == 0 || !rects || rects->empty()) return -1;
    int gx = cx(x), gy = cy(y);
    for (int r = 0; r <= side; ++r) {
        if (r > 0) {
            // every bucket outside the visited block is at least b away
            double b = 1e300;
            if (gx - r + 1 > 0) b = std::min(b, x - (x0 + (gx - r + 1) * cell));
            if (gy - r + 1 > 0) b = std::min(b, y - (y0 + (gy - r + 1) * cell));
            if (gx + r - 1 < side - 1) b = std::min(b, (x0 + (gx + r) * cell) - x);
            if (gy + r - 1 < side - 1) b = std::min(b, (y0 + (gy + r) * cell) - y);
            if (b >= best) break;
            if (gx - r + 1 <= 0 && gy - r + 1 <= 0 && gx + r - 1 >= side - 1 && gy + r - 1 >= side - 1) break;
        }
        auto visit = [&](int X, int Y) {
            if (X < 0 || Y < 0 || X >= side || Y >= side) return;
            size_t b = (size_t)Y * side + X;
            for (int q = start[b]; q < start[b + 1]; ++q) {
                int i = ids[q];
                double d = box_dist((*rects)[i], x, y);
                if (d < best || (d == best && arg >= 0 && i < arg)) { best = d; arg = i; }
            }
        };
        if (r == 0) visit(gx, gy);
        else {
            for (int X = gx - r; X <= gx + r; ++X) { visit(X, gy - r); visit(X, gy + r); }
            for (int Y = gy - r + 1; Y <= gy + r - 1; ++Y) { visit(gx - r, Y); visit(gx + r, Y); }
        }
    }
    return arg;
}
void RectGrid::within(double x, double y, double rad, std::vector<int>& out) const {
    out.clear();
    if (side == 0 || !rects) return;
    int a = cx(x - rad), b = cx(x + rad), c = cy(y - rad), d = cy(y + rad);
    for (int Y = c; Y <= d; ++Y)
        for (int X = a; X <= b; ++X) {
            size_t bk = (size_t)Y * side + X;
            for (int q = start[bk]; q < start[bk + 1]; ++q) {
                int i = ids[q];
                if (box_dist((*rects)[i], x, y) <= rad) out.push_back(i);
            }
        }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

void Scene::ensure_rgrid() {
    if (!rgrid_ready) { rgrid.build(rects, X0, Y0, X1, Y1); rgrid_ready = true; }
}
double Scene::clearance(double x, double y) const {
    double bx = std::min(std::min(x - X0, X1 - x), std::min(y - Y0, Y1 - y));
    double best;
    rgrid.nearest(x, y, bx, best);
    return std::min(best, bx);
}

int Scene::add_obstacle(const float* xy, int n) {
    if (n < 2) return -1;   // Environment.cpp:288-292
    std::vector<float> r(xy, xy + 2 * n);
    rings.push_back(r);
    // axis-aligned rectangle?
    if (n == 4) {
        float xs[4], ys[4];
        for (int i = 0; i < 4; ++i) { xs[i] = xy[2 * i]; ys[i] = xy[2 * i + 1]; }
        float a = *std::min_element(xs, xs + 4), b = *std::max_element(xs, xs + 4);
        float c = *std::min_element(ys, ys + 4), d = *std::max_element(ys, ys + 4);
        bool ok = a < b && c < d;
        for (int i = 0; i < 4 && ok; ++i) ok = (xs[i] == a || xs[i] == b) && (ys[i] == c || ys[i] == d);
        for (int i = 0; i < 4 && ok; ++i) {
            int j = (i + 1) % 4;
            ok = (xs[i] == xs[j]) != (ys[i] == ys[j]);
        }
        if (ok) {
            // the ECM built so far stays in use until the next build_ecm
            // (Environment::AddObstacle with updateECM false, Environment.cpp:285-315)
            rects.push_back({a, c, b, d});
            rgrid_ready = false;
            return 0;
        }
    }
    rect_only = false;
    return 0;
}

bool Scene::inside_obstacle(float x, float y) const {
    // Environment::InsideObstacle (Environment.cpp:275-283) with MathUtility::Contains
    // on each obstacle ring (UtilityFunctions.cpp:88-126).
    ecm::V2 p = ecm::mk(x, y);
    for (const auto& r : rings) {
        int m = (int)r.size() / 2;
        bool in = false;
        bool on_vertex = false;
        for (int i = 0; i < m && !on_vertex; ++i) {
            ecm::V2 a = ecm::mk(r[2 * i], r[2 * i + 1]);
            ecm::V2 b = ecm::mk(r[2 * ((i + 1) % m)], r[2 * ((i + 1) % m) + 1]);
            if (ecm::approx(p, a) || ecm::approx(p, b)) on_vertex = true;
            else in ^= ecm::ring_edge_toggles(p, a, b);
        }
        if (!on_vertex && in) return true;
    }
    return false;
}

int Scene::build_ecm(double h, int threads) {
    auto t_start = std::chrono::steady_clock::now();
    if (!rect_only) { report.message = "ECM construction supports axis-aligned rectangles only"; return -1; }
    if (!(X1 > X0) || !(Y1 > Y0)) { report.message = "empty box"; return -1; }
    for (const RectD& q : rects)
        if (q.x0 <= X0 || q.y0 <= Y0 || q.x1 >= X1 || q.y1 >= Y1) {
            report.message = "obstacles must lie strictly inside the walkable box";
            return -1;
        }
    if (threads <= 0) threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    threads = std::min(threads, 64);
    ensure_rgrid();
    Builder B(*this);
    B.make_features();
    B.rg.build(rects, X0, Y0, X1, Y1);
    const RectGrid& rg = B.rg;
    const int nF = (int)B.F.size();
    report.n_features = nF;

    // 1-2. candidate triples.  For a grid cell with centre c and half
    // diagonal rho, every feature that is nearest somewhere in the cell has
    // dist(c) <= dmin(c) + 2 rho (distances are 1-Lipschitz), so all triples of
    // such features cover every Voronoi vertex inside the cell.
    if (h <= 0) h = 1.0;
    const int nx = std::max(1, (int)std::ceil((X1 - X0) / h)), ny = std::max(1, (int)std::ceil((Y1 - Y0) / h));
    const double hx = (X1 - X0) / nx, hy = (Y1 - Y0) / ny;
    const double rho = 0.5 * std::hypot(hx, hy);
    std::vector<uint64_t> triples;
    {
        std::vector<std::vector<uint64_t>> part(threads);
        std::vector<std::thread> th;
        std::atomic<int> row{0};
        for (int t = 0; t < threads; ++t)
            th.emplace_back([&, t]() {
                auto& out = part[t];
                std::vector<int> near, fs;
                for (int j; (j = row++) < ny;) {
                    double y = Y0 + (j + 0.5) * hy;
                    for (int i = 0; i < nx; ++i) {
                        double x = X0 + (i + 0.5) * hx;
                        double dmin = B.clearance(x, y);
                        if (dmin < -rho) continue;
                        double lim = std::max(dmin, 0.0) + 2 * rho + 1e-9;
                        fs.clear();
                        for (int f = 0; f < 4; ++f) {
                            double d = B.closed_dist(f, x, y);
                            if (d <= lim) fs.push_back(f);
                        }
                        rg.within(x, y, lim, near);
                        for (int ri : near)
                            for (int k = 0; k < 8; ++k) {
                                int f = 4 + 8 * ri + k;
                                if (B.closed_dist(f, x, y) <= lim) fs.push_back(f);
                            }
                        const int n = (int)fs.size();
                        if (n < 3) continue;
                        for (int a = 0; a < n; ++a)
                            for (int b = a + 1; b < n; ++b)
                                for (int c = b + 1; c < n; ++c) out.push_back(key3(fs[a], fs[b], fs[c]));
                    }
                    if ((j & 63) == 0 && out.size() > (1u << 22)) {
                        std::sort(out.begin(), out.end());
                        out.erase(std::unique(out.begin(), out.end()), out.end());
                    }
                }
                std::sort(out.begin(), out.end());
                out.erase(std::unique(out.begin(), out.end()), out.end());
            });
        for (auto& t : th) t.join();
        for (auto& p : part) triples.insert(triples.end(), p.begin(), p.end());
        std::sort(triples.begin(), triples.end());
        triples.erase(std::unique(triples.begin(), triples.end()), triples.end());
    }
    // the four box corners: distance-0 vertices on the bisector of two sides
    struct Cand { double x, y, r; };
    std::vector<Cand> cand;
    // 3. solve + validate
    {
        std::vector<std::vector<Cand>> part(threads);
        std::vector<std::thread> th;
        std::atomic<size_t> next{0};
        for (int t = 0; t < threads; ++t)
            th.emplace_back([&, t]() {
                std::vector<std::pair<double, double>> sol;
                for (size_t q; (q = next.fetch_add(64)) < triples.size();) {
                    for (size_t w = q; w < std::min(triples.size(), q + 64); ++w) {
                        uint64_t k = triples[w];
                        int f[3] = {(int)(k >> 42), (int)((k >> 21) & 0x1fffff), (int)(k & 0x1fffff)};
                        B.solve(f[0], f[1], f[2], sol);
                        for (auto& s : sol) {
                            double x = s.first, y = s.second;
                            if (!std::isfinite(x) || !std::isfinite(y)) continue;
                            if (!B.in_free_space(x, y)) continue;
                            double r = -1;
                            bool ok = true;
                            for (int i = 0; i < 3 && ok; ++i) {
                                bool v;
                                double d = B.F[f[i]].dist(x, y, v, B.tol_ext);
                                ok = v && d >= -B.tol_d;
                                if (r < 0) r = d;
                                else ok = ok && std::fabs(d - r) <= B.tol_d * std::max(1.0, r);
                            }
                            if (!ok || r <= B.tol_d) continue;
                            if (B.clearance(x, y) < r - B.tol_d) continue;
                            part[t].push_back({x, y, r});
                        }
                    }
                }
            });
        for (auto& t : th) t.join();
        for (auto& p : part) cand.insert(cand.end(), p.begin(), p.end());
    }
    report.n_candidates = (int)cand.size();
    cand.push_back({X0, Y0, 0.0});
    cand.push_back({X1, Y0, 0.0});
    cand.push_back({X1, Y1, 0.0});
    cand.push_back({X0, Y1, 0.0});
    // merge coincident candidates
    std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
    std::vector<Cand> V;
    {
        std::vector<int> parent(cand.size());
        for (size_t i = 0; i < cand.size(); ++i) parent[i] = (int)i;
        std::function<int(int)> findp = [&](int i) { while (parent[i] != i) i = parent[i] = parent[parent[i]]; return i; };
        const double mt = 1e-6;
        for (size_t i = 0; i < cand.size(); ++i)
            for (size_t j = i + 1; j < cand.size() && cand[j].x - cand[i].x <= mt; ++j)
                if (std::fabs(cand[j].y - cand[i].y) <= mt) parent[findp((int)j)] = findp((int)i);
        std::vector<int> rep(cand.size(), -1);
        for (size_t i = 0; i < cand.size(); ++i) {
            int p = findp((int)i);
            if (rep[p] < 0) { rep[p] = (int)V.size(); V.push_back(cand[p]); }
        }
    }
    // vertex feature sets and pair lists
    const int nV = (int)V.size();
    std::vector<std::vector<int>> VF(nV);
    for (int v = 0; v < nV; ++v) {
        if (V[v].r == 0.0) {   // box corner
            int a = (V[v].y == Y0) ? 0 : 2;
            int b = (V[v].x == X0) ? 3 : 1;
            VF[v] = {std::min(a, b), std::max(a, b)};
        } else {
            B.features_at(V[v].x, V[v].y, V[v].r, VF[v]);
        }
    }
    std::unordered_map<uint64_t, std::vector<int>> pairs;
    pairs.reserve((size_t)nV * 3);
    for (int v = 0; v < nV; ++v) {
        auto& fs = VF[v];
        for (size_t i = 0; i < fs.size(); ++i)
            for (size_t j = i + 1; j < fs.size(); ++j) {
                int a = fs[i], b = fs[j];
                if (B.is_endpoint(a, b) || B.is_endpoint(b, a)) continue;       // secondary
                if (B.F[a].rect >= 0 && B.F[a].rect == B.F[b].rect) continue;   // same obstacle
                pairs[key2(a, b)].push_back(v);
            }
    }
    // 4. edges along each bisector
    struct E { int v0, v1, L, R; };
    std::vector<E> edges;
    for (auto& kv : pairs) {
        int a = (int)(kv.first >> 32), b = (int)(kv.first & 0xffffffffu);
        auto& vs = kv.second;
        if (vs.size() < 2) continue;
        Builder::Bis bs;
        if (!B.bisector(a, b, bs)) continue;
        std::vector<std::pair<double, int>> ord;
        for (int v : vs) ord.push_back({Builder::bparam(bs, V[v].x, V[v].y), v});
        std::sort(ord.begin(), ord.end());
        for (size_t i = 0; i + 1 < ord.size(); ++i) {
            double u0 = ord[i].first, u1 = ord[i + 1].first;
            if (u1 - u0 < 1e-9) continue;
            double um = 0.5 * (u0 + u1), mx, my;
            Builder::bpoint(bs, um, mx, my);
            bool va, vb;
            double da = B.F[a].dist(mx, my, va, B.tol_ext), db = B.F[b].dist(mx, my, vb, B.tol_ext);
            if (!va || !vb || da < 0 || std::fabs(da - db) > 1e-6 * std::max(1.0, da)) continue;
            if (!B.in_free_space(mx, my)) continue;
            if (B.clearance(mx, my) < da - B.tol_d) continue;
            double tx, ty, qx, qy;
            Builder::btangent(bs, um, tx, ty);
            B.F[a].closest(mx, my, qx, qy);
            bool a_left = tx * (qy - my) - ty * (qx - mx) > 0;
            edges.push_back({ord[i].second, ord[i + 1].second, a_left ? a : b, a_left ? b : a});
        }
    }
    // 5. keep vertices with edges; order vertices by (x, y), edges by vertices
    std::vector<int> deg(nV, 0);
    for (auto& e : edges) { deg[e.v0]++; deg[e.v1]++; }
    std::vector<int> keep;
    for (int v = 0; v < nV; ++v) if (deg[v] > 0) keep.push_back(v);
    std::sort(keep.begin(), keep.end(), [&](int a, int b) { return V[a].x < V[b].x || (V[a].x == V[b].x && V[a].y < V[b].y); });
    std::vector<int> newid(nV, -1);
    for (size_t i = 0; i < keep.size(); ++i) newid[keep[i]] = (int)i;
    for (auto& e : edges) { e.v0 = newid[e.v0]; e.v1 = newid[e.v1]; }
    std::sort(edges.begin(), edges.end(), [](const E& p, const E& q) {
        int a = std::min(p.v0, p.v1), b = std::min(q.v0, q.v1);
        if (a != b) return a < b;
        int c = std::max(p.v0, p.v1), d = std::max(q.v0, q.v1);
        return c < d;
    });
    report.deg_anomalies = 0;
    for (int v : keep) {
        bool corner = V[v].r == 0.0;
        if ((corner && deg[v] != 1) || (!corner && deg[v] < 2)) report.deg_anomalies++;
    }
    EcmGraph& G = ecm;
    const int nv = (int)keep.size(), ne = (int)edges.size();
    G.vxy.resize(2 * (size_t)nv);
    G.vclr.resize(nv);
    G.vhe.assign(nv, -1);
    G.he_target.resize(2 * (size_t)ne);
    G.he_next.assign(2 * (size_t)ne, -1);
    G.he_pts.resize(8 * (size_t)ne);
    for (int i = 0; i < nv; ++i) {
        G.vxy[2 * i] = (float)V[keep[i]].x;
        G.vxy[2 * i + 1] = (float)V[keep[i]].y;
        G.vclr[i] = (float)V[keep[i]].r;
    }
    for (int e = 0; e < ne; ++e) {
        const E& q = edges[e];
        const Cand& A = V[keep[q.v0]];
        const Cand& Bv = V[keep[q.v1]];
        double lx0, ly0, rx0, ry0, lx1, ly1, rx1, ry1;
        B.F[q.L].closest(A.x, A.y, lx0, ly0);
        B.F[q.R].closest(A.x, A.y, rx0, ry0);
        B.F[q.L].closest(Bv.x, Bv.y, lx1, ly1);
        B.F[q.R].closest(Bv.x, Bv.y, rx1, ry1);
        G.he_target[2 * e] = q.v1;       // he0: v0 -> v1, closest points at v0
        G.he_target[2 * e + 1] = q.v0;   // he1: v1 -> v0, closest points at v1
        float* P = &G.he_pts[8 * (size_t)e];
        P[0] = (float)lx0; P[1] = (float)ly0;   // he0.closest_left
        P[2] = (float)rx0; P[3] = (float)ry0;   // he0.closest_right
        P[4] = (float)rx1; P[5] = (float)ry1;   // he1.closest_left  (right feature)
        P[6] = (float)lx1; P[7] = (float)ly1;   // he1.closest_right (left feature)
    }
    // outgoing half-edges per vertex in counter-clockwise order
    std::vector<std::vector<std::pair<double, int>>> out(nv);
    for (int e = 0; e < ne; ++e) {
        for (int sd = 0; sd < 2; ++sd) {
            int src = G.he_target[2 * e + 1 - sd], dst = G.he_target[2 * e + sd];
            double ang = std::atan2((double)G.vxy[2 * dst + 1] - G.vxy[2 * src + 1], (double)G.vxy[2 * dst] - G.vxy[2 * src]);
            out[src].push_back({ang, 2 * e + sd});
        }
    }
    for (int v = 0; v < nv; ++v) {
        auto& o = out[v];
        std::sort(o.begin(), o.end());
        for (size_t i = 0; i < o.size(); ++i) G.he_next[o[i].second] = o[(i + 1) % o.size()].second;
        G.vhe[v] = o.empty() ? -1 : o[0].second;
    }
    index.build(G);
    has_ecm = true;
    report.n_vertices = nv;
    report.n_edges = ne;
    report.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return 0;
}

// ---------------------------------------------------------------- locate
void CellIndex::build(const EcmGraph& g) {
    const int ne = g.ne();
    ncell = 2 * ne;
    quad.resize(8 * (size_t)ncell);
    for (int e = 0; e < ne; ++e) {
        const float* v0 = &g.vxy[2 * g.he_target[2 * e + 1]];
        const float* v1 = &g.vxy[2 * g.he_target[2 * e]];
        const float* h = &g.he_pts[8 * (size_t)e];
        float L[8] = {v0[0], v0[1], h[0], h[1], h[6], h[7], v1[0], v1[1]};
        float R[8] = {v0[0], v0[1], h[2], h[3], h[4], h[5], v1[0], v1[1]};
        std::memcpy(&quad[16 * (size_t)e], L, sizeof(L));
        std::memcpy(&quad[16 * (size_t)e + 8], R, sizeof(R));
    }
    float bx0 = 1e30f, by0 = 1e30f, bx1 = -1e30f, by1 = -1e30f;
    for (int c = 0; c < ncell; ++c)
        for (int k = 0; k < 4; ++k) {
            bx0 = std::min(bx0, quad[8 * c + 2 * k]); bx1 = std::max(bx1, quad[8 * c + 2 * k]);
            by0 = std::min(by0, quad[8 * c + 2 * k + 1]); by1 = std::max(by1, quad[8 * c + 2 * k + 1]);
        }
    if (ncell == 0) { side = 0; return; }
    cell = std::max(1.0f, std::max(bx1 - bx0, by1 - by0) / 1024.0f);
    inv = 1.0f / cell;
    x0 = bx0 - 1.0f; y0 = by0 - 1.0f;
    side = (int)std::ceil((std::max(bx1 - bx0, by1 - by0) + 2.0f) * inv) + 1;
    std::vector<std::vector<int>> l((size_t)side * side);
    const float m = 0.01f;
    for (int c = 0; c < ncell; ++c) {
        float lx = 1e30f, ly = 1e30f, hx = -1e30f, hy = -1e30f;
        for (int k = 0; k < 4; ++k) {
            lx = std::min(lx, quad[8 * c + 2 * k]); hx = std::max(hx, quad[8 * c + 2 * k]);
            ly = std::min(ly, quad[8 * c + 2 * k + 1]); hy = std::max(hy, quad[8 * c + 2 * k + 1]);
        }
        int a = ecm::grid_cell(lx - m, x0, inv, side), b = ecm::grid_cell(hx + m, x0, inv, side);
        int cc = ecm::grid_cell(ly - m, y0, inv, side), d = ecm::grid_cell(hy + m, y0, inv, side);
        for (int y = cc; y <= d; ++y)
            for (int x = a; x <= b; ++x) l[(size_t)y * side + x].push_back(c);
    }
    start.assign(l.size() + 1, 0);
    for (size_t i = 0; i < l.size(); ++i) start[i + 1] = start[i] + (int)l[i].size();
    cells.clear();
    for (auto& v : l) cells.insert(cells.end(), v.begin(), v.end());
    std::vector<std::pair<uint32_t, int>> yc;
    for (int c = 0; c < ncell; ++c)
        for (int k = 0; k < 4; ++k) {
            float y = quad[8 * c + 2 * k + 1];
            if (y == 0.0f) y = 0.0f;
            uint32_t u;
            std::memcpy(&u, &y, 4);
            yc.push_back({u, c});
        }
    std::sort(yc.begin(), yc.end());
    yc.erase(std::unique(yc.begin(), yc.end()), yc.end());
    ykey.clear(); yoff.clear(); ycells.clear();
    for (size_t i = 0; i < yc.size();) {
        ykey.push_back(yc[i].first);
        yoff.push_back((int)ycells.size());
        size_t j = i;
        while (j < yc.size() && yc[j].first == yc[i].first) ycells.push_back(yc[j++].second);
        i = j;
    }
    yoff.push_back((int)ycells.size());
}

int CellIndex::locate(float x, float y) const {
    if (ncell == 0) return -1;
    ecm::V2 p = ecm::mk(x, y);
    int bx = ecm::grid_cell(x, x0, inv, side), by = ecm::grid_cell(y, y0, inv, side);
    size_t b = (size_t)by * side + bx;
    int i = start[b], ie = start[b + 1], j = 0, je = 0;
    float yy = (y == 0.0f) ? 0.0f : y;
    uint32_t u;
    std::memcpy(&u, &yy, 4);
    auto it = std::lower_bound(ykey.begin(), ykey.end(), u);
    if (y == y && it != ykey.end() && *it == u) {
        size_t k = it - ykey.begin();
        j = yoff[k]; je = yoff[k + 1];
    }
    while (i < ie || j < je) {
        int ci = i < ie ? cells[i] : 0x7fffffff, cj = j < je ? ycells[j] : 0x7fffffff, c;
        if (ci < cj) { c = ci; ++i; }
        else if (cj < ci) { c = cj; ++j; }
        else { c = ci; ++i; ++j; }
        const float* q = &quad[8 * (size_t)c];
        if (ecm::quad_contains(p, ecm::mk(q[0], q[1]), ecm::mk(q[2], q[3]), ecm::mk(q[4], q[5]), ecm::mk(q[6], q[7])))
            return c;
    }
    return -1;
}

}  // namespace ecm_host
