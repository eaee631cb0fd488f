// scene_plan.cpp -- path planning on the ECM, used at spawn and on replan
// requests of the step (Simulator.cpp:97-124, 579-587).
//
// Restates ECMPathPlanner::FindPath (ECMPathPlanner.cpp:22-136): locate and
// retract start and goal, A* over ECM vertices with a clearance filter
// (AStar.cpp:44-160), corridor (:146-157), shrink (:159-215), portals with
// arc sampling (:217-279), portal range fit (:281-321) and the simple
// stupid funnel (:323-411).  The A* open list is a binary heap with lazy
// deletion; the reference mutates keys of nodes already inside its
// std::priority_queue, which is not reproducible.
#include <algorithm>
#include <cmath>
#include <queue>
#include <vector>

#include "../csrc/ecm_geom.h"
#include "scene.h"

namespace ecm_host {

using ecm::V2;
using ecm::mk;

namespace {

inline V2 vtx(const EcmGraph& g, int i) { return mk(g.vxy[2 * i], g.vxy[2 * i + 1]); }
inline V2 hpt(const EcmGraph& g, int he, int which) {  // which: 0 closest_left, 1 closest_right
    int e = he >> 1, s = he & 1;
    const float* p = &g.he_pts[8 * (size_t)e + 4 * s + 2 * which];
    return mk(p[0], p[1]);
}
inline float dist(V2 a, V2 b) {    // MathUtility::Distance (UtilityFunctions.cpp:15-22)
    float dx = b.x - a.x, dy = b.y - a.y;
    return sqrtf(dx * dx + dy * dy);
}
// Point::operator== (ECMDataTypes.h:99-103)
inline bool peq(V2 p, V2 a) {
    return p.x > a.x - ecm::kEps && p.x < a.x + ecm::kEps && p.y > a.y - ecm::kEps && p.y < a.y + ecm::kEps;
}
// MathUtility::TriangleArea (UtilityFunctions.cpp:199-206)
inline float tri_area(V2 p1, V2 p2, V2 p3) {
    float ax = p2.x - p1.x, ay = p2.y - p1.y, bx = p3.x - p1.x, by = p3.y - p1.y;
    return bx * ay - ax * by;
}

struct Seg { V2 p0, p1; };

// ECM::RetractPoint (ECM.cpp:20-96) on the host, through the bucketed locate.
bool retract(const Scene& sc, V2 loc, V2& out, int& edge) {
    int cell = sc.index.locate(loc.x, loc.y);
    if (cell < 0) return false;
    const EcmGraph& g = sc.ecm;
    edge = cell >> 1;
    int e = edge;
    V2 p1 = vtx(g, g.he_target[2 * e + 1]), p2 = vtx(g, g.he_target[2 * e]);
    const float* h = &g.he_pts[8 * (size_t)e];
    V2 o1, o2, ray;
    if (ecm::left_of_segment(p1, p2, loc)) {
        o1 = mk(h[0], h[1]); o2 = mk(h[6], h[7]);
        if (ecm::approx(o1, o2)) ray = (p1 - o1) + (p2 - o1);
        else { V2 v = o2 - o1; ray = mk(v.y, -v.x); }
    } else {
        o1 = mk(h[2], h[3]); o2 = mk(h[4], h[5]);
        if (ecm::approx(o1, o2)) ray = (p1 - o1) + (p2 - o1);
        else { V2 v = o2 - o1; ray = mk(-v.y, v.x); }
    }
    ray = ecm::normalize(ray);
    return ecm::ray_segment(loc, ray, p1, p2, out);
}

// AStar::FindPath (AStar.cpp:44-160)
bool astar(const Scene& sc, V2 start, V2 goal, int startEdge, int goalEdge, float clearance, std::vector<int>& path) {
    const EcmGraph& g = sc.ecm;
    const int nv = g.nv();
    std::vector<float> gc(nv, ecm::kMaxFloat);
    std::vector<int> parent(nv, nv);
    std::vector<char> visited(nv, 0);
    typedef std::pair<float, int> QE;
    std::priority_queue<QE, std::vector<QE>, std::greater<QE>> open;
    int sa = g.he_target[2 * startEdge], sb = g.he_target[2 * startEdge + 1];
    gc[sa] = dist(start, vtx(g, sa));
    gc[sb] = dist(start, vtx(g, sb));
    open.push({gc[sa] + dist(vtx(g, sa), goal), sa});
    open.push({gc[sb] + dist(vtx(g, sb), goal), sb});
    int ga = g.he_target[2 * goalEdge], gb = g.he_target[2 * goalEdge + 1];
    while (!open.empty()) {
        int cur = open.top().second;
        open.pop();
        if (visited[cur]) continue;
        visited[cur] = 1;
        if (g.vclr[cur] < clearance) continue;
        if (cur == ga || cur == gb) {
            // AStar::ConstructPath (AStar.cpp:184-213)
            std::vector<int> rev;
            if (cur == ga) { rev.push_back(gb); rev.push_back(ga); }
            else { rev.push_back(ga); rev.push_back(gb); }
            for (int n = parent[cur]; n < nv; n = parent[n]) rev.push_back(n);
            path.assign(rev.rbegin(), rev.rend());
            return true;
        }
        int h0 = g.vhe[cur];
        if (h0 < 0) continue;
        int he = h0;
        V2 cp = vtx(g, cur);
        do {
            int nb = g.he_target[he];
            if (!visited[nb]) {
                V2 np = vtx(g, nb);
                float ng = gc[cur] + dist(cp, np);
                if (ng < gc[nb]) {
                    gc[nb] = ng;
                    parent[nb] = cur;
                    open.push({ng + dist(np, goal), nb});
                }
            }
            he = g.he_next[he];
        } while (he != h0 && he >= 0);
    }
    return false;
}

// ECMPathPlanner::SampleCorridorArc (ECMPathPlanner.cpp:255-279)
void sample_arc(V2 p1, V2 p2, V2 o1, V2 o2, V2 c, float radius, bool leftArc, std::vector<Seg>& portals) {
    const float maxLen = 10.0f;
    portals.push_back(leftArc ? Seg{p1, o1} : Seg{o1, p1});
    float el = ecm::len(p2 - p1);
    int ns = (int)std::ceil(el / maxLen);
    float sl = el / (float)ns;
    V2 dir = (p2 - p1) / el;
    for (int i = 0; i < ns; ++i) {
        V2 p = p1 + dir * sl * (float)i;
        V2 ad = ecm::normalize(p - c);
        p = c + ad * radius;
        portals.push_back(leftArc ? Seg{p, o2} : Seg{o2, p});
    }
}

}  // namespace

bool Scene::plan(float sx, float sy, float gx, float gy, float clearance, std::vector<float>& out) const {
    out.clear();
    if (!has_ecm) return false;
    V2 start = mk(sx, sy), goal = mk(gx, gy);
    if (index.locate(sx, sy) < 0 || index.locate(gx, gy) < 0) return false;
    V2 rs, rg;
    int es = -1, eg = -1;
    if (!retract(*this, start, rs, es)) return false;
    if (!retract(*this, goal, rg, eg)) return false;
    if (es == eg) {                               // ECMPathPlanner.cpp:144-150
        out = {sx, sy, gx, gy};
        return true;
    }
    std::vector<int> ap;
    if (!astar(*this, rs, rg, es, eg, clearance, ap)) return false;
    const EcmGraph& g = ecm;
    // edge path: outgoing half-edge from ap[i] to ap[i+1] (ECMPathPlanner.cpp:163-183)
    std::vector<int> hes;
    for (size_t i = 0; i + 1 < ap.size(); ++i) {
        int h0 = g.vhe[ap[i]], he = h0;
        if (h0 < 0) return false;
        do {
            if (g.he_target[he] == ap[i + 1]) { hes.push_back(he); break; }
            he = g.he_next[he];
        } while (he != h0);
    }
    if (hes.empty()) return false;
    // CreateCorridor (:146-157)
    const size_t n = hes.size();
    std::vector<V2> centers(n), lb(n), rb(n), lcb, rcb;
    std::vector<float> radii(n);
    for (size_t i = 0; i < n; ++i) {
        int src = g.he_target[hes[i] ^ 1];
        centers[i] = vtx(g, src);
        radii[i] = g.vclr[src];
        lb[i] = hpt(g, hes[i], 0);
        rb[i] = hpt(g, hes[i], 1);
    }
    // ShrinkCorridor (:159-215)
    enum { LINEAR, LEFT_ARC, RIGHT_ARC };
    std::vector<int> curve;
    auto shrink = [&](size_t i) {
        if (radii[i] < clearance) { lcb.push_back(centers[i]); rcb.push_back(centers[i]); }
        else {
            V2 ml = ecm::normalize(centers[i] - lb[i]), mr = ecm::normalize(centers[i] - rb[i]);
            lcb.push_back(lb[i] + ml * clearance);
            rcb.push_back(rb[i] + mr * clearance);
        }
    };
    for (size_t i = 0; i + 1 < n; ++i) {
        shrink(i);
        if (peq(lb[i], lb[i + 1])) curve.push_back(LEFT_ARC);
        else if (peq(rb[i], rb[i + 1])) curve.push_back(RIGHT_ARC);
        else curve.push_back(LINEAR);
    }
    shrink(n - 1);
    // TriangulateCorridor (:217-253)
    std::vector<Seg> portals;
    for (size_t i = 0; i + 1 < n; ++i) {
        switch (curve[i]) {
            case LINEAR:
                portals.push_back({lcb[i], rcb[i]});
                portals.push_back({lcb[i + 1], rcb[i]});
                break;
            case LEFT_ARC: sample_arc(lcb[i], lcb[i + 1], rcb[i], rcb[i + 1], lb[i], clearance, true, portals); break;
            default: sample_arc(rcb[i], rcb[i + 1], lcb[i], lcb[i + 1], rb[i], clearance, false, portals); break;
        }
    }
    portals.push_back({lcb.back(), rcb.back()});
    // FitPortalRange (:281-321)
    int first = 0;
    for (size_t i = 0; i < portals.size(); ++i)
        if (!ecm::left_of_segment(portals[i].p0, portals[i].p1, start)) { first = (int)i; break; }
    portals.erase(portals.begin(), portals.begin() + first);
    int last = (int)portals.size() - 1;
    for (int i = (int)portals.size() - 1; i >= 0; --i)
        if (ecm::left_of_segment(portals[i].p0, portals[i].p1, goal)) { last = i; break; }
    int toRemove = (int)portals.size() - last;
    for (int i = 0; i < toRemove && !portals.empty(); ++i) portals.pop_back();
    portals.push_back({goal, goal});
    // Funnel (:323-411)
    std::vector<V2> sp;
    V2 apex = start, pl = start, pr = start;
    int leftIdx = 0, rightIdx = 0, apexIdx = 0;
    sp.push_back(start);
    int guard = 0;
    for (int i = 0; i < (int)portals.size(); ++i) {
        if (++guard > 100000) return false;
        V2 l = portals[i].p0, r = portals[i].p1;
        if (tri_area(apex, pr, r) <= 0.0f) {
            if (ecm::approx(apex, pr) || tri_area(apex, pl, r) > 0.0f) { pr = r; rightIdx = i; }
            else {
                sp.push_back(pl);
                apex = pl; apexIdx = leftIdx;
                pl = apex; pr = apex; leftIdx = apexIdx; rightIdx = apexIdx;
                i = apexIdx;
                continue;
            }
        }
        if (tri_area(apex, pl, l) >= 0.0f) {
            if (ecm::approx(apex, pl) || tri_area(apex, pr, l) < 0.0f) { pl = l; leftIdx = i; }
            else {
                sp.push_back(pr);
                apex = pr; apexIdx = rightIdx;
                pl = apex; pr = apex; leftIdx = apexIdx; rightIdx = apexIdx;
                i = apexIdx;
                continue;
            }
        }
    }
    if (!ecm::approx(sp.back(), goal)) sp.push_back(goal);
    out.reserve(2 * sp.size());
    for (V2 p : sp) { out.push_back(p.x); out.push_back(p.y); }
    return true;
}

}  // namespace ecm_host
