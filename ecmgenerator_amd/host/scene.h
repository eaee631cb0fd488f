// scene.h -- host-side environment: walkable box, rectangular obstacles, the
// Explicit Corridor Map built from them, and the path planner over it.
// Mirrors ECMGenerator/Environment.h:18-85, ECM.h:52-138 and
// ECMPathPlanner.h:46-69 (the parts the crowd step consumes).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ecm_host {

struct RectD {
    double x0, y0, x1, y1;
};

// ECM graph in the reference's layout (ECM.h:20-50).
struct EcmGraph {
    std::vector<float> vxy;          // [2 nv] vertex positions
    std::vector<float> vclr;         // [nv]   vertex clearance
    std::vector<int> vhe;            // [nv]   an outgoing half-edge (edge*2 + side)
    std::vector<int> he_target;      // [2 ne] v_target of he0, he1
    std::vector<int> he_next;        // [2 ne] next outgoing half-edge around the source
    std::vector<float> he_pts;       // [8 ne] he0.cl, he0.cr, he1.cl, he1.cr
    int nv() const { return (int)vclr.size(); }
    int ne() const { return (int)(he_target.size() / 2); }
};

// Bucketed first-match cell locate (same answer as the linear scan of
// ECMCellCollection::PointLocationQueryLinear, ECMCellCollection.cpp:57-90).
struct CellIndex {
    float x0 = 0, y0 = 0, cell = 1, inv = 1;
    int side = 0;
    std::vector<int> start, cells;          // bucket CSR (ascending cell ids)
    std::vector<uint32_t> ykey;             // sorted distinct vertex y bit patterns
    std::vector<int> yoff, ycells;          // CSR per ykey (ascending cell ids)
    std::vector<float> quad;                // [8 nc] ring q0..q3 per cell
    int ncell = 0;
    void build(const EcmGraph& g);
    int locate(float x, float y) const;
};

// Uniform bucket grid over rectangles: nearest box distance queries.
struct RectGrid {
    double x0 = 0, y0 = 0, cell = 1, inv = 1;
    int side = 0;
    std::vector<int> start, ids;
    const std::vector<RectD>* rects = nullptr;
    void build(const std::vector<RectD>& r, double X0, double Y0, double X1, double Y1);
    int cx(double x) const;
    int cy(double y) const;
    static double box_dist(const RectD& q, double x, double y);
    // nearest rectangle closer than `limit` (-1 if none); best = its distance
    int nearest(double x, double y, double limit, double& best) const;
    // ids of rectangles whose box is within `rad` of (x, y), ascending
    void within(double x, double y, double rad, std::vector<int>& out) const;
};

struct BuildReport {
    int n_features = 0, n_candidates = 0, n_vertices = 0, n_edges = 0;
    int deg_anomalies = 0;      // vertices whose degree is 0 or an unexpected 1
    double seconds = 0;
    std::string message;
};

class Scene {
   public:
    double X0, Y0, X1, Y1;                  // walkable box
    std::vector<std::vector<float>> rings;  // obstacle rings (CCW), as given
    std::vector<RectD> rects;               // the same, as rectangles
    bool rect_only = true;
    EcmGraph ecm;
    CellIndex index;
    bool has_ecm = false;
    BuildReport report;

    Scene(double x0, double y0, double x1, double y1) : X0(x0), Y0(y0), X1(x1), Y1(y1) {}
    int add_obstacle(const float* xy, int n);
    int build_ecm(double resolution, int threads);
    bool inside_obstacle(float x, float y) const;   // Environment::InsideObstacle (Environment.cpp:275-283)
    // distance to the nearest rectangle or box side (rectangle scenes)
    double clearance(double x, double y) const;
    RectGrid rgrid;
    bool rgrid_ready = false;
    void ensure_rgrid();

    // ECMPathPlanner::FindPath (ECMPathPlanner.cpp:22-136); false when no path.
    bool plan(float sx, float sy, float gx, float gy, float clearance, std::vector<float>& out) const;
};

}  // namespace ecm_host
