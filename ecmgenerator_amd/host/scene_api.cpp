// scene_api.cpp -- C ABI of the host scene (include/ecm_scene.h) and the
// seeded benchmark scenarios.
#include "../../include/ecm_scene.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "scene.h"

using ecm_host::RectD;
using ecm_host::Scene;

struct ecm_scene {
    Scene sc;
    ecm_scene(float a, float b, float c, float d) : sc(a, b, c, d) {}
};

struct ecm_scenario {
    std::vector<float> start, goal, radius, speed, path;
    std::vector<int> off, cnt;
};

namespace {
thread_local std::string g_scene_err;
int sfail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_scene_err = buf;
    std::fprintf(stderr, "ecm_scene: %s\n", buf);
    return code;
}
}  // namespace

extern "C" int ecm_scene_create(float x0, float y0, float x1, float y1, ecm_scene** out) {
    if (!out || !(x1 > x0) || !(y1 > y0)) return sfail(ECM_E_ARG, "bad box");
    *out = new ecm_scene(x0, y0, x1, y1);
    return 0;
}

extern "C" void ecm_scene_destroy(ecm_scene* s) { delete s; }

extern "C" int ecm_scene_add_obstacle(ecm_scene* s, const float* xy, int n) {
    if (!s || !xy || n < 2) return sfail(ECM_E_ARG, "obstacles need at least 2 vertices (Environment.cpp:288)");
    return s->sc.add_obstacle(xy, n) == 0 ? 0 : sfail(ECM_E_ARG, "add_obstacle failed");
}

extern "C" int ecm_scene_build_ecm(ecm_scene* s, float resolution, int threads) {
    if (!s) return sfail(ECM_E_ARG, "null scene");
    int rc = s->sc.build_ecm(resolution, threads);
    if (rc) return sfail(ECM_E_STATE, "ECM build failed: %s", s->sc.report.message.c_str());
    return 0;
}

extern "C" int ecm_scene_ecm_counts(ecm_scene* s, int* nv, int* ne) {
    if (!s || !nv || !ne) return sfail(ECM_E_ARG, "null argument");
    if (!s->sc.has_ecm) return sfail(ECM_E_STATE, "no ECM built");
    *nv = s->sc.ecm.nv();
    *ne = s->sc.ecm.ne();
    return 0;
}

extern "C" int ecm_scene_ecm_export(ecm_scene* s, float* vxy, float* vclr, int* tgt, int* nxt, float* pts) {
    if (!s) return sfail(ECM_E_ARG, "null scene");
    if (!s->sc.has_ecm) return sfail(ECM_E_STATE, "no ECM built");
    const auto& g = s->sc.ecm;
    if (vxy) std::memcpy(vxy, g.vxy.data(), g.vxy.size() * sizeof(float));
    if (vclr) std::memcpy(vclr, g.vclr.data(), g.vclr.size() * sizeof(float));
    if (tgt) std::memcpy(tgt, g.he_target.data(), g.he_target.size() * sizeof(int));
    if (nxt) std::memcpy(nxt, g.he_next.data(), g.he_next.size() * sizeof(int));
    if (pts) std::memcpy(pts, g.he_pts.data(), g.he_pts.size() * sizeof(float));
    return 0;
}

extern "C" int ecm_scene_obstacle_counts(ecm_scene* s, int* nr, int* nv) {
    if (!s || !nr || !nv) return sfail(ECM_E_ARG, "null argument");
    *nr = (int)s->sc.rings.size();
    int t = 0;
    for (auto& r : s->sc.rings) t += (int)r.size() / 2;
    *nv = t;
    return 0;
}

extern "C" int ecm_scene_obstacle_export(ecm_scene* s, float* xy, int* ring_off) {
    if (!s || !xy || !ring_off) return sfail(ECM_E_ARG, "null argument");
    int o = 0;
    ring_off[0] = 0;
    for (size_t i = 0; i < s->sc.rings.size(); ++i) {
        const auto& r = s->sc.rings[i];
        std::memcpy(xy + 2 * o, r.data(), r.size() * sizeof(float));
        o += (int)r.size() / 2;
        ring_off[i + 1] = o;
    }
    return 0;
}

extern "C" int ecm_scene_locate(ecm_scene* s, float x, float y) {
    if (!s || !s->sc.has_ecm) return -1;
    return s->sc.index.locate(x, y);
}

extern "C" int ecm_scene_inside_obstacle(ecm_scene* s, float x, float y) {
    if (!s) return 0;
    return s->sc.inside_obstacle(x, y) ? 1 : 0;
}

extern "C" int ecm_scene_plan_path(ecm_scene* s, float sx, float sy, float gx, float gy, float clearance,
                                   float* out_xy, int cap, int* n_out) {
    if (!s || !out_xy || !n_out) return sfail(ECM_E_ARG, "null argument");
    *n_out = 0;
    std::vector<float> p;
    if (!s->sc.plan(sx, sy, gx, gy, clearance, p)) return 1;
    int n = (int)p.size() / 2;
    if (n > cap) return sfail(ECM_E_ARG, "path of %d points exceeds capacity %d", n, cap);
    std::memcpy(out_xy, p.data(), p.size() * sizeof(float));
    *n_out = n;
    return 0;
}

extern "C" int ecm_scene_plan_paths(ecm_scene* s, int n, const float* st, const float* go, const float* clr,
                                    int threads, int* cnt, float* pts, long long cap, int* off, int) {
    if (!s || n < 0 || (n > 0 && (!st || !go || !clr || !cnt || !pts || !off))) return sfail(ECM_E_ARG, "null argument");
    if (threads <= 0) threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::vector<float>> res(n);
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t)
        th.emplace_back([&]() {
            for (int i; (i = next++) < n;)
                if (!s->sc.plan(st[2 * i], st[2 * i + 1], go[2 * i], go[2 * i + 1], clr[i], res[i])) res[i].clear();
        });
    for (auto& t : th) t.join();
    long long o = 0;
    for (int i = 0; i < n; ++i) {
        int c = (int)res[i].size() / 2;
        if (o + c > cap) return sfail(ECM_E_ARG, "path buffer too small");
        off[i] = (int)o;
        cnt[i] = c;
        std::memcpy(pts + 2 * o, res[i].data(), res[i].size() * sizeof(float));
        o += c;
    }
    return 0;
}

namespace {
int replan_cb(void* user, int, float x, float y, float gx, float gy, float clearance, float* out, int cap, int* n) {
    ecm_scene* s = (ecm_scene*)user;
    *n = 0;
    std::vector<float> p;
    if (!s || !s->sc.plan(x, y, gx, gy, clearance, p)) return 1;
    int m = (int)p.size() / 2;
    if (m > cap) return 1;
    std::memcpy(out, p.data(), p.size() * sizeof(float));
    *n = m;
    return 0;
}
}  // namespace

extern "C" ecm_replan_fn ecm_scene_replan_callback(void) { return &replan_cb; }

// ------------------------------------------------------------- scenarios
namespace {

struct Cfg {
    double half;
    int n_rects;
    int side_lo, side_hi;
    int n_agents;
    bool barrier;
};

Cfg config_of(int c) {
    switch (c) {
        case 0: return {100.0, 40, 2, 15, 5000, false};
        case 1: return {500.0, 400, 2, 25, 100000, false};
        case 2: return {1000.0, 1437, 2, 15, 1000000, true};
        default: return {30.0, 6, 2, 8, 500, false};
    }
}

bool overlaps(const RectD& a, const RectD& b, double gap) {
    return !(a.x1 + gap <= b.x0 || b.x1 + gap <= a.x0 || a.y1 + gap <= b.y0 || b.y1 + gap <= a.y0);
}

void add_rect(Scene& sc, const RectD& r) {
    // CCW ring as the editor builds it (Simulator.cpp:404-407)
    float xy[8] = {(float)r.x1, (float)r.y1, (float)r.x0, (float)r.y1, (float)r.x0, (float)r.y0, (float)r.x1, (float)r.y0};
    sc.add_obstacle(xy, 4);
}

// Dense bottleneck: two opposing flows start in bands against the barrier
// (x in [-2 - depth, -2] and [2, 2 + depth]) at kBandDensity agents/m^2 and
// walk to goals in the band across (22 to 22 + depth m on the other side, y
// within 100 m of the start), so every one of the 64 gaps of 8 m carries a
// counterflow and the crowd in front of the barrier compresses within the
// first seconds of the run.  Agents alternate between the flows.  A start
// keeps the reference rule (no earlier agent within r, Simulator.cpp:295-311)
// and, stricter, overlaps no earlier agent (distance >= r_i + r_j), so the
// run starts without collisions; start and goal are at least 2r from every
// obstacle.  Candidates are drawn one at a time against a 1 m grid of the
// accepted starts; paths are planned afterwards on up to 16 threads and an
// agent without a path is replaced in the next round.
constexpr double kBandDensity = 2.0;   // 1M agents: bands 125 m deep

int spawn_bottleneck_flows(const Scene& es, const Cfg& cf, std::mt19937_64& rng, int n_agents, ecm_scenario& sc) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> Nspd(1.3, 0.2);
    const double H = cf.half;
    const double kBand = std::min(H - 30.0, std::max(10.0, n_agents / (2.0 * 2.0 * H * kBandDensity)));
    const double acell = 1.0;
    const int aside = (int)std::ceil(2 * H / acell) + 1;
    std::vector<std::vector<int>> agrid((size_t)aside * aside);
    auto aidx = [&](double v) { return std::min(aside - 1, std::max(0, (int)std::floor((v + H) / acell))); };
    std::vector<float> st, go, rr, sp;
    std::vector<char> dead;
    auto free_start = [&](double x, double y, double r) {
        int gx = aidx(x), gy = aidx(y);
        for (int yy = std::max(0, gy - 1); yy <= std::min(aside - 1, gy + 1); ++yy)
            for (int xx = std::max(0, gx - 1); xx <= std::min(aside - 1, gx + 1); ++xx)
                for (int j : agrid[(size_t)yy * aside + xx]) {
                    if (dead[j]) continue;
                    double dx = x - st[2 * j], dy = y - st[2 * j + 1], rs = r + rr[j];
                    if (dx * dx + dy * dy < rs * rs) return false;
                }
        return true;
    };
    int have = 0;
    for (int round = 0; round < 20 && have < n_agents; ++round) {
        const size_t first = rr.size();
        for (int i = 0; i < n_agents - have; ++i) {
            const int side = (int)((first + i) & 1);      // 0: left flow walking +x, 1: right flow walking -x
            const double sgn = side == 0 ? -1.0 : 1.0;
            double r = 0.2 + 0.1 * U(rng);
            double s = std::min(1.8, std::max(0.8, Nspd(rng)));
            double sx = 0, sy = 0, gx = 0, gy = 0;
            bool ok = false;
            for (int t = 0; t < 20000 && !ok; ++t) {
                sx = sgn * (2.0 + kBand * U(rng));
                sy = -H + 2 * H * U(rng);
                ok = es.clearance(sx, sy) >= 2 * r && free_start(sx, sy, r);
            }
            if (!ok) return sfail(ECM_E_STATE, "bottleneck band full after %zu agents", rr.size());
            ok = false;
            for (int t = 0; t < 1000 && !ok; ++t) {
                gx = -sgn * (22.0 + kBand * U(rng));
                gy = std::min(H - 2.0, std::max(-H + 2.0, sy + 200.0 * (U(rng) - 0.5)));
                ok = es.clearance(gx, gy) >= 2 * r;
            }
            if (!ok) return sfail(ECM_E_STATE, "no free goal for agent %zu", rr.size());
            agrid[(size_t)aidx(sy) * aside + aidx(sx)].push_back((int)rr.size());
            st.push_back((float)sx); st.push_back((float)sy);
            go.push_back((float)gx); go.push_back((float)gy);
            rr.push_back((float)r);
            sp.push_back((float)s);
            dead.push_back(0);
        }
        const int m = (int)(rr.size() - first);
        std::vector<std::vector<float>> paths(m);
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&]() {
                for (int i; (i = next++) < m;) {
                    size_t a = first + i;
                    if (!es.plan(st[2 * a], st[2 * a + 1], go[2 * a], go[2 * a + 1], rr[a], paths[i])) paths[i].clear();
                }
            });
        for (auto& t : th) t.join();
        for (int i = 0; i < m; ++i) {
            size_t a = first + i;
            if (paths[i].size() < 4) {
                dead[a] = 1;
                continue;
            }
            sc.start.push_back(st[2 * a]); sc.start.push_back(st[2 * a + 1]);
            sc.goal.push_back(go[2 * a]); sc.goal.push_back(go[2 * a + 1]);
            sc.radius.push_back(rr[a]);
            sc.speed.push_back(sp[a]);
            sc.off.push_back((int)(sc.path.size() / 2));
            sc.cnt.push_back((int)paths[i].size() / 2);
            sc.path.insert(sc.path.end(), paths[i].begin(), paths[i].end());
            ++have;
        }
    }
    if (have < n_agents) return sfail(ECM_E_STATE, "could only place %d of %d agents", have, n_agents);
    return 0;
}

}  // namespace

static int generate(Cfg cf, uint64_t seed, int n_agents, ecm_scene** scene_out, ecm_scenario** out);

extern "C" int ecm_scenario_generate(int config, uint64_t seed, int n_agents, ecm_scene** scene_out, ecm_scenario** out) {
    if (!scene_out || !out) return sfail(ECM_E_ARG, "null argument");
    return generate(config_of(config), seed, n_agents, scene_out, out);
}

extern "C" int ecm_scenario_generate_box(float half, int n_rects, int side_lo, int side_hi, int n_agents, uint64_t seed,
                                         ecm_scene** scene_out, ecm_scenario** out) {
    if (!scene_out || !out) return sfail(ECM_E_ARG, "null argument");
    if (!(half >= 10.0f) || n_rects < 0 || side_lo < 1 || side_hi < side_lo || n_agents < 1)
        return sfail(ECM_E_ARG, "bad scenario arguments");
    Cfg cf{(double)half, n_rects, side_lo, side_hi, n_agents, false};
    return generate(cf, seed, n_agents, scene_out, out);
}

static int generate(Cfg cf, uint64_t seed, int n_agents, ecm_scene** scene_out, ecm_scenario** out) {
    if (n_agents <= 0) n_agents = cf.n_agents;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double H = cf.half, gap = 2.0;
    std::vector<RectD> rects;
    // barrier: 63 blocks 4 m thick on x in [-2, 2], 64 gaps of 8 m
    // (blocks alternate 24 m and 23 m so that 64*8 + 39*24 + 24*23 = 2000 m)
    if (cf.barrier) {
        double y = -H + 8.0;
        for (int b = 0; b < 63; ++b) {
            double len = (b % 8 < 5) ? 24.0 : 23.0;
            if (b == 62) len = H - 8.0 - y;
            rects.push_back({-2.0, y, 2.0, y + len});
            y += len + 8.0;
        }
    }
    const size_t n_fixed = rects.size();
    // uniform grid over rectangles for the overlap test
    const double gcell = 32.0;
    const int gside = (int)std::ceil(2 * H / gcell) + 1;
    std::vector<std::vector<int>> grid((size_t)gside * gside);
    auto gidx = [&](double v) { return std::min(gside - 1, std::max(0, (int)std::floor((v + H) / gcell))); };
    auto insert = [&](int i) {
        for (int y = gidx(rects[i].y0 - gap); y <= gidx(rects[i].y1 + gap); ++y)
            for (int x = gidx(rects[i].x0 - gap); x <= gidx(rects[i].x1 + gap); ++x) grid[(size_t)y * gside + x].push_back(i);
    };
    for (size_t i = 0; i < rects.size(); ++i) insert((int)i);
    long tries = 0;
    while ((int)(rects.size() - n_fixed) < cf.n_rects && tries < 5000000) {
        ++tries;
        int w = cf.side_lo + (int)std::floor(U(rng) * (cf.side_hi - cf.side_lo + 1));
        int h = cf.side_lo + (int)std::floor(U(rng) * (cf.side_hi - cf.side_lo + 1));
        double x0 = std::floor(-H + gap + U(rng) * (2 * H - 2 * gap - w));
        double y0 = std::floor(-H + gap + U(rng) * (2 * H - 2 * gap - h));
        RectD r{x0, y0, x0 + w, y0 + h};
        if (r.x1 > H - gap || r.y1 > H - gap || r.x0 < -H + gap || r.y0 < -H + gap) continue;
        if (cf.barrier && r.x1 > -10.0 && r.x0 < 10.0) continue;   // keep the corridors 8 m wide
        bool bad = false;
        for (int y = gidx(r.y0 - gap); y <= gidx(r.y1 + gap) && !bad; ++y)
            for (int x = gidx(r.x0 - gap); x <= gidx(r.x1 + gap) && !bad; ++x)
                for (int q : grid[(size_t)y * gside + x])
                    if (overlaps(r, rects[q], gap)) { bad = true; break; }
        if (bad) continue;
        rects.push_back(r);
        insert((int)rects.size() - 1);
    }
    auto* es = new ecm_scene((float)-H, (float)-H, (float)H, (float)H);
    for (auto& r : rects) add_rect(es->sc, r);
    if (es->sc.build_ecm(1.0, 0) != 0) {
        std::string m = es->sc.report.message;
        delete es;
        return sfail(ECM_E_STATE, "ECM build failed: %s", m.c_str());
    }
    es->sc.ensure_rgrid();
    if (cf.barrier) {
        auto* sc = new ecm_scenario();
        int rc = spawn_bottleneck_flows(es->sc, cf, rng, n_agents, *sc);
        if (rc) {
            delete sc;
            delete es;
            return rc;
        }
        *scene_out = es;
        *out = sc;
        return 0;
    }
    // agents
    auto* sc = new ecm_scenario();
    std::normal_distribution<double> Nspd(1.3, 0.2);
    const double acell = 1.0;
    const int aside = (int)std::ceil(2 * H / acell) + 1;
    std::vector<std::vector<int>> agrid((size_t)aside * aside);
    auto aidx = [&](double v) { return std::min(aside - 1, std::max(0, (int)std::floor((v + H) / acell))); };
    int have = 0, rounds = 0;
    while (have < n_agents && rounds < 50) {
        ++rounds;
        int want = n_agents - have;
        int batch = std::max(want + want / 8 + 16, 64);
        std::vector<float> st, go, rr, sp;
        for (int i = 0; i < batch; ++i) {
            double r = 0.2 + 0.1 * U(rng);
            double s = std::min(1.8, std::max(0.8, Nspd(rng)));
            double sx, sy, gx, gy;
            auto draw = [&](double& x, double& y) {
                for (int t = 0; t < 1000; ++t) {
                    x = -H + 2 * H * U(rng);
                    y = -H + 2 * H * U(rng);
                    if (es->sc.clearance(x, y) >= 2 * r) return true;
                }
                return false;
            };
            if (!(draw(sx, sy) && draw(gx, gy))) continue;
            st.push_back((float)sx); st.push_back((float)sy);
            go.push_back((float)gx); go.push_back((float)gy);
            rr.push_back((float)r);
            sp.push_back((float)s);
        }
        int m = (int)rr.size();
        std::vector<std::vector<float>> paths(m);
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        int nt = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&]() {
                for (int i; (i = next++) < m;)
                    if (!es->sc.plan(st[2 * i], st[2 * i + 1], go[2 * i], go[2 * i + 1], rr[i], paths[i])) paths[i].clear();
            });
        for (auto& t : th) t.join();
        for (int i = 0; i < m && have < n_agents; ++i) {
            if (paths[i].size() < 4) continue;
            float x = st[2 * i], y = st[2 * i + 1], r = rr[i];
            // ValidSpawnLocation (Simulator.cpp:295-311): no earlier agent within r
            bool ok = true;
            int gx = aidx(x), gy = aidx(y);
            for (int yy = std::max(0, gy - 1); yy <= std::min(aside - 1, gy + 1) && ok; ++yy)
                for (int xx = std::max(0, gx - 1); xx <= std::min(aside - 1, gx + 1) && ok; ++xx)
                    for (int j : agrid[(size_t)yy * aside + xx]) {
                        float dx = x - sc->start[2 * j], dy = y - sc->start[2 * j + 1];
                        if (dx * dx + dy * dy < r * r) { ok = false; break; }
                    }
            if (!ok) continue;
            agrid[(size_t)gy * aside + gx].push_back(have);
            sc->start.push_back(x); sc->start.push_back(y);
            sc->goal.push_back(go[2 * i]); sc->goal.push_back(go[2 * i + 1]);
            sc->radius.push_back(r);
            sc->speed.push_back(sp[i]);
            sc->off.push_back((int)(sc->path.size() / 2));
            sc->cnt.push_back((int)paths[i].size() / 2);
            sc->path.insert(sc->path.end(), paths[i].begin(), paths[i].end());
            ++have;
        }
    }
    if (have < n_agents) {
        delete sc;
        delete es;
        return sfail(ECM_E_STATE, "could only place %d of %d agents", have, n_agents);
    }
    *scene_out = es;
    *out = sc;
    return 0;
}

extern "C" int ecm_scenario_counts(ecm_scenario* sc, int* n, long long* np) {
    if (!sc || !n || !np) return sfail(ECM_E_ARG, "null argument");
    *n = (int)sc->radius.size();
    *np = (long long)(sc->path.size() / 2);
    return 0;
}

extern "C" int ecm_scenario_export(ecm_scenario* sc, float* st, float* go, float* r, float* s, int* off, int* cnt,
                                   float* path) {
    if (!sc) return sfail(ECM_E_ARG, "null argument");
    if (st) std::memcpy(st, sc->start.data(), sc->start.size() * sizeof(float));
    if (go) std::memcpy(go, sc->goal.data(), sc->goal.size() * sizeof(float));
    if (r) std::memcpy(r, sc->radius.data(), sc->radius.size() * sizeof(float));
    if (s) std::memcpy(s, sc->speed.data(), sc->speed.size() * sizeof(float));
    if (off) std::memcpy(off, sc->off.data(), sc->off.size() * sizeof(int));
    if (cnt) std::memcpy(cnt, sc->cnt.data(), sc->cnt.size() * sizeof(int));
    if (path) std::memcpy(path, sc->path.data(), sc->path.size() * sizeof(float));
    return 0;
}

extern "C" void ecm_scenario_destroy(ecm_scenario* sc) { delete sc; }
