// Spawn and goal areas of the C++17 Simulator facade, used as the reference's
// editor uses them (EnvironmentEditor.cpp, Command.cpp: AddSpawnArea,
// AddGoalArea, ConnectSpawnGoalAreas, then Update per frame), checked against
// the rule of UpdateSpawnAreas (Simulator.cpp:494-536): per connection,
// int(time * rate) agents per step with the time carried over, starts inside
// the spawn area, goals inside the goal area.  Exit code 0 when every check
// holds.
#include <cmath>
#include <cstdio>

#include "../../ecmgenerator_amd/host/Simulator.hpp"

using namespace ecm_amd::Simulation;

static int fails = 0;
static void expect(bool ok, const char* what) {
    std::printf("%s: %s\n", ok ? "ok  " : "FAIL", what);
    if (!ok) ++fails;
}

// agents the reference's accumulator spawns in `steps` steps at `rate`
static int expected_spawns(float rate, float dt, int steps) {
    float t = 0.0f;
    int total = 0;
    for (int i = 0; i < steps; ++i) {
        t += dt;
        const int n = (int)(t * rate);
        total += n;
        t -= (float)n / rate;
    }
    return total;
}

int main() {
    ecm_scene* scene = nullptr;
    ecm_scenario* sc = nullptr;
    if (ecm_scenario_generate_box(30.0f, 0, 1, 1, 1, 11, &scene, &sc) != 0) {
        std::printf("scenario failed: %s\n", ecm_last_error());
        return 2;
    }
    ecm_scenario_destroy(sc);
    const float dt = 0.1f;
    Simulator sim(scene, -30, -30, 30, 30, 512, dt);
    sim.Initialize();
    sim.SetAreaSeed(7);

    SpawnConfiguration cfg;
    cfg.clearanceMin = 0.25f;
    cfg.preferredSpeedMin = 1.3f;
    const int sa = sim.AddSpawnArea(-20.0f, -20.0f, 6.0f, 4.0f, cfg);
    const int sb = sim.AddSpawnArea(20.0f, -20.0f, 3.0f, 3.0f, cfg);
    const int ga = sim.AddGoalArea(20.0f, 20.0f, 5.0f, 5.0f);
    const int gb = sim.AddGoalArea(-20.0f, 20.0f, 2.0f, 2.0f);
    expect(sa == 0 && sb == 1 && ga == 0 && gb == 1, "area ids count from 0 per kind");
    expect(sim.GetSpawnArea(sa) && sim.GetGoalArea(gb) && !sim.GetSpawnArea(9), "area lookup");
    sim.ConnectSpawnGoalAreas(sa, ga, 10.0f);
    sim.ConnectSpawnGoalAreas(sa, ga, 99.0f);          // already connected: ignored
    sim.ConnectSpawnGoalAreas(sb, gb, 3.0f);
    expect(sim.GetSpawnArea(sa)->connectedGoalAreas.size() == 1 && sim.GetSpawnArea(sa)->spawnRate[0] == 10.0f,
           "a second connection to the same goal area is ignored");

    const int steps = 20;
    int last_n = 0, outside = 0;
    for (int f = 0; f < steps; ++f) {
        sim.Update(dt);
        // new agents start inside their spawn area (moved at most one step)
        const int n = sim.GetNumAgents();
        PositionComponent* p = sim.GetPositionData();
        bool* act = sim.GetActiveFlags();
        for (int i = last_n; i < n; ++i) {
            if (!act[i]) continue;
            const bool in_a = std::fabs(p[i].x + 20.0f) <= 6.5f && std::fabs(p[i].y + 20.0f) <= 4.5f;
            const bool in_b = std::fabs(p[i].x - 20.0f) <= 3.5f && std::fabs(p[i].y + 20.0f) <= 3.5f;
            outside += !(in_a || in_b);
        }
        last_n = n;
    }
    const int want = expected_spawns(10.0f, dt, steps) + expected_spawns(3.0f, dt, steps);
    std::printf("spawned %d agents in %d steps, accumulator rule gives %d\n", sim.GetNumAgents(), steps, want);
    expect(sim.GetNumAgents() == want, "agents spawned per the accumulator rule");
    expect(outside == 0, "every agent starts inside its spawn area");
    // goals: the attraction point heads for the goal area (straight paths in
    // an empty world), so the preferred velocity points up
    VelocityComponent* pv = sim.GetPreferredVelocityData();
    int up = 0, live = 0;
    bool* act = sim.GetActiveFlags();
    for (int i = 0; i < sim.GetLastIndex() + 1; ++i)
        if (act[i]) { ++live; up += pv[i].dy > 0.0f; }
    expect(live > 0 && up == live, "agents head for the goal areas");

    sim.DeconnectSpawnGoalAreas(sa, ga);
    expect(sim.GetSpawnArea(sa)->connectedGoalAreas.empty() && sim.GetSpawnArea(sa)->spawnRate.empty(),
           "deconnect removes the connection with its rate");
    sim.RemoveArea(SPAWN, sb);
    const int before = sim.GetNumAgents();
    for (int f = 0; f < 10; ++f) sim.Update(dt);
    expect(sim.GetNumAgents() == before, "no spawns after deconnect and remove");

    // a connection to a missing goal area, or with rate 0, spawns nothing
    const int sc2 = sim.AddSpawnArea(0.0f, -20.0f, 2.0f, 2.0f, cfg);
    sim.ConnectSpawnGoalAreas(sc2, 42, 50.0f);
    sim.ConnectSpawnGoalAreas(sc2, ga, 0.0f);
    for (int f = 0; f < 5; ++f) sim.Update(dt);
    expect(sim.GetNumAgents() == before, "missing goal area or zero rate: no spawns");

    // obstacle area with an ECM update (Simulator.cpp:383-412)
    int nr0 = 0, nv0 = 0, ev0 = 0, ee0 = 0;
    ecm_scene_obstacle_counts(scene, &nr0, &nv0);
    ecm_scene_ecm_counts(scene, &ev0, &ee0);
    const int oa = sim.AddObstacleArea(0.0f, 12.0f, 3.0f, 2.0f, true);
    const int ob = sim.AddObstacleArea(-12.0f, 12.0f, 1.0f, 1.0f, false);
    int nr1 = 0, nv1 = 0, ev1 = 0, ee1 = 0;
    ecm_scene_obstacle_counts(scene, &nr1, &nv1);
    ecm_scene_ecm_counts(scene, &ev1, &ee1);
    expect(oa == 0 && ob == 1 && sim.GetObstacleAreas().size() == 2, "obstacle area ids");
    expect(nr1 == nr0 + 2 && nv1 == nv0 + 8, "obstacle areas join the scene as 4-vertex rings");
    expect(ee1 != ee0, "the ECM was rebuilt with the first obstacle area");
    expect(ecm_scene_inside_obstacle(scene, 0.0f, 12.0f) == 1 && ecm_scene_inside_obstacle(scene, 0.0f, 16.0f) == 0,
           "the new obstacle covers its rectangle");
    int bad = 0;
    for (int f = 0; f < 20; ++f) sim.Update(dt);
    PositionComponent* p = sim.GetPositionData();
    act = sim.GetActiveFlags();
    for (int i = 0; i <= sim.GetLastIndex(); ++i)
        if (act[i] && !(std::isfinite(p[i].x) && std::isfinite(p[i].y))) ++bad;
    expect(bad == 0, "the simulation steps on after the obstacle areas");

    std::printf(fails ? "areas FAILED\n" : "areas ok\n");
    ecm_scene_destroy(scene);
    return fails ? 1 : 0;
}
