// Simulator.hpp -- C++17 drop-in shape of ECM::Simulation::Simulator
// (ECMAgentSimulator/Simulator.h:59-188) over the C ABI of libecmcrowd.so.
//
// Header-only.  Method names, component structs and getter semantics follow
// the reference: SpawnAgent returns the slot or -1, DestroyAgent frees the
// slot (LIFO pool), Update(dt) runs one step (dt ignored, the step time is
// fixed at construction, Simulator.cpp:313), getters return pointers to
// host mirrors of the SoA components, refreshed lazily after Update.
// Differences: the ECM, planner and environment arrive as an ecm_scene
// (include/ecm_scene.h) instead of the reference's Boost-backed classes, and
// component writes go through the explicit Set* calls.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ecm_scene.h"
#include "../../include/ecm_sim.h"

namespace ecm_amd {
namespace Simulation {

typedef int Entity;

// Simulator.h:29-57
struct PositionComponent { float x; float y; };
struct VelocityComponent { float dx; float dy; };
struct ClearanceComponent { float clearance; };
struct SpeedComponent { float speed; };

class Simulator {
   public:
    // Simulator(ecm, planner, environment, maxAgents, simStepTime) (Simulator.h:63-74)
    Simulator(ecm_scene* scene, float xmin, float ymin, float xmax, float ymax, int maxAgents, float simStepTime,
              const ecm_sim_params* params = nullptr)
        : m_Scene(scene), m_Box{xmin, ymin, xmax, ymax}, m_MaxNumEntities(maxAgents) {
        if (params) m_Params = *params;
        else ecm_sim_default_params(&m_Params);
        m_Params.max_agents = maxAgents;
        m_Params.dt = simStepTime;
    }
    ~Simulator() { ClearSimulator(); }
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    // Simulator::Initialize (Simulator.cpp:21-60): device buffers, ECM and
    // obstacles from the scene, replanning through the scene's planner.
    void Initialize() {
        check(ecm_sim_create(&m_Params, m_Box[0], m_Box[1], m_Box[2], m_Box[3], &m_Sim), "create");
        int nr = 0, nv = 0;
        check(ecm_scene_obstacle_counts(m_Scene, &nr, &nv), "obstacle_counts");
        std::vector<float> xy(2 * (size_t)nv + 2);
        std::vector<int> off(nr + 1);
        check(ecm_scene_obstacle_export(m_Scene, xy.data(), off.data()), "obstacle_export");
        check(ecm_sim_set_obstacles(m_Sim, xy.data(), off.data(), nr), "set_obstacles");
        int ev = 0, ee = 0;
        check(ecm_scene_ecm_counts(m_Scene, &ev, &ee), "ecm_counts");
        std::vector<float> vxy(2 * (size_t)ev), pts(8 * (size_t)ee);
        std::vector<int> tgt(2 * (size_t)ee);
        check(ecm_scene_ecm_export(m_Scene, vxy.data(), nullptr, tgt.data(), nullptr, pts.data()), "ecm_export");
        check(ecm_sim_set_ecm(m_Sim, ev, vxy.data(), ee, tgt.data(), pts.data()), "set_ecm");
        check(ecm_sim_set_replan_callback(m_Sim, ecm_scene_replan_callback(), m_Scene), "replan_callback");
        check(ecm_sim_set_replan_threads(m_Sim, 8), "replan_threads");
        const size_t N = (size_t)m_MaxNumEntities;
        m_Positions.resize(N); m_Velocities.resize(N); m_PreferredVelocities.resize(N);
        m_AttractionPoints.resize(N); m_Clearances.resize(N); m_Active.resize(N);
        m_Dirty = true;
    }

    // Simulator::SpawnAgent (Simulator.cpp:168-200): plans the path with the
    // scene's planner (UpdatePath, :97-124); -1 when the pool is full, the
    // location invalid or no path exists.
    int SpawnAgent(float sx, float sy, float gx, float gy, float clearance, float preferredSpeed) {
        std::vector<float> path(2 * 4096);
        int n = 0;
        int rc = ecm_scene_plan_path(m_Scene, sx, sy, gx, gy, clearance, path.data(), 4096, &n);
        if (rc != 0 || n < 1) return -1;
        float st[2] = {sx, sy};
        int off = 0, id = -1;
        check(ecm_sim_spawn(m_Sim, 1, st, &clearance, &preferredSpeed, &off, &n, path.data(), 1, &id), "spawn");
        m_Dirty = true;
        return id;
    }
    // DestroyAgent (Simulator.cpp:202-208)
    void DestroyAgent(int idx) { check(ecm_sim_destroy_agent(m_Sim, idx), "destroy"); m_Dirty = true; }
    // Update (Simulator.cpp:314-323)
    void Update(float /*dt*/) { check(ecm_sim_step(m_Sim, 1), "step"); m_Dirty = true; }
    // Reset (Simulator.cpp:325-333)
    void Reset() { check(ecm_sim_reset(m_Sim), "reset"); m_Dirty = true; }

    // getters (Simulator.h:114-126)
    int GetNumAgents() const { int n = 0; check(ecm_sim_num_agents(m_Sim, &n), "num_agents"); return n; }
    int GetLastIndex() const { int n = 0; check(ecm_sim_last_index(m_Sim, &n), "last_index"); return n; }
    PositionComponent* GetPositionData() { sync(); return m_Positions.data(); }
    VelocityComponent* GetVelocityData() { sync(); return m_Velocities.data(); }
    VelocityComponent* GetPreferredVelocityData() { sync(); return m_PreferredVelocities.data(); }
    PositionComponent* GetAttractionPointData() { sync(); return m_AttractionPoints.data(); }
    ClearanceComponent* GetClearanceData() { sync(); return m_Clearances.data(); }
    bool* GetActiveFlags() { sync(); return reinterpret_cast<bool*>(m_Active.data()); }
    float GetSimulationStepTime() const { return m_Params.dt; }
    ecm_sim* Handle() const { return m_Sim; }

    // Simulator::AddPosition (Simulator.h:87-90)
    void AddPosition(Entity e, float x, float y) {
        sync();
        m_Positions[e].x = x;
        m_Positions[e].y = y;
        check(ecm_sim_write(m_Sim, ECM_C_POSITION, m_Positions.data(), m_Positions.size() * sizeof(PositionComponent)),
              "write position");
    }

   private:
    static void check(int rc, const char* what) {
        if (rc != 0) throw std::runtime_error(std::string("ecm_sim ") + what + ": " + ecm_last_error());
    }
    void sync() {
        if (!m_Dirty || !m_Sim) return;
        const size_t N = (size_t)m_MaxNumEntities;
        check(ecm_sim_read(m_Sim, ECM_C_POSITION, m_Positions.data(), N * 8), "read");
        check(ecm_sim_read(m_Sim, ECM_C_VELOCITY, m_Velocities.data(), N * 8), "read");
        check(ecm_sim_read(m_Sim, ECM_C_PREF_VELOCITY, m_PreferredVelocities.data(), N * 8), "read");
        check(ecm_sim_read(m_Sim, ECM_C_ATTRACTION, m_AttractionPoints.data(), N * 8), "read");
        check(ecm_sim_read(m_Sim, ECM_C_CLEARANCE, m_Clearances.data(), N * 4), "read");
        check(ecm_sim_read(m_Sim, ECM_C_ACTIVE, m_Active.data(), N), "read");
        m_Dirty = false;
    }
    void ClearSimulator() {
        if (m_Sim) ecm_sim_destroy(m_Sim);
        m_Sim = nullptr;
    }

    ecm_scene* m_Scene;
    float m_Box[4];
    int m_MaxNumEntities;
    ecm_sim_params m_Params{};
    ecm_sim* m_Sim = nullptr;
    bool m_Dirty = true;
    std::vector<PositionComponent> m_Positions, m_AttractionPoints;
    std::vector<VelocityComponent> m_Velocities, m_PreferredVelocities;
    std::vector<ClearanceComponent> m_Clearances;
    std::vector<unsigned char> m_Active;
};

}  // namespace Simulation
}  // namespace ecm_amd
