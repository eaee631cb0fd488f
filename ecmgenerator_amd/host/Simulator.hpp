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
//
// Spawn and goal areas (Simulator.h:92-104, Area.h) follow the reference's
// UpdateSpawnAreas (Simulator.cpp:494-536), with these differences: random
// positions come from a seeded std::mt19937 (SetAreaSeed) instead of rand();
// a connection to a goal area that does not exist, or with a rate <= 0,
// spawns nothing (the reference default-constructs the goal area with
// uninitialised extents, and divides 0 by a zero rate); Deconnect removes
// the connection's rate and timer with it (the reference keeps them, which
// shifts the later connections' rates).  Obstacle areas
// (Simulator.cpp:383-412) add a rectangle to the scene; the simulator sees
// it at once and, with updateECM, on the rebuilt ECM (the scene's own
// builder, not Boost's Voronoi).
#pragma once

#include <cstdint>
#include <cstring>
#include <map>
#include <random>
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
// Simulator.h:51-57.  x and y point into the facade's host mirror of the
// path pool; they stay valid until the next call that changes agents or
// steps the simulation.  currentIndex is 0, as the reference never
// advances it (Simulator.cpp:104,118).
struct PathComponent { int currentIndex; int numPoints; float* x; float* y; };

// A host mirror of one component array in page-locked memory
// (ecm_host_alloc), so that reading it back every step copies at full PCIe
// speed.  Trivially copyable elements only; not copyable itself.
template <typename T>
class PinnedArray {
public:
    PinnedArray() = default;
    PinnedArray(const PinnedArray&) = delete;
    PinnedArray& operator=(const PinnedArray&) = delete;
    ~PinnedArray() { ecm_host_free(m_Data); }
    void resize(size_t n) {
        if (n == m_Size) return;
        void* p = nullptr;
        if (ecm_host_alloc(n * sizeof(T), &p) != 0)
            throw std::runtime_error(std::string("ecm_host_alloc: ") + ecm_last_error());
        if (m_Data) std::memcpy(p, m_Data, (n < m_Size ? n : m_Size) * sizeof(T));
        if (n > m_Size) std::memset(static_cast<T*>(p) + m_Size, 0, (n - m_Size) * sizeof(T));
        ecm_host_free(m_Data);
        m_Data = static_cast<T*>(p);
        m_Size = n;
    }
    T* data() { return m_Data; }
    const T* data() const { return m_Data; }
    size_t size() const { return m_Size; }
    T& operator[](size_t i) { return m_Data[i]; }
    const T& operator[](size_t i) const { return m_Data[i]; }

private:
    T* m_Data = nullptr;
    size_t m_Size = 0;
};

// Area.h:11-103
enum SimAreaType { NONE, WALKABLE, SPAWN, GOAL, OBSTACLE };
struct SpawnConfiguration {
    float preferredSpeedMin = 5.0f;
    float preferredSpeedMax = 10.0f;
    float clearanceMin = 5.0f;
    float clearanceMax = 10.0f;
};
struct Area {
    int ID = -1;
    float x = 0.0f, y = 0.0f;          // Position
    float HalfHeight = 0.0f, HalfWidth = 0.0f;
    SimAreaType Type = NONE;
    bool Intersects(float px, float py) const {
        return px <= x + HalfWidth && px >= x - HalfWidth && py <= y + HalfHeight && py >= y - HalfHeight;
    }
};
struct GoalArea : Area { GoalArea() { Type = GOAL; } };
struct ObstacleArea : Area {
    ObstacleArea() { Type = OBSTACLE; }
    std::vector<float> obstacleVerts;       // x, y pairs, CCW
};
struct SpawnArea : Area {
    SpawnArea() { Type = SPAWN; }
    SpawnConfiguration spawnConfiguration;
    std::vector<int> connectedGoalAreas;
    std::vector<float> spawnRate;           // agents per second, per connected goal area
    std::vector<float> timeSinceLastSpawn;
};

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
        UploadObstacles();
        UploadEcm();
        check(ecm_sim_set_replan_callback(m_Sim, ecm_scene_replan_callback(), m_Scene), "replan_callback");
        check(ecm_sim_set_replan_threads(m_Sim, 8), "replan_threads");
        const size_t N = (size_t)m_MaxNumEntities;
        m_Positions.resize(N); m_Velocities.resize(N); m_PreferredVelocities.resize(N);
        m_AttractionPoints.resize(N); m_Clearances.resize(N); m_Active.resize(N);
        m_Stale = kAllMirrors;
        m_PathsDirty = true;
    }

    // Simulator::SpawnAgent (Simulator.cpp:168-200): plans the path with the
    // scene's planner (UpdatePath, :97-124); -1 when the pool is full, the
    // location invalid or no path exists.
    int SpawnAgent(float sx, float sy, float gx, float gy, float clearance, float preferredSpeed) {
        const int id = SpawnOne(sx, sy, gx, gy, clearance, preferredSpeed, 1);
        m_Stale = kAllMirrors;
        m_PathsDirty = true;
        return id;
    }
    // DestroyAgent (Simulator.cpp:202-208)
    void DestroyAgent(int idx) {
        check(ecm_sim_destroy_agent(m_Sim, idx), "destroy");
        m_Stale = kAllMirrors;
        m_PathsDirty = true;
    }
    // Update (Simulator.cpp:314-323): spawn areas first, as the reference
    void Update(float /*dt*/) {
        UpdateSpawnAreas();
        check(ecm_sim_step(m_Sim, 1), "step");
        m_Stale = kAllMirrors;
        m_PathsDirty = true;
    }

    // ValidSpawnLocation (Simulator.cpp:295-311): no active agent closer than
    // the clearance
    bool ValidSpawnLocation(float x, float y, float clearance) {
        sync(kPosMirror | kActiveMirror);
        return ValidOnMirror(x, y, clearance);
    }

    // Areas (Simulator.cpp:336-381, 414-470)
    int AddSpawnArea(float px, float py, float halfW, float halfH, const SpawnConfiguration& config, int ID = -1) {
        SpawnArea sa;
        sa.HalfWidth = halfW;
        sa.HalfHeight = halfH;
        sa.ID = ID == -1 ? m_NextSpawnID++ : ID;
        sa.x = px;
        sa.y = py;
        sa.spawnConfiguration = config;
        m_SpawnAreas.emplace(sa.ID, sa);
        return sa.ID;
    }
    int AddGoalArea(float px, float py, float halfW, float halfH, int ID = -1) {
        GoalArea ga;
        ga.ID = ID == -1 ? m_NextGoalID++ : ID;
        ga.x = px;
        ga.y = py;
        ga.HalfWidth = halfW;
        ga.HalfHeight = halfH;
        m_GoalAreas.emplace(ga.ID, ga);
        return ga.ID;
    }
    void RemoveArea(SimAreaType type, int ID) {
        if (type == SPAWN) m_SpawnAreas.erase(ID);
        if (type == GOAL) m_GoalAreas.erase(ID);
    }
    void ConnectSpawnGoalAreas(int spawnID, int goalID, float spawnRate = 0.0f) {
        SpawnArea& sa = m_SpawnAreas[spawnID];
        for (int ga : sa.connectedGoalAreas)
            if (ga == goalID) return;
        sa.connectedGoalAreas.push_back(goalID);
        sa.timeSinceLastSpawn.push_back(0.0f);
        sa.spawnRate.push_back(spawnRate);
    }
    void DeconnectSpawnGoalAreas(int spawnID, int goalID) {
        auto it = m_SpawnAreas.find(spawnID);
        if (it == m_SpawnAreas.end()) return;
        SpawnArea& sa = it->second;
        for (size_t i = 0; i < sa.connectedGoalAreas.size(); ++i)
            if (sa.connectedGoalAreas[i] == goalID) {
                sa.connectedGoalAreas.erase(sa.connectedGoalAreas.begin() + i);
                sa.spawnRate.erase(sa.spawnRate.begin() + i);
                sa.timeSinceLastSpawn.erase(sa.timeSinceLastSpawn.begin() + i);
                return;
            }
    }
    std::map<int, SpawnArea>& GetSpawnAreas() { return m_SpawnAreas; }
    std::map<int, GoalArea>& GetGoalAreas() { return m_GoalAreas; }
    SpawnArea* GetSpawnArea(int ID) {
        auto it = m_SpawnAreas.find(ID);
        return it == m_SpawnAreas.end() ? nullptr : &it->second;
    }
    GoalArea* GetGoalArea(int ID) {
        auto it = m_GoalAreas.find(ID);
        return it == m_GoalAreas.end() ? nullptr : &it->second;
    }
    void SetAreaSeed(uint32_t seed) { m_Rng.seed(seed); }

    // AddObstacleArea (Simulator.cpp:383-412): the rectangle's CCW ring goes
    // to the scene (Environment::AddObstacle) and to the simulator's
    // obstacles; updateECM rebuilds the scene's ECM and hands it over.
    int AddObstacleArea(float px, float py, float halfW, float halfH, bool updateECM = false) {
        ObstacleArea oa;
        oa.ID = m_ObstacleAreas.empty() ? 0 : m_ObstacleAreas.back().ID + 1;
        oa.x = px;
        oa.y = py;
        oa.HalfWidth = halfW;
        oa.HalfHeight = halfH;
        oa.obstacleVerts = {px + halfW, py + halfH, px - halfW, py + halfH, px - halfW, py - halfH, px + halfW, py - halfH};
        check(ecm_scene_add_obstacle(m_Scene, oa.obstacleVerts.data(), 4), "add_obstacle");
        if (m_Sim) {
            UploadObstacles();
            if (updateECM) {
                check(ecm_scene_build_ecm(m_Scene, 0.0f, 0), "build_ecm");
                UploadEcm();
            }
        }
        m_ObstacleAreas.push_back(oa);
        return oa.ID;
    }
    std::vector<ObstacleArea>& GetObstacleAreas() { return m_ObstacleAreas; }

    // UpdateSpawnAreas (Simulator.cpp:494-536): per connection, the time since
    // the last spawn grows by the step time; int(time * rate) agents spawn,
    // each at the first of up to 10 random positions in the area where
    // ValidSpawnLocation holds, with the minimum clearance and preferred
    // speed of the area's configuration and a random goal in the goal area;
    // then time -= spawned / rate.
    void UpdateSpawnAreas() {
        // the host mirror is read once; agents spawned here are added to its
        // positions and flags as they are placed
        bool spawned = false;
        if (!m_SpawnAreas.empty()) sync(kPosMirror | kActiveMirror);
        for (auto& kv : m_SpawnAreas) {
            SpawnArea& area = kv.second;
            for (size_t ga = 0; ga < area.connectedGoalAreas.size(); ++ga) {
                area.timeSinceLastSpawn[ga] += m_Params.dt;
                const float rate = area.spawnRate[ga];
                auto goal = m_GoalAreas.find(area.connectedGoalAreas[ga]);
                if (!(rate > 0.0f) || goal == m_GoalAreas.end()) continue;
                const int agentsToSpawn = (int)(area.timeSinceLastSpawn[ga] * rate);
                for (int i = 0; i < agentsToSpawn; ++i) {
                    const float clearance = area.spawnConfiguration.clearanceMin;
                    const float speed = area.spawnConfiguration.preferredSpeedMin;
                    float sx = 0.0f, sy = 0.0f;
                    bool found = false;
                    for (int attempt = 0; attempt < 10 && !found; ++attempt) {
                        RandomPositionInArea(area, sx, sy);
                        found = ValidOnMirror(sx, sy, clearance);
                    }
                    if (!found) continue;
                    float gx = 0.0f, gy = 0.0f;
                    RandomPositionInArea(goal->second, gx, gy);
                    const int id = SpawnOne(sx, sy, gx, gy, clearance, speed, 0);
                    if (id < 0) continue;
                    m_Positions[id].x = sx;
                    m_Positions[id].y = sy;
                    m_Active[id] = 1;
                    m_LastIdx = id > m_LastIdx ? id : m_LastIdx;
                    spawned = true;
                }
                area.timeSinceLastSpawn[ga] -= (float)agentsToSpawn / rate;
            }
        }
        if (spawned) {
            m_Stale = kAllMirrors;
            m_PathsDirty = true;
        }
    }
    // Reset (Simulator.cpp:325-333)
    void Reset() {
        check(ecm_sim_reset(m_Sim), "reset");
        m_Stale = kAllMirrors;
        m_PathsDirty = true;
    }

    // getters (Simulator.h:114-126)
    int GetNumAgents() const { int n = 0; check(ecm_sim_num_agents(m_Sim, &n), "num_agents"); return n; }
    int GetLastIndex() const { int n = 0; check(ecm_sim_last_index(m_Sim, &n), "last_index"); return n; }
    // each getter reads back only its own component, once per step (a
    // renderer that draws positions every step moves 8 B per slot, not 37)
    PositionComponent* GetPositionData() { sync(kPosMirror); return m_Positions.data(); }
    VelocityComponent* GetVelocityData() { sync(kVelMirror); return m_Velocities.data(); }
    VelocityComponent* GetPreferredVelocityData() { sync(kPrefVelMirror); return m_PreferredVelocities.data(); }
    PositionComponent* GetAttractionPointData() { sync(kAttrMirror); return m_AttractionPoints.data(); }
    ClearanceComponent* GetClearanceData() { sync(kClearMirror); return m_Clearances.data(); }
    bool* GetActiveFlags() { sync(kActiveMirror); return reinterpret_cast<bool*>(m_Active.data()); }
    // GetPathData (Simulator.h:119): one PathComponent per slot, read back
    // with ecm_sim_read_paths when a spawn, replan or removal may have
    // changed a path since the last read.
    PathComponent* GetPathData() {
        if (m_PathsDirty && m_Sim) {
            const size_t N = (size_t)m_MaxNumEntities;
            std::vector<int> off(N), cnt(N);
            size_t used = 0;
            check(ecm_sim_read_paths(m_Sim, off.data(), cnt.data(), nullptr, 0, &used), "read_paths");
            std::vector<float> pool(2 * used + 2);
            check(ecm_sim_read_paths(m_Sim, off.data(), cnt.data(), pool.data(), used, &used), "read_paths");
            m_PathX.resize(used + 1);
            m_PathY.resize(used + 1);
            for (size_t j = 0; j < used; ++j) {
                m_PathX[j] = pool[2 * j];
                m_PathY[j] = pool[2 * j + 1];
            }
            m_Paths.resize(N);
            for (size_t i = 0; i < N; ++i) {
                const bool ok = cnt[i] > 0 && off[i] >= 0 && (size_t)off[i] + (size_t)cnt[i] <= used;
                m_Paths[i].currentIndex = 0;
                m_Paths[i].numPoints = ok ? cnt[i] : 0;
                m_Paths[i].x = ok ? m_PathX.data() + off[i] : nullptr;
                m_Paths[i].y = ok ? m_PathY.data() + off[i] : nullptr;
            }
            m_PathsDirty = false;
        }
        return m_Paths.data();
    }
    float GetSimulationStepTime() const { return m_Params.dt; }
    ecm_sim* Handle() const { return m_Sim; }

    // Simulator::AddPosition (Simulator.h:87-90)
    void AddPosition(Entity e, float x, float y) {
        sync(kPosMirror);
        m_Positions[e].x = x;
        m_Positions[e].y = y;
        check(ecm_sim_write(m_Sim, ECM_C_POSITION, m_Positions.data(), m_Positions.size() * sizeof(PositionComponent)),
              "write position");
    }

   private:
    static void check(int rc, const char* what) {
        if (rc != 0) throw std::runtime_error(std::string("ecm_sim ") + what + ": " + ecm_last_error());
    }
    // Reads back the components of `want` whose host mirror is stale.
    void sync(unsigned want) {
        const unsigned need = want & m_Stale;
        if (!need || !m_Sim) return;
        const size_t N = (size_t)m_MaxNumEntities;
        if (need & kPosMirror) check(ecm_sim_read(m_Sim, ECM_C_POSITION, m_Positions.data(), N * 8), "read");
        if (need & kVelMirror) check(ecm_sim_read(m_Sim, ECM_C_VELOCITY, m_Velocities.data(), N * 8), "read");
        if (need & kPrefVelMirror)
            check(ecm_sim_read(m_Sim, ECM_C_PREF_VELOCITY, m_PreferredVelocities.data(), N * 8), "read");
        if (need & kAttrMirror) check(ecm_sim_read(m_Sim, ECM_C_ATTRACTION, m_AttractionPoints.data(), N * 8), "read");
        if (need & kClearMirror) check(ecm_sim_read(m_Sim, ECM_C_CLEARANCE, m_Clearances.data(), N * 4), "read");
        if (need & kActiveMirror) check(ecm_sim_read(m_Sim, ECM_C_ACTIVE, m_Active.data(), N), "read");
        if (m_Stale == kAllMirrors) check(ecm_sim_last_index(m_Sim, &m_LastIdx), "last_index");
        m_Stale &= ~need;
    }
    void ClearSimulator() {
        if (m_Sim) ecm_sim_destroy(m_Sim);
        m_Sim = nullptr;
    }
    void UploadObstacles() {
        int nr = 0, nv = 0;
        check(ecm_scene_obstacle_counts(m_Scene, &nr, &nv), "obstacle_counts");
        std::vector<float> xy(2 * (size_t)nv + 2);
        std::vector<int> off(nr + 1);
        check(ecm_scene_obstacle_export(m_Scene, xy.data(), off.data()), "obstacle_export");
        check(ecm_sim_set_obstacles(m_Sim, xy.data(), off.data(), nr), "set_obstacles");
    }
    void UploadEcm() {
        int ev = 0, ee = 0;
        check(ecm_scene_ecm_counts(m_Scene, &ev, &ee), "ecm_counts");
        std::vector<float> vxy(2 * (size_t)ev), pts(8 * (size_t)ee);
        std::vector<int> tgt(2 * (size_t)ee);
        check(ecm_scene_ecm_export(m_Scene, vxy.data(), nullptr, tgt.data(), nullptr, pts.data()), "ecm_export");
        check(ecm_sim_set_ecm(m_Sim, ev, vxy.data(), ee, tgt.data(), pts.data()), "set_ecm");
    }
    // plans the path with the scene's planner, then spawns; -1 when no path
    // exists, the pool is full or (validate) the location is taken
    int SpawnOne(float sx, float sy, float gx, float gy, float clearance, float speed, int validate) {
        std::vector<float> path(2 * 4096);
        int n = 0;
        int rc = ecm_scene_plan_path(m_Scene, sx, sy, gx, gy, clearance, path.data(), 4096, &n);
        if (rc != 0 || n < 1) return -1;
        float st[2] = {sx, sy};
        int off = 0, id = -1;
        check(ecm_sim_spawn(m_Sim, 1, st, &clearance, &speed, &off, &n, path.data(), validate, &id), "spawn");
        return id;
    }
    bool ValidOnMirror(float x, float y, float clearance) const {
        const float c2 = clearance * clearance;
        for (int i = 0; i <= m_LastIdx; ++i) {
            if (!m_Active[i]) continue;
            const float dx = x - m_Positions[i].x, dy = y - m_Positions[i].y;
            if (dx * dx + dy * dy < c2) return false;
        }
        return true;
    }
    // Area::GetRandomPositionInArea (Area.h:39-53), uniform in the rectangle
    void RandomPositionInArea(const Area& a, float& px, float& py) {
        const float u = (float)(m_Rng() >> 8) * (1.0f / 16777216.0f);
        const float v = (float)(m_Rng() >> 8) * (1.0f / 16777216.0f);
        px = (a.x - a.HalfWidth) + u * (2.0f * a.HalfWidth);
        py = (a.y - a.HalfHeight) + v * (2.0f * a.HalfHeight);
    }

    ecm_scene* m_Scene;
    float m_Box[4];
    int m_MaxNumEntities;
    ecm_sim_params m_Params{};
    ecm_sim* m_Sim = nullptr;
    // host mirrors not read back since the simulation last changed
    static constexpr unsigned kPosMirror = 1u, kVelMirror = 2u, kPrefVelMirror = 4u, kAttrMirror = 8u,
                              kClearMirror = 16u, kActiveMirror = 32u, kAllMirrors = 63u;
    unsigned m_Stale = kAllMirrors;
    bool m_PathsDirty = true;
    std::vector<PathComponent> m_Paths;
    std::vector<float> m_PathX, m_PathY;
    PinnedArray<PositionComponent> m_Positions, m_AttractionPoints;
    PinnedArray<VelocityComponent> m_Velocities, m_PreferredVelocities;
    PinnedArray<ClearanceComponent> m_Clearances;
    PinnedArray<unsigned char> m_Active;
    std::map<int, SpawnArea> m_SpawnAreas;
    std::map<int, GoalArea> m_GoalAreas;
    std::vector<ObstacleArea> m_ObstacleAreas;
    int m_NextSpawnID = 0, m_NextGoalID = 0;
    int m_LastIdx = -1;
    std::mt19937 m_Rng{12345u};
};

}  // namespace Simulation
}  // namespace ecm_amd
