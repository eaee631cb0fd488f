/*
 * ecm_sim.h -- C ABI of the MI355X crowd step (libecmcrowd.so).
 *
 * Drop-in for the per-timestep update behind the reference's
 * ECM::Simulation::Simulator (ECMAgentSimulator/Simulator.h:59-188).  Plain
 * pointers and sizes only.  The library owns device buffers; the caller owns
 * host buffers.  One handle per host thread; calls on a handle are
 * serialised on the handle's own HIP stream.  Every entry point returns 0 on
 * success or a negative ECM_E* code; ecm_last_error() gives the message.
 */
#ifndef ECM_SIM_H
#define ECM_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ECM_OK 0
#define ECM_E_ARG -1        /* bad argument                        */
#define ECM_E_HIP -2        /* HIP runtime error                   */
#define ECM_E_FULL -3       /* no free agent slot (Simulator.cpp:170-171) */
#define ECM_E_STATE -4      /* call out of order (e.g. step before set_ecm) */
#define ECM_E_NOGPU -5      /* no usable gfx950 device             */

typedef struct ecm_sim ecm_sim;

/* Parameters.  Defaults (ecm_sim_default_params) are the benchmark values:
 * k = 10, T = T_obst = 2 s, maxSpeed = 2 m/s, dt = 0.1 s.  The reference's
 * literals (Simulator.cpp:55, ORCA.h:102-103, Simulator.cpp:671) are
 * k = 5, T = T_obst = 10 s and maxSpeed = the agent's preferred speed
 * (ecm_sim_reference_params). */
typedef struct {
    int   max_agents;           /* slot pool size (Simulator.h:63)            */
    int   k;                    /* max ORCA neighbours, 1..16 (Simulator.cpp:55) */
    float time_horizon;         /* ORCA.h:102                                 */
    float time_horizon_obst;    /* ORCA.h:103                                 */
    float max_speed;            /* LP speed limit when !max_speed_from_pref   */
    int   max_speed_from_pref;  /* 1: limit = own preferred speed (Simulator.cpp:671-674) */
    float dt;                   /* simStepTime (Simulator.h:63)               */
    float mass;                 /* Simulator.cpp:622                          */
    float arrival_radius;       /* Simulator.cpp:543                          */
    float delete_radius;        /* Simulator.cpp:542                          */
    float grid_cell;            /* agent grid cell (m); <= 0: automatic       */
    int   keep_orca_velocity;   /* 1: keep the ORCA velocity per agent (diagnostic) */
} ecm_sim_params;

void ecm_sim_default_params(ecm_sim_params* p);
void ecm_sim_reference_params(ecm_sim_params* p);

/* Simulator(ecm, planner, env, maxAgents, simStepTime) + Initialize()
 * (Simulator.h:63-74, Simulator.cpp:21-60).  world_* is the walkable box
 * (Environment::AddWalkableArea, Environment.cpp:274-281). */
int ecm_sim_create(const ecm_sim_params* p, float world_xmin, float world_ymin,
                   float world_xmax, float world_ymax, ecm_sim** out);
/* ~Simulator / ClearSimulator (Simulator.cpp:62-95) */
void ecm_sim_destroy(ecm_sim* s);

/* Static obstacles as CCW rings, read by FindNearestObstacles
 * (Simulator.cpp:259-292) through Environment::GetObstacles
 * (Environment.h:55); rings built as Obstacle::Initialize
 * (ECMDataTypes.cpp:23-61).  xy: [2 * ring_off[n_rings]]. */
int ecm_sim_set_obstacles(ecm_sim* s, const float* xy, const int* ring_off, int n_rings);

/* ECM graph, query side (ECM.h:20-50): vertex positions and, per edge, the
 * two half-edge targets and closest points he0.cl, he0.cr, he1.cl, he1.cr.
 * Cells are derived as ECMCellCollection::Construct (ECMCellCollection.cpp:18-46). */
int ecm_sim_set_ecm(ecm_sim* s, int n_vert, const float* vert_xy, int n_edge,
                    const int* he_target, const float* he_pts);

/* SpawnAgent (Simulator.cpp:168-200), batched.  Each agent takes the next
 * free slot of the LIFO pool (Simulator.h:66-69); a start closer than its own
 * clearance to an active agent is rejected (ValidSpawnLocation,
 * Simulator.cpp:295-311) and gets id -1, as does any agent once the pool is
 * empty.  Paths are CSR: agent i owns path_xy[2*path_off[i] ..
 * 2*(path_off[i]+path_cnt[i])), path_cnt[i] >= 1.  Velocities start at 0.
 * validate = 0 skips the spawn-location check (for state injection). */
int ecm_sim_spawn(ecm_sim* s, int n, const float* start_xy, const float* radius,
                  const float* pref_speed, const int* path_off, const int* path_cnt,
                  const float* path_xy, int validate, int* out_ids);

/* DestroyAgent (Simulator.cpp:202-208) */
int ecm_sim_destroy_agent(ecm_sim* s, int id);
/* Reset (Simulator.cpp:325-333) */
int ecm_sim_reset(ecm_sim* s);
/* UpdatePath result (Simulator.cpp:97-124): replace the path of agent id. */
int ecm_sim_set_path(ecm_sim* s, int id, const float* path_xy, int n_points);

/* Simulator::Update (Simulator.cpp:314-323), n_steps times.  After each
 * step the agents whose IRM attraction failed (Simulator.cpp:579-587) are
 * handed to the replan callback, when one is set, before the next step. */
typedef int (*ecm_replan_fn)(void* user, int id, float x, float y, float goal_x, float goal_y,
                             float clearance, float* out_xy, int cap, int* out_n);
int ecm_sim_set_replan_callback(ecm_sim* s, ecm_replan_fn fn, void* user);
/* Threads that may call the replan callback concurrently within one step
 * (default 1).  Only raise it for a thread-safe callback, such as
 * ecm_scene_replan_callback().  With n > 1 the handle keeps n - 1 worker
 * threads (made at the first step that replans, remade when n changes,
 * joined by ecm_sim_destroy); the stepping thread is the n-th. */
int ecm_sim_set_replan_threads(ecm_sim* s, int n);

/* Step graphs: with on, the device work of a step is captured into one
 * hipGraph (one per position-buffer parity; the host reads the replan
 * events from host-mapped memory while the graph still runs) and
 * replayed, recaptured whenever buffers, parameters or the launch shape
 * change.  Steps with timing on run eagerly.  If capture fails the handle
 * falls back to eager launches; ecm_sim_graph_status reports whether
 * graphs are in use and, if not, why. */
int ecm_sim_set_graph(ecm_sim* s, int on);
int ecm_sim_graph_status(ecm_sim* s, int* active, char* why, int cap);

/* Launch shape; no choice changes any result.  cell_scale multiplies the
 * agent grid cell edge (about one agent per cell at 1; default 1.25),
 * knn_block is the workgroup size of the k-nearest kernel (64,
 * 128, 256; default 256; 16: 16 agents per 64-lane wavefront of the fused
 * kernel below, for crowds of a few thousand agents, and 64 otherwise),
 * and orca_obst_slots sizes the main ORCA pass (1,
 * 2, 4; default 2): the plain step holds 4 obstacle constraints per agent
 * in registers (8 with orca_obst_slots 4 and k <= 10) and sends agents with
 * more to the overflow pass; with orca_obst_slots 2 and k 5 or 10 that pass
 * runs inside the k-nearest kernel; the learned and logging steps
 * give each agent orca_obst_slots slots beyond k in an LDS tile, agents
 * with more constraints borrow from their workgroup's small spill area, and
 * those with more obstacle constraints than the tile holds, or whose
 * workgroup's spill area is used up, take the overflow pass.  best_ms: step
 * time the autotuner measured (0 before). */
typedef struct ecm_sim_tuning {
    float cell_scale;
    int knn_block;
    int orca_obst_slots;
    float best_ms;
} ecm_sim_tuning;
int ecm_sim_get_tuning(ecm_sim* s, ecm_sim_tuning* out);
int ecm_sim_set_tuning(ecm_sim* s, const ecm_sim_tuning* t);
/* Coordinate search of the launch shape on the current workload: the
 * options of each coordinate are sampled round robin, `steps` rounds of one
 * timed step each (after two untimed ones), and compared by their median;
 * the simulation advances through every sampled step.  The best shape is
 * kept and returned. */
int ecm_sim_autotune(ecm_sim* s, int steps, ecm_sim_tuning* out);

/* Batched multi-scenario stepping.  A handle holds one or more
 * environments: independent scenes (world box, obstacles, ECM, agents)
 * whose agents never interact across environments, advanced together by
 * the same launches of every step.  ecm_sim_create makes environment 0;
 * the calls without an environment argument act on it.  Agents keep their
 * environment for life; slot ids are shared by all environments.
 * (No reference counterpart: the reference simulates one environment,
 * Simulator.h:59-188.) */
int ecm_sim_add_env(ecm_sim* s, float x0, float y0, float x1, float y1, int* env);
int ecm_sim_num_envs(ecm_sim* s, int* n);
int ecm_sim_env_set_obstacles(ecm_sim* s, int env, const float* xy, const int* ring_off, int n_rings);
int ecm_sim_env_set_ecm(ecm_sim* s, int env, int n_vert, const float* vert_xy, int n_edge, const int* he_target,
                        const float* he_pts);
int ecm_sim_env_spawn(ecm_sim* s, int env, int n, const float* start_xy, const float* radius,
                      const float* pref_speed, const int* path_off, const int* path_cnt, const float* path_xy,
                      int validate, int* out_ids);
int ecm_sim_env_set_replan_callback(ecm_sim* s, int env, ecm_replan_fn fn, void* user);
/* Environment of every slot (max_agents entries, -1 for free slots). */
int ecm_sim_agent_envs(ecm_sim* s, int* env);
/* Runs n_steps steps (Simulator::Update, above).  Returns once the host part
 * of the last step is done (events read, replans planned and
 * uploaded, destroyed agents booked); its last kernels may still be running
 * on the handle's stream.  Every other call that reads or replaces device
 * data waits for the stream first. */
int ecm_sim_step(ecm_sim* s, int n_steps);
/* Wait for the handle's stream. */
int ecm_sim_sync(ecm_sim* s);

/* Getters (Simulator.h:114-126).  Components: */
#define ECM_C_POSITION 0        /* float[2 * max_agents]                     */
#define ECM_C_VELOCITY 1        /* float[2 * max_agents]                     */
#define ECM_C_PREF_VELOCITY 2   /* float[2 * max_agents]                     */
#define ECM_C_ATTRACTION 3      /* float[2 * max_agents]                     */
#define ECM_C_CLEARANCE 4       /* float[max_agents]                         */
#define ECM_C_PREF_SPEED 5      /* float[max_agents]                         */
#define ECM_C_ACTIVE 6          /* uint8[max_agents]                         */
#define ECM_C_ORCA_VELOCITY 7   /* float[2 * max_agents] (keep_orca_velocity)*/
/* Copies one component of every slot into dst (bytes must match).  Waits
 * for the last step; reading positions alone brings only the positions
 * into slot order. */
int ecm_sim_read(ecm_sim* s, int component, void* dst, size_t bytes);
/* Page-locked host memory for read-back buffers (no counterpart in the
 * reference, whose getters return its own host arrays).  ecm_sim_read into
 * such a buffer copies at full PCIe speed: 8 MB of positions in 0.15 ms
 * against 0.47 ms into pageable memory on MI355X.  Free with
 * ecm_host_free; the memory is not tied to a handle. */
int ecm_host_alloc(size_t bytes, void** out);
int ecm_host_free(void* p);
/* Overwrite a component (snapshot restore / state injection). */
int ecm_sim_write(ecm_sim* s, int component, const void* src, size_t bytes);
/* Device address of a component, up to date at the call (between steps
 * the state lives in the last step's sorted arrays; the slot-indexed
 * component arrays are refreshed when the host asks).  The array is not
 * updated by later steps until the next read / device_ptr / sync call, and
 * writing through it does not change the simulation: use ecm_sim_write. */
int ecm_sim_device_ptr(ecm_sim* s, int component, void** out);

int ecm_sim_num_agents(ecm_sim* s, int* out);    /* GetNumAgents (Simulator.h:114) */
int ecm_sim_last_index(ecm_sim* s, int* out);    /* GetLastIndex (Simulator.h:115) */

/* Path CSR of every slot: off[max_agents], cnt[max_agents] (points) and the
 * pool of points (2 floats each); *pool_used receives the pool length in
 * points; pool may be NULL to query the length. */
int ecm_sim_read_paths(ecm_sim* s, int* off, int* cnt, float* pool, size_t pool_cap, size_t* pool_used);

/* Neighbour lists of the last step (parity / debug, the KD-tree's
 * outNeighbors, Simulator.cpp:211-227): nbr[id*k + j], -1 past the count. */
int ecm_sim_get_neighbors(ecm_sim* s, int* nbr, int* cnt);
/* Agents whose IRM failed in the last step. */
int ecm_sim_get_replans(ecm_sim* s, int* ids, int cap, int* n);

/* Counters since create: [0] steps, [1] agent updates, [2] 2D-LP failures
 * (3D LP run), [3] exact-y locate fallbacks, [4] replans, [5] destroyed,
 * [6] agents with more obstacle constraints than the main ORCA tile holds,
 * [7] agents the learned proposal was tried on, [8] of those, accepted by
 * the verification, [9] constraints proposed over the tried agents, [10] of
 * the 3D-LP agents ([2]), those with more obstacle constraints than the
 * 3D LP's register tile holds (solved from an LDS tile instead).
 * Copies min(n, ECM_SIM_N_STATS). */
#define ECM_SIM_N_STATS 11
int ecm_sim_stats(ecm_sim* s, uint64_t* out, int n);
/* Device time of the phases of the last step in ms (HIP events on the
 * handle's stream): [0] grid build, [1] attraction + k-nearest, [2] learned
 * proposal (features + network; 0 without one), [3] ORCA + LP +
 * integration, [4] of [3], the main ORCA pass alone (the rest is the
 * passes over its queues: obstacle overflow and the 3D LP), [5] of [1], the
 * attraction pass alone (goal checks, IRM, steering; the rest is the
 * k-nearest pass).  In the plain step with k 5 or 10 and orca_obst_slots 2
 * the main ORCA pass runs inside the k-nearest kernel: [1] includes it and
 * [4] is empty.  Enable with ecm_sim_set_timing(s, 1); copies min(n, 6). */
int ecm_sim_set_timing(ecm_sim* s, int on);
int ecm_sim_last_timing(ecm_sim* s, float* out, int n);

/* Snapshot of the whole state (SoA + paths + params) to / from a file. */
int ecm_sim_snapshot_save(ecm_sim* s, const char* path);
int ecm_sim_snapshot_load(ecm_sim* s, const char* path);

const char* ecm_last_error(void);
/* 1 when a gfx950 device is usable by this process. */
int ecm_device_available(void);

#ifdef __cplusplus
}
#endif
#endif
