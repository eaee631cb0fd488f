/*
 * ecm_scene.h -- C ABI of the host-side scene in libecmcrowd.so: walkable
 * box and obstacles (ECMGenerator/Environment.h:18-85), construction of the
 * Explicit Corridor Map (ECMGenerator::GenerateECM, ECMGenerator.cpp:235-256),
 * point location, path planning (ECMPathPlanner::FindPath,
 * ECMPathPlanner.cpp:22-136) and the seeded benchmark scenarios.
 * These run on the CPU; they produce the inputs of the GPU step (ecm_sim.h)
 * and answer its replan requests.  Return codes as in ecm_sim.h.
 */
#ifndef ECM_SCENE_H
#define ECM_SCENE_H

#include <stdint.h>

#include "ecm_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ecm_scene ecm_scene;

/* Environment::AddWalkableArea with a box (Environment.cpp:274-281). */
int ecm_scene_create(float xmin, float ymin, float xmax, float ymax, ecm_scene** out);
void ecm_scene_destroy(ecm_scene* s);
/* Environment::AddObstacle (Environment.cpp:285-315): CCW ring of n >= 2
 * vertices.  The ECM builder accepts axis-aligned rectangles.  An ECM built
 * before stays in use (planning, export) until ecm_scene_build_ecm runs
 * again, as with updateECM false in the reference. */
int ecm_scene_add_obstacle(ecm_scene* s, const float* xy, int n);
/* ComputeECM / GenerateECM (Environment.cpp:232-239, ECMGenerator.cpp:235-256).
 * resolution: label-grid spacing in m (<= 0: 0.5); threads <= 0: all cores. */
int ecm_scene_build_ecm(ecm_scene* s, float resolution, int threads);
int ecm_scene_ecm_counts(ecm_scene* s, int* n_vert, int* n_edge);
/* ECM.h:20-50 arrays: vert_xy[2nv], vert_clearance[nv], he_target[2ne],
 * he_next[2ne], he_pts[8ne] (he0.cl, he0.cr, he1.cl, he1.cr).  Any output
 * pointer may be NULL. */
int ecm_scene_ecm_export(ecm_scene* s, float* vert_xy, float* vert_clearance, int* he_target,
                         int* he_next, float* he_pts);
int ecm_scene_obstacle_counts(ecm_scene* s, int* n_rings, int* n_verts);
int ecm_scene_obstacle_export(ecm_scene* s, float* xy, int* ring_off);
/* ECMGraph::FindCell (ECM.cpp:191-194): first containing cell or -1. */
int ecm_scene_locate(ecm_scene* s, float x, float y);
/* Environment::InsideObstacle (Environment.cpp:275-283): 1 inside. */
int ecm_scene_inside_obstacle(ecm_scene* s, float x, float y);
/* ECMPathPlanner::FindPath: 0 and *n_out points on success, 1 when no path. */
int ecm_scene_plan_path(ecm_scene* s, float sx, float sy, float gx, float gy, float clearance,
                        float* out_xy, int cap, int* n_out);
/* Batch of n queries on `threads` threads.  Query i writes path_cnt[i]
 * points (0 = no path) at path_xy[2*path_off[i]...]; offsets are packed in
 * query order into a buffer of cap_points points. */
int ecm_scene_plan_paths(ecm_scene* s, int n, const float* start_xy, const float* goal_xy,
                         const float* clearance, int threads, int* path_cnt, float* path_xy,
                         long long cap_points, int* path_off, int reserved);
/* A replan callback for ecm_sim_set_replan_callback; `user` is the scene. */
ecm_replan_fn ecm_scene_replan_callback(void);

/* Seeded scenarios: config 0 = 200x200 m, 40 rectangles with sides in
 * [2, 15] m; 1 = 1000x1000 m, 400 rectangles in [2, 25] m; 2 = 2000x2000 m,
 * a barrier of 63 blocks leaving 64 gaps of 8 m plus 1,437 scattered
 * rectangles, two opposing flows; 3 = 60x60 m, 6 rectangles (tests).
 * Agents: radius U[0.2, 0.3] m, preferred speed N(1.3, 0.2) clipped to
 * [0.8, 1.8] m/s, start and goal in free space at least 2r from obstacles,
 * no start within r of an earlier agent (Simulator.cpp:295-311), paths from
 * the planner.  n_agents <= 0 takes the config's count (5k / 100k / 1M / 500).
 * The scene (with its ECM built) is returned in *scene_out. */
typedef struct ecm_scenario ecm_scenario;
int ecm_scenario_generate(int config, uint64_t seed, int n_agents, ecm_scene** scene_out,
                          ecm_scenario** out);
/* The same generator for a (2 half) x (2 half) m square with n_rects
 * rectangles of integer sides in [side_lo, side_hi] m and n_agents agents
 * (one environment of the batched multi-scenario workload). */
int ecm_scenario_generate_box(float half, int n_rects, int side_lo, int side_hi, int n_agents, uint64_t seed,
                              ecm_scene** scene_out, ecm_scenario** out);
int ecm_scenario_counts(ecm_scenario* sc, int* n_agents, long long* n_points);
int ecm_scenario_export(ecm_scenario* sc, float* start_xy, float* goal_xy, float* radius,
                        float* pref_speed, int* path_off, int* path_cnt, float* path_xy);
void ecm_scenario_destroy(ecm_scenario* sc);

#ifdef __cplusplus
}
#endif
#endif
