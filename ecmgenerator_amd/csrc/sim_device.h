// sim_device.h -- device-side data layout of the crowd step (shared by the
// kernels and the host runtime that owns the buffers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ecm {

// Uniform grid over the world box.  Cell (cx, cy) of a point is
// floor((p - origin) / cell), clamped to [0, side); NaN maps to 0.
struct GridDesc {
    float x0, y0;       // origin
    float cell;         // cell edge (m)
    float inv_cell;     // 1 / cell
    int side;           // cells per axis
    int bits;           // Morton bits per axis (side <= 1 << bits)
};

// One environment: an independent scene (obstacles, ECM, world box) whose
// agents only ever interact with each other.  A simulation holds one or
// more; batched multi-scenario stepping advances all of them in the same
// launches.
struct EnvDesc {
    // agent grid: cells [cell_base, cell_base + 4^ag.bits) of the shared key space
    GridDesc ag;
    uint32_t cell_base;
    // obstacles (Environment::GetObstacles, edge per vertex in scan order)
    int n_obst_edges;
    const float4* oe0;      // (L.x, L.y, R.x, R.y),  R = next(L)
    const float4* oe1;      // (prev(L).x, prev(L).y, next(R).x, next(R).y)
    const uint8_t* oflags;  // bit0 convex(L), bit1 convex(R)
    GridDesc og;
    const int* obucket_start;   // [side^2 + 1]
    const int* obucket_edges;   // ascending edge ids
    // ECM cells (two per edge, ECMCellCollection.cpp:18-46)
    int n_cells;
    const float4* cellA;    // (q0, q1) = (v(he1.target), boundary.p0)
    const float4* cellB;    // (q2, q3) = (boundary.p1, v(he0.target))
    GridDesc eg;
    const int* ebucket_start;   // [side^2 + 1]
    const int* ebucket_cells;   // ascending cell ids
    // exact-y fallback: cells having a vertex whose y equals a given float
    int yhash_mask;             // table size - 1 (power of two), -1 if empty
    const uint32_t* yhash_key;  // y bit pattern (+0.0 for zero), 0xffffffff empty
    const int* yhash_off;       // [table] start into ycells
    const int* yhash_cnt;       // [table]
    const int* ycells;          // ascending cell ids per key
};

// Everything the step kernels read or write.  Agent arrays are indexed by
// slot id (entity == slot, Simulator.cpp:40); "sorted" arrays by the
// position of the agent in this step's Morton-ordered grid.
//
// Between steps the state lives in the sorted outputs of the last step
// (positions, velocities, attraction points by its sorted position); the
// slot-indexed arrays are brought up to date only when the host asks
// (k_unsort).  A step builds its grid from those sorted outputs (the order
// barely changes, so every access is close to coalesced), or from the
// slot-indexed arrays after the host changed them (spawn, destroy, write,
// snapshot load).
struct DevState {
    // ---- agent components by slot (Simulator.h:179-187) ----
    int n_slots;
    float2* pos_in;                              // positions
    float2* vel_in;                              // velocities
    const float* radius;                         // ClearanceComponent
    const float* pref_speed;                     // SpeedComponent
    float2* attraction;                          // attraction points
    float2* pref_vel;                            // preferred velocities
    float2* v_orca;                              // ORCA output velocities (diagnostic, may be null)
    uint8_t* active;                             // active flags
    const int* path_off;                         // PathComponent as CSR
    const int* path_cnt;
    const float2* path_pts;
    int sorted_in;                               // 1: grid from the last step's sorted outputs

    // ---- the last step's sorted outputs (read when sorted_in) ----
    const float4* sorted_pr_old;   // (x, y, radius, id bits) of its step start
    const float4* sorted_v_old;    // (vx, vy, preferred speed, environment bits)
    const float2* pos_s_old;       // its end-of-step positions
    const float2* vel_s_old;       // its end-of-step velocities
    const float2* attr_s_old;      // its attraction points
    const uint8_t* dead_s_old;     // destroyed during it (its active count is acc[13])
    const float2* pref_vel_s_old;  // its preferred velocities
    const float2* v_orca_s_old;    // its ORCA velocities (may be null)
    const int2* path_s_old;        // its path (offset, count) per agent

    // ---- this step's state by sorted position ----
    float2* attr_s;                // attraction point (carried in, updated by k_attr)
    float2* pref_vel_s;            // preferred velocity
    uint8_t* dead_s;               // destroyed this step
    float2* pos_s;                 // end-of-step position
    float2* vel_s;                 // end-of-step velocity
    float2* v_orca_s;              // ORCA output velocity (diagnostic, may be null)
    int2* path_s;                  // path (offset, count): path_off / path_cnt carried in Morton order

    // ---- environments ----
    int n_env;
    const EnvDesc* envs;         // [n_env]
    const uint16_t* agent_env;   // [n_slots] environment of each slot, null: all in environment 0

    // ---- agent grid (rebuilt every step) ----
    uint32_t n_keys;        // cells of all environments
    uint32_t* key;          // [n_slots] cell key or 0xffffffff
    uint32_t* rank;         // [n_slots] arrival order of the agent in its cell
    uint32_t* cell_count;   // [n_keys]   zero between steps (the scan clears it)
    uint32_t* cell_start;   // [n_keys + 1] exclusive scan; [n_keys] = active count
    float4* sorted_pr;      // [n_slots] (x, y, radius, id bits)
    float4* sorted_v;       // [n_slots] (vx, vy, preferred speed, environment bits)
    int* nbr;               // [k][n_slots] neighbour sorted positions (SoA)
    int* nbr_cnt;           // [n_slots]   by sorted position
    // neighbour-search seeds: each agent's neighbours of the previous step
    // bound its k-th distance before the search starts
    const int* nbr_prev;      // the previous step's nbr and nbr_cnt (by its sorted positions)
    const int* nbr_cnt_prev;
    int* fwd;                 // [n_slots] previous sorted position -> this step's, -1: left the grid
    int* prev_of;             // [n_slots] this step's sorted position -> previous, -1: none
    int* pos_of;              // [n_slots] agent id -> its sorted position in the last step (k_unsort)
    const int* new_ids;       // agents spawned since the last step, appended to a sorted-entry build
                              // (their count in acc[15], cleared by k_attr)

    // ---- events ----
    // One list per step, read back by the host after the neighbour pass:
    // (id, x bits, y bits, 0) asks UpdatePath from the step-start position
    // (x, y); (id, 0, 0, 1) is a DestroyAgent.
    int4* events;         uint32_t* events_n;
    uint32_t* stats;      // per step: [0] 3D-LP queue length, [1] obstacle-overflow queue
                          // length, [2] locate fallbacks to the exact-y list, [3] blocks of
                          // k_attr done, [4] 3D-LP entries with more obstacle constraints
                          // than the register tile holds, [5] unused
    uint32_t* wstats;     // per step, [64][4] slots: learned proposals tried, accepted,
                          // constraints proposed, unused
    unsigned long long* acc;   // [16] totals of the finished steps; [12] steps started,
                               // [13] the last grid's active count, [14] the previous one
    // host-mapped [1 + 8 + 4 * event_peek] words: [0] number of the last
    // step that published, then its counters and first events
    uint32_t* host_xfer;
    uint32_t event_peek;
    uint32_t* blocks_done;     // blocks of k_attr finished this step
    int* ovf_agent;         // overflow queue: sorted positions
    // 3D-LP queue: agents whose 2D LP failed, with their constraints
    int* q3d_agent;         // sorted position
    int4* q3d_info;         // (n constraints, n obstacle constraints, failed index, 0)
    float2* q3d_vel;        // 2D LP result at the failure
    float4* q3d_con;        // constraints (nx, ny, px, py), slot-major: constraint c of entry q
                            // at c * q3d_cap + q (q3d_slot)
    int con_stride;         // constraints per entry: MAXO + k
    uint32_t q3d_cap;       // entries the queue holds (max_agents)
    int* q3d_big;           // entries with more obstacle constraints than the register tile

    // ---- learned proposal ----
    uint4* feat;                // BF16 network inputs by sorted position t: chunk c (8 BF16) at
                                // ((t / 32) * 8 + c) * 32 + t % 32; n_slots rounded up to 32
    const uint32_t* mlp_mask;   // [n_slots] predicted active slots by sorted position
    // training-sample log (one plain step that records its LP solutions)
    uint4* log_x;               // [log_cap][8] BF16 features
    uint32_t* log_label;        // [log_cap] label words
    float2* log_v;              // [log_cap] LP velocity
    uint32_t* log_n;            // append counter
    uint32_t log_cap;

    // ---- parameters (ORCA.h:85-103, Simulator.cpp:538-686) ----
    int k;
    float time_horizon;
    float inv_time_horizon;       // 1 / time_horizon when that is exact (a power of two), else 0
    float time_horizon_obst;
    float inv_time_horizon_obst;  // 1.0f / time_horizon_obst (ORCA.cpp:241)
    float max_speed;
    int max_speed_from_pref;
    float dt;
    float mass_recip;             // 1.0f / mass (Simulator.cpp:623)
    float arrival_r2;             // 20 * 20
    float delete_r2;              // 2 * 2
};

}  // namespace ecm
