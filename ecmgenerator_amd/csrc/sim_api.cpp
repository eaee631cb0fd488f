// sim_api.cpp -- host runtime behind include/ecm_sim.h: device buffers,
// static spatial indices (obstacles, ECM cells), the slot pool and the step
// loop.  Compiled by hipcc for gfx950; the kernels live in sim_kernels.hip.
#include "../../include/ecm_sim.h"
#include "ecm_geom.h"
#include "host_util.h"
#include "bucket_geom.h"
#include "mlp_host.h"
#include "sim_device.h"
#include "replan_pool.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace ecm {
hipError_t launch_grid_build(const DevState& s, uint32_t* bsum, hipStream_t st);
hipError_t launch_unsort(const DevState& s, hipStream_t st);
hipError_t launch_unsort_pos(const DevState& s, hipStream_t st);
hipError_t launch_mark_dead(const int* ids, int n, const int* pos_of, const float4* sorted_pr_old, uint8_t* dead_old,
                            const unsigned long long* acc, hipStream_t st);
hipError_t launch_attr_knn(const DevState& s, int block, int mode, int oslots, hipStream_t st, hipEvent_t mid);
hipError_t launch_orca(const DevState& s, int maxo, int mode, int oslots, int n_cu, hipStream_t st, hipEvent_t mid);
hipError_t launch_features(const DevState& s, hipStream_t st);
hipError_t launch_export_nbrs(const DevState& s, int* out, int* out_cnt, hipStream_t st);
hipError_t launch_spawn_scatter(int n, const int* ids, const float2* start, const float* rad, const float* spd,
                                const int* off, const int* cnt, float2* pos, float2* vel, float* radius,
                                float* pref_speed, float2* pref_vel, uint8_t* active, int* path_off,
                                int* path_cnt, hipStream_t st);
hipError_t launch_path_upload(int n, const int* stage, int npts, const float2* pts, float2* pool_dst, int* path_off,
                              int* path_cnt, int2* path_s, hipStream_t st);
hipError_t launch_path_repack(int n_slots, const unsigned long long* acc, const float4* sorted_pr_old,
                              const uint8_t* dead_s_old, int2* path_s_old, int* path_off, const float2* pool,
                              float2* out, uint32_t cap, uint32_t* ctr, hipStream_t st);
hipError_t launch_path_resort(int n_slots, const unsigned long long* acc, const float4* sorted_pr_old,
                              const int* path_off, const int* path_cnt, int2* path_s_old, hipStream_t st);
}  // namespace ecm

using namespace ecm;

static thread_local std::string g_err;
int ecm::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char* ecm_last_error(void) { return g_err.c_str(); }

extern "C" int ecm_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) != hipSuccess) return 0;
    return std::strncmp(pr.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

static GridDesc make_grid(float x0, float y0, float x1, float y1, float cell, int max_side) {
    GridDesc g;
    float ext = std::max(x1 - x0, y1 - y0);
    if (cell * max_side < ext) cell = ext / max_side;
    g.x0 = x0;
    g.y0 = y0;
    g.cell = cell;
    g.inv_cell = 1.0f / cell;
    g.side = std::max(1, (int)std::ceil(ext / cell));
    int bits = 0;
    while ((1 << bits) < g.side) ++bits;
    g.bits = bits;
    return g;
}

// One environment on the host: its scene, device copies and agent grid.
// Longest obstacle bucket list a scene may have: obstacle edges within the
// obstacle range of one point.  It sizes the 3D-LP queue's constraint stride
// (kMaxBucketEdges + k constraints of 16 B per agent slot at most).
constexpr int kMaxBucketEdges = 512;

struct EnvHost {
    float x0 = 0, y0 = 0, x1 = 0, y1 = 0;       // world box
    // obstacles
    std::vector<float> obs_xy;
    std::vector<int> obs_ring_off;
    int n_obst_edges = 0;
    DBuf<float4> oe0, oe1;
    DBuf<uint8_t> oflags;
    GridDesc og{};
    DBuf<int> obucket_start, obucket_edges;
    int maxo = 16;
    float obst_range_built = -1.0f;
    // ECM
    std::vector<float> ecm_vert;
    std::vector<int> ecm_target;
    std::vector<float> ecm_pts;
    int n_cells = 0;
    DBuf<float4> cellA, cellB;
    GridDesc eg{};
    DBuf<int> ebucket_start, ebucket_cells;
    int yhash_mask = -1;
    DBuf<uint32_t> yhash_key;
    DBuf<int> yhash_off, yhash_cnt, ycells;
    bool ecm_set = false;
    // agents and their grid
    int num_agents = 0;
    GridDesc ag{};
    uint32_t cell_base = 0;
    int grid_for_n = -1;
    // replanning (UpdatePath through this environment's planner)
    ecm_replan_fn replan_fn = nullptr;
    void* replan_user = nullptr;
    void release() {
        oe0.release(); oe1.release(); oflags.release(); obucket_start.release(); obucket_edges.release();
        cellA.release(); cellB.release(); ebucket_start.release(); ebucket_cells.release();
        yhash_key.release(); yhash_off.release(); yhash_cnt.release(); ycells.release();
    }
};

struct ecm_sim {
    ecm_sim_params prm{};
    hipStream_t stream = nullptr;
    int dev = 0;

    // agent SoA
    // slot-indexed state, current when id_current (k_unsort brings it up to
    // date from the last step's sorted outputs when the host needs it)
    DBuf<float2> pos, vel, attraction, pref_vel, v_orca;
    DBuf<float> radius, pref_speed;
    DBuf<uint8_t> active;
    DBuf<int> path_off, path_cnt;
    DBuf<float2> path_pool;
    size_t path_used = 0;
    uint64_t path_compactions = 0;           // times compact_paths rebuilt the pool
    // Morton-order repacks of the path pool (repack_paths): the second pool
    // they pack into, the packed total counted on the device and read back
    // into pinned memory, the totals the host expected
    DBuf<float2> path_alt;
    DBuf<uint32_t> repack_ctr;
    uint32_t* h_repack = nullptr;
    uint64_t repack_expect = 0;
    bool repack_pending_check = false;
    int steps_since_repack = 0;
    bool repack_soon = false;                // the pool was laid out in slot order (load, compaction)
    int cur = 0;

    // agent grid
    DBuf<uint32_t> key, rank, cell_count, cell_start, bsum;
    // sorted state by step parity: this step's [cur], the last step's [cur^1]
    DBuf<float4> sorted_pr[2], sorted_v[2];
    DBuf<float2> attr_s[2], pos_s[2], vel_s[2];
    DBuf<uint8_t> dead_s[2];
    DBuf<float2> pref_vel_s[2], v_orca_s[2];
    DBuf<int2> path_s[2];                       // path (offset, count) by sorted position
    bool sorted_valid = false;                  // the last step's sorted outputs hold the state
    bool paths_resort = false;                  // the host rewrote path_off / path_cnt since the last step
    bool id_current = true;                     // the slot-indexed arrays hold it
    bool pos_current = true;                    // the slot-indexed positions do (implied by id_current)
    DBuf<int> nbr[2], nbr_cnt[2];      // by step parity: the previous step's seed the search
    DBuf<int> fwd, prev_of, pos_of;    // previous <-> this step's sorted positions; id -> last one
    // host changes kept out of the slot-entry build: agents spawned and
    // agents removed since the last step (uploaded before the next one)
    std::vector<int> pending_new, pending_dead, new_stage, dead_stage;
    std::vector<uint8_t> h_pending;
    unsigned long long n_new_stage = 0;
    DBuf<int> new_ids, dead_ids;
    int last_grid_n = 0;               // agents in the last step's grid

    // Per-step counters and the event list share one device block, so one
    // copy brings both back: words [0..7] are the counters ([0] events,
    // [2..4] stats), then one int4 per event.  The per-wavefront slots and
    // the 64-bit totals of finished steps live apart.
    static constexpr int kEventPeek = 1024;      // events read back while k_orca runs
    DBuf<uint32_t> xfer;                     // 8 + 4 * max_agents words
    DBuf<uint32_t> wstats;                   // [64][4]
    DBuf<unsigned long long> acc;            // [16] totals (folded in by k_grid_count); [12] steps run
    // host-mapped, 1 + 8 + 4 * kEventPeek words: [0] the number of the last
    // step that published, then its counters and first events (written by
    // the last block of k_attr)
    uint32_t* h_xfer = nullptr;
    uint32_t* h_xfer_dev = nullptr;          // the same memory as the kernels see it
    uint32_t step_seq = 0;                   // steps launched (matches acc[12])
    // Replan upload staging in mapped host memory, which the upload kernel
    // reads directly; reused step after step, ev_stage marks the end of the
    // last upload that read it.
    int* h_stage = nullptr;                  // pinned ids, offsets, counts, sorted positions
    int* h_stage_dev = nullptr;              // the same memory as the kernels see it
    size_t h_stage_cap = 0;
    float2* h_pts = nullptr;                 // pinned path points
    float2* h_pts_dev = nullptr;
    size_t h_pts_cap = 0;
    hipEvent_t ev_stage = nullptr;

    // environments (0 is created with the handle)
    std::vector<std::unique_ptr<EnvHost>> envs;
    DBuf<EnvDesc> d_envs;
    bool envs_dirty = true;
    uint32_t n_keys = 0;
    int maxo = 16;                 // longest obstacle bucket list of any environment (16 / 32 / 64, then
                                   // multiples of 16 up to kMaxBucketEdges)
    int n_cu = 256;                // compute units of the device (persistent queue grids)
    float max_radius = 0.0f, max_pref_speed = 0.0f;   // over every agent spawned
    DBuf<uint16_t> agent_env;      // environment of each slot (with more than one environment)
    std::vector<uint16_t> h_env;
    bool agent_env_dirty = false;

    // 3D-LP queue (capacity max_agents)
    DBuf<int> q3d_agent;
    DBuf<int> ovf_agent;
    DBuf<int4> q3d_info;
    DBuf<float2> q3d_vel;
    DBuf<float4> q3d_con;
    DBuf<int> q3d_big;
    int q3d_stride = 0;

    // host mirrors of the slot pool
    std::vector<int> free_stack;         // top = back()
    std::vector<uint8_t> h_active;
    std::vector<float> h_radius, h_speed;
    std::vector<float2> h_goal;          // last path point per slot
    std::vector<int> h_path_cnt;
    int num_agents = 0;
    int last_idx = -1;

    // callbacks, stats, timing
    int replan_threads = 1;
    ReplanPool replan_pool;
    uint64_t stat[ECM_SIM_N_STATS] = {};
    bool timing = false;
    // [5]: after the main ORCA pass, [6]: after the attraction pass
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float last_ms[6] = {0, 0, 0, 0, 0, 0};

    // learned proposal (features + network + verification) and sample log
    ecm_mlp* mlp = nullptr;
    float mlp_theta = 0.0f;
    DBuf<uint4> feat;
    DBuf<uint32_t> mlp_mask;
    ecm_mlp* log_mlp = nullptr;     // set for the duration of ecm_sim_log_step

    // launch shape (results do not depend on it; ecm_sim_autotune picks it)
    ecm_sim_tuning tune{1.25f, 256, 2, 0.0f};   // the shape the autotuner picks at 100k-2M agents
    // the device part of a step captured as two hipGraphs (up to the replan
    // read-back, and the rest), one pair per position buffer parity,
    // recaptured when anything they were captured with changes
    struct GraphKey {
        DevState d;
        int maxo, mode, oslots, knn_block, n_cu;
        float theta;
        const void* frags;
        const void* pbias;
    };
    struct StepGraph {
        bool valid = false;
        GraphKey key;
        hipGraph_t graph[2] = {nullptr, nullptr};
        hipGraphExec_t exec[2] = {nullptr, nullptr};
    };
    bool use_graph = false;
    StepGraph graphs[2][2];                     // [parity][entry from the sorted outputs]
    std::string graph_error;
    std::vector<int> h_replans;
};

extern "C" void ecm_sim_default_params(ecm_sim_params* p) {
    if (!p) return;
    p->max_agents = 10000;
    p->k = 10;
    p->time_horizon = 2.0f;
    p->time_horizon_obst = 2.0f;
    p->max_speed = 2.0f;
    p->max_speed_from_pref = 0;
    p->dt = 0.1f;
    p->mass = 0.8f;
    p->arrival_radius = 20.0f;
    p->delete_radius = 2.0f;
    p->grid_cell = 0.0f;
    p->keep_orca_velocity = 0;
}

extern "C" void ecm_sim_reference_params(ecm_sim_params* p) {
    ecm_sim_default_params(p);
    if (!p) return;
    p->k = 5;                       // Simulator.cpp:55
    p->time_horizon = 10.0f;        // ORCA.h:102
    p->time_horizon_obst = 10.0f;   // ORCA.h:103
    p->max_speed_from_pref = 1;     // Simulator.cpp:671-674
}

static void free_all(ecm_sim* s) {
    s->pos.release(); s->vel.release();
    s->attraction.release(); s->pref_vel.release(); s->v_orca.release();
    for (int b = 0; b < 2; ++b) {
        s->sorted_pr[b].release(); s->sorted_v[b].release(); s->attr_s[b].release();
        s->pos_s[b].release(); s->vel_s[b].release(); s->dead_s[b].release(); s->path_s[b].release();
    }
    for (int b = 0; b < 2; ++b) { s->pref_vel_s[b].release(); s->v_orca_s[b].release(); }
    s->radius.release(); s->pref_speed.release(); s->active.release();
    s->path_off.release(); s->path_cnt.release(); s->path_pool.release();
    s->key.release(); s->rank.release(); s->cell_count.release(); s->cell_start.release(); s->bsum.release();
    for (int b = 0; b < 2; ++b) { s->nbr[b].release(); s->nbr_cnt[b].release(); }
    s->fwd.release(); s->prev_of.release(); s->pos_of.release();
    s->new_ids.release(); s->dead_ids.release();
    s->xfer.release(); s->wstats.release(); s->acc.release();
    s->q3d_agent.release(); s->ovf_agent.release(); s->q3d_info.release(); s->q3d_vel.release(); s->q3d_con.release(); s->q3d_big.release();
    for (auto& e : s->envs) e->release();
    s->d_envs.release();
    s->agent_env.release();
    if (s->h_xfer) (void)hipHostFree(s->h_xfer);
    s->h_xfer = nullptr;
    if (s->h_repack) (void)hipHostFree(s->h_repack);
    s->h_repack = nullptr;
    s->path_alt.release();
    s->repack_ctr.release();
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    if (s->h_pts) (void)hipHostFree(s->h_pts);
    s->h_stage = s->h_stage_dev = nullptr;
    s->h_pts = s->h_pts_dev = nullptr;
    s->h_stage_cap = s->h_pts_cap = 0;
    s->feat.release();
    s->mlp_mask.release();
    for (auto& row : s->graphs)
        for (auto& g : row)
            for (int h = 0; h < 2; ++h) {
                if (g.exec[h]) (void)hipGraphExecDestroy(g.exec[h]);
                if (g.graph[h]) (void)hipGraphDestroy(g.graph[h]);
                g.exec[h] = nullptr;
                g.graph[h] = nullptr;
                g.valid = false;
            }
    for (auto& e : s->ev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (s->ev_stage) (void)hipEventDestroy(s->ev_stage);
    s->ev_stage = nullptr;
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->stream = nullptr;
}

extern "C" void ecm_sim_destroy(ecm_sim* s) {
    if (!s) return;
    (void)hipSetDevice(s->dev);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
}

static float bits_to_float(int v) {
    float f;
    std::memcpy(&f, &v, sizeof(f));
    return f;
}

// Waits for the work still queued on the handle's stream.  A step returns
// once its host part is done while its last kernels may still run, so
// every entry point that reads or replaces device data settles first.
static int settle(ecm_sim* s) {
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}
// settle + bring the slot-indexed agent arrays up to date (defined after
// make_state); for the entry points that read or change agents by slot.
static int sync_state(ecm_sim* s);

static int zero_dev(void* p, size_t bytes, hipStream_t st) {
    HIPCHK(hipMemsetAsync(p, 0, bytes, st));
    return 0;
}
static int fill_dev_bytes(void* p, int byte, size_t bytes, hipStream_t st) {
    HIPCHK(hipMemsetAsync(p, byte, bytes, st));
    return 0;
}

extern "C" int ecm_sim_create(const ecm_sim_params* p, float x0, float y0, float x1, float y1, ecm_sim** out) {
    if (!p || !out) return fail(ECM_E_ARG, "null argument");
    *out = nullptr;
    if (p->max_agents <= 0 || p->k < 1 || p->k > 16) return fail(ECM_E_ARG, "max_agents > 0 and 1 <= k <= 16 required");
    if (!(x1 > x0) || !(y1 > y0)) return fail(ECM_E_ARG, "empty world box");
    if (!(p->dt > 0) || !(p->time_horizon > 0) || !(p->time_horizon_obst > 0) || !(p->mass > 0))
        return fail(ECM_E_ARG, "dt, time horizons and mass must be positive");
    if (!ecm_device_available()) return fail(ECM_E_NOGPU, "no gfx950 device available");
    ecm_sim* s = new ecm_sim();
    s->prm = *p;
    {
        std::unique_ptr<EnvHost> e(new EnvHost());
        e->x0 = x0; e->y0 = y0; e->x1 = x1; e->y1 = y1;
        s->envs.push_back(std::move(e));
    }
    const size_t N = (size_t)p->max_agents;
    int rc = 0;
    if (hipGetDevice(&s->dev) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(ECM_E_HIP, "stream creation failed");
    }
    {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, s->dev) == hipSuccess && pr.multiProcessorCount > 0)
            s->n_cu = pr.multiProcessorCount;
    }
    if (!rc) rc |= s->pos.alloc(N);
    if (!rc) rc |= s->vel.alloc(N);
    if (!rc) rc |= s->attraction.alloc(N);
    if (!rc) rc |= s->pref_vel.alloc(N);
    if (!rc && p->keep_orca_velocity) rc |= s->v_orca.alloc(N);
    for (int b = 0; b < 2 && !rc; ++b) {
        rc |= s->sorted_pr[b].alloc(N); rc |= s->sorted_v[b].alloc(N); rc |= s->attr_s[b].alloc(N);
        rc |= s->pos_s[b].alloc(N); rc |= s->vel_s[b].alloc(N); rc |= s->dead_s[b].alloc(N);
        rc |= s->path_s[b].alloc(N);
    }
    for (int b = 0; b < 2 && !rc; ++b) {
        rc |= s->pref_vel_s[b].alloc(N);
        if (p->keep_orca_velocity) rc |= s->v_orca_s[b].alloc(N);
    }
    if (!rc) rc |= s->radius.alloc(N);
    if (!rc) rc |= s->pref_speed.alloc(N);
    if (!rc) rc |= s->active.alloc(N);
    if (!rc) rc |= s->path_off.alloc(N);
    if (!rc) rc |= s->path_cnt.alloc(N);
    if (!rc) rc |= s->path_pool.alloc(std::max<size_t>(N * 8, 1024));
    if (!rc) rc |= s->key.alloc(N);
    if (!rc) rc |= s->rank.alloc(N);
    for (int b = 0; b < 2 && !rc; ++b) {
        rc |= s->nbr[b].alloc(N * (size_t)p->k);
        rc |= s->nbr_cnt[b].alloc(N);
    }
    if (!rc) rc |= s->fwd.alloc(N);
    if (!rc) rc |= s->prev_of.alloc(N);
    if (!rc) rc |= s->pos_of.alloc(N);
    if (!rc) rc |= s->new_ids.alloc(N);
    if (!rc) rc |= s->dead_ids.alloc(N);
    if (!rc) s->h_pending.assign(N, 0);
    if (!rc) rc |= s->xfer.alloc(8 + 4 * N);
    if (!rc) rc |= s->wstats.alloc(256);
    if (!rc) rc |= s->acc.alloc(16);
    if (!rc) rc |= s->agent_env.alloc(N);
    if (!rc && hipHostMalloc((void**)&s->h_xfer, (1 + 8 + 4 * (size_t)ecm_sim::kEventPeek) * sizeof(uint32_t),
                             hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        rc = fail(ECM_E_HIP, "hipHostMalloc failed");
    if (!rc && hipHostGetDevicePointer((void**)&s->h_xfer_dev, s->h_xfer, 0) != hipSuccess)
        rc = fail(ECM_E_HIP, "hipHostGetDevicePointer failed");
    if (!rc) s->h_xfer[0] = 0u;
    if (!rc) rc |= s->repack_ctr.alloc(2);
    if (!rc && hipHostMalloc((void**)&s->h_repack, 2 * sizeof(uint32_t)) != hipSuccess)
        rc = fail(ECM_E_HIP, "hipHostMalloc failed");
    if (!rc && hipEventCreateWithFlags(&s->ev_stage, hipEventDisableTiming) != hipSuccess)
        rc = fail(ECM_E_HIP, "event creation failed");
    for (int i = 0; i < 7 && !rc; ++i)
        if (hipEventCreate(&s->ev[i]) != hipSuccess) rc = fail(ECM_E_HIP, "event creation failed");
    if (!rc) {
        rc |= zero_dev(s->pos.p, N * sizeof(float2), s->stream);
        rc |= zero_dev(s->vel.p, N * sizeof(float2), s->stream);
        for (int b = 0; b < 2 && !rc; ++b) rc |= zero_dev(s->dead_s[b].p, N, s->stream);
        rc |= zero_dev(s->attraction.p, N * sizeof(float2), s->stream);
        rc |= zero_dev(s->pref_vel.p, N * sizeof(float2), s->stream);
        rc |= zero_dev(s->active.p, N, s->stream);
        rc |= zero_dev(s->path_cnt.p, N * sizeof(int), s->stream);
        for (int b = 0; b < 2; ++b) {      // no seeds before the first step
            rc |= zero_dev(s->nbr_cnt[b].p, N * sizeof(int), s->stream);
        }
        rc |= fill_dev_bytes(s->fwd.p, 0xff, N * sizeof(int), s->stream);
        rc |= fill_dev_bytes(s->prev_of.p, 0xff, N * sizeof(int), s->stream);
        rc |= fill_dev_bytes(s->pos_of.p, 0xff, N * sizeof(int), s->stream);
        rc |= zero_dev(s->path_off.p, N * sizeof(int), s->stream);
        rc |= zero_dev(s->radius.p, N * sizeof(float), s->stream);
        rc |= zero_dev(s->pref_speed.p, N * sizeof(float), s->stream);
        rc |= zero_dev(s->xfer.p, 8 * sizeof(uint32_t), s->stream);
        rc |= zero_dev(s->wstats.p, 256 * sizeof(uint32_t), s->stream);
        rc |= zero_dev(s->acc.p, 16 * sizeof(unsigned long long), s->stream);
        if (s->v_orca.p) rc |= zero_dev(s->v_orca.p, N * sizeof(float2), s->stream);
        if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(ECM_E_HIP, "sync failed");
    }
    if (rc) { free_all(s); delete s; return rc; }
    // LIFO pool, top = slot 0 (Simulator.h:66-69)
    s->free_stack.reserve(N);
    for (int i = (int)N - 1; i >= 0; --i) s->free_stack.push_back(i);
    s->h_active.assign(N, 0);
    s->h_radius.assign(N, 0.0f);
    s->h_speed.assign(N, 0.0f);
    s->h_goal.assign(N, make_float2(0.0f, 0.0f));
    s->h_path_cnt.assign(N, 0);
    s->h_env.assign(N, 0);
    *out = s;
    return 0;
}

// ------------------------------------------------------------- obstacles
// Largest obstacle query range of any agent ever spawned (ORCA.cpp:27),
// a conservative bound for the obstacle buckets.
static float obstacle_range(const ecm_sim* s) {
    float vmax = s->prm.max_speed_from_pref ? s->max_pref_speed : s->prm.max_speed;
    return s->prm.time_horizon_obst * vmax + s->max_radius;
}

// squared distance from segment ab to the box [bx0,bx1]x[by0,by1] (double)
static double seg_box_dist2(double ax, double ay, double bx, double by, double x0, double y0, double x1, double y1) {
    auto inside = [&](double x, double y) { return x >= x0 && x <= x1 && y >= y0 && y <= y1; };
    if (inside(ax, ay) || inside(bx, by)) return 0.0;
    auto pt_seg = [](double px, double py, double sx, double sy, double tx, double ty) {
        double dx = tx - sx, dy = ty - sy, l2 = dx * dx + dy * dy;
        double u = l2 > 0 ? ((px - sx) * dx + (py - sy) * dy) / l2 : 0.0;
        u = std::min(1.0, std::max(0.0, u));
        double qx = sx + u * dx - px, qy = sy + u * dy - py;
        return qx * qx + qy * qy;
    };
    auto segs_cross = [](double ax, double ay, double bx, double by, double cx, double cy, double dx, double dy) {
        auto orient = [](double px, double py, double qx, double qy, double rx, double ry) {
            double v = (qx - px) * (ry - py) - (qy - py) * (rx - px);
            return (v > 0) - (v < 0);
        };
        int o1 = orient(ax, ay, bx, by, cx, cy), o2 = orient(ax, ay, bx, by, dx, dy);
        int o3 = orient(cx, cy, dx, dy, ax, ay), o4 = orient(cx, cy, dx, dy, bx, by);
        return o1 * o2 <= 0 && o3 * o4 <= 0;
    };
    double ex[4][4] = {{x0, y0, x1, y0}, {x1, y0, x1, y1}, {x1, y1, x0, y1}, {x0, y1, x0, y0}};
    double best = 1e300;
    for (auto& e : ex) {
        if (segs_cross(ax, ay, bx, by, e[0], e[1], e[2], e[3])) return 0.0;
        best = std::min(best, pt_seg(e[0], e[1], ax, ay, bx, by));
        best = std::min(best, pt_seg(e[2], e[3], ax, ay, bx, by));
        best = std::min(best, pt_seg(ax, ay, e[0], e[1], e[2], e[3]));
        best = std::min(best, pt_seg(bx, by, e[0], e[1], e[2], e[3]));
    }
    return best;
}

// Bucket every obstacle edge into the cells within `range` of it (capsule
// test), lists ascending by edge id = (obstacle, vertex) scan order.
static int build_obstacle_buckets(ecm_sim* s, EnvHost* E, float range) {
    const float margin = 0.05f;
    const double R = (double)range + margin;
    float cell = 2.0f;
    size_t prev_maxlen = SIZE_MAX;
    for (int attempt = 0; attempt < 6; ++attempt, cell *= 0.5f) {
        GridDesc g = make_grid(E->x0, E->y0, E->x1, E->y1, cell, 4096);
        const int side = g.side;
        std::vector<std::vector<int>> lists((size_t)side * side);
        // edge endpoints from the host copy
        const std::vector<int>& ro = E->obs_ring_off;
        int nr = (int)ro.size() - 1;
        for (int r = 0; r < nr; ++r) {
            int b = ro[r], m = ro[r + 1] - b;
            for (int i = 0; i < m; ++i) {
                int e = b + i, f = b + (i + 1) % m;
                double ax = E->obs_xy[2 * e], ay = E->obs_xy[2 * e + 1];
                double bx = E->obs_xy[2 * f], by = E->obs_xy[2 * f + 1];
                double lx = std::min(ax, bx) - R, hx = std::max(ax, bx) + R;
                double ly = std::min(ay, by) - R, hy = std::max(ay, by) + R;
                int cx0 = grid_cell((float)lx, g.x0, g.inv_cell, side), cx1 = grid_cell((float)hx, g.x0, g.inv_cell, side);
                int cy0 = grid_cell((float)ly, g.y0, g.inv_cell, side), cy1 = grid_cell((float)hy, g.y0, g.inv_cell, side);
                for (int cy = cy0; cy <= cy1; ++cy)
                    for (int cx = cx0; cx <= cx1; ++cx) {
                        double bx0 = g.x0 + (double)cx * g.cell, by0 = g.y0 + (double)cy * g.cell;
                        double bx1 = bx0 + g.cell, by1 = by0 + g.cell;
                        // cells on the grid border also hold points beyond the box
                        if (cx == 0) bx0 = -1e30;
                        if (cy == 0) by0 = -1e30;
                        if (cx == side - 1) bx1 = 1e30;
                        if (cy == side - 1) by1 = 1e30;
                        if (seg_box_dist2(ax, ay, bx, by, bx0, by0, bx1, by1) <= R * R)
                            lists[(size_t)cy * side + cx].push_back(e);
                    }
            }
        }
        size_t maxlen = 0, total = 0;
        for (auto& l : lists) { maxlen = std::max(maxlen, l.size()); total += l.size(); }
        // Finer cells shorten the lists while the cells are large against
        // the obstacle range.  Once halving the cell takes less than a tenth
        // off the longest list, the edges within range of single points
        // dominate it and finer cells only cost memory: the overflow passes
        // then take a tile sized for the longest list (up to kMaxBucketEdges).
        const bool shrinking = prev_maxlen == SIZE_MAX || maxlen * 10 < prev_maxlen * 9;
        prev_maxlen = maxlen;
        if (maxlen > 64 && attempt < 5 && shrinking) continue;
        if (maxlen > (size_t)kMaxBucketEdges)
            return fail(ECM_E_STATE, "more than %d obstacle edges within range of one point", kMaxBucketEdges);
        std::vector<int> start((size_t)side * side + 1), edges;
        edges.reserve(total);
        for (size_t c = 0; c < lists.size(); ++c) {
            start[c] = (int)edges.size();
            edges.insert(edges.end(), lists[c].begin(), lists[c].end());
        }
        start[lists.size()] = (int)edges.size();
        int rc = E->obucket_start.alloc(start.size());
        if (!rc) rc = E->obucket_edges.alloc(std::max<size_t>(edges.size(), 1));
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(E->obucket_start.p, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
        if (!edges.empty())
            HIPCHK(hipMemcpyAsync(E->obucket_edges.p, edges.data(), edges.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        E->og = g;
        s->envs_dirty = true;
        E->maxo = maxlen <= 16 ? 16 : (maxlen <= 32 ? 32 : (maxlen <= 64 ? 64 : (int)((maxlen + 15) & ~(size_t)15)));
        E->obst_range_built = range;
        return 0;
    }
    return fail(ECM_E_STATE, "obstacle bucketing failed");
}

extern "C" int ecm_sim_env_set_obstacles(ecm_sim* s, int env, const float* xy, const int* ring_off, int n_rings) {
    if (!s || n_rings < 0 || (n_rings > 0 && (!xy || !ring_off))) return fail(ECM_E_ARG, "bad obstacle arguments");
    if (env < 0 || env >= (int)s->envs.size()) return fail(ECM_E_ARG, "no environment %d", env);
    EnvHost* E = s->envs[env].get();
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    std::vector<int> ro(ring_off, ring_off + n_rings + 1);
    if (n_rings == 0) ro.assign(1, 0);
    if (ro[0] != 0) return fail(ECM_E_ARG, "ring_off[0] must be 0");
    for (int r = 0; r < n_rings; ++r)
        if (ro[r + 1] - ro[r] < 2) return fail(ECM_E_ARG, "obstacle %d has fewer than 2 vertices (Environment.cpp:288)", r);
    const int nv = ro[n_rings];
    // unchanged rings (a snapshot of the same scene): keep the buckets
    if (E->obs_ring_off == ro && E->obs_xy.size() == 2 * (size_t)nv &&
        (nv == 0 || std::memcmp(E->obs_xy.data(), xy, 2 * (size_t)nv * sizeof(float)) == 0))
        return 0;
    E->obs_xy.assign(xy, xy + 2 * (size_t)nv);
    E->obs_ring_off = ro;
    E->n_obst_edges = nv;
    if (nv == 0) { E->obst_range_built = -1.0f; s->envs_dirty = true; return 0; }
    // Obstacle::Initialize (ECMDataTypes.cpp:23-61): prev/next in the ring,
    // isConvex = det(prev - next, p - prev) >= 0 (true for 2-vertex rings).
    std::vector<float4> e0(nv), e1(nv);
    std::vector<uint8_t> fl(nv);
    std::vector<uint8_t> convex(nv);
    std::vector<int> nxt(nv), prv(nv);
    for (int r = 0; r < n_rings; ++r) {
        int b = ro[r], m = ro[r + 1] - b;
        for (int i = 0; i < m; ++i) {
            int v = b + i;
            prv[v] = (i == 0) ? b + m - 1 : b + i - 1;
            nxt[v] = (i == 0) ? b + 1 : b + (i + 1) % m;
        }
        for (int i = 0; i < m; ++i) {
            int v = b + i;
            if (m == 2) { convex[v] = 1; continue; }
            V2 P = mk(xy[2 * v], xy[2 * v + 1]);
            V2 A = mk(xy[2 * prv[v]], xy[2 * prv[v] + 1]);
            V2 B = mk(xy[2 * nxt[v]], xy[2 * nxt[v] + 1]);
            convex[v] = det(A - B, P - A) >= 0.0f;
        }
    }
    for (int v = 0; v < nv; ++v) {
        int R = nxt[v];
        e0[v] = make_float4(xy[2 * v], xy[2 * v + 1], xy[2 * R], xy[2 * R + 1]);
        e1[v] = make_float4(xy[2 * prv[v]], xy[2 * prv[v] + 1], xy[2 * nxt[R]], xy[2 * nxt[R] + 1]);
        fl[v] = (uint8_t)(convex[v] | (convex[R] << 1));
    }
    int rc = E->oe0.alloc(nv);
    if (!rc) rc = E->oe1.alloc(nv);
    if (!rc) rc = E->oflags.alloc(nv);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(E->oe0.p, e0.data(), nv * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(E->oe1.p, e1.data(), nv * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(E->oflags.p, fl.data(), nv, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    E->obst_range_built = -1.0f;   // (re)bucketed at the next step
    s->envs_dirty = true;
    return 0;
}

extern "C" int ecm_sim_set_obstacles(ecm_sim* s, const float* xy, const int* ring_off, int n_rings) {
    return ecm_sim_env_set_obstacles(s, 0, xy, ring_off, n_rings);
}

// ------------------------------------------------------------------- ECM
static uint32_t hash_u32_host(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

extern "C" int ecm_sim_env_set_ecm(ecm_sim* s, int env, int n_vert, const float* vert_xy, int n_edge,
                                   const int* he_target, const float* he_pts) {
    if (!s || n_vert < 0 || n_edge < 0 || (n_edge > 0 && (!vert_xy || !he_target || !he_pts)))
        return fail(ECM_E_ARG, "bad ECM arguments");
    if (env < 0 || env >= (int)s->envs.size()) return fail(ECM_E_ARG, "no environment %d", env);
    EnvHost* E = s->envs[env].get();
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    for (int e = 0; e < 2 * n_edge; ++e)
        if (he_target[e] < 0 || he_target[e] >= n_vert) return fail(ECM_E_ARG, "half-edge target out of range");
    E->ecm_vert.assign(vert_xy, vert_xy + 2 * (size_t)n_vert);
    E->ecm_target.assign(he_target, he_target + 2 * (size_t)n_edge);
    E->ecm_pts.assign(he_pts, he_pts + 8 * (size_t)n_edge);
    const int nc = 2 * n_edge;
    E->n_cells = nc;
    E->ecm_set = true;
    s->envs_dirty = true;
    if (nc == 0) { E->yhash_mask = -1; return 0; }
    // cells: left (he0.cl, he1.cr), right (he0.cr, he1.cl); ring
    // (v(he1.target), b.p0, b.p1, v(he0.target)) (ECMCellCollection.cpp:18-46, 57-90)
    std::vector<float4> A(nc), B(nc);
    for (int e = 0; e < n_edge; ++e) {
        const float* v0 = &vert_xy[2 * he_target[2 * e + 1]];
        const float* v1 = &vert_xy[2 * he_target[2 * e + 0]];
        const float* h = &he_pts[8 * e];
        A[2 * e] = make_float4(v0[0], v0[1], h[0], h[1]);
        B[2 * e] = make_float4(h[6], h[7], v1[0], v1[1]);
        A[2 * e + 1] = make_float4(v0[0], v0[1], h[2], h[3]);
        B[2 * e + 1] = make_float4(h[4], h[5], v1[0], v1[1]);
    }
    // bucket grid: a quarter of the root of the mean cell box area, in
    // [1, 4] m, so that a bucket lists few cells
    double area = 0.0;
    float bx0 = E->x0, by0 = E->y0, bx1 = E->x1, by1 = E->y1;
    std::vector<float> box(4 * (size_t)nc);
    const float margin = 0.01f;
    for (int c = 0; c < nc; ++c) {
        float xs[4] = {A[c].x, A[c].z, B[c].x, B[c].z}, ys[4] = {A[c].y, A[c].w, B[c].y, B[c].w};
        float lx = *std::min_element(xs, xs + 4), hx = *std::max_element(xs, xs + 4);
        float ly = *std::min_element(ys, ys + 4), hy = *std::max_element(ys, ys + 4);
        box[4 * c] = lx - margin; box[4 * c + 1] = ly - margin; box[4 * c + 2] = hx + margin; box[4 * c + 3] = hy + margin;
        area += (double)(hx - lx) * (hy - ly);
        bx0 = std::min(bx0, lx); by0 = std::min(by0, ly); bx1 = std::max(bx1, hx); by1 = std::max(by1, hy);
    }
    float cell = (float)std::sqrt(std::max(area / nc, 1e-4));
    cell = std::min(std::max(cell * 0.25f, 1.0f), 4.0f);
    GridDesc g = make_grid(bx0 - 1.0f, by0 - 1.0f, bx1 + 1.0f, by1 + 1.0f, cell, 4096);
    const int side = g.side;
    std::vector<int> cnt((size_t)side * side + 1, 0);
    auto range_of = [&](int c, int& cx0, int& cx1, int& cy0, int& cy1) {
        cx0 = grid_cell(box[4 * c], g.x0, g.inv_cell, side);
        cy0 = grid_cell(box[4 * c + 1], g.y0, g.inv_cell, side);
        cx1 = grid_cell(box[4 * c + 2], g.x0, g.inv_cell, side);
        cy1 = grid_cell(box[4 * c + 3], g.y0, g.inv_cell, side);
    };
    // A bucket lists a cell only if the cell's convex hull comes within
    // `hm` of the bucket square (bucket squares as grid_cell assigns points:
    // [x0 + i cell, x0 + (i + 1) cell), the border ones open outward).
    const double hm = 0.05 + 1e-5 * std::max(std::max(std::fabs((double)bx0), std::fabs((double)bx1)),
                                             std::max(std::fabs((double)by0), std::fabs((double)by1)));
    auto listed = [&](int c, int x, int y) {
        const double sx0 = x == 0 ? -1e30 : (double)g.x0 + (double)x * g.cell;
        const double sx1 = x == side - 1 ? 1e30 : (double)g.x0 + (double)(x + 1) * g.cell;
        const double sy0 = y == 0 ? -1e30 : (double)g.y0 + (double)y * g.cell;
        const double sy1 = y == side - 1 ? 1e30 : (double)g.y0 + (double)(y + 1) * g.cell;
        const double q[4][2] = {{A[c].x, A[c].y}, {A[c].z, A[c].w}, {B[c].x, B[c].y}, {B[c].z, B[c].w}};
        return hull_may_touch(q, sx0, sy0, sx1, sy1, hm);
    };
    for (int c = 0; c < nc; ++c) {
        int cx0, cx1, cy0, cy1;
        range_of(c, cx0, cx1, cy0, cy1);
        for (int y = cy0; y <= cy1; ++y)
            for (int x = cx0; x <= cx1; ++x)
                if (listed(c, x, y)) cnt[(size_t)y * side + x]++;
    }
    std::vector<int> start((size_t)side * side + 1, 0);
    for (size_t i = 0; i < (size_t)side * side; ++i) start[i + 1] = start[i] + cnt[i];
    std::vector<int> cells(start.back());
    std::vector<int> fillp(start.begin(), start.end() - 1);
    for (int c = 0; c < nc; ++c) {            // ascending c -> ascending lists
        int cx0, cx1, cy0, cy1;
        range_of(c, cx0, cx1, cy0, cy1);
        for (int y = cy0; y <= cy1; ++y)
            for (int x = cx0; x <= cx1; ++x)
                if (listed(c, x, y)) cells[fillp[(size_t)y * side + x]++] = c;
    }
    // exact-y lists: every distinct vertex y of every cell
    std::vector<std::pair<uint32_t, int>> yc;
    yc.reserve(4 * (size_t)nc);
    for (int c = 0; c < nc; ++c) {
        float ys[4] = {A[c].y, A[c].w, B[c].y, B[c].w};
        for (float y : ys) {
            uint32_t k;
            float yy = (y == 0.0f) ? 0.0f : y;
            std::memcpy(&k, &yy, 4);
            if (y != y) continue;
            yc.push_back({k, c});
        }
    }
    std::sort(yc.begin(), yc.end());
    yc.erase(std::unique(yc.begin(), yc.end()), yc.end());
    std::vector<uint32_t> keys;
    std::vector<int> koff, kcnt, ycells;
    for (size_t i = 0; i < yc.size();) {
        size_t j = i;
        keys.push_back(yc[i].first);
        koff.push_back((int)ycells.size());
        while (j < yc.size() && yc[j].first == yc[i].first) ycells.push_back(yc[j++].second);
        kcnt.push_back((int)(j - i));
        i = j;
    }
    size_t tsize = 1;
    while (tsize < 2 * keys.size() + 2) tsize <<= 1;
    std::vector<uint32_t> tkey(tsize, 0xffffffffu);
    std::vector<int> toff(tsize, 0), tcnt(tsize, 0);
    for (size_t i = 0; i < keys.size(); ++i) {
        uint32_t h = hash_u32_host(keys[i]) & (uint32_t)(tsize - 1);
        while (tkey[h] != 0xffffffffu) h = (h + 1) & (uint32_t)(tsize - 1);
        tkey[h] = keys[i]; toff[h] = koff[i]; tcnt[h] = kcnt[i];
    }
    int rc = E->cellA.alloc(nc);
    if (!rc) rc = E->cellB.alloc(nc);
    if (!rc) rc = E->ebucket_start.alloc(start.size());
    if (!rc) rc = E->ebucket_cells.alloc(cells.size());
    if (!rc) rc = E->yhash_key.alloc(tsize);
    if (!rc) rc = E->yhash_off.alloc(tsize);
    if (!rc) rc = E->yhash_cnt.alloc(tsize);
    if (!rc) rc = E->ycells.alloc(ycells.size());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(E->cellA.p, A.data(), nc * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(E->cellB.p, B.data(), nc * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(E->ebucket_start.p, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
    if (!cells.empty())
        HIPCHK(hipMemcpyAsync(E->ebucket_cells.p, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(E->yhash_key.p, tkey.data(), tsize * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(E->yhash_off.p, toff.data(), tsize * sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(E->yhash_cnt.p, tcnt.data(), tsize * sizeof(int), hipMemcpyHostToDevice, s->stream));
    if (!ycells.empty())
        HIPCHK(hipMemcpyAsync(E->ycells.p, ycells.data(), ycells.size() * sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    E->eg = g;
    E->yhash_mask = (int)tsize - 1;
    return 0;
}

extern "C" int ecm_sim_set_ecm(ecm_sim* s, int n_vert, const float* vert_xy, int n_edge, const int* he_target,
                               const float* he_pts) {
    return ecm_sim_env_set_ecm(s, 0, n_vert, vert_xy, n_edge, he_target, he_pts);
}

// ------------------------------------------------------------- agents
// Spawns, replans and set_path append to the path pool; the points of
// replaced paths and of destroyed agents stay behind (the reference frees
// them, Simulator.cpp:100-105).  compact_paths rebuilds the pool with the
// active agents' paths only, in slot order; the other slots get no path.
static size_t live_path_points(const ecm_sim* s) {
    size_t live = 0;
    for (int i = 0; i < s->prm.max_agents; ++i)
        if (s->h_active[i]) live += (size_t)std::max(0, s->h_path_cnt[i]);
    return live;
}

static int compact_paths(ecm_sim* s) {
    if (int rc = settle(s)) return rc;
    const size_t N = (size_t)s->prm.max_agents;
    std::vector<int> off(N), cnt(N);
    std::vector<float2> pool(s->path_used), out;
    HIPCHK(hipMemcpy(off.data(), s->path_off.p, N * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cnt.data(), s->path_cnt.p, N * sizeof(int), hipMemcpyDeviceToHost));
    if (s->path_used) HIPCHK(hipMemcpy(pool.data(), s->path_pool.p, s->path_used * sizeof(float2), hipMemcpyDeviceToHost));
    out.reserve(live_path_points(s));
    for (size_t i = 0; i < N; ++i) {
        const bool keep = s->h_active[i] && cnt[i] > 0 && off[i] >= 0 && (size_t)off[i] + (size_t)cnt[i] <= s->path_used;
        if (!keep) {
            off[i] = 0;
            cnt[i] = 0;
            s->h_path_cnt[i] = 0;
            continue;
        }
        const size_t o = out.size();
        out.insert(out.end(), pool.begin() + off[i], pool.begin() + off[i] + cnt[i]);
        off[i] = (int)o;
    }
    if (!out.empty()) HIPCHK(hipMemcpy(s->path_pool.p, out.data(), out.size() * sizeof(float2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->path_off.p, off.data(), N * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->path_cnt.p, cnt.data(), N * sizeof(int), hipMemcpyHostToDevice));
    s->path_used = out.size();
    ++s->path_compactions;
    s->paths_resort = true;        // the last step's sorted copy holds the old offsets
    s->repack_soon = true;         // the pool is in slot order now
    return 0;
}

// Room for `extra` more points.  Path offsets are int: the pool never
// passes INT_MAX points.  Before growing, a pool whose dead points
// outnumber the live ones is compacted instead.
static int ensure_path_capacity(ecm_sim* s, size_t extra) {
    const size_t kMaxPoints = (size_t)INT_MAX;
    if (s->path_used + extra > kMaxPoints ||
        (s->path_used + extra > s->path_pool.n && s->path_used > 2 * live_path_points(s) + extra)) {
        int rc = compact_paths(s);
        if (rc) return rc;
    }
    if (s->path_used + extra > kMaxPoints)
        return fail(ECM_E_STATE, "path pool would pass %zu points (%zu live, %zu more asked)", kMaxPoints,
                    s->path_used, extra);
    if (s->path_used + extra <= s->path_pool.n) return 0;
    size_t cap = std::min(kMaxPoints, std::max(s->path_pool.n * 2, s->path_used + extra + 1024));
    float2* np = nullptr;
    if (hipMalloc((void**)&np, cap * sizeof(float2)) != hipSuccess) return fail(ECM_E_HIP, "path pool growth failed");
    if (s->path_used)
        HIPCHK(hipMemcpyAsync(np, s->path_pool.p, s->path_used * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->path_pool.release();
    s->path_pool.p = np;
    s->path_pool.n = cap;
    return 0;
}

struct HostGrid {
    float x0, y0, inv;
    int side;
    std::vector<std::vector<int>> cells;
    void init(float ax0, float ay0, float ax1, float ay1, float cell) {
        x0 = ax0; y0 = ay0; inv = 1.0f / cell;
        side = std::max(1, std::min(4096, (int)std::ceil(std::max(ax1 - ax0, ay1 - ay0) / cell)));
        cells.assign((size_t)side * side, {});
    }
    int cx(float x) const { return grid_cell(x, x0, inv, side); }
    int cy(float y) const { return grid_cell(y, y0, inv, side); }
    void add(int id, float x, float y) { cells[(size_t)cy(y) * side + cx(x)].push_back(id); }
};

extern "C" int ecm_sim_env_spawn(ecm_sim* s, int env, int n, const float* start_xy, const float* radius,
                                 const float* pref_speed, const int* path_off, const int* path_cnt, const float* path_xy,
                                 int validate, int* out_ids) {
    if (!s || n < 0 || (n > 0 && (!start_xy || !radius || !pref_speed || !path_off || !path_cnt || !path_xy || !out_ids)))
        return fail(ECM_E_ARG, "bad spawn arguments");
    if (env < 0 || env >= (int)s->envs.size()) return fail(ECM_E_ARG, "no environment %d", env);
    EnvHost* E = s->envs[env].get();
    (void)hipSetDevice(s->dev);
    if (int rc_ = sync_state(s)) return rc_;
    if (n == 0) return 0;
    for (int i = 0; i < n; ++i)
        if (path_cnt[i] < 1 || path_off[i] < 0)
            return fail(ECM_E_ARG, "agent %d: empty path (the reference reads path.x[-1], Simulator.cpp:554)", i);
    // ValidSpawnLocation (Simulator.cpp:295-311) against the active agents
    // of the environment, including those spawned earlier in this batch.
    HostGrid hg;
    float rmax_new = 0.0f;
    for (int i = 0; i < n; ++i) rmax_new = std::max(rmax_new, radius[i]);
    const float gcell = std::max(1.0f, rmax_new);
    std::vector<float2> cur_pos;
    if (validate) {
        hg.init(E->x0 - 1, E->y0 - 1, E->x1 + 1, E->y1 + 1, gcell);
        if (s->num_agents > 0) {
            cur_pos.resize(s->prm.max_agents);
            HIPCHK(hipMemcpyAsync(cur_pos.data(), s->pos.p, cur_pos.size() * sizeof(float2), hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            for (int i = 0; i <= s->last_idx; ++i)
                if (s->h_active[i] && s->h_env[i] == env) hg.add(i, cur_pos[i].x, cur_pos[i].y);
        }
    }
    std::vector<float2> placed(n);
    std::vector<int> ids(n, -1);
    std::vector<int> offs(n);
    size_t total_pts = 0;
    for (int i = 0; i < n; ++i) total_pts += (size_t)path_cnt[i];
    int rc = ensure_path_capacity(s, total_pts);
    if (rc) return rc;
    std::vector<float2> pts;
    pts.reserve(total_pts);
    for (int i = 0; i < n; ++i) {
        out_ids[i] = -1;
        if (s->free_stack.empty()) continue;                       // Simulator.cpp:170
        float x = start_xy[2 * i], y = start_xy[2 * i + 1];
        if (validate) {
            const float c2 = radius[i] * radius[i];
            bool ok = true;
            int gx = hg.cx(x), gy = hg.cy(y);
            int reach = (int)std::ceil(radius[i] * hg.inv) + 1;
            for (int yy = std::max(0, gy - reach); ok && yy <= std::min(hg.side - 1, gy + reach); ++yy)
                for (int xx = std::max(0, gx - reach); ok && xx <= std::min(hg.side - 1, gx + reach); ++xx)
                    for (int j : hg.cells[(size_t)yy * hg.side + xx]) {
                        float px = (j >= 0) ? cur_pos[j].x : placed[-j - 1].x;
                        float py = (j >= 0) ? cur_pos[j].y : placed[-j - 1].y;
                        float dx = x - px, dy = y - py;
                        if (dx * dx + dy * dy < c2) { ok = false; break; }
                    }
            if (!ok) continue;                                     // Simulator.cpp:172
        }
        int id = s->free_stack.back();
        s->free_stack.pop_back();
        ++s->num_agents;
        ++E->num_agents;
        if (s->h_env[id] != (uint16_t)env) {
            s->h_env[id] = (uint16_t)env;
            s->agent_env_dirty = true;
        }
        s->last_idx = std::max(s->last_idx, id);
        ids[i] = id;
        out_ids[i] = id;
        placed[i] = make_float2(x, y);
        if (validate) hg.add(-i - 1, x, y);
        s->h_active[id] = 1;
        s->h_radius[id] = radius[i];
        s->h_speed[id] = pref_speed[i];
        s->max_radius = std::max(s->max_radius, radius[i]);
        s->max_pref_speed = std::max(s->max_pref_speed, pref_speed[i]);
        int c = path_cnt[i];
        offs[i] = (int)(s->path_used + pts.size());
        for (int q = 0; q < c; ++q) pts.push_back(make_float2(path_xy[2 * ((size_t)path_off[i] + q)], path_xy[2 * ((size_t)path_off[i] + q) + 1]));
        s->h_goal[id] = pts.back();
        s->h_path_cnt[id] = c;
    }
    if (!pts.empty())
        HIPCHK(hipMemcpyAsync(s->path_pool.p + s->path_used, pts.data(), pts.size() * sizeof(float2), hipMemcpyHostToDevice, s->stream));
    s->path_used += pts.size();
    // upload the batch and scatter it into the SoA
    std::vector<float2> st(n);
    for (int i = 0; i < n; ++i) st[i] = make_float2(start_xy[2 * i], start_xy[2 * i + 1]);
    DBuf<int> d_ids, d_off, d_cnt;
    DBuf<float2> d_st;
    DBuf<float> d_r, d_s;
    rc = d_ids.alloc(n);
    if (!rc) rc = d_off.alloc(n);
    if (!rc) rc = d_cnt.alloc(n);
    if (!rc) rc = d_st.alloc(n);
    if (!rc) rc = d_r.alloc(n);
    if (!rc) rc = d_s.alloc(n);
    if (!rc) {
        HIPCHK(hipMemcpyAsync(d_ids.p, ids.data(), n * sizeof(int), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(d_off.p, offs.data(), n * sizeof(int), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(d_cnt.p, path_cnt, n * sizeof(int), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(d_st.p, st.data(), n * sizeof(float2), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(d_r.p, radius, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(d_s.p, pref_speed, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
        HIPCHK(launch_spawn_scatter(n, d_ids.p, d_st.p, d_r.p, d_s.p, d_off.p, d_cnt.p, s->pos.p,
                                    s->vel.p, s->radius.p, s->pref_speed.p, s->pref_vel.p, s->active.p,
                                    s->path_off.p, s->path_cnt.p, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    d_ids.release(); d_off.release(); d_cnt.release(); d_st.release(); d_r.release(); d_s.release();
    // the next step keeps entering from the last step's sorted outputs and
    // appends the new agents (read from the slot arrays)
    if (s->sorted_valid)
        for (int i = 0; i < n; ++i)
            if (ids[i] >= 0 && !s->h_pending[ids[i]]) {
                s->h_pending[ids[i]] = 1;
                s->pending_new.push_back(ids[i]);
            }
    return rc;
}

extern "C" int ecm_sim_spawn(ecm_sim* s, int n, const float* start_xy, const float* radius, const float* pref_speed,
                             const int* path_off, const int* path_cnt, const float* path_xy, int validate,
                             int* out_ids) {
    return ecm_sim_env_spawn(s, 0, n, start_xy, radius, pref_speed, path_off, path_cnt, path_xy, validate, out_ids);
}

// The next step enters from the slot arrays: host changes pending for a
// sorted-entry build are dropped (the slot arrays already hold them).
static void invalidate_sorted(ecm_sim* s) {
    s->sorted_valid = false;
    s->paths_resort = false;
    for (int id : s->pending_new) s->h_pending[id] = 0;
    s->pending_new.clear();
    s->pending_dead.clear();
}

// The path points are appended as agents spawn and replan, in no spatial
// order.  Every kPathRepackSteps steps the live paths are packed in the last
// step's Morton order (into path_alt, then copied back so the kernels keep
// the pool's address): an agent's points then share cache lines with its
// neighbours' instead of costing one or two lines of their own.  The packed
// total the device counts is checked against the host's at the next repack.
static constexpr int kPathRepackSteps = 64;

static int repack_paths(ecm_sim* s) {
    if (s->repack_pending_check) {
        // long finished: the stream has run 64 steps since, unless this
        // repack comes early (after a load or a compaction): then wait
        if (s->steps_since_repack < kPathRepackSteps)
            if (int rc = settle(s)) return rc;
        if (s->h_repack[1] != 0u || s->h_repack[0] != (uint32_t)s->repack_expect)
            return fail(ECM_E_STATE, "path repack packed %u points (expected %llu, skipped %u)", s->h_repack[0],
                        (unsigned long long)s->repack_expect, s->h_repack[1]);
        s->repack_pending_check = false;
    }
    const size_t live = live_path_points(s);
    if (live == 0 || live > s->path_pool.n || s->path_pool.n > 0xffffffffu) return 0;
    if (s->path_alt.n < s->path_pool.n) {
        if (int rc = settle(s)) return rc;
        s->path_alt.release();
        if (int rc = s->path_alt.alloc(s->path_pool.n)) return rc;
    }
    const int o = s->cur ^ 1;
    HIPCHK(hipMemsetAsync(s->repack_ctr.p, 0, 2 * sizeof(uint32_t), s->stream));
    HIPCHK(launch_path_repack(s->prm.max_agents, s->acc.p, s->sorted_pr[o].p, s->dead_s[o].p, s->path_s[o].p,
                              s->path_off.p, s->path_pool.p, s->path_alt.p, (uint32_t)s->path_pool.n,
                              s->repack_ctr.p, s->stream));
    HIPCHK(hipMemcpyAsync(s->path_pool.p, s->path_alt.p, live * sizeof(float2), hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->h_repack, s->repack_ctr.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
    s->path_used = live;
    s->repack_expect = live;
    s->repack_pending_check = true;
    return 0;
}

// Before a step that enters from the last step's sorted outputs: agents the
// host removed are marked destroyed at their last sorted position, and the
// ids of agents spawned since then go to the device with their count
// (acc[15]).  Falls back to the slot-entry build when the appended agents
// would not fit the grid build's n_slots elements.
static int upload_pending(ecm_sim* s) {
    if (!s->sorted_valid) { invalidate_sorted(s); return 0; }
    if ((size_t)s->last_grid_n + s->pending_new.size() > (size_t)s->prm.max_agents) {
        invalidate_sorted(s);
        return 0;
    }
    const int o = s->cur ^ 1;
    if (s->paths_resort) {
        HIPCHK(launch_path_resort(s->prm.max_agents, s->acc.p, s->sorted_pr[o].p, s->path_off.p, s->path_cnt.p,
                                  s->path_s[o].p, s->stream));
        s->paths_resort = false;
    }
    if (!s->pending_dead.empty()) {
        s->dead_stage.swap(s->pending_dead);
        s->pending_dead.clear();
        const int n = (int)s->dead_stage.size();
        HIPCHK(hipMemcpyAsync(s->dead_ids.p, s->dead_stage.data(), n * sizeof(int), hipMemcpyHostToDevice, s->stream));
        HIPCHK(launch_mark_dead(s->dead_ids.p, n, s->pos_of.p, s->sorted_pr[o].p, s->dead_s[o].p, s->acc.p, s->stream));
    }
    // every live agent is in the last step's sorted outputs unless some were
    // spawned since
    // also at the first step after a snapshot load or a compaction, which
    // leave the pool in slot order
    if ((++s->steps_since_repack >= kPathRepackSteps || s->repack_soon) && s->pending_new.empty()) {
        if (int rc = repack_paths(s)) return rc;
        s->steps_since_repack = 0;
        s->repack_soon = false;
    }
    if (!s->pending_new.empty()) {
        s->new_stage.swap(s->pending_new);
        s->pending_new.clear();
        for (int id : s->new_stage) s->h_pending[id] = 0;
        s->n_new_stage = s->new_stage.size();
        HIPCHK(hipMemcpyAsync(s->new_ids.p, s->new_stage.data(), s->new_stage.size() * sizeof(int),
                              hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->acc.p + 15, &s->n_new_stage, sizeof(unsigned long long), hipMemcpyHostToDevice,
                              s->stream));
    }
    return 0;
}

static void mark_destroyed(ecm_sim* s, int id) {
    // DestroyAgent (Simulator.cpp:202-208)
    --s->num_agents;
    --s->envs[s->h_env[id]]->num_agents;
    s->h_active[id] = 0;
    s->free_stack.push_back(id);
}

extern "C" int ecm_sim_destroy_agent(ecm_sim* s, int id) {
    if (!s || id < 0 || id >= s->prm.max_agents) return fail(ECM_E_ARG, "bad agent id");
    if (!s->h_active[id]) return fail(ECM_E_ARG, "agent %d is not active", id);
    (void)hipSetDevice(s->dev);
    if (int rc_ = sync_state(s)) return rc_;
    uint8_t z = 0;
    HIPCHK(hipMemcpyAsync(s->active.p + id, &z, 1, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    mark_destroyed(s, id);
    // an agent spawned since the last step leaves the list of new ones; any
    // other leaves the last step's sorted order (marked destroyed there
    // before the next step)
    if (s->sorted_valid) {
        if (s->h_pending[id]) {
            s->h_pending[id] = 0;
            s->pending_new.erase(std::find(s->pending_new.begin(), s->pending_new.end(), id));
        } else {
            s->pending_dead.push_back(id);
        }
    }
    return 0;
}

extern "C" int ecm_sim_reset(ecm_sim* s) {
    if (!s) return fail(ECM_E_ARG, "null handle");
    for (int i = 0; i <= s->last_idx; ++i)        // Simulator.cpp:325-333
        if (s->h_active[i]) {
            int rc = ecm_sim_destroy_agent(s, i);
            if (rc) return rc;
        }
    return 0;
}

extern "C" int ecm_sim_set_path(ecm_sim* s, int id, const float* xy, int n) {
    if (!s || id < 0 || id >= s->prm.max_agents || n < 1 || !xy) return fail(ECM_E_ARG, "bad set_path arguments");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    int rc = ensure_path_capacity(s, (size_t)n);
    if (rc) return rc;
    int off = (int)s->path_used;
    HIPCHK(hipMemcpyAsync(s->path_pool.p + off, xy, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->path_off.p + id, &off, sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->path_cnt.p + id, &n, sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->path_used += (size_t)n;
    s->h_goal[id] = make_float2(xy[2 * (n - 1)], xy[2 * (n - 1) + 1]);
    s->h_path_cnt[id] = n;
    s->paths_resort = true;
    return 0;
}

extern "C" int ecm_sim_env_set_replan_callback(ecm_sim* s, int env, ecm_replan_fn fn, void* user) {
    if (!s) return fail(ECM_E_ARG, "null handle");
    if (env < 0 || env >= (int)s->envs.size()) return fail(ECM_E_ARG, "no environment %d", env);
    s->envs[env]->replan_fn = fn;
    s->envs[env]->replan_user = user;
    return 0;
}

extern "C" int ecm_sim_set_replan_callback(ecm_sim* s, ecm_replan_fn fn, void* user) {
    return ecm_sim_env_set_replan_callback(s, 0, fn, user);
}

// ---------------------------------------------------------- environments
extern "C" int ecm_sim_add_env(ecm_sim* s, float x0, float y0, float x1, float y1, int* env) {
    if (!s || !env) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    if (!(x1 > x0) || !(y1 > y0)) return fail(ECM_E_ARG, "empty world box");
    if (s->envs.size() >= 65535) return fail(ECM_E_FULL, "too many environments");
    std::unique_ptr<EnvHost> e(new EnvHost());
    e->x0 = x0; e->y0 = y0; e->x1 = x1; e->y1 = y1;
    s->envs.push_back(std::move(e));
    s->envs_dirty = true;
    for (auto& q : s->envs) q->grid_for_n = -1;       // new key layout at the next step
    *env = (int)s->envs.size() - 1;
    return 0;
}

extern "C" int ecm_sim_num_envs(ecm_sim* s, int* n) {
    if (!s || !n) return fail(ECM_E_ARG, "null argument");
    *n = (int)s->envs.size();
    return 0;
}

extern "C" int ecm_sim_agent_envs(ecm_sim* s, int* env) {
    if (!s || !env) return fail(ECM_E_ARG, "null argument");
    for (int i = 0; i < s->prm.max_agents; ++i) env[i] = s->h_active[i] ? (int)s->h_env[i] : -1;
    return 0;
}

extern "C" int ecm_sim_set_replan_threads(ecm_sim* s, int n) {
    if (!s || n < 1) return fail(ECM_E_ARG, "bad replan thread count");
    s->replan_threads = n;
    return 0;
}

// ------------------------------------------------------------------ step
// Agent grid of every environment, sized by its agent count (about one
// agent per cell unless grid_cell fixes the edge), laid out one after the
// other in a shared key space.  Rebuilt when an environment's count leaves
// [n / 2, 2 n] of the count its grid was made for.
static int ensure_agent_grid(ecm_sim* s) {
    bool rebuild = false;
    for (auto& e : s->envs) {
        const int n = std::max(1, e->num_agents);
        if (e->grid_for_n < 0) rebuild = true;
        else if (s->prm.grid_cell <= 0.0f && (n > 2 * e->grid_for_n || 2 * n < e->grid_for_n)) rebuild = true;
    }
    if (!rebuild) return 0;
    int rc = settle(s);            // the previous step may still read the grid
    if (rc) return rc;
    uint64_t base = 0;
    for (auto& e : s->envs) {
        const int n = std::max(1, e->num_agents);
        float cell = s->prm.grid_cell;
        if (cell <= 0.0f) {
            double area = (double)(e->x1 - e->x0) * (e->y1 - e->y0);
            cell = (float)std::sqrt(area / n) * s->tune.cell_scale;
            cell = std::min(std::max(cell, 0.5f), 32.0f);
        }
        e->ag = make_grid(e->x0, e->y0, e->x1, e->y1, cell, 2048);
        e->cell_base = (uint32_t)base;
        e->grid_for_n = n;
        base += (uint64_t)1 << (2 * e->ag.bits);
    }
    if (base >= 0xffffffffull) return fail(ECM_E_STATE, "agent grids of the environments exceed 2^32 cells");
    const size_t G = (size_t)base;
    rc = s->cell_count.alloc(G);
    if (!rc) rc = s->cell_start.alloc(G + 1);
    if (!rc) rc = s->bsum.alloc((G + 1023) / 1024 + 1);
    if (rc) return rc;
    rc = zero_dev(s->cell_count.p, G * sizeof(uint32_t), s->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->n_keys = (uint32_t)G;
    s->envs_dirty = true;
    return 0;
}

// Device copy of the environment descriptors and of the slot -> environment
// map, refreshed when the host side changed.
static int upload_envs(ecm_sim* s) {
    if (s->envs_dirty || s->agent_env_dirty) {
        int rc = settle(s);        // the previous step may still read them
        if (rc) return rc;
    }
    if (s->envs_dirty) {
        std::vector<EnvDesc> h(s->envs.size());
        for (size_t i = 0; i < s->envs.size(); ++i) {
            const EnvHost& E = *s->envs[i];
            EnvDesc& d = h[i];
            std::memset(&d, 0, sizeof(d));
            d.ag = E.ag; d.cell_base = E.cell_base;
            d.n_obst_edges = E.n_obst_edges; d.oe0 = E.oe0.p; d.oe1 = E.oe1.p; d.oflags = E.oflags.p; d.og = E.og;
            d.obucket_start = E.obucket_start.p; d.obucket_edges = E.obucket_edges.p;
            d.n_cells = E.n_cells; d.cellA = E.cellA.p; d.cellB = E.cellB.p; d.eg = E.eg;
            d.ebucket_start = E.ebucket_start.p; d.ebucket_cells = E.ebucket_cells.p;
            d.yhash_mask = E.yhash_mask; d.yhash_key = E.yhash_key.p; d.yhash_off = E.yhash_off.p;
            d.yhash_cnt = E.yhash_cnt.p; d.ycells = E.ycells.p;
        }
        if (s->d_envs.n < h.size()) {
            int rc = s->d_envs.alloc(h.size());
            if (rc) return rc;
        }
        HIPCHK(hipMemcpyAsync(s->d_envs.p, h.data(), h.size() * sizeof(EnvDesc), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        s->envs_dirty = false;
    }
    if (s->agent_env_dirty) {
        HIPCHK(hipMemcpyAsync(s->agent_env.p, s->h_env.data(), s->h_env.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                              s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        s->agent_env_dirty = false;
    }
    return 0;
}

static DevState make_state(ecm_sim* s) {
    DevState d;
    std::memset(&d, 0, sizeof(d));
    d.n_slots = s->prm.max_agents;
    d.pos_in = s->pos.p;
    d.vel_in = s->vel.p;
    d.radius = s->radius.p; d.pref_speed = s->pref_speed.p;
    d.attraction = s->attraction.p; d.pref_vel = s->pref_vel.p; d.v_orca = s->v_orca.p; d.active = s->active.p;
    d.path_off = s->path_off.p; d.path_cnt = s->path_cnt.p; d.path_pts = s->path_pool.p;
    d.sorted_in = s->sorted_valid ? 1 : 0;
    const int o = s->cur ^ 1;
    d.sorted_pr_old = s->sorted_pr[o].p; d.sorted_v_old = s->sorted_v[o].p;
    d.pos_s_old = s->pos_s[o].p; d.vel_s_old = s->vel_s[o].p;
    d.attr_s_old = s->attr_s[o].p; d.dead_s_old = s->dead_s[o].p;
    d.pref_vel_s_old = s->pref_vel_s[o].p; d.v_orca_s_old = s->v_orca_s[o].p; d.path_s_old = s->path_s[o].p;
    d.attr_s = s->attr_s[s->cur].p; d.pref_vel_s = s->pref_vel_s[s->cur].p; d.dead_s = s->dead_s[s->cur].p;
    d.pos_s = s->pos_s[s->cur].p; d.vel_s = s->vel_s[s->cur].p; d.v_orca_s = s->v_orca_s[s->cur].p;
    d.path_s = s->path_s[s->cur].p;
    d.n_env = (int)s->envs.size();
    d.envs = s->d_envs.p;
    d.agent_env = s->envs.size() > 1 ? s->agent_env.p : nullptr;
    d.n_keys = s->n_keys;
    d.key = s->key.p; d.rank = s->rank.p; d.cell_count = s->cell_count.p; d.cell_start = s->cell_start.p;
    d.sorted_pr = s->sorted_pr[s->cur].p; d.sorted_v = s->sorted_v[s->cur].p;
    d.nbr = s->nbr[s->cur].p; d.nbr_cnt = s->nbr_cnt[s->cur].p;
    d.nbr_prev = s->nbr[s->cur ^ 1].p; d.nbr_cnt_prev = s->nbr_cnt[s->cur ^ 1].p;
    d.fwd = s->fwd.p; d.prev_of = s->prev_of.p; d.pos_of = s->pos_of.p; d.new_ids = s->new_ids.p;
    d.events = reinterpret_cast<int4*>(s->xfer.p + 8); d.events_n = s->xfer.p;
    d.stats = s->xfer.p + 2;
    d.wstats = s->wstats.p;
    d.acc = s->acc.p;
    d.host_xfer = s->h_xfer_dev;
    d.event_peek = (uint32_t)std::min(ecm_sim::kEventPeek, s->prm.max_agents);
    d.blocks_done = s->xfer.p + 5;
    d.ovf_agent = s->ovf_agent.p;
    d.feat = s->feat.p;
    d.mlp_mask = s->mlp_mask.p;
    if (s->log_mlp) {
        d.log_x = reinterpret_cast<uint4*>(s->log_mlp->ds_x.p);
        d.log_label = s->log_mlp->ds_label.p;
        d.log_v = s->log_mlp->ds_v.p;
        d.log_n = s->log_mlp->ds_count.p;
        d.log_cap = (uint32_t)s->log_mlp->ds_cap;
    } else {
        d.log_x = nullptr; d.log_label = nullptr; d.log_v = nullptr; d.log_n = nullptr; d.log_cap = 0;
    }
    d.q3d_agent = s->q3d_agent.p; d.q3d_info = s->q3d_info.p; d.q3d_vel = s->q3d_vel.p;
    d.q3d_con = s->q3d_con.p; d.con_stride = s->q3d_stride;
    d.q3d_cap = (uint32_t)s->prm.max_agents; d.q3d_big = s->q3d_big.p;
    d.k = s->prm.k;
    d.time_horizon = s->prm.time_horizon;
    {
        // x / T == x * (1 / T) bit for bit when 1 / T is exact: both round
        // the same real number (T a power of two, 1 / T not subnormal)
        int e = 0;
        const float m = std::frexp(s->prm.time_horizon, &e);
        const float inv = 1.0f / s->prm.time_horizon;
        d.inv_time_horizon = (m == 0.5f && std::isnormal(inv) && inv * s->prm.time_horizon == 1.0f) ? inv : 0.0f;
    }
    d.time_horizon_obst = s->prm.time_horizon_obst;
    d.inv_time_horizon_obst = 1.0f / s->prm.time_horizon_obst;
    d.max_speed = s->prm.max_speed;
    d.max_speed_from_pref = s->prm.max_speed_from_pref;
    d.dt = s->prm.dt;
    d.mass_recip = 1.0f / s->prm.mass;
    d.arrival_r2 = s->prm.arrival_radius * s->prm.arrival_radius;
    d.delete_r2 = s->prm.delete_radius * s->prm.delete_radius;
    return d;
}

static int sync_state(ecm_sim* s) {
    if (int rc = settle(s)) return rc;
    if (s->id_current) return 0;
    const DevState d = make_state(s);        // its "old" arrays are the last step's outputs
    HIPCHK(launch_unsort(d, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->id_current = s->pos_current = true;
    return 0;
}

// The slot-indexed arrays brought up to date on the stream, without waiting:
// for a read-back that queues its copy behind the pass and then waits once.
static int sync_state_queued(ecm_sim* s) {
    if (s->id_current) return 0;
    const DevState d = make_state(s);
    HIPCHK(launch_unsort(d, s->stream));
    s->id_current = s->pos_current = true;
    return 0;
}

// UpdatePath (Simulator.cpp:97-124, 579-587) for the agents whose IRM
// failed: planned from their step-start position to their goal.  Runs on
// replan_threads threads while the GPU finishes the step; the new paths
// are uploaded stream-ordered, before the next step reads them.
static int run_replans(ecm_sim* s, const std::vector<int>& rid, const std::vector<int>& rat,
                       const std::vector<float2>& rpos) {
    const int n = (int)rid.size();
    const int cap = 4096;
    std::vector<std::vector<float>> res(n);
    std::vector<int> cnt(n, 0);
    auto work = [&](int i) {
        // the planner writes into a per-thread buffer of cap points, made
        // once per thread; only the points it returns are kept
        thread_local std::vector<float> buf;
        if (buf.size() < 2 * (size_t)cap) buf.resize(2 * (size_t)cap);
        int id = rid[i];
        float2 g = s->h_goal[id];
        int np = 0;
        const EnvHost* E = s->envs[s->h_env[id]].get();
        int r = E->replan_fn ? E->replan_fn(E->replan_user, id, rpos[i].x, rpos[i].y, g.x, g.y, s->h_radius[id],
                                            buf.data(), cap, &np)
                             : 1;
        // an empty result, or one that claims more than cap points, keeps the old path
        cnt[i] = (r == 0 && np >= 1 && np <= cap) ? np : 0;
        if (cnt[i]) res[i].assign(buf.begin(), buf.begin() + 2 * (size_t)cnt[i]);
    };
    // an item that threw (host memory exhausted) keeps its old path, as a
    // planner failure does
    (void)s->replan_pool.run(std::max(1, s->replan_threads), n, work);
    std::vector<int> ids, offs, cnts, ats;
    std::vector<float2> pts;
    for (int i = 0; i < n; ++i) {
        if (!cnt[i]) continue;
        ids.push_back(rid[i]);
        ats.push_back(rat[i]);
        offs.push_back((int)pts.size());
        cnts.push_back(cnt[i]);
        for (int q = 0; q < cnt[i]; ++q) pts.push_back(make_float2(res[i][2 * q], res[i][2 * q + 1]));
    }
    if (ids.empty()) return 0;
    // the capacity call may compact the pool (and waits for the stream then):
    // offsets are taken after it
    int rc = ensure_path_capacity(s, pts.size());
    if (rc) return rc;
    const size_t m = ids.size();
    for (size_t j = 0; j < m; ++j) {
        const int id = ids[j];
        s->h_goal[id] = pts[(size_t)offs[j] + (size_t)cnts[j] - 1];
        s->h_path_cnt[id] = cnts[j];
        offs[j] += (int)s->path_used;
    }
    // The pinned staging is reused from step to step without a stream
    // sync: wait for the copies of the last upload that read it (long done
    // in the usual case), and grow it only when the stream is idle.
    HIPCHK(hipEventSynchronize(s->ev_stage));
    if (s->h_stage_cap < 4 * m || s->h_pts_cap < pts.size()) {
        if ((rc = settle(s)) != 0) return rc;
        const unsigned flags = hipHostMallocMapped | hipHostMallocCoherent;
        if (s->h_stage_cap < 4 * m) {
            if (s->h_stage) (void)hipHostFree(s->h_stage);
            s->h_stage = s->h_stage_dev = nullptr;
            s->h_stage_cap = 0;
            const size_t cap = std::max<size_t>(4 * m, 4 * 1024);
            if (hipHostMalloc((void**)&s->h_stage, cap * sizeof(int), flags) != hipSuccess ||
                hipHostGetDevicePointer((void**)&s->h_stage_dev, s->h_stage, 0) != hipSuccess)
                return fail(ECM_E_HIP, "hipHostMalloc (replan staging) failed");
            s->h_stage_cap = cap;
        }
        if (s->h_pts_cap < pts.size()) {
            if (s->h_pts) (void)hipHostFree(s->h_pts);
            s->h_pts = s->h_pts_dev = nullptr;
            s->h_pts_cap = 0;
            const size_t cap = std::max<size_t>(pts.size(), 32 * 1024);
            if (hipHostMalloc((void**)&s->h_pts, cap * sizeof(float2), flags) != hipSuccess ||
                hipHostGetDevicePointer((void**)&s->h_pts_dev, s->h_pts, 0) != hipSuccess)
                return fail(ECM_E_HIP, "hipHostMalloc (replan staging) failed");
            s->h_pts_cap = cap;
        }
    }
    std::memcpy(s->h_stage, ids.data(), m * sizeof(int));
    std::memcpy(s->h_stage + m, offs.data(), m * sizeof(int));
    std::memcpy(s->h_stage + 2 * m, cnts.data(), m * sizeof(int));
    std::memcpy(s->h_stage + 3 * m, ats.data(), m * sizeof(int));
    std::memcpy(s->h_pts, pts.data(), pts.size() * sizeof(float2));
    // the step that asked for them is this one ([cur]): its sorted copy is
    // what the next step carries in
    HIPCHK(launch_path_upload((int)m, s->h_stage_dev, (int)pts.size(), s->h_pts_dev, s->path_pool.p + s->path_used,
                              s->path_off.p, s->path_cnt.p, s->path_s[s->cur].p, s->stream));
    HIPCHK(hipEventRecord(s->ev_stage, s->stream));
    s->path_used += pts.size();
    return 0;
}

// Host work before a step: agent grids, obstacle buckets, environment
// descriptors and queue capacity; then the kernel arguments.
static int prepare_step(ecm_sim* s, DevState* out) {
    // UpdateMaxAgentIndex (Simulator.cpp:481-492)
    while (s->last_idx >= 0 && !s->h_active[s->last_idx]) --s->last_idx;
    int rc = ensure_agent_grid(s);
    if (rc) return rc;
    if ((rc = upload_pending(s)) != 0) return rc;
    {
        const float range = obstacle_range(s);
        int maxo = 16;
        for (auto& e : s->envs) {
            if (e->n_obst_edges > 0 && e->obst_range_built < range) {
                if ((rc = settle(s)) != 0) return rc;
                rc = build_obstacle_buckets(s, e.get(), range * 1.25f + 0.5f);
                if (rc) return rc;
            }
            if (e->n_obst_edges > 0) maxo = std::max(maxo, e->maxo);
        }
        s->maxo = maxo;
    }
    rc = upload_envs(s);
    if (rc) return rc;
    {
        int stride = s->maxo + s->prm.k;
        if (stride != s->q3d_stride) {
            if ((rc = settle(s)) != 0) return rc;
            const size_t N = (size_t)s->prm.max_agents;
            rc = s->q3d_agent.alloc(N);
            if (!rc) rc = s->ovf_agent.alloc(N);
            if (!rc) rc = s->q3d_info.alloc(N);
            if (!rc) rc = s->q3d_vel.alloc(N);
            if (!rc) rc = s->q3d_con.alloc(N * (size_t)stride);
            if (!rc) rc = s->q3d_big.alloc(N);
            if (rc) return rc;
            s->q3d_stride = stride;
        }
    }
    *out = make_state(s);
    return 0;
}

static int step_mode(const ecm_sim* s) { return s->log_mlp ? 1 : (s->mlp ? 2 : 0); }

// The device operations of one step, in stream order, in two halves: up to
// the replan read-back (half 0) and the rest (half 1).  No memset and no
// host synchronization: each half is what one step graph captures.
static int enqueue_device_half(ecm_sim* s, const DevState& d, int mode, int half) {
    if (half == 0) {
        if (s->timing) HIPCHK(hipEventRecord(s->ev[0], s->stream));
        // block 0 of the count pass also folds the previous step's counters
        // into the totals and clears them for this step
        HIPCHK(launch_grid_build(d, s->bsum.p, s->stream));
        if (s->timing) HIPCHK(hipEventRecord(s->ev[1], s->stream));
        // the plain step's main ORCA pass may run inside this kernel (k 5
        // and 10): its events are published once both are done
        HIPCHK(launch_attr_knn(d, s->tune.knn_block, mode, s->tune.orca_obst_slots, s->stream,
                               s->timing ? s->ev[6] : nullptr));
        // replan and destroy events are final once this kernel is done; its
        // last block writes them to host memory, and the host handles them
        // while the ORCA kernels (or, fused, the queue passes) run
        return 0;
    }
    if (s->timing) HIPCHK(hipEventRecord(s->ev[2], s->stream));
    if (mode == 2) {
        // learned proposal: features, network, then the verifying ORCA pass
        HIPCHK(launch_features(d, s->stream));
        HIPCHK(mlp::launch_mlp_infer(s->feat.p, 1, s->cell_start.p + s->n_keys, 0, s->mlp->frags.p, s->mlp->pbias.p,
                                     s->mlp_theta, s->mlp_mask.p, nullptr, nullptr, s->mlp->n_cu, s->stream));
    }
    if (s->timing) HIPCHK(hipEventRecord(s->ev[3], s->stream));
    HIPCHK(launch_orca(d, s->maxo, mode, s->tune.orca_obst_slots, s->n_cu, s->stream, s->timing ? s->ev[5] : nullptr));
    if (s->timing) HIPCHK(hipEventRecord(s->ev[4], s->stream));
    return 0;
}

// Waits until step `seq` has published its events (k_attr writes the
// step number last, after a system-scope fence).  The stream is polled
// now and then so that a step that ends without publishing is an error,
// not a hang.
static int wait_published(ecm_sim* s, uint32_t seq) {
    const volatile uint32_t* flag = s->h_xfer;
    for (uint64_t spin = 1; *flag != seq; ++spin) {
        if ((spin & 255u) == 0) {
            const hipError_t q = hipStreamQuery(s->stream);
            if (q == hipSuccess) {
                if (*flag == seq) break;
                s->step_seq = *flag;           // resynchronise with what the device counted
                return fail(ECM_E_STATE, "step %u finished without publishing its events", seq);
            }
            if (q != hipErrorNotReady) return fail(ECM_E_HIP, "hipStreamQuery: %s", hipGetErrorString(q));
            std::this_thread::yield();
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return 0;
}

static void drop_graph(ecm_sim::StepGraph& g) {
    for (int h = 0; h < 2; ++h) {
        if (g.exec[h]) (void)hipGraphExecDestroy(g.exec[h]);
        if (g.graph[h]) (void)hipGraphDestroy(g.graph[h]);
        g.exec[h] = nullptr;
        g.graph[h] = nullptr;
    }
    g.valid = false;
}

// Captures both halves of the step for this buffer parity when missing or
// stale.  Returns 1 when capture is not possible, after recording why in
// graph_error.
static int ensure_step_graph(ecm_sim* s, const DevState& d, int mode) {
    ecm_sim::GraphKey key;
    std::memset(&key, 0, sizeof(key));
    key.d = d;
    key.maxo = s->maxo;
    key.mode = mode;
    key.oslots = s->tune.orca_obst_slots;
    key.knn_block = s->tune.knn_block;
    key.theta = s->mlp_theta;
    key.frags = mode == 2 ? (const void*)s->mlp->frags.p : nullptr;
    key.pbias = mode == 2 ? (const void*)s->mlp->pbias.p : nullptr;
    key.n_cu = mode == 2 ? s->mlp->n_cu : 0;
    ecm_sim::StepGraph& g = s->graphs[s->cur][d.sorted_in];
    if (g.valid && std::memcmp(&g.key, &key, sizeof(key)) == 0) return 0;
    if (int rc = settle(s)) return rc;      // an earlier launch of g may still run
    drop_graph(g);
    {
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            s->graph_error = std::string("hipStreamBeginCapture: ") + hipGetErrorString(e);
            drop_graph(g);
            return 1;
        }
        int rc = enqueue_device_half(s, d, mode, 0);
        if (!rc) rc = enqueue_device_half(s, d, mode, 1);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(s->stream, &graph);
        hipGraphExec_t exec = nullptr;
        if (!rc && e == hipSuccess && graph) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (rc || e != hipSuccess || !graph || !exec) {
            if (graph) (void)hipGraphDestroy(graph);
            s->graph_error = rc ? std::string("capture: ") + ecm_last_error()
                                : std::string("graph capture / instantiate: ") + hipGetErrorString(e);
            (void)hipGetLastError();
            drop_graph(g);
            return 1;
        }
        g.graph[0] = graph;
        g.exec[0] = exec;
    }
    g.key = key;
    g.valid = true;
    return 0;
}

static int step_once(ecm_sim* s) {
    DevState d;
    int rc = prepare_step(s, &d);
    if (rc) return rc;
    const int mode = step_mode(s);
    bool graph = false;
    if (s->use_graph && !s->timing) {
        rc = ensure_step_graph(s, d, mode);
        if (rc < 0) return rc;
        if (rc == 0) graph = true;
        else s->use_graph = false;           // capture refused: eager from now on (see graph_error)
    }
    // the whole step; its first half ends by publishing the events
    if (graph) HIPCHK(hipGraphLaunch(s->graphs[s->cur][d.sorted_in].exec[0], s->stream));
    else if ((rc = enqueue_device_half(s, d, mode, 0)) != 0 || (rc = enqueue_device_half(s, d, mode, 1)) != 0)
        return rc;
    const uint32_t peek = (uint32_t)std::min(ecm_sim::kEventPeek, s->prm.max_agents);
    if ((rc = wait_published(s, ++s->step_seq)) != 0) return rc;
    const uint32_t ne = s->h_xfer[1];
    std::vector<int4> ev(std::min(ne, peek));          // events start at word 9 (not 16-byte aligned)
    if (!ev.empty()) std::memcpy(ev.data(), s->h_xfer + 1 + 8, ev.size() * sizeof(int4));
    if (ne > peek) {
        ev.resize(ne);
        HIPCHK(hipStreamSynchronize(s->stream));
        HIPCHK(hipMemcpy(ev.data() + peek, s->xfer.p + 8 + 4 * (size_t)peek, (ne - peek) * sizeof(int4),
                         hipMemcpyDeviceToHost));
    }
    // (id, x, y, sorted position) asks for a replan, (id, 0, 0, -1) is a destroy
    std::vector<int> rid, rat, dead;
    std::vector<float2> rpos;
    for (const int4& e : ev) {
        if (e.w < 0) dead.push_back(e.x);
        else {
            rid.push_back(e.x);
            rat.push_back(e.w);
            rpos.push_back(make_float2(bits_to_float(e.y), bits_to_float(e.z)));
        }
    }
    const uint32_t nr = (uint32_t)rid.size(), nd = (uint32_t)dead.size();
    // order as the reference visits slots (Simulator.cpp:545)
    {
        std::vector<int> ord(nr);
        for (uint32_t i = 0; i < nr; ++i) ord[i] = (int)i;
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return rid[a] < rid[b]; });
        std::vector<int> r2(nr), a2(nr);
        std::vector<float2> p2(nr);
        for (uint32_t i = 0; i < nr; ++i) { r2[i] = rid[ord[i]]; a2[i] = rat[ord[i]]; p2[i] = rpos[ord[i]]; }
        rid.swap(r2);
        rat.swap(a2);
        rpos.swap(p2);
    }
    s->h_replans = rid;
    bool any_planner = false;
    for (auto& e : s->envs) any_planner |= e->replan_fn != nullptr;
    if (nr && any_planner) {
        rc = run_replans(s, rid, rat, rpos);
        if (rc) return rc;
    }
    // The step returns with its ORCA kernels possibly still running: every
    // entry point that touches device data waits for the stream first, and
    // the next step is ordered behind this one on the same stream.
    s->cur ^= 1;
    s->sorted_valid = true;                     // the state now lives in this step's sorted outputs
    s->id_current = s->pos_current = false;
    const int n_before = s->num_agents;
    s->last_grid_n = n_before;                // this step's grid held every agent active at its start
    std::sort(dead.begin(), dead.end());      // reference visits slots in ascending order
    for (int id : dead) mark_destroyed(s, id);
    s->stat[0] += 1;
    s->stat[1] += (uint64_t)n_before;
    s->stat[4] += nr;
    s->stat[5] += nd;
    if (s->timing) {
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int i = 0; i < 4; ++i) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
            s->last_ms[i] = ms;
        }
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, s->ev[3], s->ev[5]));
        s->last_ms[4] = ms;
        HIPCHK(hipEventElapsedTime(&ms, s->ev[1], s->ev[6]));
        s->last_ms[5] = ms;
    }
    return 0;
}

extern "C" int ecm_sim_step(ecm_sim* s, int n_steps) {
    if (!s || n_steps < 0) return fail(ECM_E_ARG, "bad step arguments");
    for (size_t e = 0; e < s->envs.size(); ++e)
        if (!s->envs[e]->ecm_set) return fail(ECM_E_STATE, "set_ecm must be called before step (environment %zu)", e);
    (void)hipSetDevice(s->dev);
    for (int i = 0; i < n_steps; ++i) {
        int rc = step_once(s);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int ecm_sim_set_graph(ecm_sim* s, int on) {
    if (!s) return fail(ECM_E_ARG, "null handle");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    s->use_graph = on != 0;
    if (on) s->graph_error.clear();
    return 0;
}

extern "C" int ecm_sim_graph_status(ecm_sim* s, int* active, char* why, int cap) {
    if (!s || !active) return fail(ECM_E_ARG, "null argument");
    *active = s->use_graph ? 1 : 0;
    if (why && cap > 0) {
        std::strncpy(why, s->graph_error.c_str(), (size_t)cap - 1);
        why[cap - 1] = 0;
    }
    return 0;
}

static int check_tuning(const ecm_sim_tuning& t) {
    if (!(t.cell_scale >= 0.25f && t.cell_scale <= 4.0f)) return fail(ECM_E_ARG, "cell_scale must be in [0.25, 4]");
    if (t.knn_block != 16 && t.knn_block != 64 && t.knn_block != 128 && t.knn_block != 256)
        return fail(ECM_E_ARG, "knn_block must be 16, 64, 128 or 256");
    if (t.orca_obst_slots != 1 && t.orca_obst_slots != 2 && t.orca_obst_slots != 4)
        return fail(ECM_E_ARG, "orca_obst_slots must be 1, 2 or 4");
    return 0;
}

static void apply_tuning(ecm_sim* s, const ecm_sim_tuning& t) {
    if (t.cell_scale != s->tune.cell_scale)
        for (auto& e : s->envs) e->grid_for_n = -1;     // new agent grids at the next step
    const float best = s->tune.best_ms;
    s->tune = t;
    s->tune.best_ms = best;
}

extern "C" int ecm_sim_get_tuning(ecm_sim* s, ecm_sim_tuning* out) {
    if (!s || !out) return fail(ECM_E_ARG, "null argument");
    *out = s->tune;
    return 0;
}

extern "C" int ecm_sim_set_tuning(ecm_sim* s, const ecm_sim_tuning* t) {
    if (!s || !t) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    int rc = check_tuning(*t);
    if (rc) return rc;
    apply_tuning(s, *t);
    return 0;
}

// Coordinate search over the launch shape on the live workload: the agent
// grid cell scale, then the k-nearest block size, then the obstacle slots
// of the main ORCA tile.  Every trial runs one untimed step (rebuilds) and
// `steps` timed steps; the simulation advances through all of them.
extern "C" int ecm_sim_autotune(ecm_sim* s, int steps, ecm_sim_tuning* out) {
    if (!s || steps < 1) return fail(ECM_E_ARG, "bad autotune arguments");
    for (size_t e = 0; e < s->envs.size(); ++e)
        if (!s->envs[e]->ecm_set) return fail(ECM_E_STATE, "set_ecm must be called before autotune (environment %zu)", e);
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    // One sample of a shape: two untimed steps after switching to it (they
    // recapture the step graphs of both buffer parities), then one step
    // timed from an idle stream to an idle stream (the device work and the
    // host's part: events, replans).  The options of a coordinate are
    // sampled round robin, `steps` rounds, and compared by their median: the
    // crowd moves on while the search runs, and round robin spreads that
    // drift over every option.
    auto sample = [&](const ecm_sim_tuning& t, float* ms) -> int {
        apply_tuning(s, t);
        int rc = 0;
        for (int w = 0; w < 2; ++w)
            if ((rc = step_once(s)) != 0) return rc;
        HIPCHK(hipStreamSynchronize(s->stream));
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = step_once(s)) != 0) return rc;
        HIPCHK(hipStreamSynchronize(s->stream));
        *ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    };
    ecm_sim_tuning best = s->tune;
    float best_ms = 0.0f;
    int rc = 0;
    const float scales[] = {0.6f, 0.8f, 1.0f, 1.25f, 1.6f, 2.0f};
    const int blocks[] = {64, 128, 256, 16};
    const int slots[] = {1, 2, 4};
    for (int dim = 0; dim < 3; ++dim) {
        std::vector<ecm_sim_tuning> opts{best};
        const int nopt = dim == 0 ? 6 : (dim == 1 ? 4 : 3);
        for (int o = 0; o < nopt; ++o) {
            ecm_sim_tuning t = best;
            if (dim == 0) t.cell_scale = scales[o];
            else if (dim == 1) t.knn_block = blocks[o];
            else t.orca_obst_slots = slots[o];
            if (std::memcmp(&t, &best, sizeof(t)) != 0) opts.push_back(t);
        }
        std::vector<std::vector<float>> ms(opts.size());
        for (int r = 0; r < steps; ++r)
            for (size_t o = 0; o < opts.size(); ++o) {
                float m = 0.0f;
                if ((rc = sample(opts[o], &m)) != 0) return rc;
                ms[o].push_back(m);
            }
        for (size_t o = 0; o < opts.size(); ++o) {
            std::sort(ms[o].begin(), ms[o].end());
            const float med = ms[o][ms[o].size() / 2];
            if (o == 0 || med < best_ms) { best_ms = med; best = opts[o]; }
        }
    }
    apply_tuning(s, best);
    s->tune.best_ms = best_ms;
    if (out) *out = s->tune;
    return 0;
}

extern "C" int ecm_sim_sync(ecm_sim* s) {
    if (!s) return fail(ECM_E_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

static int component_buf(ecm_sim* s, int c, void** p, size_t* elem) {
    switch (c) {
        case ECM_C_POSITION: *p = s->pos.p; *elem = 8; return 0;
        case ECM_C_VELOCITY: *p = s->vel.p; *elem = 8; return 0;
        case ECM_C_PREF_VELOCITY: *p = s->pref_vel.p; *elem = 8; return 0;
        case ECM_C_ATTRACTION: *p = s->attraction.p; *elem = 8; return 0;
        case ECM_C_CLEARANCE: *p = s->radius.p; *elem = 4; return 0;
        case ECM_C_PREF_SPEED: *p = s->pref_speed.p; *elem = 4; return 0;
        case ECM_C_ACTIVE: *p = s->active.p; *elem = 1; return 0;
        case ECM_C_ORCA_VELOCITY:
            if (!s->v_orca.p) return fail(ECM_E_STATE, "keep_orca_velocity was not set");
            *p = s->v_orca.p; *elem = 8; return 0;
        default: return fail(ECM_E_ARG, "unknown component %d", c);
    }
}

extern "C" int ecm_sim_read(ecm_sim* s, int c, void* dst, size_t bytes) {
    if (!s || !dst) return fail(ECM_E_ARG, "null argument");
    void* p;
    size_t el;
    int rc = component_buf(s, c, &p, &el);
    if (rc) return rc;
    if (bytes != el * (size_t)s->prm.max_agents) return fail(ECM_E_ARG, "component %d needs %zu bytes", c, el * (size_t)s->prm.max_agents);
    (void)hipSetDevice(s->dev);
    // the slot-order pass and the copy queue behind the step; one wait.
    // Positions alone need only the positions' pass.
    if (c == ECM_C_POSITION && !s->id_current) {
        if (!s->pos_current) {
            HIPCHK(launch_unsort_pos(make_state(s), s->stream));
            s->pos_current = true;
        }
    } else if (int rc_ = sync_state_queued(s)) {
        return rc_;
    }
    HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return 0;
}

extern "C" int ecm_sim_write(ecm_sim* s, int c, const void* src, size_t bytes) {
    if (!s || !src) return fail(ECM_E_ARG, "null argument");
    void* p;
    size_t el;
    int rc = component_buf(s, c, &p, &el);
    if (rc) return rc;
    const size_t N = (size_t)s->prm.max_agents;
    if (bytes != el * N) return fail(ECM_E_ARG, "component %d needs %zu bytes", c, el * N);
    (void)hipSetDevice(s->dev);
    if (int rc_ = sync_state(s)) return rc_;
    HIPCHK(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const float* f = (const float*)src;
    if (c == ECM_C_CLEARANCE)
        for (size_t i = 0; i < N; ++i) { s->h_radius[i] = f[i]; s->max_radius = std::max(s->max_radius, f[i]); }
    if (c == ECM_C_PREF_SPEED)
        for (size_t i = 0; i < N; ++i) { s->h_speed[i] = f[i]; s->max_pref_speed = std::max(s->max_pref_speed, f[i]); }
    if (c == ECM_C_ACTIVE) {
        // rebuild the pool: inactive slots, lowest on top
        const uint8_t* a = (const uint8_t*)src;
        s->free_stack.clear();
        s->num_agents = 0;
        s->last_idx = -1;
        for (int i = (int)N - 1; i >= 0; --i) {
            s->h_active[i] = a[i] ? 1 : 0;
            if (a[i]) { ++s->num_agents; s->last_idx = std::max(s->last_idx, i); }
            else s->free_stack.push_back(i);
        }
    }
    invalidate_sorted(s);                       // the next step starts from the slot arrays
    return 0;
}

extern "C" int ecm_host_alloc(size_t bytes, void** out) {
    if (!out) return fail(ECM_E_ARG, "null argument");
    *out = nullptr;
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(ECM_E_HIP, "hipHostMalloc of %zu bytes failed", bytes);
    }
    return 0;
}

extern "C" int ecm_host_free(void* p) {
    if (p && hipHostFree(p) != hipSuccess) return fail(ECM_E_HIP, "hipHostFree failed");
    return 0;
}

extern "C" int ecm_sim_device_ptr(ecm_sim* s, int c, void** out) {
    if (!s || !out) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(s->dev);
    if (int rc_ = sync_state(s)) return rc_;
    size_t el;
    return component_buf(s, c, out, &el);
}

extern "C" int ecm_sim_num_agents(ecm_sim* s, int* out) {
    if (!s || !out) return fail(ECM_E_ARG, "null argument");
    *out = s->num_agents;
    return 0;
}
extern "C" int ecm_sim_last_index(ecm_sim* s, int* out) {
    if (!s || !out) return fail(ECM_E_ARG, "null argument");
    *out = s->last_idx;
    return 0;
}

extern "C" int ecm_sim_get_neighbors(ecm_sim* s, int* nbr, int* cnt) {
    if (!s || !nbr || !cnt) return fail(ECM_E_ARG, "null argument");
    if (s->n_keys == 0) return fail(ECM_E_STATE, "no step has run");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    const size_t N = (size_t)s->prm.max_agents;
    DBuf<int> dn, dc;
    int rc = dn.alloc(N * s->prm.k);
    if (!rc) rc = dc.alloc(N);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(dn.p, 0xff, N * s->prm.k * sizeof(int), s->stream));
    HIPCHK(hipMemsetAsync(dc.p, 0, N * sizeof(int), s->stream));
    // the last step ran with parity cur^1; sorted arrays are still those of that step
    DevState d = make_state(s);
    d.nbr = s->nbr[s->cur ^ 1].p;
    d.nbr_cnt = s->nbr_cnt[s->cur ^ 1].p;
    d.sorted_pr = s->sorted_pr[s->cur ^ 1].p;
    HIPCHK(launch_export_nbrs(d, dn.p, dc.p, s->stream));
    HIPCHK(hipMemcpyAsync(nbr, dn.p, N * s->prm.k * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(cnt, dc.p, N * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    dn.release();
    dc.release();
    return 0;
}

extern "C" int ecm_sim_read_paths(ecm_sim* s, int* off, int* cnt, float* pool, size_t pool_cap, size_t* pool_used) {
    if (!s || !pool_used) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    *pool_used = s->path_used;
    const size_t N = (size_t)s->prm.max_agents;
    if (off) HIPCHK(hipMemcpy(off, s->path_off.p, N * sizeof(int), hipMemcpyDeviceToHost));
    if (cnt) HIPCHK(hipMemcpy(cnt, s->path_cnt.p, N * sizeof(int), hipMemcpyDeviceToHost));
    if (pool) {
        if (pool_cap < s->path_used) return fail(ECM_E_ARG, "pool buffer too small (%zu points needed)", s->path_used);
        if (s->path_used) HIPCHK(hipMemcpy(pool, s->path_pool.p, s->path_used * sizeof(float2), hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int ecm_sim_get_replans(ecm_sim* s, int* ids, int cap, int* n) {
    if (!s || !n || (cap > 0 && !ids)) return fail(ECM_E_ARG, "null argument");
    *n = (int)s->h_replans.size();
    for (int i = 0; i < cap && i < *n; ++i) ids[i] = s->h_replans[i];
    return 0;
}

extern "C" int ecm_sim_stats(ecm_sim* s, uint64_t* out, int n) {
    if (!s || !out || n < 0) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(s->dev);
    if (int rc = settle(s)) return rc;
    // device totals of the finished steps + the counters of the last step,
    // which the next step's grid build has not folded in yet
    unsigned long long acc[12];
    uint32_t cur[8], ws[256];
    HIPCHK(hipMemcpy(acc, s->acc.p, sizeof(acc), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cur, s->xfer.p, sizeof(cur), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ws, s->wstats.p, sizeof(ws), hipMemcpyDeviceToHost));
    uint64_t col[4] = {acc[8], acc[9], acc[10], acc[11]};
    for (int i = 0; i < 256; ++i) col[i & 3] += ws[i];
    uint64_t st[ECM_SIM_N_STATS];
    for (int i = 0; i < ECM_SIM_N_STATS; ++i) st[i] = s->stat[i];
    st[2] = acc[2] + cur[2];        // 3D-LP queue
    st[3] = acc[4] + cur[4];        // locate fallbacks
    st[6] = acc[3] + cur[3];        // obstacle overflow queue
    st[10] = acc[6] + cur[6];       // 3D-LP entries solved from the LDS tile
    st[7] = col[0];
    st[8] = col[1];
    st[9] = col[2];
    for (int i = 0; i < n && i < ECM_SIM_N_STATS; ++i) out[i] = st[i];
    return 0;
}

extern "C" int ecm_sim_set_timing(ecm_sim* s, int on) {
    if (!s) return fail(ECM_E_ARG, "null handle");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    s->timing = on != 0;
    return 0;
}
extern "C" int ecm_sim_last_timing(ecm_sim* s, float* out, int n) {
    if (!s || !out || n < 0) return fail(ECM_E_ARG, "null argument");
    for (int i = 0; i < n && i < 6; ++i) out[i] = s->last_ms[i];
    return 0;
}

// ------------------------------------------------------ learned proposal
static int ensure_learned_buffers(ecm_sim* s) {
    const size_t N = (size_t)s->prm.max_agents;
    const size_t NF = (N + 31) / 32 * 32 * 8;      // tiles of 32 agents
    if (s->feat.n < NF || s->mlp_mask.n < N) {
        int rc = s->feat.alloc(NF);
        if (!rc) rc = s->mlp_mask.alloc(N);
        if (rc) return rc;
    }
    return 0;
}

extern "C" int ecm_sim_set_learned(ecm_sim* s, ecm_mlp* m, float theta) {
    if (!s) return fail(ECM_E_ARG, "null handle");
    if (!std::isfinite(theta)) return fail(ECM_E_ARG, "theta must be finite");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    if (m) {
        if (m->dev != s->dev) return fail(ECM_E_ARG, "network and simulator on different devices");
        int rc = ensure_learned_buffers(s);
        if (rc) return rc;
        rc = ecm_mlp_pack(m);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(m->stream));      // packed weights ready before the sim stream reads them
    }
    s->mlp = m;
    s->mlp_theta = theta;
    return 0;
}

extern "C" int ecm_sim_log_step(ecm_sim* s, ecm_mlp* m, int* n_logged) {
    if (!s || !m) return fail(ECM_E_ARG, "null argument");
    if (m->dev != s->dev) return fail(ECM_E_ARG, "network and simulator on different devices");
    if (m->ds_cap < 1) return fail(ECM_E_STATE, "reserve dataset capacity first (ecm_mlp_dataset_reserve)");
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    int64_t before = 0;
    int rc = ecm_mlp_dataset_size(m, &before);
    if (rc) return rc;
    s->log_mlp = m;
    rc = step_once(s);
    s->log_mlp = nullptr;
    if (rc) return rc;
    if ((rc = settle(s)) != 0) return rc;
    int64_t after = 0;
    rc = ecm_mlp_dataset_size(m, &after);
    if (rc) return rc;
    if (n_logged) *n_logged = (int)(after - before);
    return 0;
}

// ------------------------------------------------------------- snapshot
static const uint32_t kSnapMagic = 0x45434d53u;   // "ECMS"

template <typename T>
static bool wr(FILE* f, const T* p, size_t n) { return fwrite(p, sizeof(T), n, f) == n; }
template <typename T>
static bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

// Snapshot layout (version 2): header, parameters, counts, the agent SoA,
// the path pool, the free-slot stack, the slot -> environment map, then per
// environment its world box, obstacles, ECM and agent count.
extern "C" int ecm_sim_snapshot_save(ecm_sim* s, const char* path) {
    if (!s || !path) return fail(ECM_E_ARG, "null argument");
    (void)hipSetDevice(s->dev);
    if (int rc_ = sync_state(s)) return rc_;
    const size_t N = (size_t)s->prm.max_agents;
    std::vector<float2> pos(N), vel(N), attr(N), pv(N);
    std::vector<float> rad(N), spd(N);
    std::vector<uint8_t> act(N);
    std::vector<int> off(N), cnt(N);
    std::vector<float2> pool(s->path_used);
    int rc = ecm_sim_read(s, ECM_C_POSITION, pos.data(), N * 8);
    if (!rc) rc = ecm_sim_read(s, ECM_C_VELOCITY, vel.data(), N * 8);
    if (!rc) rc = ecm_sim_read(s, ECM_C_ATTRACTION, attr.data(), N * 8);
    if (!rc) rc = ecm_sim_read(s, ECM_C_PREF_VELOCITY, pv.data(), N * 8);
    if (!rc) rc = ecm_sim_read(s, ECM_C_CLEARANCE, rad.data(), N * 4);
    if (!rc) rc = ecm_sim_read(s, ECM_C_PREF_SPEED, spd.data(), N * 4);
    if (!rc) rc = ecm_sim_read(s, ECM_C_ACTIVE, act.data(), N);
    if (rc) return rc;
    HIPCHK(hipMemcpy(off.data(), s->path_off.p, N * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cnt.data(), s->path_cnt.p, N * sizeof(int), hipMemcpyDeviceToHost));
    if (s->path_used) HIPCHK(hipMemcpy(pool.data(), s->path_pool.p, s->path_used * sizeof(float2), hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) return fail(ECM_E_ARG, "cannot open %s", path);
    uint32_t hdr[4] = {kSnapMagic, 2u, (uint32_t)N, (uint32_t)sizeof(ecm_sim_params)};
    uint64_t pu = s->path_used, fs = s->free_stack.size();
    int32_t ints[3] = {s->num_agents, s->last_idx, (int32_t)s->envs.size()};
    bool ok = wr(f, hdr, 4) && wr(f, &s->prm, 1) && wr(f, ints, 3) && wr(f, &pu, 1) && wr(f, &fs, 1) &&
              wr(f, pos.data(), N) && wr(f, vel.data(), N) && wr(f, attr.data(), N) && wr(f, pv.data(), N) &&
              wr(f, rad.data(), N) && wr(f, spd.data(), N) && wr(f, act.data(), N) && wr(f, off.data(), N) &&
              wr(f, cnt.data(), N) && wr(f, pool.data(), pool.size()) && wr(f, s->free_stack.data(), fs) &&
              wr(f, s->h_env.data(), N);
    for (size_t e = 0; ok && e < s->envs.size(); ++e) {
        const EnvHost& E = *s->envs[e];
        float box[4] = {E.x0, E.y0, E.x1, E.y1};
        int32_t c[4] = {(int32_t)E.obs_ring_off.size(), (int32_t)(E.ecm_vert.size() / 2),
                        (int32_t)(E.ecm_target.size() / 2), E.num_agents};
        ok = wr(f, box, 4) && wr(f, c, 4) && wr(f, E.obs_ring_off.data(), E.obs_ring_off.size()) &&
             wr(f, E.obs_xy.data(), E.obs_xy.size()) && wr(f, E.ecm_vert.data(), E.ecm_vert.size()) &&
             wr(f, E.ecm_target.data(), E.ecm_target.size()) && wr(f, E.ecm_pts.data(), E.ecm_pts.size());
    }
    fclose(f);
    if (!ok) return fail(ECM_E_ARG, "write to %s failed", path);
    return 0;
}

static int snapshot_load_impl(ecm_sim* s, const char* path);

// The file is untrusted: every count is bounded by the file's size and
// every index checked before anything is allocated from it or sent to the
// device, and nothing thrown inside crosses the C ABI.
extern "C" int ecm_sim_snapshot_load(ecm_sim* s, const char* path) {
    if (!s || !path) return fail(ECM_E_ARG, "null argument");
    try {
        return snapshot_load_impl(s, path);
    } catch (const std::bad_alloc&) {
        return fail(ECM_E_ARG, "%s: snapshot too large for this host", path);
    } catch (...) {
        return fail(ECM_E_ARG, "%s: unreadable snapshot", path);
    }
}

static int snapshot_load_impl(ecm_sim* s, const char* path) {
    (void)hipSetDevice(s->dev);
    if (int rc_ = settle(s)) return rc_;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(ECM_E_ARG, "cannot open %s", path);
    uint64_t fsize = 0;
    if (fseek(f, 0, SEEK_END) == 0) {
        long e = ftell(f);
        fsize = e > 0 ? (uint64_t)e : 0;
    }
    if (fseek(f, 0, SEEK_SET) != 0) {
        fclose(f);
        return fail(ECM_E_ARG, "%s: not seekable", path);
    }
    uint32_t hdr[4];
    ecm_sim_params prm;
    int32_t ints[3];
    uint64_t pu, fs;
    bool ok = rd(f, hdr, 4) && hdr[0] == kSnapMagic && hdr[1] == 2u && hdr[3] == sizeof(ecm_sim_params) &&
              rd(f, &prm, 1) && rd(f, ints, 3) && rd(f, &pu, 1) && rd(f, &fs, 1);
    if (!ok || (int)hdr[2] != s->prm.max_agents || prm.k != s->prm.k || ints[2] < 1 ||
        ints[2] < (int)s->envs.size()) {
        fclose(f);
        return fail(ECM_E_ARG, "%s: not a snapshot for this handle (max_agents / k / environments differ)", path);
    }
    const size_t N = hdr[2];
    // counts the rest of the file must be able to hold
    if (pu > fsize / sizeof(float2) || pu > (uint64_t)INT_MAX || fs > N || (uint64_t)ints[2] > fsize / 32 ||
        ints[0] < 0 || (size_t)ints[0] > N || ints[1] < -1 || ints[1] >= (int)N) {
        fclose(f);
        return fail(ECM_E_ARG, "%s: corrupt snapshot header", path);
    }
    std::vector<float2> pos(N), vel(N), attr(N), pv(N), pool(pu);
    std::vector<float> rad(N), spd(N);
    std::vector<uint8_t> act(N);
    std::vector<int> off(N), cnt(N), stack(fs);
    std::vector<uint16_t> henv(N);
    ok = rd(f, pos.data(), N) && rd(f, vel.data(), N) && rd(f, attr.data(), N) && rd(f, pv.data(), N) &&
         rd(f, rad.data(), N) && rd(f, spd.data(), N) && rd(f, act.data(), N) && rd(f, off.data(), N) &&
         rd(f, cnt.data(), N) && rd(f, pool.data(), pu) && rd(f, stack.data(), fs) && rd(f, henv.data(), N);
    struct EnvData {
        float box[4];
        int32_t c[4];
        std::vector<int> ro;
        std::vector<float> oxy, vxy, pts;
        std::vector<int> tgt;
    };
    std::vector<EnvData> ed(ints[2]);
    for (int e = 0; ok && e < ints[2]; ++e) {
        EnvData& d = ed[e];
        ok = rd(f, d.box, 4) && rd(f, d.c, 4) && d.c[0] >= 1 && d.c[1] >= 0 && d.c[2] >= 0 && d.c[3] >= 0 &&
             (uint64_t)d.c[0] <= fsize / 4 && (uint64_t)d.c[1] <= fsize / 8 && (uint64_t)d.c[2] <= fsize / 40;
        if (!ok) break;
        d.ro.resize(d.c[0]);
        ok = rd(f, d.ro.data(), d.ro.size()) && d.ro[0] == 0;
        for (size_t r = 1; ok && r < d.ro.size(); ++r) ok = d.ro[r] >= d.ro[r - 1] + 2;   // rings of >= 2 vertices
        if (!ok || (uint64_t)d.ro.back() > fsize / 8) { ok = false; break; }
        d.oxy.resize(2 * (size_t)d.ro.back());
        d.vxy.resize(2 * (size_t)d.c[1]);
        d.tgt.resize(2 * (size_t)d.c[2]);
        d.pts.resize(8 * (size_t)d.c[2]);
        ok = rd(f, d.oxy.data(), d.oxy.size()) && rd(f, d.vxy.data(), d.vxy.size()) && rd(f, d.tgt.data(), d.tgt.size()) &&
             rd(f, d.pts.data(), d.pts.size());
        for (size_t q = 0; ok && q < d.tgt.size(); ++q) ok = d.tgt[q] >= 0 && d.tgt[q] < d.c[1];
    }
    fclose(f);
    if (!ok) return fail(ECM_E_ARG, "%s: truncated or corrupt snapshot", path);
    // every active slot has a path inside the pool and an environment of
    // the file; the free stack holds distinct inactive slots
    {
        std::vector<uint8_t> on_stack(N, 0);
        for (uint64_t i = 0; i < fs; ++i) {
            const int id = stack[i];
            if (id < 0 || (size_t)id >= N || act[id] || on_stack[id])
                return fail(ECM_E_ARG, "%s: corrupt free-slot stack", path);
            on_stack[id] = 1;
        }
        for (size_t i = 0; i < N; ++i) {
            if (cnt[i] < 0 || off[i] < 0 || (uint64_t)off[i] + (uint64_t)cnt[i] > pu ||
                (act[i] && (cnt[i] < 1 || henv[i] >= (uint16_t)ints[2])))
                return fail(ECM_E_ARG, "%s: corrupt path or environment of slot %zu", path, i);
        }
    }
    const ecm_sim_params prm_before = s->prm;
    s->prm = prm;
    int rc = 0;
    for (int e = 0; !rc && e < ints[2]; ++e) {
        const EnvData& d = ed[e];
        if (e >= (int)s->envs.size()) {
            int id = 0;
            rc = ecm_sim_add_env(s, d.box[0], d.box[1], d.box[2], d.box[3], &id);
            if (rc) break;
        }
        EnvHost* E = s->envs[e].get();
        if (E->x0 != d.box[0] || E->y0 != d.box[1] || E->x1 != d.box[2] || E->y1 != d.box[3]) {
            E->x0 = d.box[0]; E->y0 = d.box[1]; E->x1 = d.box[2]; E->y1 = d.box[3];
            E->obst_range_built = -1.0f;
        }
        rc = ecm_sim_env_set_obstacles(s, e, d.oxy.data(), d.ro.data(), (int)d.ro.size() - 1);
        if (!rc) rc = ecm_sim_env_set_ecm(s, e, d.c[1], d.vxy.data(), d.c[2], d.tgt.data(), d.pts.data());
        E->num_agents = d.c[3];
        E->grid_for_n = -1;
    }
    if (!rc) rc = ensure_path_capacity(s, pu > s->path_used ? pu - s->path_used : 0);
    if (rc) {
        s->prm = prm_before;
        return rc;
    }
    (void)hipSetDevice(s->dev);
    if (pu) HIPCHK(hipMemcpy(s->path_pool.p, pool.data(), pu * sizeof(float2), hipMemcpyHostToDevice));
    s->path_used = pu;
    HIPCHK(hipMemcpy(s->path_off.p, off.data(), N * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->path_cnt.p, cnt.data(), N * sizeof(int), hipMemcpyHostToDevice));
    rc = ecm_sim_write(s, ECM_C_POSITION, pos.data(), N * 8);
    if (!rc) rc = ecm_sim_write(s, ECM_C_VELOCITY, vel.data(), N * 8);
    if (!rc) rc = ecm_sim_write(s, ECM_C_ATTRACTION, attr.data(), N * 8);
    if (!rc) rc = ecm_sim_write(s, ECM_C_PREF_VELOCITY, pv.data(), N * 8);
    if (!rc) rc = ecm_sim_write(s, ECM_C_CLEARANCE, rad.data(), N * 4);
    if (!rc) rc = ecm_sim_write(s, ECM_C_PREF_SPEED, spd.data(), N * 4);
    if (!rc) rc = ecm_sim_write(s, ECM_C_ACTIVE, act.data(), N);
    if (rc) return rc;
    s->free_stack = stack;
    s->num_agents = ints[0];
    s->last_idx = ints[1];
    s->h_env = henv;
    s->agent_env_dirty = true;
    for (size_t i = 0; i < N; ++i) {
        s->h_path_cnt[i] = cnt[i];
        s->h_goal[i] = cnt[i] > 0 ? pool[off[i] + cnt[i] - 1] : make_float2(0.0f, 0.0f);
    }
    s->repack_soon = true;                      // the file's pool is in slot order
    return 0;
}
