// sim_kernels.hip -- gfx950 kernels for one Simulator::Update step
// (Simulator.cpp:314-323).
//
// Step pipeline (all on the handle's stream, no host round trip):
//   1. k_grid_count / scan / k_grid_scatter: counting sort of the active
//      agents by Morton cell (replaces KDTree::Construct, KDTree.cpp:22-83).
//   2. k_attr: goal checks + IRM attraction point on the ECM
//      (Simulator.cpp:538-590, IRMPathFollower.cpp:14-115, ECM.cpp:20-96,
//      ECMCellCollection.cpp:57-90) and steering (Simulator.cpp:638-657); its
//      last block publishes the step's replan and destroy events.
//   3. k_knn<K>: the exact k-nearest query (replaces KDTree::KNearestAgents,
//      KDTree.cpp:85-202).
//   4. k_orca_reg<K> (plain step; or fused into k_knn_orca) / k_orca<K>
//      (learned and logging steps): obstacle query (Simulator.cpp:259-292),
//      ORCA half-planes (ORCA.cpp:60-424), 2D LP (ORCA.cpp:428-587), velocity
//      and position integration (Simulator.cpp:592-635, 676-677) fused.
//      Compacted queues take the rest: k_orca_ovf (more obstacle constraints
//      than the main pass holds), k_lp3d_reg and k_lp3d (2D LP infeasible,
//      RandomizedLP3D, ORCA.cpp:592-669).
// One agent per lane; lanes are ordered along the Morton curve so a
// wavefront covers a compact patch of the world.
#include "ecm_geom.h"
#include "lp_reg.h"
#include "mlp.h"
#include "sim_device.h"

#include <limits.h>

#include <algorithm>

namespace ecm {

static constexpr uint32_t kInvalidKey = 0xffffffffu;

__device__ __forceinline__ uint32_t part1by1(uint32_t x) {
    x &= 0x0000ffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
__device__ __forceinline__ uint32_t morton2(uint32_t cx, uint32_t cy) {
    return part1by1(cx) | (part1by1(cy) << 1);
}
__device__ __forceinline__ int slot_env(const DevState& s, int id) {
    return s.agent_env ? (int)s.agent_env[id] : 0;
}
// Velocity of a sorted agent (the first half of its sorted_v entry).
__device__ __forceinline__ float2 sorted_vel(const DevState& s, int t) {
    return *reinterpret_cast<const float2*>(&s.sorted_v[t]);
}
// Environment of a sorted agent from its sorted_v entry (w = environment bits).
__device__ __forceinline__ const EnvDesc& env_of_sorted(const DevState& s, float4 v) {
    return s.envs[__float_as_int(v.w)];
}
__device__ __forceinline__ uint32_t cell_key(const EnvDesc& E, float2 p) {
    const int cx = grid_cell(p.x, E.ag.x0, E.ag.inv_cell, E.ag.side);
    const int cy = grid_cell(p.y, E.ag.y0, E.ag.inv_cell, E.ag.side);
    return E.cell_base + morton2((uint32_t)cx, (uint32_t)cy);
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i -> XCD i % 8).
// Remap so that XCD x processes the x-th contiguous eighth of the Morton
// order: neighbouring agents then share one L2.  gridDim.x is a multiple of 8.
__device__ __forceinline__ int xcd_block() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// ------------------------------------------------------------ grid build
// Start of a step (block 0 of k_grid_count, 256 threads): add the
// previous step's counters (8 words) and the column sums of the
// per-wavefront slots ([64][4] words) to the 64-bit totals (acc[0..7] and
// acc[8..11]), count the step (acc[12]), then clear both for this step.
// No other block of the grid touches them.
__device__ __forceinline__ void step_begin(uint32_t* counters, uint32_t* wstats, unsigned long long* acc) {
    __shared__ uint32_t col[4];
    const int t = threadIdx.x;
    if (t < 4) col[t] = 0u;
    __syncthreads();
    const uint32_t w = wstats[t];
    if (w) atomicAdd(&col[t & 3], w);
    const uint32_t c = t < 8 ? counters[t] : 0u;
    __syncthreads();
    if (t < 8) { acc[t] += c; counters[t] = 0u; }
    if (t < 4) acc[8 + t] += col[t];
    if (t == 0) acc[12] += 1ull;
    wstats[t] = 0u;
}

// Histogram pass, entered from the slot-indexed arrays: one element per
// slot id; the atomic's return value is the agent's arrival order in its cell.
__global__ __launch_bounds__(256) void k_grid_count(DevState s) {
    const uint32_t n_old = (uint32_t)s.acc[13];
    if (blockIdx.x == 0) {
        step_begin(s.events_n, s.wstats, s.acc);
        if (threadIdx.x == 0) s.acc[14] = n_old;
    }
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n_slots) return;
    if (i < (int)n_old) s.fwd[i] = -1;        // set again by the scatter for agents found there
    if (!s.active[i]) { s.key[i] = kInvalidKey; return; }
    const uint32_t m = cell_key(s.envs[slot_env(s, i)], s.pos_in[i]);
    s.key[i] = m;
    s.rank[i] = atomicAdd(&s.cell_count[m], 1u);
}

// Histogram pass, entered from the last step's sorted outputs: one element
// per previous sorted position (in the previous Morton order, so the cells
// of a wavefront are few and close); agents destroyed during the last step
// leave the grid.
// The last step's active count is acc[13] (its scan); it is kept in
// acc[14] for the scatter, since this step's scan overwrites acc[13].
// Agents spawned since the last step (acc[15] of them, ids in new_ids) follow
// as elements n_old, n_old + 1, ... read from the slot arrays; the host
// keeps n_old + their count within n_slots.
__global__ __launch_bounds__(256) void k_grid_count_sorted(DevState s) {
    const uint32_t n_old = (uint32_t)s.acc[13];
    const uint32_t n_new = (uint32_t)s.acc[15];
    if (blockIdx.x == 0) {
        step_begin(s.events_n, s.wstats, s.acc);
        if (threadIdx.x == 0) s.acc[14] = n_old;
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t m = kInvalidKey;
    if (t < (int)n_old) {
        if (s.dead_s_old[t]) { s.key[t] = kInvalidKey; s.fwd[t] = -1; }
        else {
            // one environment: no need to read the agent's environment
            const EnvDesc& E = s.agent_env ? env_of_sorted(s, s.sorted_v_old[t]) : s.envs[0];
            m = cell_key(E, s.pos_s_old[t]);
        }
    } else if (t < (int)(n_old + n_new)) {
        const int id = s.new_ids[t - n_old];
        if (!s.active[id]) s.key[t] = kInvalidKey;
        else m = cell_key(s.envs[slot_env(s, id)], s.pos_in[id]);
    }
    // Lanes in Morton order mostly share cells with their neighbours: each
    // run of equal keys in consecutive lanes takes one atomic, and its lanes
    // take consecutive arrival orders from the returned base.  Every lane of
    // the wavefront takes part (no lane has returned).
    const int lane = (int)__lane_id();
    const uint32_t mp = __shfl_up(m, 1u);
    const bool start = lane == 0 || m != mp;
    const unsigned long long starts = __ballot(start);
    const unsigned long long upto = (2ull << lane) - 1ull;          // lanes 0..lane
    const int leader = 63 - __clzll((long long)(starts & upto));
    const unsigned long long later = starts & ~upto;
    const int next = later ? __ffsll((long long)later) - 1 : 64;
    uint32_t base = 0;
    if (start && m != kInvalidKey) base = atomicAdd(&s.cell_count[m], (uint32_t)(next - lane));
    base = __shfl(base, leader);
    if (m != kInvalidKey) {
        s.key[t] = m;
        s.rank[t] = base + (uint32_t)(lane - leader);
    }
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}
// exclusive scan over a 256-thread block
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t& total) {
    __shared__ uint32_t wsum[4];
    uint32_t inc = wave_incl_scan(v);
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int i = 0; i < w; ++i) off += wsum[i];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + inc - v;
}
// 1024 counts per block -> local exclusive offsets + block sums; the
// counts are cleared for the next step
__global__ __launch_bounds__(256) void k_scan_local(uint32_t* cnt, uint32_t* start, uint32_t* bsum, uint32_t n) {
    uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) a[q] = (base + q < n) ? cnt[base + q] : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < n) cnt[base + q] = 0u;
    uint32_t loc = a[0] + a[1] + a[2] + a[3];
    uint32_t tot;
    uint32_t ex = block_excl_scan256(loc, tot);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (base + q < n) start[base + q] = ex;
        ex += a[q];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
// single block: exclusive scan of the block sums, grand total to start[n]
__global__ __launch_bounds__(256) void k_scan_top(uint32_t* bsum, uint32_t nb, uint32_t* start, uint32_t n,
                                                  unsigned long long* n_active) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nb; c += 1024u) {
        uint32_t base = c + threadIdx.x * 4u;
        uint32_t a[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = (base + q < nb) ? bsum[base + q] : 0u;
        uint32_t loc = a[0] + a[1] + a[2] + a[3];
        uint32_t tot;
        uint32_t ex = block_excl_scan256(loc, tot) + carry;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (base + q < nb) bsum[base + q] = ex;
            ex += a[q];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) { start[n] = carry; *n_active = carry; }
}
__global__ __launch_bounds__(256) void k_scan_add(uint32_t* start, const uint32_t* bsum, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) start[i] += bsum[i >> 10];
}

// Counting-sort scatter from the slot-indexed arrays: the sorted position
// is the cell's start plus the arrival order the count pass recorded.  An
// agent that was at sorted position u of the last step (pos_of, written by
// k_unsort before the host changed anything; checked against the id stored
// there) keeps its neighbour seeds: prev_of and fwd link u and its new
// position.  Any k distinct agents bound the k-th distance, so seeds stay
// safe after spawns, removals, writes and loads.
__global__ void k_grid_scatter(DevState s) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n_slots) return;
    uint32_t m = s.key[i];
    if (m == kInvalidKey) return;
    uint32_t slot = s.cell_start[m] + s.rank[i];
    const float2 p = s.pos_in[i], v = s.vel_in[i];
    s.sorted_pr[slot] = make_float4(p.x, p.y, s.radius[i], __int_as_float(i));
    s.sorted_v[slot] = make_float4(v.x, v.y, s.pref_speed[i], __int_as_float(slot_env(s, i)));
    s.attr_s[slot] = s.attraction[i];
    s.path_s[slot] = make_int2(s.path_off[i], s.path_cnt[i]);
    const int u = s.pos_of[i];
    const bool was = u >= 0 && u < (int)(uint32_t)s.acc[14] && __float_as_int(s.sorted_pr_old[u].w) == i;
    s.prev_of[slot] = was ? u : -1;
    if (was) s.fwd[u] = (int)slot;
}

// Counting-sort scatter from the last step's sorted outputs: reads and
// writes both follow the Morton order, which barely changes from step to
// step.  Also records the maps between the two orders for the neighbour
// seeds.
__global__ void k_grid_scatter_sorted(DevState s) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_old = (int)(uint32_t)s.acc[14];
    if (t >= n_old + (int)(uint32_t)s.acc[15]) return;
    const uint32_t m = s.key[t];
    if (m == kInvalidKey) return;
    const uint32_t slot = s.cell_start[m] + s.rank[t];
    if (t >= n_old) {                                  // spawned since the last step: no seeds
        const int id = s.new_ids[t - n_old];
        const float2 p = s.pos_in[id], v = s.vel_in[id];
        s.sorted_pr[slot] = make_float4(p.x, p.y, s.radius[id], __int_as_float(id));
        s.sorted_v[slot] = make_float4(v.x, v.y, s.pref_speed[id], __int_as_float(slot_env(s, id)));
        s.attr_s[slot] = s.attraction[id];
        s.path_s[slot] = make_int2(s.path_off[id], s.path_cnt[id]);
        s.prev_of[slot] = -1;
        return;
    }
    const float2 p = s.pos_s_old[t], v = s.vel_s_old[t];
    const float4 pr = s.sorted_pr_old[t], sv = s.sorted_v_old[t];
    s.sorted_pr[slot] = make_float4(p.x, p.y, pr.z, pr.w);
    s.sorted_v[slot] = make_float4(v.x, v.y, sv.z, sv.w);
    s.attr_s[slot] = s.attr_s_old[t];
    s.path_s[slot] = s.path_s_old[t];
    s.fwd[t] = (int)slot;
    s.prev_of[slot] = t;
}

// Agents the host removed since the last step leave a sorted-entry build:
// their last sorted position (pos_of, checked against the id there) is
// marked destroyed.
__global__ void k_mark_dead(const int* ids, int n, const int* pos_of, const float4* sorted_pr_old, uint8_t* dead_old,
                            const unsigned long long* acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int id = ids[i];
    const int u = pos_of[id];
    if (u >= 0 && u < (int)(uint32_t)acc[13] && __float_as_int(sorted_pr_old[u].w) == id) dead_old[u] = 1;
}

// Brings the slot-indexed arrays up to date from the last step's sorted
// outputs (the host asks before it reads or changes agents).  Agents
// destroyed during a step had their slots written then (k_attr).
__global__ void k_unsort(DevState s) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int)(uint32_t)s.acc[13]) return;         // the last step's active count
    if (s.dead_s_old[t]) return;                       // written when it was destroyed
    const int id = __float_as_int(s.sorted_pr_old[t].w);
    s.pos_of[id] = t;
    s.pos_in[id] = s.pos_s_old[t];
    s.vel_in[id] = s.vel_s_old[t];
    s.attraction[id] = s.attr_s_old[t];
    s.pref_vel[id] = s.pref_vel_s_old[t];
    if (s.v_orca) s.v_orca[id] = s.v_orca_s_old[t];
}

// ------------------------------------------------------------- ECM query
__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}

// First cell, in collection order, whose ring contains p.  Equivalent to the
// linear scan of ECMCellCollection::PointLocationQueryLinear: a cell can only
// contain p if p lies in its (margin-expanded) box, except when p.y equals
// the y of one of its vertices exactly; those cells are listed per y value.
__device__ int locate_cell(const EnvDesc& s, V2 p, uint32_t& fallback) {
    if (s.n_cells == 0) return -1;
    int bx = grid_cell(p.x, s.eg.x0, s.eg.inv_cell, s.eg.side);
    int by = grid_cell(p.y, s.eg.y0, s.eg.inv_cell, s.eg.side);
    int b = by * s.eg.side + bx;
    int i = s.ebucket_start[b], ie = s.ebucket_start[b + 1];
    int j = 0, je = 0;
    if (s.yhash_mask >= 0) {
        uint32_t key = (p.y == 0.0f) ? 0u : __float_as_uint(p.y);
        uint32_t h = hash_u32(key) & (uint32_t)s.yhash_mask;
        while (true) {
            uint32_t kk = s.yhash_key[h];
            if (kk == 0xffffffffu) break;
            if (kk == key) { j = s.yhash_off[h]; je = j + s.yhash_cnt[h]; fallback = 1; break; }
            h = (h + 1u) & (uint32_t)s.yhash_mask;
        }
    }
    while (i < ie || j < je) {
        int ci = i < ie ? s.ebucket_cells[i] : INT_MAX;
        int cj = j < je ? s.ycells[j] : INT_MAX;
        int c;
        if (ci < cj) { c = ci; ++i; }
        else if (cj < ci) { c = cj; ++j; }
        else { c = ci; ++i; ++j; }
        float4 A = s.cellA[c], B = s.cellB[c];
        if (quad_contains(p, mk(A.x, A.y), mk(A.z, A.w), mk(B.x, B.y), mk(B.z, B.w))) return c;
    }
    return -1;
}

// ECM::RetractPoint (ECM.cpp:20-96) given the located cell.  Cell 2e holds
// (p1, he0.cl, he1.cr, p2), cell 2e+1 holds (p1, he0.cr, he1.cl, p2).
__device__ bool retract(const EnvDesc& s, V2 loc, int cell, V2& out, V2& oa, V2& ob) {
    int e2 = cell & ~1;
    float4 A0 = s.cellA[e2], B0 = s.cellB[e2];
    float4 A1 = s.cellA[e2 + 1], B1 = s.cellB[e2 + 1];
    V2 p1 = mk(A0.x, A0.y), p2 = mk(B0.z, B0.w);
    oa = mk(A0.z, A0.w);            // he0.closest_left
    ob = mk(B0.x, B0.y);            // he1.closest_right
    V2 o1, o2, ray;
    if (left_of_segment(p1, p2, loc)) {
        o1 = oa; o2 = ob;
        if (approx(o1, o2)) ray = (p1 - o1) + (p2 - o1);
        else { V2 v = o2 - o1; ray = mk(v.y, -v.x); }
    } else {
        o1 = mk(A1.z, A1.w);        // he0.closest_right
        o2 = mk(B1.x, B1.y);        // he1.closest_left
        if (approx(o1, o2)) ray = (p1 - o1) + (p2 - o1);
        else { V2 v = o2 - o1; ray = mk(-v.y, v.x); }
    }
    ray = normalize(ray);
    return ray_segment(loc, ray, p1, p2, out);
}

// IRMPathFollower::FindAttractionPoint (IRMPathFollower.cpp:14-115)
__device__ bool irm_attraction(const EnvDesc& s, V2 pos, const float2* path, int np, V2& out, uint32_t& fb) {
    int cell = locate_cell(s, pos, fb);
    if (cell < 0) return false;
    V2 rl = mk(0.0f, 0.0f), oa, ob;
    if (!retract(s, pos, cell, rl, oa, ob)) return false;
    V2 co = closest_on_segment(rl, oa, ob);
    float clearance = len(rl - co);
    float clr2 = clearance * clearance;
    float2 g = path[np - 1];
    V2 goal = mk(g.x, g.y);
    if (len2(goal - rl) < clr2) { out = goal; return true; }
    // The forward loop of the reference keeps the point of the LAST segment
    // whose circle intersection lies on the segment, and succeeds when any
    // segment's supporting line meets the circle (disc >= eps, or NaN).
    // Windows of 32 segments are taken from the end: a cheap pass marks the
    // segments whose line meets the circle, then the marked ones are solved
    // from the highest down until one yields a point.  Each segment's
    // arithmetic is the forward loop's, so the point is the same, and a
    // lane rarely solves more than one or two segments.
    bool success = false;
    for (int hi = np - 1; hi > 0; hi -= 32) {
        const int lo = max(hi - 32, 0);
        uint32_t mask = 0;
        float2 a = path[lo];
#pragma unroll 4
        for (int i = lo; i < hi; ++i) {
            const float2 b = path[i + 1];
            const V2 q1 = mk(a.x, a.y) - rl;
            const V2 q2 = mk(b.x, b.y) - rl;
            a = b;
            const V2 ed = q2 - q1;
            const float el2 = len2(ed);
            const float dm = det(q1, q2);
            const float disc = clr2 * el2 - dm * dm;
            if (!(disc < kEps)) mask |= 1u << (i - lo);
        }
        success |= mask != 0u;
        while (mask) {
            const int bit = 31 - __clz(mask);
            mask &= ~(1u << bit);
            const int i = lo + bit;
            const float2 pa = path[i], pb = path[i + 1];
            const V2 q1 = mk(pa.x, pa.y) - rl;
            const V2 q2 = mk(pb.x, pb.y) - rl;
            const V2 ed = q2 - q1;
            const float el2 = len2(ed);
            const float dm = det(q1, q2);
            const float disc = clr2 * el2 - dm * dm;
            const float sgn = ed.y < 0.0f ? -1.0f : 1.0f;
            const float sq = sqrtf(disc);
            const V2 i1 = mk((dm * ed.y + sgn * ed.x * sq) / el2, (-dm * ed.x + fabsf(ed.y) * sq) / el2);
            const V2 i2 = mk((dm * ed.y - sgn * ed.x * sq) / el2, (-dm * ed.x - fabsf(ed.y) * sq) / el2);
            const V2 g1 = q1 + rl, g2 = q2 + rl;
            const V2 gi1 = i1 + rl, gi2 = i2 + rl;
            const V2 edg = g2 - g1;
            const float t1 = dot(gi1 - g1, edg) / el2;
            const float t2 = dot(gi2 - g1, edg) / el2;
            const bool v1 = t1 >= 0.0f && t1 <= 1.0f;
            const bool v2 = t2 >= 0.0f && t2 <= 1.0f && t2 > (v1 ? t1 : -1.0f);
            if (v1 || v2) {
                out = v2 ? gi2 : gi1;
                return true;
            }
        }
    }
    return success;
}

// ---------------------------------------------------------- exact k-NN
// (d2, id) lexicographic order; d2 as KDTree.cpp:106-108.
__device__ __forceinline__ bool nb_better(float d, int id, float d2, int id2) {
    return (d < d2) | ((d == d2) & (id < id2));      // bitwise: no branches
}

// Agent id at a sorted position (INT_MAX for an empty list entry, -1).
__device__ __forceinline__ int id_at(const DevState& s, int slot) {
    const int v = __float_as_int(s.sorted_pr[slot < 0 ? 0 : slot].w);
    return slot < 0 ? INT_MAX : v;
}

// Insert candidate (dd, id ii, sorted position ss) into the ascending list
// (d, sl) of length K, ordered by (d2, id): one comparison per entry (the
// flags are monotone along the list), then a predicated shift; no dynamic
// indexing.  The list keeps sorted positions, not ids: ids are loaded only
// when dd ties an entry's distance.
template <int K>
__device__ __forceinline__ void nb_insert(const DevState& s, float (&d)[K], int (&sl)[K], float dd, int ii, int ss) {
    bool tie = false;
#pragma unroll
    for (int i = 0; i < K; ++i) tie |= dd == d[i];
    bool b[K];
#pragma unroll
    for (int i = 0; i < K; ++i) b[i] = dd < d[i];
    if (tie) {
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (dd == d[i]) b[i] = ii < id_at(s, sl[i]);
    }
#pragma unroll
    for (int i = K - 1; i > 0; --i) {
        d[i] = b[i - 1] ? d[i - 1] : (b[i] ? dd : d[i]);
        sl[i] = b[i - 1] ? sl[i - 1] : (b[i] ? ss : sl[i]);
    }
    d[0] = b[0] ? dd : d[0];
    sl[0] = b[0] ? ss : sl[0];
}

// Upper bound on the squared k-th neighbour distance of the agent at sorted
// position t (id, at p): the largest squared distance to the k agents it
// had as neighbours in the previous step, taken to this step's sorted
// positions through fwd, when all of them are still on the grid of its
// environment and none is the agent itself (any k distinct candidates bound
// the k-th distance from above; fwd is one-to-one).  INFINITY when there
// is no such seed (spawned since the last step, or a neighbour left).  The
// seeds (this step's sorted positions) and their squared distances are left
// in ss / sd, unordered.
template <int K>   // K == k
__device__ __forceinline__ float knn_seed_bound(const DevState& s, int t, V2 p, float (&sd)[K], int (&ss)[K]) {
    const int q = s.prev_of[t];
    const int env = s.agent_env ? __float_as_int(s.sorted_v[t].w) : 0;
    if (q < 0 || q >= s.n_slots || s.nbr_cnt_prev[q] < K) return INFINITY;
    int* ns = ss;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int os = s.nbr_prev[(size_t)j * s.n_slots + q];
        const bool in = os >= 0 && os < s.n_slots;
        const int f = s.fwd[in ? os : 0];          // out-of-range entries read slot 0
        ns[j] = in ? f : -1;
    }
#pragma unroll
    for (int j = 0; j < K; ++j) ok &= ns[j] >= 0 && ns[j] < s.n_slots;
    if (!ok) return INFINITY;
    float bound = 0.0f;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const float4 c = s.sorted_pr[ns[j]];
        const float dx = c.x - p.x, dy = c.y - p.y;
        const float d2 = dx * dx + dy * dy;
        ok &= d2 > kEps;
        if (s.agent_env) ok &= __float_as_int(s.sorted_v[ns[j]].w) == env;
        bound = fmaxf(bound, d2);
        sd[j] = d2;
    }
    return ok ? bound : INFINITY;
}

// Merge the nb buffered candidates (sorted positions, entry e at
// buf[e * bstride]) into the ordered list (d, sl) holding n entries, and
// tighten the keep threshold thr (and the squared cell-pruning radius
// prune2) once the list holds k.
template <int K>
__device__ __forceinline__ void knn_merge(const DevState& s, V2 p, const int* buf, int bstride, int& nb,
                                          float (&d)[K], int (&sl)[K], int& n, float& thr, float& prune2) {
    for (int e = 0; e < nb; ++e) {
        const int j = buf[e * bstride];
        const float4 c = s.sorted_pr[j];
        const float dx = c.x - p.x, dy = c.y - p.y;
        const float d2 = dx * dx + dy * dy;
        const int cid = __float_as_int(c.w);
        bool better = d2 < d[K - 1];
        if (!better && d2 == d[K - 1]) better = cid < id_at(s, sl[K - 1]);
        if (better) {
            nb_insert<K>(s, d, sl, d2, cid, j);
            n = min(n + 1, K);
        }
    }
    nb = 0;
    if (n >= K) {
        thr = fminf(thr, d[K - 1]);
        const float rr = __fsqrt_rn(thr) * 1.0001f + 2e-3f;
        prune2 = rr * rr;
    }
}

// Squared distance from v to the band [lo, hi) of grid row / column i
// along one axis; the border bands are open outward (agents outside the
// grid box are clamped into them).
__device__ __forceinline__ float band_d2(float v, float v0, float cell, int i, int side) {
    const float b0 = i == 0 ? -INFINITY : v0 + (float)i * cell;
    const float b1 = i == side - 1 ? INFINITY : v0 + (float)(i + 1) * cell;
    const float dd = fmaxf(fmaxf(b0 - v, v - b1), 0.0f);
    return dd * dd;
}

// Exact k nearest by ring search over the agent grid.  Rings are walked as
// one loop over their 8r perimeter cells; a cell is skipped when its box
// (open toward the outside on the grid border, where clamped agents live)
// is farther than the current bound on the k-th distance, and the search
// stops when the whole unvisited region is.  A 2 mm guard covers rounding.
// Candidates within the bound are appended to a per-lane LDS buffer
// (`buf`, entry e at buf[e * bstride]) and merged into the ordered register
// list at the end of each ring or when the buffer is full, so the lanes of
// a wave rarely wait on each other's list inserts.  The list is ordered by
// (d2, id), so the result does not depend on visiting order; `bound` (from
// knn_seed_bound) only prunes.  The list holds sorted positions.
template <int K, int B>   // K == k
__device__ __forceinline__ int knn_query(const DevState& s, const EnvDesc& E, V2 p, float bound,
                                         int* buf, int bstride, int (&sl)[K]) {
    const GridDesc& g = E.ag;
    const int cx = grid_cell(p.x, g.x0, g.inv_cell, g.side);
    const int cy = grid_cell(p.y, g.y0, g.inv_cell, g.side);
    float d[K];
#pragma unroll
    for (int i = 0; i < K; ++i) { d[i] = INFINITY; sl[i] = -1; }
    int n = 0, nb = 0;
    const float guard = 2e-3f;
    // thr: candidates with d2 <= thr are kept; it is the bound until the
    // list holds k entries, then the smaller of the bound and the k-th
    float thr = bound;
    // a cell box at squared distance dd2 is skipped when
    // sqrt(dd2) - guard > sqrt(thr)  <=>  dd2 > (sqrt(thr) + guard)^2
    float prune2 = INFINITY;
    if (thr < INFINITY) {
        const float rr = __fsqrt_rn(thr) * 1.0001f + guard;
        prune2 = rr * rr;
    }
    auto merge = [&]() { knn_merge<K>(s, p, buf, bstride, nb, d, sl, n, thr, prune2); };
    for (int r = 0;; ++r) {
        const int perim = r == 0 ? 1 : 8 * r;
        for (int q = 0; q < perim; ++q) {
            int x, y;
            if (r == 0) { x = cx; y = cy; }
            else if (q < 2 * r + 1) { x = cx - r + q; y = cy + r; }
            else if (q < 4 * r + 2) { x = cx - r + (q - (2 * r + 1)); y = cy - r; }
            else if (q < 6 * r + 1) { x = cx - r; y = cy - r + 1 + (q - (4 * r + 2)); }
            else { x = cx + r; y = cy - r + 1 + (q - (6 * r + 1)); }
            if (x < 0 || y < 0 || x >= g.side || y >= g.side) continue;
            if (prune2 < INFINITY) {
                // distance from p to the cell box, border cells open outward
                float bx0 = x == 0 ? -INFINITY : g.x0 + (float)x * g.cell;
                float bx1 = x == g.side - 1 ? INFINITY : g.x0 + (float)(x + 1) * g.cell;
                float by0 = y == 0 ? -INFINITY : g.y0 + (float)y * g.cell;
                float by1 = y == g.side - 1 ? INFINITY : g.y0 + (float)(y + 1) * g.cell;
                float ddx = fmaxf(fmaxf(bx0 - p.x, p.x - bx1), 0.0f);
                float ddy = fmaxf(fmaxf(by0 - p.y, p.y - by1), 0.0f);
                if (ddx * ddx + ddy * ddy > prune2) continue;
            }
            const uint32_t m = E.cell_base + morton2((uint32_t)x, (uint32_t)y);
            const uint32_t a = s.cell_start[m], b = s.cell_start[m + 1];
            for (uint32_t j = a; j < b; ++j) {
                const float2 c = *reinterpret_cast<const float2*>(&s.sorted_pr[j]);
                const float dx = c.x - p.x, dy = c.y - p.y;
                const float d2 = dx * dx + dy * dy;
                if (d2 > kEps && d2 <= thr) {
                    buf[nb * bstride] = (int)j;
                    if (++nb == B) merge();
                }
            }
        }
        if (nb > 0) merge();
        const int xl = cx - r, xh = cx + r, yl = cy - r, yh = cy + r;
        if (xl <= 0 && yl <= 0 && xh >= g.side - 1 && yh >= g.side - 1) break;
        if (thr < INFINITY) {
            float b = INFINITY;
            if (xl > 0) b = fminf(b, p.x - (g.x0 + (float)xl * g.cell));
            if (yl > 0) b = fminf(b, p.y - (g.y0 + (float)yl * g.cell));
            if (xh < g.side - 1) b = fminf(b, (g.x0 + (float)(xh + 1) * g.cell) - p.x);
            if (yh < g.side - 1) b = fminf(b, (g.y0 + (float)(yh + 1) * g.cell) - p.y);
            b -= guard;
            if (b > 0.0f && b * b > thr) break;
        }
    }
    return n;
}

// Orders (d, sl) by (d2, id) in place: odd-even transposition sort, K
// rounds of compare-exchanges; ids are loaded only when two distances tie.
template <int K>
__device__ __forceinline__ void knn_sort(const DevState& s, float (&d)[K], int (&sl)[K]) {
#pragma unroll
    for (int r = 0; r < K; ++r) {
#pragma unroll
        for (int i = r & 1; i + 1 < K; i += 2) {
            bool sw = d[i + 1] < d[i];
            if (d[i + 1] == d[i]) sw = id_at(s, sl[i + 1]) < id_at(s, sl[i]);
            const float da = d[i], db = d[i + 1];
            const int sa = sl[i], sb = sl[i + 1];
            d[i] = sw ? db : da; d[i + 1] = sw ? da : db;
            sl[i] = sw ? sb : sa; sl[i + 1] = sw ? sa : sb;
        }
    }
}

// Exact k nearest for an agent with a finite seed bound.  Every candidate
// within the bound lies in the rectangle of cells [grid_cell(p - rr),
// grid_cell(p + rr)] on both axes (rr = the bound's radius with the 2 mm
// guard; grid_cell is monotone), so the cells are walked row by row in one
// pass, without rings.  The walk appends the non-empty, unpruned cells as
// ranges of sorted positions (consecutive ranges joined) to a per-lane LDS
// list of R entries (`rbuf`, range r at rbuf[2r * bstride], rbuf[(2r + 1) *
// bstride]); when it is full, or the walk is over, one flat loop visits all
// the listed candidates, so the lanes of a wave step through candidates
// together instead of waiting for each other at every cell.  Candidates
// within the threshold go to the candidate buffer as in knn_query, merged
// when it fills and at the end.  The list starts as the k seeds (d, sl from
// knn_seed_bound), sorted: they are candidates within the bound, so a
// candidate the walk meets that is in the list is skipped, and only the few
// new ones are inserted.  The result is again the k smallest (d2, id) among
// the candidates with eps < d2 <= bound.
template <int K, int B, int R>   // K == k
__device__ __forceinline__ int knn_query_seeded(const DevState& s, const EnvDesc& E, V2 p, float bound,
                                                int* buf, int* rbuf, int bstride, float (&d)[K], int (&sl)[K]) {
    const GridDesc& g = E.ag;
    knn_sort<K>(s, d, sl);
    int n = K, nb = 0;
    float thr = bound;
    const float rr = __fsqrt_rn(thr) * 1.0001f + 2e-3f;
    float prune2 = rr * rr;
    const int xl = grid_cell(p.x - rr, g.x0, g.inv_cell, g.side);
    const int xh = grid_cell(p.x + rr, g.x0, g.inv_cell, g.side);
    const int yl = grid_cell(p.y - rr, g.y0, g.inv_cell, g.side);
    const int yh = grid_cell(p.y + rr, g.y0, g.inv_cell, g.side);
    const uint32_t mx0 = part1by1((uint32_t)xl);
    int x = xl, y = yl;
    uint32_t mx = mx0, my = part1by1((uint32_t)yl) << 1;
    float dy2 = band_d2(p.y, g.y0, g.cell, y, g.side);
    bool walking = true;
    uint32_t ca = 0, cb = 0;             // open range [ca, cb), empty when ca == cb
    while (true) {
        int nr = 0;
        auto add = [&](uint32_t a, uint32_t b) {
            if (a < b) {
                if (a == cb) cb = b;
                else {
                    if (ca < cb) {
                        rbuf[(2 * nr) * bstride] = (int)ca;
                        rbuf[(2 * nr + 1) * bstride] = (int)cb;
                        ++nr;
                    }
                    ca = a; cb = b;
                }
            }
        };
        while (walking && nr < R) {
            // this cell and, when the row has one, the next: the range loads
            // of both go out together
            const bool two = x < xh;
            const uint32_t mxn = ((mx | 0xaaaaaaaau) + 1u) & 0x55555555u;
            const uint32_t m0 = E.cell_base + (mx | my);
            const uint32_t m1 = two ? E.cell_base + (mxn | my) : m0;
            const bool in0 = band_d2(p.x, g.x0, g.cell, x, g.side) + dy2 <= prune2;
            const bool in1 = two && band_d2(p.x, g.x0, g.cell, x + 1, g.side) + dy2 <= prune2;
            const uint32_t a0 = s.cell_start[m0], b0 = s.cell_start[m0 + 1];
            const uint32_t a1 = s.cell_start[m1], b1 = s.cell_start[m1 + 1];
            if (in0) add(a0, b0);
            const bool took_two = two && nr < R;
            if (took_two && in1) add(a1, b1);
            const int nx = x + (took_two ? 2 : 1);
            if (nx <= xh) { x = nx; mx = took_two ? ((mxn | 0xaaaaaaaau) + 1u) & 0x55555555u : mxn; }
            else if (y < yh) {
                ++y; x = xl; mx = mx0; my = part1by1((uint32_t)y) << 1;
                dy2 = band_d2(p.y, g.y0, g.cell, y, g.side);
            } else walking = false;
        }
        if (!walking && nr < R && ca < cb) {
            rbuf[(2 * nr) * bstride] = (int)ca;
            rbuf[(2 * nr + 1) * bstride] = (int)cb;
            ++nr;
            ca = cb;
        }
        if (nr > 0) {
            int r = 0;
            uint32_t j = (uint32_t)rbuf[0], e = (uint32_t)rbuf[bstride];
            auto take = [&](uint32_t jj, float2 c) {
                const float dx = c.x - p.x, dy = c.y - p.y;
                const float d2 = dx * dx + dy * dy;
                if (d2 > kEps && d2 <= thr) {
                    bool listed = false;
#pragma unroll
                    for (int i = 0; i < K; ++i) listed |= sl[i] == (int)jj;
                    if (!listed) {
                        buf[nb * bstride] = (int)jj;
                        if (++nb == B) knn_merge<K>(s, p, buf, bstride, nb, d, sl, n, thr, prune2);
                    }
                }
            };
            while (true) {
                // two candidates of the range per trip, their loads together
                const bool two = j + 1u < e;
                const float2 c0 = *reinterpret_cast<const float2*>(&s.sorted_pr[j]);
                const float2 c1 = *reinterpret_cast<const float2*>(&s.sorted_pr[two ? j + 1u : j]);
                take(j, c0);
                if (two) take(j + 1u, c1);
                j += two ? 2u : 1u;
                if (j == e) {
                    if (++r == nr) break;
                    j = (uint32_t)rbuf[(2 * r) * bstride];
                    e = (uint32_t)rbuf[(2 * r + 1) * bstride];
                }
            }
        }
        if (!walking && ca == cb) break;
    }
    if (nb > 0) knn_merge<K>(s, p, buf, bstride, nb, d, sl, n, thr, prune2);
    return n;
}

// --------------------------------------- attraction + steering + neighbours
// LDS of the neighbour search, per lane: a candidate buffer of kKnnBuf
// sorted positions, then kKnnRanges ranges of sorted positions (two ints
// each) for the seeded walk; (kKnnBuf + 2 kKnnRanges) * blockDim.x ints per
// block, lane-interleaved.
static constexpr int kKnnBuf = 16;
// Seeded agents walk their bound's cell rectangle when its radius is at most
// this many cells.
static constexpr float kKnnWalkCells = 6.0f;
static constexpr int kKnnRanges = 4;

// One sorted agent's goal check, attraction point and steering (with the
// destroy or replan request they may make).  Returns whether it appended an
// event, after fencing it device-wide.
__device__ __forceinline__ bool attr_agent(const DevState& s) {
    const uint32_t n_active = s.cell_start[s.n_keys];
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= (int)n_active) return false;
    const float4 pr = s.sorted_pr[t];
    const float4 sv = s.sorted_v[t];
    const int2 pc = s.path_s[t];
    const int id = __float_as_int(pr.w);
    const V2 pos = mk(pr.x, pr.y);
    const EnvDesc& E = env_of_sorted(s, sv);

    // UpdateAttractionPointSystem (Simulator.cpp:538-590)
    const int np = pc.y;
    const float2* path = s.path_pts + pc.x;
    float2 gg = path[np - 1];
    float gx = pos.x - gg.x, gy = pos.y - gg.y;
    float dGoal = gx * gx + gy * gy;
    float2 a0 = s.attr_s[t];
    V2 attr = mk(a0.x, a0.y);
    bool alive = true, event = false;
    uint32_t fb = 0;
    if (dGoal < s.arrival_r2) {
        attr = mk(gg.x, gg.y);
        if (dGoal < s.delete_r2) alive = false;
    } else {
        V2 ap = mk(0.0f, 0.0f);
        if (irm_attraction(E, pos, path, np, ap, fb)) attr = ap;
        else {
            // UpdatePath(e, currentPosition, goal) with the step-start position
            uint32_t q = atomicAdd(s.events_n, 1u);
            s.events[q] = make_int4(id, __float_as_int(pos.x), __float_as_int(pos.y), t);
            event = true;
        }
    }
    s.attr_s[t] = make_float2(attr.x, attr.y);
    if (fb) atomicAdd(&s.stats[2], 1u);
    s.dead_s[t] = alive ? 0 : 1;
    if (!alive) {
        // the slot keeps the state it is destroyed with: step-start position
        // and velocity, this step's attraction point, the previous step's
        // preferred and ORCA velocities (already there after a step that
        // started from the slot arrays)
        s.pos_in[id] = make_float2(pos.x, pos.y);
        s.vel_in[id] = make_float2(sv.x, sv.y);
        s.attraction[id] = make_float2(attr.x, attr.y);
        const int qp = s.prev_of[t];
        if (s.sorted_in && qp >= 0) {
            s.pref_vel[id] = s.pref_vel_s_old[qp];
            if (s.v_orca) s.v_orca[id] = s.v_orca_s_old[qp];
        }
        s.active[id] = 0;                                   // DestroyAgent
        uint32_t q = atomicAdd(s.events_n, 1u);
        s.events[q] = make_int4(id, 0, 0, -1);
        __threadfence();
        return true;
    }
    // ApplySteeringForce (Simulator.cpp:638-657)
    const float speed = sv.z;
    V2 des = normalize(attr - pos) * speed;
    s.pref_vel_s[t] = make_float2(des.x, des.y);
    if (event) __threadfence();
    return event;
}

// The k nearest of sorted agent t (replaces KDTree::KNearestAgents); an
// agent destroyed this step gets an empty list.
template <int K>   // K == k: the k-th neighbour is the last register of the list
__device__ __forceinline__ void knn_agent(const DevState& s, int* knn_lds, int t) {
    if (s.dead_s[t]) {
        s.nbr_cnt[t] = 0;
        return;
    }
    const float4 pr = s.sorted_pr[t];
    const float4 sv = s.sorted_v[t];
    const V2 pos = mk(pr.x, pr.y);
    const EnvDesc& E = env_of_sorted(s, sv);
    int slot[K];
    float sd[K];
    const float bound = knn_seed_bound<K>(s, t, pos, sd, slot);
    int* buf = knn_lds + threadIdx.x;
    // the one-pass walk covers the bound's whole cell rectangle: a loose
    // bound (seeds from far away, a snapshot from another time) goes to the
    // ring search instead, which stops at the k-th candidate
    const float wmax = kKnnWalkCells * E.ag.cell;
    const int cnt = bound <= wmax * wmax
        ? knn_query_seeded<K, kKnnBuf, kKnnRanges>(s, E, pos, bound, buf, buf + kKnnBuf * blockDim.x, blockDim.x,
                                                   sd, slot)
        : knn_query<K, kKnnBuf>(s, E, pos, bound, buf, blockDim.x, slot);
#pragma unroll
    for (int i = 0; i < K; ++i)
        if (i < cnt) s.nbr[(size_t)i * s.n_slots + t] = slot[i];
    s.nbr_cnt[t] = cnt;
}

// The block that finishes last copies the step's counters and its first
// event_peek events (words [0, 8 + 4 * min(events, peek)) of the block that
// events_n heads) to host_xfer[1..], then, after a system-scope fence, the
// step number (acc[12]) to host_xfer[0]: the host polls that word and reads
// the rest once it matches.  Only the threads that appended an event fence
// (device scope, before the block barrier), so the rare events cost the
// rare fences; an agent-scope fence in every block would write back its
// XCD's L2 each time.  The last block reads with device-scope atomic loads.
__device__ __forceinline__ void publish_if_last(const DevState& s) {
    __shared__ uint32_t last;
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(s.blocks_done, 1u) == gridDim.x - 1u ? 1u : 0u;
    __syncthreads();
    if (!last) return;
    __threadfence();
    const uint32_t* xfer = s.events_n;
    const uint32_t ne = __hip_atomic_load(xfer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t n = 8u + 4u * min(ne, s.event_peek);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        s.host_xfer[1 + i] = __hip_atomic_load(xfer + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        *s.blocks_done = 0u;
        __threadfence_system();
        *(volatile uint32_t*)s.host_xfer = (uint32_t)s.acc[12];
    }
}

// Goal checks, attraction points and steering of every agent; the block
// that finishes last publishes the step's events, so the host plans the
// replans while the neighbour search and the ORCA passes run.  No LDS: the
// IRM's dependent loads are hidden by as many waves as the registers allow.
__global__ __launch_bounds__(256) void k_attr(DevState s) {
    if (blockIdx.x == 0 && threadIdx.x == 0) s.acc[15] = 0;   // the grid build has used it
    attr_agent(s);
    publish_if_last(s);
}

// The k nearest of every agent (kKnnBuf + 2 kKnnRanges ints of LDS per lane).
template <int K>
__global__ __launch_bounds__(256) void k_knn(DevState s) {
    extern __shared__ int knn_lds[];
    const uint32_t n_active = s.cell_start[s.n_keys];
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t < (int)n_active) knn_agent<K>(s, knn_lds, t);
}

// ------------------------------------------------------------------- ORCA
struct Con { V2 n, p; };

// Where a lane's constraint slot c lives in its block's LDS tile (float4
// units): slot-major, c * 64 + tid, for c < main; the lane's slots from
// `main` on are contiguous at ext + (c - main), its share of the block's
// spill area.  Tiles without a spill area have main = INT_MAX.
struct Tile {
    int tid, main, ext;
    __device__ __forceinline__ int at(int c) const { return c < main ? c * 64 + tid : ext + (c - main); }
};

// Spill area of the main ORCA pass: float4 slots per 64-lane block.
static constexpr int kOrcaSpill = 63;

__device__ __forceinline__ Con con_ld(const float4* lds, const Tile& T, int slot) {
    float4 c = lds[T.at(slot)];
    Con r; r.n = mk(c.x, c.y); r.p = mk(c.z, c.w);
    return r;
}
__device__ __forceinline__ void con_st(float4* lds, const Tile& T, int slot, V2 p, V2 n) {
    lds[T.at(slot)] = make_float4(n.x, n.y, p.x, p.y);
}
// Constraint c of 3D-LP queue entry q (slot-major, so consecutive entries
// held by consecutive lanes are consecutive float4s).
__device__ __forceinline__ float4* q3d_slot(const DevState& s, uint32_t q, int c) {
    return s.q3d_con + (size_t)c * s.q3d_cap + q;
}

// Constraint::Contains, methodB (ORCA.h:49-60)
__device__ __forceinline__ bool con_contains(const Con& c, V2 q) { return det(right(c.n), c.p - q) <= 0.0f; }

// Obstacle half-plane for edge e (ORCA.cpp:70-333).  Returns false when the
// reference emits nothing for it (`continue` without push_back).
__device__ bool obstacle_constraint(const DevState& s, const EnvDesc& E, int e, V2 position, V2 velocity, float rad,
                                    V2& outP, V2& outN) {
    float4 a = E.oe0[e], b = E.oe1[e];
    uint8_t fl = E.oflags[e];
    V2 pL = mk(a.x, a.y), pR = mk(a.z, a.w);
    V2 pPrevL = mk(b.x, b.y), pNextR = mk(b.z, b.w);
    bool cvL = fl & 1, cvR = (fl >> 1) & 1;

    V2 rp1 = pL - position;
    V2 rp2 = pR - position;
    V2 seg = pR - pL;
    const float sp = dot(neg(rp1), seg) / len2(seg);
    const float distSqLine = len2(neg(rp1) - seg * sp);
    const float distSq1 = len2(rp1);
    const float distSq2 = len2(rp2);
    seg = normalize(seg);
    const float radiusSq = rad * rad;

    if (sp < 0.0f && distSq1 <= radiusSq) {                         // case 1
        if (!cvL) return false;
        outP = mk(0.0f, 0.0f); outN = normalized(neg(rp1));
        return true;
    } else if (sp > 1.0f && distSq2 <= radiusSq) {                  // case 2
        V2 rnd = normalize(pNextR - pR);
        if (cvR && det(rp2, rnd) >= 0.0f) {
            outP = mk(0.0f, 0.0f); outN = normalized(neg(rp2));
            return true;
        }
        return false;
    } else if (sp >= 0.0f && sp < 1.0f && distSqLine <= radiusSq) { // case 3
        outP = mk(0.0f, 0.0f); outN = right(seg);
        return true;
    }

    // The obstacle's (L, R) pair, possibly collapsed to one vertex.  With
    // R := L, next(R) is the original R; with L := R, prev(L) is the
    // original L.
    V2 L = pL, R = pR, prevL = pPrevL, nextR = pNextR;
    bool cL = cvL, cR = cvR;
    bool same = false;          // obstNeighborL == obstNeighborR (pointer identity)
    V2 leftLeg, rightLeg;
    if (sp < 0.0f && distSqLine <= radiusSq) {
        if (!cvL) return false;
        R = pL; nextR = pR; cR = cvL; same = true;
        const float leg1 = sqrtf(distSq1 - radiusSq);
        leftLeg = mk(rp1.x * leg1 - rp1.y * rad, rp1.x * rad + rp1.y * leg1) / distSq1;
        rightLeg = mk(rp1.x * leg1 + rp1.y * rad, -rp1.x * rad + rp1.y * leg1) / distSq1;
    } else if (sp > 1.0f && distSqLine <= radiusSq) {
        if (!cvR) return false;
        L = pR; prevL = pL; cL = cvR; same = true;
        const float leg2 = sqrtf(distSq2 - radiusSq);
        leftLeg = mk(rp2.x * leg2 - rp2.y * rad, rp2.x * rad + rp2.y * leg2) / distSq2;
        rightLeg = mk(rp2.x * leg2 + rp2.y * rad, -rp2.x * rad + rp2.y * leg2) / distSq2;
    } else {
        if (cvL) {
            const float leg1 = sqrtf(distSq1 - radiusSq);
            leftLeg = mk(rp1.x * leg1 - rp1.y * rad, rp1.x * rad + rp1.y * leg1) / distSq1;
        } else {
            leftLeg = neg(seg);
        }
        if (cvR) {
            const float leg2 = sqrtf(distSq2 - radiusSq);
            rightLeg = mk(rp2.x * leg2 + rp2.y * rad, -rp2.x * rad + rp2.y * leg2) / distSq2;
        } else {
            rightLeg = seg;
        }
    }

    bool leftForeign = false, rightForeign = false;
    V2 lnd = normalize(L - prevL);
    if (cL && det(leftLeg, neg(lnd)) >= 0.0f) { leftLeg = neg(lnd); leftForeign = true; }
    V2 rnd = normalize(nextR - R);
    if (cR && det(rightLeg, rnd) <= 0.0f) { rightLeg = rnd; rightForeign = true; }

    const float invT = s.inv_time_horizon_obst;
    const V2 leftCutoff = (L - position) * invT;
    const V2 rightCutoff = (R - position) * invT;
    const V2 cutoffVec = rightCutoff - leftCutoff;
    const float t = same ? 0.5f : dot(velocity - leftCutoff, cutoffVec) / len2(cutoffVec);
    const float tLeft = dot(velocity - leftCutoff, leftLeg);
    const float tRight = dot(velocity - rightCutoff, rightLeg);

    if ((t < 0.0f && tLeft < 0.0f) || (same && tLeft < 0.0f && tRight < 0.0f)) {
        V2 uw = normalized(velocity - leftCutoff);
        outP = leftCutoff + (uw * invT) * rad; outN = uw;
        return true;
    } else if (t > 1.0f && tRight < 0.0f) {
        V2 uw = normalized(velocity - rightCutoff);
        outP = rightCutoff + (uw * invT) * rad; outN = uw;
        return true;
    }
    const float inf = INFINITY;
    const float dCut = (t < 0.0f || t > 1.0f || same) ? inf : len2(velocity - (leftCutoff + cutoffVec * t));
    const float dLeft = (tLeft < 0.0f) ? inf : len2(velocity - (leftCutoff + leftLeg * tLeft));
    const float dRight = (tRight < 0.0f) ? inf : len2(velocity - (rightCutoff + rightLeg * tRight));
    if (dCut <= dLeft && dCut <= dRight) {
        V2 nrm = left(neg(seg));
        outP = leftCutoff + (nrm * invT) * rad; outN = nrm;
        return true;
    } else if (dLeft <= dRight) {
        if (leftForeign) return false;
        V2 nrm = left(leftLeg);
        outP = leftCutoff + (nrm * rad) * invT; outN = nrm;
        return true;
    } else {
        if (rightForeign) return false;
        V2 nrm = right(rightLeg);
        outP = rightCutoff + (nrm * rad) * invT; outN = nrm;
        return true;
    }
}

// Agent half-plane (ORCA.cpp:339-423)
__device__ void agent_constraint(const DevState& s, V2 position, V2 velocity, float rad,
                                 float4 q, float2 qv, V2& outP, V2& outN) {
    const float T = s.time_horizon;
    const float combined = q.z + rad;
    V2 VOPos;
    float VORadius;
    if (s.inv_time_horizon != 0.0f) {           // T a power of two: the same values without dividing
        const float iT = s.inv_time_horizon;
        VOPos = mk((q.x - position.x) * iT, (q.y - position.y) * iT);
        VORadius = combined * iT;
    } else {
        VOPos = mk((q.x - position.x) / T, (q.y - position.y) / T);
        VORadius = combined / T;
    }
    float VOPosLength = len(VOPos);
    V2 relVel = mk(velocity.x - qv.x, velocity.y - qv.y);
    V2 relPos = mk(q.x - position.x, q.y - position.y);
    float relPosLength = len(relPos);
    if (relPosLength < combined) {
        const V2 w = relVel - relPos / s.dt;
        const float wl = len(w);
        const V2 uw = w / wl;
        const V2 U = uw * (combined / s.dt - wl);
        outP = velocity + U * 0.5f; outN = uw;
        return;
    }
    float half = atan_f(VORadius / VOPosLength);
    V2 VOLeftLeg, VORightLeg;
    rotate_pm(VOPos, half, VOLeftLeg, VORightLeg);     // rotate(VOPos, +-half)
    float sqFromCentre = sq_dist(VOPos, relVel);
    bool below = det(VOLeftLeg - VOPos, relVel - VOPos) > 0;
    if (below) {
        float distToEdge = VORadius - sqrtf(sqFromCentre);
        V2 ln = normalize(relVel - VOPos);
        outP = velocity + (ln * distToEdge) * 0.5f; outN = ln;
    } else {
        const V2 lperp = left(VOPos) / VOPosLength;
        if (dot(lperp, relVel) >= 0) {
            const V2 ll = normalized(VOLeftLeg);
            const V2 U = ll * dot(relVel, ll) - relVel;
            outP = velocity + U * 0.5f; outN = left(ll);
        } else {
            const V2 rl = normalized(VORightLeg);
            const V2 U = rl * dot(relVel, rl) - relVel;
            outP = velocity + U * 0.5f; outN = right(rl);
        }
    }
}

// ORCA::RandomizedLP (ORCA.cpp:428-587) over a virtual list: entries
// [0, n_direct) at LDS slots [0, n_direct), entries [n_direct, n) at slots
// [proj_base, ...).  Returns the failing index or n.
__device__ int lp2d(const float4* lds, const Tile& T, int n, int n_direct, int proj_base, V2 opt, float maxSpeed,
                    bool useDir, V2& out) {
    if (useDir) out = opt * maxSpeed;
    else if (len(opt) > maxSpeed) out = normalized(opt) * maxSpeed;
    else out = opt;
    if (n == 0) return n;
    for (int i = 0; i < n; ++i) {
        Con h = con_ld(lds, T, i < n_direct ? i : proj_base + (i - n_direct));
        if (con_contains(h, out)) continue;
        V2 dir = right(h.n);
        float dpd = dot(dir, h.p);
        float disc = dpd * dpd + maxSpeed * maxSpeed - dot(h.p, h.p);
        if (disc <= 0.0f) return i;
        float sq = sqrtf(disc);
        float lft = -dpd - sq;
        float rgt = -dpd + sq;
        for (int j = 0; j < i; ++j) {
            Con hj = con_ld(lds, T, j < n_direct ? j : proj_base + (j - n_direct));
            float den = det(right(h.n), right(hj.n));
            float num = det(right(hj.n), h.p - hj.p);
            if (fabsf(den) <= kEps) {
                if (num < 0.0f) return i;
                continue;
            }
            const float t = num / den;
            if (den >= 0.0f) rgt = std_min(rgt, t);
            else lft = std_max(lft, t);
            if (lft > rgt) return i;
        }
        if (useDir) {
            if (dot(opt, right(h.n)) > 0) out = h.p + right(h.n) * rgt;
            else out = h.p + right(h.n) * lft;
        } else {
            float t = dot(right(h.n), opt - h.p);
            if (t < lft) out = h.p + right(h.n) * lft;
            else if (t > rgt) out = h.p + right(h.n) * rgt;
            else out = h.p + right(h.n) * t;
        }
    }
    return n;
}

// ORCA::RandomizedLP3D (ORCA.cpp:592-669); projected agent constraints are
// written to LDS slots [proj_base, proj_base + i - nObst).
__device__ void lp3d(float4* lds, const Tile& T, int n, int nObst, int proj_base, float maxSpeed, int failed, V2& out) {
    float maxPen = 0.0f;
    for (int i = failed; i < n; ++i) {
        const Con ci = con_ld(lds, T, i);
        V2 dir = right(ci.n);
        if (det(dir, ci.p - out) <= maxPen) continue;
        int np = 0;
        for (int j = nObst; j < i; ++j) {
            const Con cj = con_ld(lds, T, j);
            float d = det(right(ci.n), right(cj.n));
            V2 newP;
            if (fabsf(d) <= kEps) {
                if (dot(ci.n, cj.n) > 0) continue;
                newP = (ci.p + cj.p) * 0.5f;
            } else {
                float t = det(right(cj.n), ci.p - cj.p) / d;
                newP = ci.p + right(ci.n) * t;
            }
            con_st(lds, T, proj_base + np, newP, normalized(cj.n - ci.n));
            ++np;
        }
        const V2 temp = out;
        if (lp2d(lds, T, nObst + np, nObst, proj_base, ci.n, maxSpeed, true, out) < nObst + np) out = temp;
        maxPen = det(dir, ci.p - out);
    }
}

// Velocity + position systems and the force between them
// (Simulator.cpp:592-635, 676-677).
// Written by sorted position t.
__device__ __forceinline__ void integrate(const DevState& s, int t, V2 position, V2 velocity, V2 o) {
    const V2 F = mk(o.x - velocity.x, o.y - velocity.y);
    const V2 nv = mk(velocity.x + F.x * s.mass_recip * s.dt, velocity.y + F.y * s.mass_recip * s.dt);
    const V2 npn = mk(position.x + (nv.x * s.dt), position.y + (nv.y * s.dt));
    s.pos_s[t] = make_float2(npn.x, npn.y);
    s.vel_s[t] = make_float2(nv.x, nv.y);
    if (s.v_orca_s) s.v_orca_s[t] = make_float2(o.x, o.y);
}

// ----------------------------------------------- learned-proposal support
// The 4 nearest obstacle edges in range (the ones the obstacle query keeps,
// Simulator.cpp:259-292), by (squared distance to the closest point, edge
// id), with the index of the constraint each emitted (-1: none).  They are
// feature slots 10..13 of the proposal network.
struct Near4 {
    float d[4];
    int e[4];
    int c[4];
};
__device__ __forceinline__ void near4_init(Near4& T) {
#pragma unroll
    for (int i = 0; i < 4; ++i) { T.d[i] = INFINITY; T.e[i] = INT_MAX; T.c[i] = -1; }
}
__device__ __forceinline__ void near4_insert(Near4& T, float dd, int e, int c) {
    bool b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) b[i] = (dd < T.d[i]) | ((dd == T.d[i]) & (e < T.e[i]));
#pragma unroll
    for (int i = 3; i > 0; --i) {
        T.d[i] = b[i - 1] ? T.d[i - 1] : (b[i] ? dd : T.d[i]);
        T.e[i] = b[i - 1] ? T.e[i - 1] : (b[i] ? e : T.e[i]);
        T.c[i] = b[i - 1] ? T.c[i - 1] : (b[i] ? c : T.c[i]);
    }
    T.d[0] = b[0] ? dd : T.d[0];
    T.e[0] = b[0] ? e : T.e[0];
    T.c[0] = b[0] ? c : T.c[0];
}

__device__ __forceinline__ float feat_clamp(float x) {
    return fminf(fmaxf(x, -mlp::kFeatClamp), mlp::kFeatClamp);
}
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    const __bf16 x = (__bf16)a, y = (__bf16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}

// The 64 BF16 features of one agent (layout in mlp.h): its 8 chunks of 8 BF16 to out[0], out[stride],
// ... out[7 * stride].
template <int K>
__device__ void agent_features(const DevState& s, const EnvDesc& E, int t, V2 position, V2 velocity, float rad,
                               float2 pv, int cnt, const Near4& T, uint4* out, int stride) {
    float f[mlp::kInPad];
#pragma unroll
    for (int i = 0; i < mlp::kInPad; ++i) f[i] = 0.0f;
#pragma unroll
    for (int j = 0; j < mlp::kAgentSlots && j < K; ++j) {
        if (j < cnt) {
            const int sj = s.nbr[(size_t)j * s.n_slots + t];
            const float4 q = s.sorted_pr[sj];
            const float2 qv = sorted_vel(s, sj);
            f[4 * j + 0] = feat_clamp((q.x - position.x) * mlp::kPosScale);
            f[4 * j + 1] = feat_clamp((q.y - position.y) * mlp::kPosScale);
            f[4 * j + 2] = feat_clamp((qv.x - velocity.x) * mlp::kVelScale);
            f[4 * j + 3] = feat_clamp((qv.y - velocity.y) * mlp::kVelScale);
        }
    }
#pragma unroll
    for (int i = 0; i < mlp::kObstFeat; ++i) {
        if (T.e[i] != INT_MAX) {
            const float4 a = E.oe0[T.e[i]];
            f[40 + 4 * i + 0] = feat_clamp((a.x - position.x) * mlp::kPosScale);
            f[40 + 4 * i + 1] = feat_clamp((a.y - position.y) * mlp::kPosScale);
            f[40 + 4 * i + 2] = feat_clamp((a.z - position.x) * mlp::kPosScale);
            f[40 + 4 * i + 3] = feat_clamp((a.w - position.y) * mlp::kPosScale);
        }
    }
    f[56] = feat_clamp(pv.x * mlp::kVelScale);
    f[57] = feat_clamp(pv.y * mlp::kVelScale);
    f[58] = feat_clamp(velocity.x * mlp::kVelScale);
    f[59] = feat_clamp(velocity.y * mlp::kVelScale);
    f[60] = feat_clamp(rad * mlp::kRadScale);
#pragma unroll
    for (int c = 0; c < 8; ++c)
        out[c * stride] = make_uint4(pack_bf16x2(f[8 * c + 0], f[8 * c + 1]), pack_bf16x2(f[8 * c + 2], f[8 * c + 3]),
                                     pack_bf16x2(f[8 * c + 4], f[8 * c + 5]), pack_bf16x2(f[8 * c + 6], f[8 * c + 7]));
}

// The obstacle bucket of position: entries [e0, e1) of obucket_edges, the
// candidate edges in ascending id order.
__device__ __forceinline__ void obstacle_bucket(const EnvDesc& s, V2 position, int& e0, int& e1) {
    int bx = grid_cell(position.x, s.og.x0, s.og.inv_cell, s.og.side);
    int by = grid_cell(position.y, s.og.y0, s.og.inv_cell, s.og.side);
    int b = by * s.og.side + bx;
    e0 = s.obucket_start[b];
    e1 = s.obucket_start[b + 1];
}
// Whether bucket entry q is an obstacle neighbour of position (the test of
// Simulator.cpp:266-287); e receives its edge, dd the squared distance to
// the edge's closest point.
__device__ __forceinline__ bool obstacle_near(const EnvDesc& s, V2 position, float range2, int q, int& e,
                                              float& dd) {
    e = s.obucket_edges[q];
    float4 a = s.oe0[e];
    V2 L = mk(a.x, a.y), R = mk(a.z, a.w);
    float sld = det(L - position, R - L);
    float sqd = sld * sld / sq_dist(L, R);
    if (!(sqd < range2 && sld < 0.0f)) return false;
    V2 c = closest_on_segment(position, L, R);
    dd = sq_dist(c, position);
    return dd < range2;
}
// Obstacle query with the reference's filter (Simulator.cpp:266-287) over
// the agent's bucket; `emit(e, dd)` runs for every edge it keeps, in scan
// order, and returns false to stop the scan.
template <typename F>
__device__ __forceinline__ void obstacle_query(const EnvDesc& s, V2 position, float range2, F&& emit) {
    int e0, e1;
    obstacle_bucket(s, position, e0, e1);
    for (int q = e0; q < e1; ++q) {
        int e;
        float dd;
        if (obstacle_near(s, position, range2, q, e, dd))
            if (!emit(e, dd)) return;
    }
}

// A queue slot for every active lane of the wavefront, with one atomic.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* ctr) {
    const unsigned long long act = __ballot(1);
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)act) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(ctr, (uint32_t)__popcll(act));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
}

// Count `pred` over the active lanes of the wavefront with one atomic.
__device__ __forceinline__ void wave_count(uint32_t* ctr, bool pred) {
    const unsigned long long b = __ballot(pred);
    const unsigned long long act = __ballot(1);
    if (b && (int)__lane_id() == __ffsll((long long)act) - 1) atomicAdd(ctr, (uint32_t)__popcll(b));
}
// Sum of v < 32 over the active lanes, bit by bit (one atomic per bit).
__device__ __forceinline__ void wave_sum5(uint32_t* ctr, uint32_t v) {
    const unsigned long long act = __ballot(1);
    const bool leader = (int)__lane_id() == __ffsll((long long)act) - 1;
    uint32_t total = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) total += (uint32_t)__popcll(__ballot((v >> b) & 1u)) << b;
    if (leader && total) atomicAdd(ctr, total);
}

// RandomizedLP (ORCA.cpp:428-587, useDirOpt false) restricted to the
// constraints in `mask`, in index order, with the same arithmetic as lp2d.
// Returns false when the restricted problem is infeasible.
__device__ bool lp2d_mask(const float4* lds, const Tile& T, uint32_t mask, V2 opt, float maxSpeed, V2& out) {
    if (len(opt) > maxSpeed) out = normalized(opt) * maxSpeed;
    else out = opt;
    uint32_t done = 0;
    for (uint32_t m = mask; m; m &= m - 1) {
        const int i = __ffs(m) - 1;
        const Con h = con_ld(lds, T, i);
        if (!con_contains(h, out)) {
            V2 dir = right(h.n);
            float dpd = dot(dir, h.p);
            float disc = dpd * dpd + maxSpeed * maxSpeed - dot(h.p, h.p);
            if (disc <= 0.0f) return false;
            float sq = sqrtf(disc);
            float lft = -dpd - sq;
            float rgt = -dpd + sq;
            for (uint32_t d = done; d; d &= d - 1) {
                const Con hj = con_ld(lds, T, __ffs(d) - 1);
                float den = det(right(h.n), right(hj.n));
                float num = det(right(hj.n), h.p - hj.p);
                if (fabsf(den) <= kEps) {
                    if (num < 0.0f) return false;
                    continue;
                }
                const float t = num / den;
                if (den >= 0.0f) rgt = std_min(rgt, t);
                else lft = std_max(lft, t);
                if (lft > rgt) return false;
            }
            float t = dot(right(h.n), opt - h.p);
            if (t < lft) out = h.p + right(h.n) * lft;
            else if (t > rgt) out = h.p + right(h.n) * rgt;
            else out = h.p + right(h.n) * t;
        }
        done |= 1u << i;
    }
    return true;
}

// ---------------------------------------------------------- ORCA passes
// Modes of the main pass: the plain step, the plain step that also logs a
// training sample per agent, and the learned proposal with verification.
enum OrcaMode { kPlain = 0, kLog = 1, kLearned = 2 };

constexpr float kActiveTol = 1e-3f;   // a constraint is labelled active within 1 mm/s of the LP velocity

// One agent of ObstacleAvoidanceForce (Simulator.cpp:659-686): obstacle
// query, constraints into the LDS tile (OSLOTS obstacle + K agent slots),
// RandomizedLP.  With OVERFLOW, an agent that needs more than OSLOTS
// obstacle constraints is parked on the overflow queue untouched and
// redone by k_orca_ovf with a full tile.  A failed 2D LP parks the agent on
// the 3D-LP queue; so does a rejected learned proposal (failed index -1:
// the queue kernel runs the full 2D LP first).
// SPILL (the main pass): the lane's tile holds OSLOTS + K slots; an agent
// with more obstacle constraints than that is parked for the overflow pass,
// and one whose obstacle and agent constraints together exceed it takes the
// rest from the block's spill area (kOrcaSpill slots, counted by *spill_used)
// or, when that is used up, is parked too.  Without SPILL the tile holds
// every constraint (OSLOTS >= the longest obstacle bucket list).
template <int K, int OSLOTS, bool SPILL, int MODE>
__device__ __forceinline__ void orca_agent(const DevState& s, int t, float4* lds, Tile TL, int* spill_used) {
    const float4 pr = s.sorted_pr[t];
    const float4 sv = s.sorted_v[t];
    const float2 v0 = make_float2(sv.x, sv.y);
    const int id = __float_as_int(pr.w);
    if (s.dead_s[t]) {                   // destroyed this step: state unchanged
        s.pos_s[t] = make_float2(pr.x, pr.y);
        s.vel_s[t] = v0;
        return;
    }
    const V2 position = mk(pr.x, pr.y);
    const V2 velocity = mk(v0.x, v0.y);
    const float rad = pr.z;
    const float speed = sv.z;
    const float maxSpeed = s.max_speed_from_pref ? speed : s.max_speed;
    const float2 pv = s.pref_vel_s[t];
    const EnvDesc& E = env_of_sorted(s, sv);

    // obstacle neighbours in scan order, constraints generated in place
    int nc = 0;
    Near4 T;
    if (MODE != kPlain) near4_init(T);
    if (E.n_obst_edges > 0) {
        const float range = s.time_horizon_obst * maxSpeed + rad;          // ORCA.cpp:27
        bool parked = false;
        obstacle_query(E, position, range * range, [&](int e, float dd) {
            V2 cp, cn;
            const bool made = obstacle_constraint(s, E, e, position, velocity, rad, cp, cn);
            if (made && SPILL && nc == OSLOTS + K) {
                s.ovf_agent[wave_alloc(&s.stats[1])] = t;
                parked = true;
                return false;
            }
            if (MODE != kPlain) near4_insert(T, dd, e, made ? nc : -1);
            if (made) {
                con_st(lds, TL, nc, cp, cn);
                ++nc;
            }
            return true;
        });
        if (parked) return;
    }
    const int nObst = nc;
    const int cnt = s.nbr_cnt[t];
    if (SPILL) {
        // the learned proposal indexes constraints by the bits of a word
        const int need = nObst + cnt - (OSLOTS + K);
        bool park = MODE == kLearned && nObst + cnt > 32;
        if (!park && need > 0) {
            const int base = atomicAdd(spill_used, need);
            if (base + need > kOrcaSpill) park = true;
            else TL.ext = (OSLOTS + K) * 64 + base;
        }
        if (park) {
            s.ovf_agent[wave_alloc(&s.stats[1])] = t;
            return;
        }
    }
    {
        // all neighbour loads issued before the first constraint
        float4 q[K];
        float2 qv[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const int sj = j < cnt ? s.nbr[(size_t)j * s.n_slots + t] : t;
            q[j] = s.sorted_pr[sj];
            qv[j] = sorted_vel(s, sj);
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (j < cnt) {
                V2 cp, cn;
                agent_constraint(s, position, velocity, rad, q[j], qv[j], cp, cn);
                con_st(lds, TL, nc, cp, cn);
                ++nc;
            }
        }
    }

    V2 o;
    if (MODE == kLearned) {
        // proposal: predicted slots -> constraint indices
        const uint32_t pm = s.mlp_mask[t];
        uint32_t cm = 0;
#pragma unroll
        for (int j = 0; j < mlp::kAgentSlots && j < K; ++j)
            if (j < cnt && ((pm >> j) & 1u)) cm |= 1u << (nObst + j);
#pragma unroll
        for (int i = 0; i < mlp::kObstFeat; ++i)
            if (((pm >> (mlp::kAgentSlots + i)) & 1u) && T.c[i] >= 0) cm |= 1u << T.c[i];
        // verification: optimum of the restricted LP, contained by the rest
        bool ok = lp2d_mask(lds, TL, cm, mk(pv.x, pv.y), maxSpeed, o);
        for (int c = 0; ok && c < nc; ++c)
            if (!((cm >> c) & 1u) && !con_contains(con_ld(lds, TL, c), o)) ok = false;
        // per-wavefront counters spread over 64 slots (one address each
        // would serialize 15k atomics per launch)
        uint32_t* ws = s.wstats + 4 * (blockIdx.x & 63);
        wave_count(ws + 0, true);
        wave_count(ws + 1, ok);
        wave_sum5(ws + 2, (uint32_t)__popc(cm));
        if (ok) {
            integrate(s, t, position, velocity, o);
            return;
        }
        // rejected: the full 2D LP, as the plain pass
    }

    int fail = lp2d(lds, TL, nc, nc, 0, mk(pv.x, pv.y), maxSpeed, false, o);
    if (fail < nc) {
        // RandomizedLP3D runs on a compacted queue: park the constraints,
        // the 2D result and the failing index
        uint32_t q = wave_alloc(&s.stats[0]);
        s.q3d_agent[q] = t;
        s.q3d_info[q] = make_int4(nc, nObst, fail, 0);
        s.q3d_vel[q] = make_float2(o.x, o.y);
        for (int c = 0; c < nc; ++c) *q3d_slot(s, q, c) = lds[TL.at(c)];
        return;
    }
    if (MODE == kLog) {
        // training sample: features, active slots at the LP velocity, v
        const uint32_t idx = wave_alloc(s.log_n);
        if (idx < s.log_cap) {
            agent_features<K>(s, E, t, position, velocity, rad, pv, cnt, T, s.log_x + (size_t)idx * 8, 1);
            uint32_t act = 0, valid = 0;
            bool repr = true;
            for (int c = 0; c < nc; ++c) {
                const Con h = con_ld(lds, TL, c);
                const bool a = det(right(h.n), h.p - o) >= -kActiveTol;
                int slot = -1;
                if (c >= nObst) {
                    if (c - nObst < mlp::kAgentSlots) slot = c - nObst;
                } else {
#pragma unroll
                    for (int i = 0; i < mlp::kObstFeat; ++i)
                        if (T.c[i] == c) slot = mlp::kAgentSlots + i;
                }
                if (slot >= 0) {
                    valid |= 1u << slot;
                    if (a) act |= 1u << slot;
                } else if (a) {
                    repr = false;
                }
            }
            s.log_label[idx] = act | (valid << mlp::kLabelValidShift) | (repr ? mlp::kLabelRepresentable : 0u);
            s.log_v[idx] = make_float2(o.x, o.y);
        }
    }
    integrate(s, t, position, velocity, o);
}

// Main ORCA pass: one agent per lane in Morton order.  The constraint tile
// is small so that more waves fit per CU: OSLOTS + K slots per lane (OSLOTS:
// 1, 2 or 4, picked by the autotuner) and a spill area of kOrcaSpill slots
// for the block's few lanes with more constraints; its last float4 holds the
// spill counter.  12 KiB at OSLOTS 1, K 10.
template <int K, int MODE, int OSLOTS>
__global__ __launch_bounds__(64) void k_orca(DevState s) {
    __shared__ float4 lds[(OSLOTS + K) * 64 + kOrcaSpill + 1];
    int* spill_used = reinterpret_cast<int*>(&lds[(OSLOTS + K) * 64 + kOrcaSpill]);
    const int tid = threadIdx.x;
    if (tid == 0) *spill_used = 0;
    __syncthreads();
    const uint32_t n_active = s.cell_start[s.n_keys];
    const int t = xcd_block() * 64 + tid;
    if (t >= (int)n_active) return;
    orca_agent<K, OSLOTS, true, MODE>(s, t, lds, Tile{tid, OSLOTS + K, 0}, spill_used);
}

// Main ORCA pass of the plain step with every constraint in registers (no
// LDS tile): up to OB obstacle constraints in a shift register (the last
// nObst slots, in scan order) and the K agent constraints, RandomizedLP
// over them (lp2_reg), then integration or the 3D-LP queue.  An agent with
// more than OB obstacle constraints is parked untouched for k_orca_ovf.
// At K 10 and OB 4 the pass is capped at 80 VGPRs for 6 waves per SIMD: a
// small spill (68 B per lane) costs less than the latency the sixth wave
// hides (ORCA phase 0.45 -> 0.42 ms at c2's dense window; 7 and 8 waves
// spill 156 and 192 B and are slower).
constexpr int orca_reg_waves(int K, int OB) { return K <= 10 && OB <= 4 ? 6 : 1; }

template <int K, int OB>
__device__ __forceinline__ void orca_reg_agent(const DevState& s, int t);

template <int K, int OB>
__global__ __launch_bounds__(64, orca_reg_waves(K, OB)) void k_orca_reg(DevState s) {
    const uint32_t n_active = s.cell_start[s.n_keys];
    const int t = xcd_block() * 64 + threadIdx.x;
    if (t >= (int)n_active) return;
    orca_reg_agent<K, OB>(s, t);
}

// The plain step's k-nearest pass followed, in the same thread, by the main
// ORCA pass of the same agent (k 5 and 10, 4 obstacle slots): the agent's
// ORCA reads only step-start state of other agents and its own new
// neighbour list and preferred velocity, so no other thread's result is
// waited for.  One kernel instead of two: the neighbour list is read back
// from the thread's own writes, and the tail of one pass overlaps the
// other's.  Capped at 80 VGPRs for 6 waves per SIMD.
// k_knn_orca for small crowds: 16 agents per 64-lane wavefront.  A crowd
// of a few thousand agents fills a few dozen wavefronts, one per SIMD, and
// the pass lasts as long as one wavefront's walk, constraints and LP; with
// 16 lanes a wavefront runs the union of 16 lanes' loops instead of 64, on
// four times as many SIMDs (c0 at cell scale 2.0: attraction + this pass
// 0.132 -> 0.115 ms; 100k agents: 8 % slower).  The launch shape's
// knn_block 16 selects it; the LDS is laid out for 64 lanes as in
// k_knn_orca.
template <int K>
__global__ __launch_bounds__(64, 6) void k_knn_orca_sparse(DevState s) {
    extern __shared__ int knn_lds[];
    const uint32_t n_active = s.cell_start[s.n_keys];
    if (threadIdx.x >= 16) return;
    const int t = xcd_block() * 16 + threadIdx.x;
    if (t >= (int)n_active) return;
    knn_agent<K>(s, knn_lds, t);
    orca_reg_agent<K, 4>(s, t);
}

template <int K>
__global__ __launch_bounds__(256, 6) void k_knn_orca(DevState s) {
    extern __shared__ int knn_lds[];
    const uint32_t n_active = s.cell_start[s.n_keys];
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= (int)n_active) return;
    knn_agent<K>(s, knn_lds, t);
    orca_reg_agent<K, 4>(s, t);
}

template <int K, int OB>
__device__ __forceinline__ void orca_reg_agent(const DevState& s, int t) {
    const float4 pr = s.sorted_pr[t];
    const float4 sv = s.sorted_v[t];
    const float2 v0 = make_float2(sv.x, sv.y);
    if (s.dead_s[t]) {                   // destroyed this step: state unchanged
        s.pos_s[t] = make_float2(pr.x, pr.y);
        s.vel_s[t] = v0;
        return;
    }
    const V2 position = mk(pr.x, pr.y);
    const V2 velocity = mk(v0.x, v0.y);
    const float rad = pr.z;
    const float maxSpeed = s.max_speed_from_pref ? sv.z : s.max_speed;
    const float2 pv = s.pref_vel_s[t];
    const EnvDesc& E = env_of_sorted(s, sv);

    V2 on[OB], op[OB];
#pragma unroll
    for (int u = 0; u < OB; ++u) { on[u] = mk(0.0f, 0.0f); op[u] = mk(0.0f, 0.0f); }
    int nObst = 0;
    if (E.n_obst_edges > 0) {
        const float range = s.time_horizon_obst * maxSpeed + rad;          // ORCA.cpp:27
        bool parked = false;
        obstacle_query(E, position, range * range, [&](int e, float) {
            V2 cp, cn;
            if (!obstacle_constraint(s, E, e, position, velocity, rad, cp, cn)) return true;
            if (nObst == OB) {
                s.ovf_agent[wave_alloc(&s.stats[1])] = t;
                parked = true;
                return false;
            }
#pragma unroll
            for (int u = 0; u + 1 < OB; ++u) { on[u] = on[u + 1]; op[u] = op[u + 1]; }
            on[OB - 1] = cn;
            op[OB - 1] = cp;
            ++nObst;
            return true;
        });
        if (parked) return;
    }
    const int cnt = s.nbr_cnt[t];
    V2 an[K], ap[K];
    uint32_t aval = 0u;
    // the neighbours' loads in batches of kNbrBatch, each batch issued before
    // its constraints are built: the loads of a batch are in flight together,
    // and the batch's 6 registers per neighbour stay within the 80-VGPR cap
    // (all 10 at once spilled 68 B per lane; two batches of 5 spill 24 B and
    // took the pass from 0.263 to 0.238 ms at c2's jam)
    constexpr int kNbrBatch = (K + 1) / 2;
#pragma unroll
    for (int h = 0; h < K; h += kNbrBatch) {
        float4 q[kNbrBatch];
        float2 qv[kNbrBatch];
#pragma unroll
        for (int u = 0; u < kNbrBatch; ++u) {
            const int j = h + u;
            const int sj = (j < K && j < cnt) ? s.nbr[(size_t)j * s.n_slots + t] : t;
            q[u] = s.sorted_pr[sj];
            qv[u] = sorted_vel(s, sj);
        }
#pragma unroll
        for (int u = 0; u < kNbrBatch; ++u) {
            const int j = h + u;
            if (j < K) {
                an[j] = mk(0.0f, 0.0f);
                ap[j] = mk(0.0f, 0.0f);
                if (j < cnt) {
                    agent_constraint(s, position, velocity, rad, q[u], qv[u], ap[j], an[j]);
                    aval |= 1u << j;
                }
            }
        }
    }
    V2 o;
    const int fs = lp2_reg<OB, K>(on, op, OB - nObst, an, ap, aval, mk(pv.x, pv.y), maxSpeed, o);
    if (fs < OB + K) {
        // RandomizedLP3D runs on the compacted queue (k_lp3d_reg)
        const uint32_t qi = wave_alloc(&s.stats[0]);
        s.q3d_agent[qi] = t;
        s.q3d_info[qi] = make_int4(nObst + cnt, nObst, slot_to_index<OB>(fs, nObst), 0);
        s.q3d_vel[qi] = make_float2(o.x, o.y);
#pragma unroll
        for (int u = 0; u < OB; ++u)
            if (u >= OB - nObst) *q3d_slot(s, qi, u - (OB - nObst)) = make_float4(on[u].x, on[u].y, op[u].x, op[u].y);
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < cnt) *q3d_slot(s, qi, nObst + j) = make_float4(an[j].x, an[j].y, ap[j].x, ap[j].y);
        return;
    }
    integrate(s, t, position, velocity, o);
}

// One overflow agent per wavefront (64-thread block): the lanes build its
// constraints together, and lane 0 runs the 2D LP over them.  The obstacle
// constraints of the bucket's edges are built one edge per lane and kept in
// bucket order (a ballot prefix gives each its slot), the order
// obstacle_query emits them in; the agent constraints are built one
// neighbour per lane.  Each constraint's arithmetic, the LP and what follows
// it are those of orca_agent in the plain mode, so the result is the same;
// a lane working through an agent alone waits out every constraint's
// loads and arithmetic in turn.  lds holds maxo + K constraints at slot c.
template <int K>
__device__ __forceinline__ void orca_agent_wave(const DevState& s, int t, float4* lds, int maxo) {
    const int lane = threadIdx.x;
    const Tile T{0, 0, 0};                           // slot c at lds[c]
    const float4 pr = s.sorted_pr[t];
    const float4 sv = s.sorted_v[t];
    const float2 v0 = make_float2(sv.x, sv.y);
    if (s.dead_s[t]) {                               // destroyed this step: state unchanged
        if (lane == 0) {
            s.pos_s[t] = make_float2(pr.x, pr.y);
            s.vel_s[t] = v0;
        }
        return;
    }
    const V2 position = mk(pr.x, pr.y);
    const V2 velocity = mk(v0.x, v0.y);
    const float rad = pr.z;
    const float maxSpeed = s.max_speed_from_pref ? sv.z : s.max_speed;
    const EnvDesc& E = env_of_sorted(s, sv);
    int nObst = 0;                                   // the same in every lane
    if (E.n_obst_edges > 0) {
        const float range = s.time_horizon_obst * maxSpeed + rad;          // ORCA.cpp:27
        int e0, e1;
        obstacle_bucket(E, position, e0, e1);
        for (int q0 = e0; q0 < e1; q0 += 64) {
            const int q = q0 + lane;
            bool made = false;
            V2 cp = mk(0.0f, 0.0f), cn = mk(0.0f, 0.0f);
            int e;
            float dd;
            if (q < e1 && obstacle_near(E, position, range * range, q, e, dd))
                made = obstacle_constraint(s, E, e, position, velocity, rad, cp, cn);
            const unsigned long long m = __ballot(made);
            const int slot = nObst + __popcll(m & ((1ull << lane) - 1ull));
            if (made && slot < maxo) con_st(lds, T, slot, cp, cn);
            nObst += __popcll(m);
        }
        nObst = min(nObst, maxo);                    // maxo covers the longest bucket
    }
    const int cnt = s.nbr_cnt[t];
    if (lane < cnt) {
        const int sj = s.nbr[(size_t)lane * s.n_slots + t];
        V2 cp, cn;
        agent_constraint(s, position, velocity, rad, s.sorted_pr[sj], sorted_vel(s, sj), cp, cn);
        con_st(lds, T, nObst + lane, cp, cn);
    }
    __syncthreads();
    if (lane == 0) {
        const int nc = nObst + cnt;
        const float2 pv = s.pref_vel_s[t];
        V2 o;
        const int fail = lp2d(lds, T, nc, nc, 0, mk(pv.x, pv.y), maxSpeed, false, o);
        if (fail < nc) {
            // RandomizedLP3D runs on the compacted queue, as from orca_agent
            const uint32_t qi = wave_alloc(&s.stats[0]);
            s.q3d_agent[qi] = t;
            s.q3d_info[qi] = make_int4(nc, nObst, fail, 0);
            s.q3d_vel[qi] = make_float2(o.x, o.y);
            for (int c = 0; c < nc; ++c) *q3d_slot(s, qi, c) = lds[c];
        } else {
            integrate(s, t, position, velocity, o);
        }
    }
    __syncthreads();                                 // lds is reused by the next agent
}

// Agents with more obstacle constraints than the main tile holds, with a tile for
// MAXO (>= the longest obstacle bucket list), plain LP.  Fixed grid,
// strided loop.  A queue no longer than the grid (the usual case: a few
// hundred agents of millions) gets one agent per wavefront
// (orca_agent_wave); a longer one, one agent per lane.
template <int K, int MAXO>
__global__ __launch_bounds__(64) void k_orca_ovf(DevState s) {
    __shared__ float4 lds[(MAXO + K) * 64];
    const uint32_t nq = s.stats[1];
    const int tid = threadIdx.x;
    if (nq <= gridDim.x) {
        if (blockIdx.x < nq) orca_agent_wave<K>(s, s.ovf_agent[blockIdx.x], lds, MAXO);
        return;
    }
    // entries go lane-major across the blocks (entry i: lane i / gridDim,
    // block i % gridDim), so a short queue spreads over many CUs instead of
    // filling the lanes of a few divergent wavefronts
    for (uint32_t i = tid * gridDim.x + blockIdx.x; i < nq; i += gridDim.x * 64)
        orca_agent<K, MAXO, false, kPlain>(s, s.ovf_agent[i], lds, Tile{tid, INT_MAX, 0}, nullptr);
}

// The overflow pass for scenes whose obstacle buckets hold more than 64
// edges: one agent per wavefront for any queue length, the tile of maxo + K
// constraints in dynamic LDS (one agent's, not 64 lanes').
template <int K>
__global__ __launch_bounds__(64) void k_orca_ovf_wide(DevState s, int maxo) {
    extern __shared__ float4 wide_lds[];
    const uint32_t nq = s.stats[1];
    for (uint32_t i = blockIdx.x; i < nq; i += gridDim.x) orca_agent_wave<K>(s, s.ovf_agent[i], wide_lds, maxo);
}

// RandomizedLP3D (ORCA.cpp:592-669) with the constraints in an LDS tile,
// for the 3D-LP queue entries with more obstacle constraints than the
// register tile of k_lp3d_reg holds (listed by it in q3d_big), then
// integration.  Fixed grid, strided loop.
template <int K, int MAXO>
__global__ __launch_bounds__(64) void k_lp3d(DevState s) {
    __shared__ float4 lds[(MAXO + 2 * K) * 64];
    const uint32_t nq = s.stats[4];
    const int tid = threadIdx.x;
    const Tile T{tid, INT_MAX, 0};
    // entries go lane-major across the blocks, as in k_orca_ovf
    for (uint32_t i = tid * gridDim.x + blockIdx.x; i < nq; i += gridDim.x * 64) {
        const uint32_t q = (uint32_t)s.q3d_big[i];
        const int t = s.q3d_agent[q];
        const int4 info = s.q3d_info[q];
        for (int c = 0; c < info.x; ++c) lds[T.at(c)] = *q3d_slot(s, q, c);
        const float4 pr = s.sorted_pr[t];
        const float4 sv = s.sorted_v[t];
        const float maxSpeed = s.max_speed_from_pref ? sv.z : s.max_speed;
        const float2 ov = s.q3d_vel[q];
        V2 o = mk(ov.x, ov.y);
        lp3d(lds, T, info.x, info.y, MAXO + K, maxSpeed, info.z, o);
        integrate(s, t, mk(pr.x, pr.y), mk(sv.x, sv.y), o);
    }
}

// k_lp3d for scenes whose obstacle buckets hold more than 64 edges: one
// queue entry per wavefront, its constraints copied by the lanes into a tile
// of maxo + 2K slots in dynamic LDS (the projected constraints from maxo + K
// on), and lane 0 runs the same lp3d.
template <int K>
__global__ __launch_bounds__(64) void k_lp3d_wide(DevState s, int maxo) {
    extern __shared__ float4 wide_lds[];
    const uint32_t nq = s.stats[4];
    const int lane = threadIdx.x;
    const Tile T{0, 0, 0};                           // slot c at wide_lds[c]
    for (uint32_t i = blockIdx.x; i < nq; i += gridDim.x) {
        const uint32_t q = (uint32_t)s.q3d_big[i];
        const int4 info = s.q3d_info[q];
        for (int c = lane; c < info.x; c += 64) wide_lds[c] = *q3d_slot(s, q, c);
        __syncthreads();
        if (lane == 0) {
            const int t = s.q3d_agent[q];
            const float4 pr = s.sorted_pr[t];
            const float4 sv = s.sorted_v[t];
            const float maxSpeed = s.max_speed_from_pref ? sv.z : s.max_speed;
            const float2 ov = s.q3d_vel[q];
            V2 o = mk(ov.x, ov.y);
            lp3d(wide_lds, T, info.x, info.y, maxo + K, maxSpeed, info.z, o);
            integrate(s, t, mk(pr.x, pr.y), mk(sv.x, sv.y), o);
        }
        __syncthreads();                             // the tile is reused by the next entry
    }
}

// Queues up to this many entries per wavefront of the grid are dealt
// lane-major in k_lp3d_reg.
static constexpr uint32_t kLp3dSpreadLanes = 4;

// The 3D-LP queue (agents whose 2D LP failed), then integration: constraints
// in registers (up to OB obstacle and K agent constraints; entries with more
// obstacle constraints go to q3d_big for k_lp3d).  A long queue: each
// wavefront takes 64 consecutive entries (coalesced slot-major loads),
// wavefronts dealt to the blocks round robin.  A short one (at most
// kLp3dSpreadLanes entries per wavefront of the grid): entry q goes to lane
// q / nw of wavefront q % nw, so that each wavefront holds a few entries and
// runs only their iterations of the candidate loop (the loop is
// wave-uniform: a full wavefront runs the union of its lanes').
template <int K, int OB>
__global__ __launch_bounds__(256, 4) void k_lp3d_reg(DevState s) {
    const uint32_t nq = s.stats[0];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t nw = gridDim.x * wpb;
    const uint32_t w0 = blockIdx.x + gridDim.x * (threadIdx.x >> 6);
    const bool spread = nq <= nw * kLp3dSpreadLanes;
    for (uint32_t w = w0; spread ? w == w0 : w * 64u < nq; w += nw) {
        const uint32_t q = spread ? lane * nw + w : w * 64u + lane;
        if (q >= nq) continue;
        const int4 info = s.q3d_info[q];
        const int nObst = info.y, na = info.x - info.y;
        if (nObst > OB) {
            s.q3d_big[wave_alloc(&s.stats[4])] = (int)q;
            continue;
        }
        const int t = s.q3d_agent[q];
        V2 on[OB], op[OB], an[K], ap[K];
#pragma unroll
        for (int c = 0; c < OB; ++c) {
            // obstacle constraint c - (OB - nObst) in slot c (the last nObst slots)
            const float4 v = c >= OB - nObst ? *q3d_slot(s, q, c - (OB - nObst)) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            on[c] = mk(v.x, v.y);
            op[c] = mk(v.z, v.w);
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const float4 v = j < na ? *q3d_slot(s, q, nObst + j) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            an[j] = mk(v.x, v.y);
            ap[j] = mk(v.z, v.w);
        }
        const float4 pr = s.sorted_pr[t];
        const float4 sv = s.sorted_v[t];
        const float maxSpeed = s.max_speed_from_pref ? sv.z : s.max_speed;
        const float2 ov = s.q3d_vel[q];
        V2 o = mk(ov.x, ov.y);
        lp3d_reg<OB, K>(on, op, an, ap, nObst, na, info.z, maxSpeed,
                        [&](int c) { return *q3d_slot(s, q, c); }, o);
        integrate(s, t, mk(pr.x, pr.y), mk(sv.x, sv.y), o);
    }
}

// Proposal network inputs of every sorted agent (zeros for agents
// destroyed this step).
template <int K>
__global__ __launch_bounds__(256) void k_features(DevState s) {
    const uint32_t n_active = s.cell_start[s.n_keys];
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= (int)n_active) return;
    const float4 pr = s.sorted_pr[t];
    const float4 sv = s.sorted_v[t];
    const float2 v0 = make_float2(sv.x, sv.y);
    // tiles of 32 agents, chunk-major: chunk c of agent t at
    // ((t / 32) * 8 + c) * 32 + t % 32, so each store (and the network's
    // loads) covers 512 contiguous bytes per 32 lanes
    uint4* out = s.feat + ((size_t)(t >> 5) * 8) * 32 + (t & 31);
    if (s.dead_s[t]) {
#pragma unroll
        for (int c = 0; c < 8; ++c) out[c * 32] = make_uint4(0, 0, 0, 0);
        return;
    }
    const V2 position = mk(pr.x, pr.y);
    const V2 velocity = mk(v0.x, v0.y);
    const float rad = pr.z;
    const float maxSpeed = s.max_speed_from_pref ? sv.z : s.max_speed;
    const EnvDesc& E = env_of_sorted(s, sv);
    Near4 T;
    near4_init(T);
    if (E.n_obst_edges > 0) {
        const float range = s.time_horizon_obst * maxSpeed + rad;
        obstacle_query(E, position, range * range, [&](int e, float dd) {
            near4_insert(T, dd, e, -1);
            return true;
        });
    }
    agent_features<K>(s, E, t, position, velocity, rad, s.pref_vel_s[t], s.nbr_cnt[t], T, out, 32);
}

// neighbour lists per slot id (debug / parity): [id*k + j] = neighbour id
__global__ void k_export_nbrs(DevState s, int* out, int* out_cnt) {
    const uint32_t n_active = s.cell_start[s.n_keys];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int)n_active) return;
    int id = __float_as_int(s.sorted_pr[t].w);
    int cnt = s.nbr_cnt[t];
    for (int j = 0; j < s.k; ++j)
        out[(size_t)id * s.k + j] = j < cnt ? __float_as_int(s.sorted_pr[s.nbr[(size_t)j * s.n_slots + t]].w) : -1;
    out_cnt[id] = cnt;
}

// ------------------------------------------------------------- launchers
static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
static inline int round8(int n) { return (n + 7) & ~7; }

// From the last step's sorted outputs (s.sorted_in) the count and scatter
// passes cover its agents (the count is read on the device); from the slot
// arrays, every slot.  The launch shapes depend only on the capacity, so
// one captured graph serves each entry for any agent count.
hipError_t launch_grid_build(const DevState& s, uint32_t* bsum, hipStream_t st) {
    const uint32_t G = s.n_keys;
    const dim3 ga(cdiv(s.n_slots, 256)), b(256);
    if (s.sorted_in) hipLaunchKernelGGL(k_grid_count_sorted, ga, b, 0, st, s);
    else hipLaunchKernelGGL(k_grid_count, ga, b, 0, st, s);
    const uint32_t nb = (G + 1023u) / 1024u;
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(256), 0, st, s.cell_count, s.cell_start, bsum, G);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, bsum, nb, s.cell_start, G, s.acc + 13);
    hipLaunchKernelGGL(k_scan_add, dim3(cdiv(G, 256)), dim3(256), 0, st, s.cell_start, (const uint32_t*)bsum, G);
    if (s.sorted_in) hipLaunchKernelGGL(k_grid_scatter_sorted, ga, b, 0, st, s);
    else hipLaunchKernelGGL(k_grid_scatter, ga, b, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_mark_dead(const int* ids, int n, const int* pos_of, const float4* sorted_pr_old, uint8_t* dead_old,
                            const unsigned long long* acc, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_mark_dead, dim3(cdiv(n, 256)), dim3(256), 0, st, ids, n, pos_of, sorted_pr_old,
                                  dead_old, acc);
    return hipGetLastError();
}

hipError_t launch_unsort(const DevState& s, hipStream_t st) {
    hipLaunchKernelGGL(k_unsort, dim3(cdiv(s.n_slots, 256)), dim3(256), 0, st, s);
    return hipGetLastError();
}

// Only the positions of k_unsort, for a read-back of positions alone (a
// renderer's read after every step): 13 B read and 8 B written per agent
// instead of the whole state.
__global__ void k_unsort_pos(DevState s) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int)(uint32_t)s.acc[13]) return;         // the last step's active count
    if (s.dead_s_old[t]) return;                       // written when it was destroyed
    const int id = __float_as_int(s.sorted_pr_old[t].w);
    s.pos_in[id] = s.pos_s_old[t];
}

hipError_t launch_unsort_pos(const DevState& s, hipStream_t st) {
    hipLaunchKernelGGL(k_unsort_pos, dim3(cdiv(s.n_slots, 256)), dim3(256), 0, st, s);
    return hipGetLastError();
}

// Queue kernels run on a fixed grid: their queues hold a few percent of
// the agents, and a full-size grid would reserve their LDS for nothing.
constexpr int kQueueBlocks = 1024;
// The plain step's main pass keeps its constraints in registers, with 4
// obstacle slots (8 for OSLOTS 4 next to at most 10 agent constraints; 8
// next to 16 would spill); the learned and logging passes use the LDS tile.
template <int K, int OSLOTS>
static void launch_orca_main(const DevState& s, int mode, dim3 g, dim3 b, hipStream_t st) {
    constexpr int OB = (OSLOTS == 4 && K <= 10) ? 8 : 4;
    if (mode == kLearned) hipLaunchKernelGGL((k_orca<K, kLearned, OSLOTS>), g, b, 0, st, s);
    else if (mode == kLog) hipLaunchKernelGGL((k_orca<K, kLog, OSLOTS>), g, b, 0, st, s);
    else hipLaunchKernelGGL((k_orca_reg<K, OB>), g, b, 0, st, s);
}

// Whether the plain step's main ORCA pass runs inside the k-nearest kernel
// (k_knn_orca): the launch shape's orca_obst_slots 2 selects it for k 5 and
// 10, and the autotuner weighs it on the live workload.
bool orca_fused(int k, int mode, int oslots) { return mode == kPlain && (k == 5 || k == 10) && oslots == 2; }

// Obstacle constraints held in registers by k_lp3d_reg; its persistent
// grid is kLp3dRegBlocks workgroups of 256 per CU, as many as its 4-wave
// cap keeps resident, so the whole grid starts at once and each wavefront
// loops over its share of the queue.  At c2's jam (560k entries) that took
// the queue passes from 0.153 to 0.147 ms; 2, 3, 6 and 8 workgroups per
// CU measured 0.208, 0.165, 0.175 and 0.153 ms, and handing the chunks out
// through an atomic counter 0.21-0.26 ms.
constexpr int kLp3dRegObst = 4;
constexpr int kLp3dRegBlocks = 4;

template <int K>
static void launch_orca_k(const DevState& s, int maxo, int mode, int oslots, int n_cu, hipStream_t st,
                          hipEvent_t mid) {
    dim3 g(round8(cdiv(s.n_slots, 64))), b(64), gq(kQueueBlocks);
    if (orca_fused(s.k, mode, oslots)) {
        // the main pass ran inside k_knn_orca
    } else if (oslots <= 1) launch_orca_main<K, 1>(s, mode, g, b, st);
    else if (oslots <= 2) launch_orca_main<K, 2>(s, mode, g, b, st);
    else launch_orca_main<K, 4>(s, mode, g, b, st);
    if (mid) (void)hipEventRecord(mid, st);
    const dim3 gr(std::max(1, n_cu) * kLp3dRegBlocks), br(256);
    if (maxo <= 16) {
        hipLaunchKernelGGL((k_orca_ovf<K, 16>), gq, b, 0, st, s);
        hipLaunchKernelGGL((k_lp3d_reg<K, kLp3dRegObst>), gr, br, 0, st, s);
        hipLaunchKernelGGL((k_lp3d<K, 16>), gq, b, 0, st, s);
    } else if (maxo <= 32) {
        hipLaunchKernelGGL((k_orca_ovf<K, 32>), gq, b, 0, st, s);
        hipLaunchKernelGGL((k_lp3d_reg<K, kLp3dRegObst>), gr, br, 0, st, s);
        hipLaunchKernelGGL((k_lp3d<K, 32>), gq, b, 0, st, s);
    } else if (maxo <= 64) {
        hipLaunchKernelGGL((k_orca_ovf<K, 64>), gq, b, 0, st, s);
        hipLaunchKernelGGL((k_lp3d_reg<K, kLp3dRegObst>), gr, br, 0, st, s);
        hipLaunchKernelGGL((k_lp3d<K, 64>), gq, b, 0, st, s);
    } else {
        // one agent per wavefront, the tile sized for maxo in dynamic LDS
        hipLaunchKernelGGL((k_orca_ovf_wide<K>), gq, b, (size_t)(maxo + K) * sizeof(float4), st, s, maxo);
        hipLaunchKernelGGL((k_lp3d_reg<K, kLp3dRegObst>), gr, br, 0, st, s);
        hipLaunchKernelGGL((k_lp3d_wide<K>), gq, b, (size_t)(maxo + 2 * K) * sizeof(float4), st, s, maxo);
    }
}

// `mid`, when not null, is recorded between the attraction pass and the
// k-nearest pass.
hipError_t launch_attr_knn(const DevState& s, int block, int mode, int oslots, hipStream_t st, hipEvent_t mid) {
    hipLaunchKernelGGL(k_attr, dim3(round8(cdiv(s.n_slots, 256))), dim3(256), 0, st, s);
    if (mid) (void)hipEventRecord(mid, st);
    if (block == 16 && orca_fused(s.k, mode, oslots)) {
        // 16 agents per wavefront of 64 threads
        const dim3 gs(round8(cdiv(s.n_slots, 16))), bs(64);
        const size_t lds64 = (size_t)(kKnnBuf + 2 * kKnnRanges) * 64 * sizeof(int);
        if (s.k == 5) hipLaunchKernelGGL((k_knn_orca_sparse<5>), gs, bs, lds64, st, s);
        else hipLaunchKernelGGL((k_knn_orca_sparse<10>), gs, bs, lds64, st, s);
        return hipGetLastError();
    }
    if (block != 64 && block != 128) block = block == 16 ? 64 : 256;
    dim3 g(round8(cdiv(s.n_slots, block))), b(block);
    const size_t lds = (size_t)(kKnnBuf + 2 * kKnnRanges) * block * sizeof(int);
    if (orca_fused(s.k, mode, oslots)) {
        if (s.k == 5) hipLaunchKernelGGL((k_knn_orca<5>), g, b, lds, st, s);
        else hipLaunchKernelGGL((k_knn_orca<10>), g, b, lds, st, s);
        return hipGetLastError();
    }
    switch (s.k) {
#define ECM_KNN_CASE(k) case k: hipLaunchKernelGGL((k_knn<k>), g, b, lds, st, s); break;
        ECM_KNN_CASE(1) ECM_KNN_CASE(2) ECM_KNN_CASE(3) ECM_KNN_CASE(4) ECM_KNN_CASE(5) ECM_KNN_CASE(6)
        ECM_KNN_CASE(7) ECM_KNN_CASE(8) ECM_KNN_CASE(9) ECM_KNN_CASE(10) ECM_KNN_CASE(11) ECM_KNN_CASE(12)
        ECM_KNN_CASE(13) ECM_KNN_CASE(14) ECM_KNN_CASE(15) ECM_KNN_CASE(16)
#undef ECM_KNN_CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// mode: 0 plain, 1 plain + sample log, 2 learned proposal (needs the
// predicted masks of launch_features + the proposal network).  `mid`, when
// not null, is recorded between the main pass and the queue passes.
hipError_t launch_orca(const DevState& s, int maxo, int mode, int oslots, int n_cu, hipStream_t st, hipEvent_t mid) {
    if (s.k <= 5) launch_orca_k<5>(s, maxo, mode, oslots, n_cu, st, mid);
    else if (s.k <= 10) launch_orca_k<10>(s, maxo, mode, oslots, n_cu, st, mid);
    else launch_orca_k<16>(s, maxo, mode, oslots, n_cu, st, mid);
    return hipGetLastError();
}

hipError_t launch_features(const DevState& s, hipStream_t st) {
    dim3 g(round8(cdiv(s.n_slots, 256))), b(256);
    if (s.k <= 5) hipLaunchKernelGGL((k_features<5>), g, b, 0, st, s);
    else if (s.k <= 10) hipLaunchKernelGGL((k_features<10>), g, b, 0, st, s);
    else hipLaunchKernelGGL((k_features<16>), g, b, 0, st, s);
    return hipGetLastError();
}

hipError_t launch_export_nbrs(const DevState& s, int* out, int* out_cnt, hipStream_t st) {
    hipLaunchKernelGGL(k_export_nbrs, dim3(cdiv(s.n_slots, 256)), dim3(256), 0, st, s, out, out_cnt);
    return hipGetLastError();
}

}  // namespace ecm

namespace ecm {

// SpawnAgent's component writes (Simulator.cpp:175-197) for a batch.
__global__ void k_spawn_scatter(int n, const int* ids, const float2* start, const float* rad, const float* spd,
                                const int* off, const int* cnt, float2* pos, float2* vel, float* radius,
                                float* pref_speed, float2* pref_vel, uint8_t* active, int* path_off,
                                int* path_cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int id = ids[i];
    if (id < 0) return;
    pos[id] = start[i];
    vel[id] = make_float2(0.0f, 0.0f);
    pref_vel[id] = make_float2(0.0f, 0.0f);
    radius[id] = rad[i];
    pref_speed[id] = spd[i];
    active[id] = 1;
    path_off[id] = off[i];
    path_cnt[id] = cnt[i];
}

// UpdatePath results for a batch of agents (Simulator.cpp:114-123), read
// directly from the host's mapped staging (one kernel instead of two copies
// and a scatter): the npts new points go to the pool at pool_dst, and each
// agent's offset and count to the slot arrays and to the copy the next step
// carries in Morton order (path_s, at the agent's position in the last
// step's sorted outputs).  stage holds n ids, n offsets, n counts, then n
// sorted positions.  Grid-stride loops: any grid covers both lists.
__global__ void k_path_upload(int n, const int* stage, int npts, const float2* pts, float2* pool_dst, int* path_off,
                              int* path_cnt, int2* path_s) {
    const int stride = gridDim.x * blockDim.x;
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = i0; i < npts; i += stride) pool_dst[i] = pts[i];
    for (int i = i0; i < n; i += stride) {
        const int id = stage[i], off = stage[n + i], cnt = stage[2 * n + i], at = stage[3 * n + i];
        path_off[id] = off;
        path_cnt[id] = cnt;
        path_s[at] = make_int2(off, cnt);
    }
}

hipError_t launch_path_upload(int n, const int* stage, int npts, const float2* pts, float2* pool_dst, int* path_off,
                              int* path_cnt, int2* path_s, hipStream_t st) {
    if (n <= 0 && npts <= 0) return hipSuccess;
    const int blocks = std::min(cdiv(std::max(n, npts), 256), 64);
    hipLaunchKernelGGL(k_path_upload, dim3(blocks), dim3(256), 0, st, n, stage, npts, pts, pool_dst, path_off,
                       path_cnt, path_s);
    return hipGetLastError();
}

// The last step's sorted copy of the paths, re-read from the slot arrays
// after the host rewrote them (set_path, pool compaction).
__global__ void k_path_resort(const unsigned long long* acc, const float4* sorted_pr_old, const int* path_off,
                              const int* path_cnt, int2* path_s_old) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int)(uint32_t)acc[13]) return;           // the last step's active count
    const int id = __float_as_int(sorted_pr_old[t].w);
    path_s_old[t] = make_int2(path_off[id], path_cnt[id]);
}

// The live paths packed into `out` in the last step's Morton order, so that
// neighbouring agents read neighbouring points: each wavefront takes one
// contiguous range (one atomic on *ctr, zero at launch) for its lanes'
// paths in lane order.  Agents destroyed in that step or removed since
// (dead_s_old) are dropped.  The paths of live agents are disjoint ranges of
// the pool, so the packed total is at most the pool's length `cap`; a lane
// whose range would pass it is counted in ctr[1] instead (never expected).
__global__ void k_path_repack(const unsigned long long* acc, const float4* sorted_pr_old, const uint8_t* dead_s_old,
                              int2* path_s_old, int* path_off, const float2* pool, float2* out, uint32_t cap,
                              uint32_t* ctr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < (int)(uint32_t)acc[13] && !dead_s_old[t];
    const int2 pc = live ? path_s_old[t] : make_int2(0, 0);
    const uint32_t cnt = live ? (uint32_t)pc.y : 0u;
    // exclusive prefix of cnt over the wavefront, then one atomic for its total
    const int lane = (int)__lane_id();
    uint32_t pre = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(pre, d);
        if (lane >= d) pre += o;
    }
    const uint32_t total = __shfl(pre, 63);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(ctr, total);
    base = __shfl(base, 0) + pre - cnt;
    if (!live || cnt == 0u) return;
    if ((uint64_t)base + cnt > cap || pc.x < 0 || (uint64_t)pc.x + cnt > cap) {
        atomicAdd(ctr + 1, 1u);                        // counted as skipped
        return;
    }
    for (uint32_t q = 0; q < cnt; ++q) out[base + q] = pool[(uint32_t)pc.x + q];
    path_s_old[t] = make_int2((int)base, pc.y);
    path_off[__float_as_int(sorted_pr_old[t].w)] = (int)base;
}

hipError_t launch_path_repack(int n_slots, const unsigned long long* acc, const float4* sorted_pr_old,
                              const uint8_t* dead_s_old, int2* path_s_old, int* path_off, const float2* pool,
                              float2* out, uint32_t cap, uint32_t* ctr, hipStream_t st) {
    if (n_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_path_repack, dim3(cdiv(n_slots, 256)), dim3(256), 0, st, acc, sorted_pr_old, dead_s_old,
                       path_s_old, path_off, pool, out, cap, ctr);
    return hipGetLastError();
}

hipError_t launch_path_resort(int n_slots, const unsigned long long* acc, const float4* sorted_pr_old,
                              const int* path_off, const int* path_cnt, int2* path_s_old, hipStream_t st) {
    if (n_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_path_resort, dim3(cdiv(n_slots, 256)), dim3(256), 0, st, acc, sorted_pr_old, path_off,
                       path_cnt, path_s_old);
    return hipGetLastError();
}

hipError_t launch_spawn_scatter(int n, const int* ids, const float2* start, const float* rad, const float* spd,
                                const int* off, const int* cnt, float2* pos, float2* vel, float* radius,
                                float* pref_speed, float2* pref_vel, uint8_t* active, int* path_off,
                                int* path_cnt, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spawn_scatter, dim3(cdiv(n, 256)), dim3(256), 0, st, n, ids, start, rad, spd, off, cnt,
                       pos, vel, radius, pref_speed, pref_vel, active, path_off, path_cnt);
    return hipGetLastError();
}

}  // namespace ecm
