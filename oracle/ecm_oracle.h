/*
 * ecm_oracle.h -- CPU oracle for the per-timestep crowd update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (ECMAgentSimulator/Simulator.cpp, ORCA.cpp, KDTree.cpp,
 * IRMPathFollower.cpp; ECMGenerator/ECM.cpp, ECMCellCollection.cpp,
 * UtilityFunctions.cpp, ECMDataTypes.h).  It is used only by the tests, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker, never as the product path.  The product (libecmcrowd.so) neither
 * links nor loads it.
 *
 * Pinning (what checks this restatement, see tests/test_oracle_*.py):
 *   - geometry primitives, IRM, retraction: closed forms on hand-built ECMs
 *     and FP64 re-derivations (tests/golden/make_golden.py);
 *   - exact k-nearest: brute force + invariants (k smallest d2 > 1e-4,
 *     ties by id);
 *   - reference KD-tree mode: hand-traced known-answer cases that expose
 *     each documented defect of KDTree.cpp:85-202;
 *   - ORCA constraints: closed-form head-on / oblique / colliding cases;
 *   - 2D LP: FP64 active-set re-solve (feasibility + KKT);
 *   - 3D LP: invariant (max penetration non-increasing) + FP64 restatement;
 *   - integration: closed forms.
 * Parity with MSVC's atanf/cosf/sinf and with Boost's ECM numbering is not
 * pinnable (the reference cannot be built here); the oracle uses its own
 * trig routines (Cody-Waite reduction + Taylor cores), pinned against libm.
 */
#ifndef ECM_ORACLE_H
#define ECM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* knn_mode */
#define ORACLE_KNN_EXACT 0      /* k smallest d2 > EPSILON, ordered (d2, id) */
#define ORACLE_KNN_REFERENCE 1  /* KDTree.cpp:22-202 bug-for-bug, shared cache */

typedef struct {
    int   k;                   /* ORCA neighbours (Simulator.cpp:55)          */
    float time_horizon;        /* m_lookAheadAgent (ORCA.h:102)               */
    float time_horizon_obst;   /* m_lookAheadObstacle (ORCA.h:103)            */
    float max_speed;           /* used when max_speed_from_pref == 0          */
    int   max_speed_from_pref; /* 1: maxSpeed = own preferred speed (Simulator.cpp:671-674) */
    float dt;                  /* m_SimStepTime                               */
    float mass;                /* Simulator.cpp:622                           */
    float arrival_radius;      /* Simulator.cpp:543 (20 m)                    */
    float delete_radius;       /* Simulator.cpp:542 (2 m)                     */
    int   knn_mode;
} oracle_params;

typedef struct {
    int n;                          /* number of slots                         */
    const unsigned char* active;    /* [n]                                     */
    const float* pos;               /* [2n]                                    */
    const float* vel;               /* [2n]                                    */
    const float* radius;            /* [n]  ClearanceComponent                 */
    const float* pref_speed;        /* [n]  SpeedComponent                     */
    const float* attraction;        /* [2n] carried over between steps         */
    const int* path_off;            /* [n]  offset into path_pts (points)      */
    const int* path_cnt;            /* [n]  numPoints                          */
    const float* path_pts;          /* [2 * total points]                      */
} oracle_state;

/* Obstacles as rings of vertices, CCW, in (obstacle, vertex) order
 * (Environment.cpp:285-315, ECMDataTypes.cpp:23-61). */
typedef struct {
    int n_rings;
    const int* ring_off;            /* [n_rings + 1]                           */
    const float* xy;                /* [2 * ring_off[n_rings]]                 */
} oracle_obstacles;

/* ECM graph: vertices and edges with two half-edges each (ECM.h:20-50). */
typedef struct {
    int n_vert;
    const float* vert_xy;           /* [2 n_vert]                              */
    int n_edge;
    const int* he_target;           /* [2 n_edge]: he0.v_target, he1.v_target  */
    const float* he_pts;            /* [8 n_edge]: he0.cl, he0.cr, he1.cl, he1.cr */
} oracle_ecm;

typedef struct {
    float* pos;                     /* [2n] new position                       */
    float* vel;                     /* [2n] new velocity                       */
    float* attraction;              /* [2n]                                    */
    float* pref_vel;                /* [2n]                                    */
    unsigned char* active;          /* [n]  after goal-destroy                 */
    unsigned char* replan;          /* [n]  IRM failed -> UpdatePath requested */
    float* v_orca;                  /* [2n] ORCA output velocity (may be NULL) */
    int* nbr;                       /* [n*k] neighbour slots (may be NULL)     */
    int* nbr_cnt;                   /* [n]                                     */
    int* diag;                      /* [n*4] cell, n_obst_constraints, n_constraints, lp_fail (may be NULL) */
    unsigned char* used_3d;         /* [n] (may be NULL)                       */
} oracle_out;

/* One Simulator::Update step (Simulator.cpp:314-323) restricted to agent
 * slots in `subset` (all active agents when subset == NULL).  Every agent
 * reads only step-start state, so any subset is exact in EXACT mode.  In
 * REFERENCE mode the neighbour cache is shared in slot order, so a subset run
 * is meant for timing only.  Outputs of slots not processed are untouched.
 * Returns 0 on success, negative on bad input. */
int oracle_step(const oracle_params* prm, const oracle_state* st,
                const oracle_obstacles* obs, const oracle_ecm* ecm,
                const int* subset, int n_subset, oracle_out* out);

/* --- individually exposed pieces, for known-answer tests --------------- */

/* ECMCellCollection::PointLocationQueryLinear (ECMCellCollection.cpp:57-90):
 * index of the first cell (2 per edge, edge order) containing (x, y), or -1. */
int oracle_locate_cell(const oracle_ecm* ecm, float x, float y);

/* ECM::RetractPoint (ECM.cpp:20-96).  Returns 1 on success. */
int oracle_retract(const oracle_ecm* ecm, float x, float y,
                   float* out_xy, int* out_edge);

/* IRMPathFollower::FindAttractionPoint (IRMPathFollower.cpp:14-115). */
int oracle_attraction(const oracle_ecm* ecm, float x, float y,
                      const float* path_xy, int n_points, float* out_xy);

/* ORCA::RandomizedLP (ORCA.cpp:428-587).  constraints: [4 n] = (nx, ny, px, py). */
int oracle_lp2d(const float* constraints, int n, const float* opt,
                float max_speed, int use_dir_opt, float* out_vel);

/* ORCA::RandomizedLP3D (ORCA.cpp:592-669). out_vel is in/out. */
/* Obstacle half-plane (ORCA.cpp:70-333) of one edge (global vertex index
 * `edge` -> its next vertex) for one agent: returns the number of
 * constraints emitted (0 or 1), the constraint in out4 (nx, ny, px, py) and
 * in *branches the ORACLE_OB_* bits of the branches taken. */
#define ORACLE_OB_CASE1 (1u << 0)          /* collision with L, L convex (:92-99)         */
#define ORACLE_OB_CASE1_SKIP (1u << 1)     /* collision with L, L concave (:102-104)      */
#define ORACLE_OB_CASE2 (1u << 2)          /* collision with R, emitted (:108-117)        */
#define ORACLE_OB_CASE2_SKIP (1u << 3)     /* collision with R, nothing (:119-120)        */
#define ORACLE_OB_CASE3 (1u << 4)          /* collision with the segment (:124-135)       */
#define ORACLE_OB_VERTEX_L (1u << 5)       /* oblique view of L alone (:146-169)          */
#define ORACLE_OB_VERTEX_L_SKIP (1u << 6)  /* ... L concave (:149-151)                    */
#define ORACLE_OB_VERTEX_R (1u << 7)       /* oblique view of R alone (:171-183)          */
#define ORACLE_OB_VERTEX_R_SKIP (1u << 8)  /* ... R concave (:173-175)                    */
#define ORACLE_OB_CONCAVE_L (1u << 9)      /* left leg along the segment, L concave (:193-199) */
#define ORACLE_OB_CONCAVE_R (1u << 10)     /* right leg along the segment, R concave (:206-211) */
#define ORACLE_OB_FOREIGN_L (1u << 11)     /* left leg clamped to the left neighbour (:224-228) */
#define ORACLE_OB_FOREIGN_R (1u << 12)     /* right leg clamped to the right neighbour (:233-237) */
#define ORACLE_OB_CIRCLE_L (1u << 13)      /* left cutoff circle (:259-268)               */
#define ORACLE_OB_CIRCLE_R (1u << 14)      /* right cutoff circle (:270-280)              */
#define ORACLE_OB_CUTOFF_LINE (1u << 15)   /* cutoff line (:289-301)                      */
#define ORACLE_OB_LEG_L (1u << 16)         /* left leg (:303-317)                         */
#define ORACLE_OB_LEG_L_SKIP (1u << 17)    /* ... foreign, nothing emitted (:305-307)     */
#define ORACLE_OB_LEG_R (1u << 18)         /* right leg (:319-332)                        */
#define ORACLE_OB_LEG_R_SKIP (1u << 19)    /* ... foreign, nothing emitted (:320-322)     */
int oracle_obstacle_halfplane(const oracle_params* prm, const oracle_state* st, const oracle_obstacles* obs,
                              int agent, int edge, float* out4, unsigned* branches);
/* FindNearestObstacles (Simulator.cpp:259-292): edges kept, in scan order. */
int oracle_nearest_obstacles(const oracle_obstacles* obs, float x, float y, float range, int* out, int cap);

void oracle_lp3d(const float* constraints, int n, int n_obst, float max_speed,
                 int failed_index, float* out_vel);

/* Per-agent constraint generation (ORCA.cpp:60-424) for one agent of a state,
 * with the given neighbour list.  Writes up to `cap` constraints, returns the
 * total count (>= 0) and sets *n_obst. */
int oracle_constraints(const oracle_params* prm, const oracle_state* st,
                       const oracle_obstacles* obs, int agent,
                       const int* nbrs, int n_nbrs,
                       float* constraints, int cap, int* n_obst);

/* The oracle's own trig (Cody-Waite reduction, Taylor cores). */
float oracle_atanf(float x);
float oracle_sinf(float x);
float oracle_cosf(float x);

/* Timing helper for the CPU baseline: build the reference KD-tree over all
 * active agents and return its node count. */
int oracle_kdtree_build_only(const oracle_state* st);

/* Proposal-network inputs (64 floats) and label word of one agent, for
 * the LP velocity v; see ecm_oracle.c.  Returns the constraint count. */
int oracle_proposal_sample(const oracle_params* prm, const oracle_state* st, const oracle_obstacles* obs,
                           int agent, const int* nbrs, int n_nbrs, const float* pref_vel, const float* v,
                           float* feat, unsigned* label);

#ifdef __cplusplus
}
#endif
#endif
