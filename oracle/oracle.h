/* CPU oracle for the batched PRRT extension path — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, on the CPU and in plain sequential C++, the
 * reference's semantics for the hot path so that the HIP path can be checked
 * against it.  Only tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it; the product library never does.
 *
 * Conventions: states are rows of `dim` doubles (AoS).  SE(3) states are
 * [qx, qy, qz, qw, tx, ty, tz] (quaternion first, as the reference's
 * std::tuple<Quaternion, Vector3>).  Index tie-break everywhere is "ascending
 * (distance, index)".
 *
 * Parity pinning of each function is listed in the header comment of the
 * .cpp that defines it.
 */
#ifndef MPTG_ORACLE_H
#define MPTG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_METRIC_L2 = 0, ORC_METRIC_L1 = 1, ORC_METRIC_SE3 = 2 };
enum { ORC_SCENE_NONE = 0, ORC_SCENE_BOXES = 1, ORC_SCENE_GRID = 2, ORC_SCENE_ARM = 3, ORC_SCENE_MESH = 4,
       ORC_SCENE_SHAPES = 5 };
enum { ORC_LINK_DMV = 0, ORC_LINK_BISECT = 1, ORC_LINK_BISECT_L2 = 2 };

typedef struct {
    int32_t metric;
    int32_t dim;
    double w_so3, w_l2;      /* SE(3) weights */
    int32_t kind;            /* ORC_SCENE_* */
    int32_t link_mode;       /* ORC_LINK_* */
    double link_param;       /* DMV step size, or bisection L-inf tolerance */
    const double* boxes;     /* nboxes x [x0, y0, x1, y1] closed boxes */
    int32_t nboxes;
    const uint8_t* grid;     /* height x width, row-major, nonzero = obstacle */
    int32_t width, height;
    const double* lengths;   /* arm link lengths */
    int32_t nlinks;
    double robot_radius;
    const double* disks;     /* ndisks x [cx, cy, r] */
    int32_t ndisks;
    const float* robot_tris; /* nrobot x 9 (robot frame) */
    int32_t nrobot;
    const float* env_tris;   /* nenv x 9 (world frame) */
    int32_t nenv;
    const void* mesh_index;  /* from orc_mesh_index_build */
    const double* circles;   /* ncircles x [cx, cy, r] (shapes scene; rectangles are `boxes`) */
    int32_t ncircles;
} orc_scene;

/* Uniform-grid index over the environment triangles of a mesh scene (the
 * oracle's own culling, independent of the device BVH).  Build once per
 * scene and store in sc->mesh_index; free with orc_mesh_index_free. */
void* orc_mesh_index_build(const orc_scene* sc);
void orc_mesh_index_free(void* index);
/* Closed triangle-triangle intersection (the project's FP32 predicate). */
int32_t orc_tri_tri(const float* p, const float* q);
/* Posed robot triangles of SE(3) state q (nrobot x 9 floats). */
void orc_pose_robot(const orc_scene* sc, const double* q, float* out);

/* --- shared project math (same code as the device) ---------------------- */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_sincos(const double* x, int64_t n, double* s, double* c);
void orc_acos(const double* x, int64_t n, double* y);
double orc_uniform_at(uint64_t seed, uint64_t index, uint32_t purpose, uint32_t j);

/* --- spaces ---------------------------------------------------------------- */
double orc_distance(const orc_scene* sc, const double* a, const double* b);
void orc_interpolate(const orc_scene* sc, const double* a, const double* b, double t, double* out);

/* --- nearest neighbours (brute force, exact FP64) -------------------------- */
/* k nearest of each query among pts[0..n), radius r (<= r kept; r = +inf for
 * plain kNN).  Outputs idx[nq*k], dist[nq*k] sorted by (dist, idx); slots
 * beyond the found count get idx -1 and dist +inf. */
void orc_knn(const orc_scene* sc, const double* pts, int64_t n, const double* q, int64_t nq,
             int32_t k, double r, int32_t* idx, double* dist, int32_t nthreads);

/* --- motion validation ----------------------------------------------------- */
int32_t orc_valid(const orc_scene* sc, const double* q);
/* DiscreteMotionValidator check order: indices (1..steps-1) in the order the
 * 256-slot BFS bisection ring visits them.  Returns count (= steps-1). */
int64_t orc_dmv_order(int64_t steps, int64_t* out, int64_t cap);
/* Waypoints of link(a,b) in the reference's check order, excluding the
 * endpoint checks.  Returns count, writes min(count, cap) states. */
int64_t orc_link_waypoints(const orc_scene* sc, const double* a, const double* b, double* out,
                           int64_t cap);
/* link(a,b): verdict and number of valid() calls made in reference order.
 * endpoints_known_valid: the caller vouches that the endpoint checks the
 * reference repeats inside link() pass (they are still counted). */
int32_t orc_link(const orc_scene* sc, const double* a, const double* b, int32_t endpoints_known_valid,
                 uint32_t* checks_done);
void orc_link_batch(const orc_scene* sc, const double* from, const double* to, int64_t n,
                    int32_t endpoints_known_valid, uint8_t* verdict, uint32_t* checks_done,
                    int32_t nthreads);
/* Work of link() over n arm edges in the reference order (bisectLink,
 * demo/link_manipulator_scenario.hpp:99-138): out[0] valid() calls evaluated
 * (endpoint checks of known-valid edges are not evaluated), out[1] links
 * whose forward kinematics ran, out[2] segment-disk tests (every disk until
 * the first hit), out[3] those of out[2] whose disk centre lies within
 * (link length + r + R) of the link's start joint.  -1 unless an arm scene
 * with bisection links. */
int32_t orc_arm_link_work(const orc_scene* sc, const double* from, const double* to, int64_t n,
                          int32_t known_valid, int32_t nthreads, uint64_t* out);
/* Number of valid() calls when the link's waypoints are checked in this
 * order, after the same endpoint checks: reference ordinals 0 .. lead-1 first,
 * then the other waypoints of bisection depth < top_levels by descending
 * score (ties: ascending reference ordinal), then the deeper ones in reference
 * order (top_levels < 0: every waypoint after the lead by score).  With
 * picks >= 0 only the first `picks` of the scored waypoints come before the
 * rest, and every other waypoint follows in reference order.  scores[] is
 * indexed by reference ordinal (pre-order of the bisection tree). */
uint32_t orc_checks_in_score_order(const orc_scene* sc, const double* a, const double* b,
                                   int32_t known_valid, const float* scores, int64_t nscores, int32_t lead,
                                   int32_t top_levels, int32_t picks);

/* --- batch-synchronous PRRT ------------------------------------------------ */
typedef struct {
    uint64_t seed;
    double range;        /* maxDistance (+inf: no steering) */
    double goal_bias;
    double goal_radius;
    int32_t batch;       /* samples per iteration; 1 == the sequential loop */
    int32_t _pad;
    const double* lo;    /* sampling bounds */
    const double* hi;
    const double* goal;
    /* a Scenario's own sampler (scenario_sampler.hpp:145-187), as a table:
       row sidx is sample index sidx's state, samples_ok[sidx] = 0 for a
       slot with no sample (skipped); NULL: the built-in box sampler */
    const double* samples;
    const uint8_t* samples_ok;
    int64_t nsamples;
} orc_prrt_params;

typedef struct {
    int64_t iterations;
    int64_t samples;
    int64_t nearest_calls;
    int64_t zero_dist;
    int64_t invalid_states;
    int64_t link_calls;
    int64_t link_ok;
    int64_t valid_calls;   /* valid() inside link() */
    int64_t goals;
} orc_prrt_stats;

/* Grows a tree from `start` until size >= target or max_iters iterations.
 * nodes[cap*dim], parents[cap]; returns the final size (or -1 if cap is too
 * small).  goal_nodes gets the ids of goal nodes (up to goal_cap). */
int64_t orc_prrt_run(const orc_scene* sc, const orc_prrt_params* p, const double* start,
                     int64_t target, int64_t max_iters, double* nodes, int32_t* parents, int64_t cap,
                     int32_t* goal_nodes, int64_t goal_cap, orc_prrt_stats* stats);

/* --- batch-synchronous PRRT* ----------------------------------------------- */
typedef struct {
    int64_t rewire_tests;    /* neighbours considered for rewiring */
    int64_t cp_links;        /* link() calls of the choose-parent walk */
    int64_t rw_links;        /* link() calls of the rewire pass */
    int64_t rewires;         /* neighbours moved under a new node */
    int64_t parent_changes;  /* survivors whose parent is not the near node */
    int64_t spec_cp;         /* walk entries a batched walk queues: those before the first
                                cost-bound or near-node stop */
    int64_t spec_rw;         /* unmarked pairs with pc + d < cost at the rewire phase's start */
    int64_t deferred;        /* pairs rewire-tested only because cost(new) fell in the phase */
} orc_star_stats;

/* Neighbour count: k_fixed if > 0, else ceil(factor*e*(1+1/dim)*ln(n+1)). */
int32_t orc_star_k(int32_t k_fixed, double factor, int32_t dim, int64_t n);
/* rewire_r_nearest's radius at tree size n for the sampling box [lo, hi]
 * (rrg_rewire_neighbors.hpp:99-128). */
double orc_star_radius(const orc_scene* sc, const double* lo, const double* hi, double factor, int64_t n);
/* Grows a PRRT* tree; costs[cap] receive the path costs.  k_fixed > 0: that
 * many nearest neighbours; 0: the reference's k formula; -1: every node within
 * the reference's r-nearest radius (rewire_r_nearest).  Returns the final
 * size, -1 if cap is too small, -2 on the reference's "bad distance" abort,
 * -3 if k exceeds 64. */
int64_t orc_prrt_star_run(const orc_scene* sc, const orc_prrt_params* p, int32_t k_fixed, double rewire_factor,
                          const double* start, int64_t target, int64_t max_iters, double* nodes,
                          int32_t* parents, double* costs, int64_t cap, int32_t* goal_nodes, int64_t goal_cap,
                          orc_prrt_stats* stats, orc_star_stats* star_stats, double* goal_cost0,
                          int32_t* goal_parent0);

/* Reference-style shared-memory worker pool (std::thread, one RNG per
 * worker, mt19937_64 + uniform_real_distribution, shared tree behind a
 * KD-tree with a reader/writer lock).  Nondeterministic by design, like the
 * reference; used only to time the CPU path. */
typedef struct {
    double seconds;
    int64_t size;
    int64_t samples;
    int64_t link_calls;
    int64_t valid_calls;
    int64_t goals;
} orc_pool_result;
int32_t orc_pool_run(const orc_scene* sc, const orc_prrt_params* p, const double* start, int64_t target,
                     double max_seconds, int32_t nthreads, orc_pool_result* out);
/* The same from a tree of n0 nodes (parents[0] = -1, parents[i] < i; NULL:
 * every node's parent is node 0), stopping at `target` nodes or max_seconds. */
int32_t orc_pool_run_tree(const orc_scene* sc, const orc_prrt_params* p, const double* nodes,
                          const int32_t* parents, int64_t n0, int64_t target, double max_seconds,
                          int32_t nthreads, orc_pool_result* out);

#ifdef __cplusplus
}
#endif
#endif
