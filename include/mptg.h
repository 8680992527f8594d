/* mptg — MI355X batched PRRT extension path, C ABI.
 *
 * The library replaces the two downcalls of the reference's PRRT/PRRT*
 * extension loop — the nigh nearest-neighbour structure and the Scenario's
 * valid()/link() motion validation — with batched gfx950 kernels, and adds a
 * device-resident batched iteration that runs sample -> nearest -> steer ->
 * valid -> link -> insert for a whole batch of samples per launch sequence.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative MPTG_E_* code; the
 *     message of the last failure is mptg_last_error(ctx) (per context;
 *     mptg_last_error(NULL) holds failures of mptg_create).  No exception
 *     crosses this boundary.
 *   - Pointers named h_* are host memory borrowed for the duration of the
 *     call (the call is synchronous).  Pointers named d_* are device memory
 *     on the context's device; such calls are asynchronous on the context's
 *     stream (mptg_stream) and the caller synchronises with mptg_sync.
 *   - States are rows of `dim` doubles.  SE(3) states are
 *     [qx, qy, qz, qw, tx, ty, tz].
 *   - A context owns one HIP stream and all device memory it allocates.  It is
 *     not thread-safe: one host thread drives a context (the reference instead
 *     runs N worker threads against one shared tree, prrt.hpp:325-346).
 *
 * Reference paths below are relative to the reference repository root.
 */
#ifndef MPTG_H
#define MPTG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPTG_OK 0
#define MPTG_E_ARG (-1)     /* invalid argument */
#define MPTG_E_HIP (-2)     /* HIP runtime failure */
#define MPTG_E_STATE (-3)   /* call out of order (no space/tree/scene yet) */
#define MPTG_E_RANGE (-4)   /* size beyond a documented limit */
#define MPTG_E_EMPTY (-5)   /* "there are no valid initial states" */

enum mptg_metric {
    MPTG_METRIC_L2 = 0,   /* L2Space<double, D>       src/mpt/lp_space.hpp */
    MPTG_METRIC_L1 = 1,   /* L1Space<double, D>       demo/link_manipulator_scenario.hpp:59 */
    MPTG_METRIC_SE3 = 2   /* SE3Space<double, w_so3>  demo/se3_rigid_body_scenario.hpp:250 */
};

enum mptg_scene_kind {
    MPTG_SCENE_NONE = 0,
    MPTG_SCENE_BOXES = 1, /* closed 2-D boxes, Rect::pointIsValid   demo/shape_hierarchy.hpp:177-182 */
    MPTG_SCENE_GRID = 2,  /* occupancy grid, PNG2dScenario::valid    demo/png_2d_scenario.hpp:104-110 */
    MPTG_SCENE_ARM = 3,   /* planar arm vs disks                     demo/link_manipulator_scenario.hpp:99-116 */
    MPTG_SCENE_MESH = 4,  /* rigid body vs triangle mesh (SE3 space) demo/se3_rigid_body_scenario.hpp:288-312 */
    MPTG_SCENE_SHAPES = 5,/* 2-D point among circles and rectangles   demo/holonomic_2d_point_scenario.hpp:95-113 */
    MPTG_SCENE_HOST = 6   /* the caller's own valid()/link() on the host (mptg_scene_host) */
};

enum mptg_link_mode {
    MPTG_LINK_DMV = 0,       /* DiscreteMotionValidator(space, step)   src/mpt/discrete_motion_validator.hpp:71-130 */
    MPTG_LINK_BISECT = 1,    /* recursive midpoints to L-inf < tol     demo/link_manipulator_scenario.hpp:118-138 */
    MPTG_LINK_BISECT_L2 = 2  /* recursive midpoints to squared L2 length < tol^2 (tol = 1 in the demos):
                                Rect::segmentIsValid / bisectSegment  demo/shape_hierarchy.hpp:184-203,
                                PNG2dScenario::link / validSegment    demo/png_2d_scenario.hpp:112-117,152-165;
                                for MPTG_SCENE_SHAPES the circles first take the closed form
                                Circle::segmentIsValid                demo/shape_hierarchy.hpp:228-231,259-270
                                and the midpoints are tested against the rectangles */
};

enum mptg_check_order {
    MPTG_ORDER_REFERENCE = 0, /* the reference's waypoint order */
    MPTG_ORDER_SURROGATE = 1  /* the 4 waypoints of highest surrogate score among the 15 of the
                                 bisection tree's top 4 levels (ties: reference order) first,
                                 then every other waypoint in reference order */
};

typedef struct mptg_ctx mptg_ctx;

/* ---- lifecycle --------------------------------------------------------- */
int mptg_create(int device, mptg_ctx** out);
void mptg_destroy(mptg_ctx* ctx);
const char* mptg_last_error(const mptg_ctx* ctx);
int mptg_sync(mptg_ctx* ctx);
void* mptg_stream(mptg_ctx* ctx); /* the context's hipStream_t */

/* ---- device memory and timing on the context's device/stream (plumbing for
 *      callers without their own HIP runtime; no reference counterpart).
 *      upload/download are synchronous; memset is asynchronous on the stream.
 *      timer_stop returns the milliseconds between the two events recorded
 *      on the context's stream by timer_start and timer_stop. */
int mptg_dev_alloc(mptg_ctx* ctx, uint64_t bytes, void** d_out);
int mptg_dev_free(mptg_ctx* ctx, void* d_ptr);
int mptg_dev_upload(mptg_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes);
int mptg_dev_download(mptg_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes);
int mptg_dev_memset(mptg_ctx* ctx, void* d_dst, int value, uint64_t bytes);
int mptg_timer_start(mptg_ctx* ctx);
int mptg_timer_stop(mptg_ctx* ctx, float* ms);

/* ---- space: replaces the Space given to nigh::Nigh(space)
 *      src/mpt/impl/prrt/prrt.hpp:122,151.  w0/w1 = SO(3)/R^3 weights for
 *      MPTG_METRIC_SE3 (ignored otherwise). dim: 2..32 (L1/L2), 7 (SE3). */
int mptg_space(mptg_ctx* ctx, int metric, int dim, double w0, double w1);

/* ---- tree: replaces nigh::Nigh::insert / size
 *      src/mpt/impl/prrt/prrt.hpp:177-187,447.  Node ids are insertion order. */
int mptg_tree_reserve(mptg_ctx* ctx, int64_t capacity);
int mptg_tree_clear(mptg_ctx* ctx);
int mptg_tree_insert(mptg_ctx* ctx, const double* h_states, const int32_t* h_parents, int64_t n);
int64_t mptg_tree_size(const mptg_ctx* ctx);
int mptg_tree_read(mptg_ctx* ctx, int64_t first, int64_t n, double* h_states, int32_t* h_parents);

/* ---- nearest neighbours: replaces nigh::Nigh::nearest(q)
 *      (src/mpt/impl/prrt/prrt.hpp:406-409) and nearest(nbh, q, k, r)
 *      (src/mpt/impl/rrg_rewire_neighbors.hpp:66,127).
 *      For each query: the k tree nodes with the smallest distance <= radius,
 *      ascending by (distance, node id); distances are the metric's FP64
 *      value.  Missing slots: idx -1, dist +inf.  k in 1..64; radius +inf for
 *      plain kNN (mptg_rnn below returns every node within a radius). */
int mptg_knn(mptg_ctx* ctx, const double* h_q, int64_t nq, int k, double radius,
             int32_t* h_idx, double* h_dist);
/* Every tree node within `radius` of each query, ascending by (distance,
 * node id): nigh's nearest(nbh, q, SIZE_MAX, r), the call behind
 * rewire_r_nearest (src/mpt/impl/rrg_rewire_neighbors.hpp:124-128), with no
 * limit on the list length.  Query i's neighbours are entries
 * [h_offsets[i], h_offsets[i + 1]) of h_idx / h_dist; h_offsets holds nq + 1
 * entries.  When the lists need more than `cap` entries the call fills
 * h_offsets only and returns MPTG_E_RANGE (call again with h_offsets[nq]
 * entries).  At most 2^24 queries per call. */
int mptg_rnn(mptg_ctx* ctx, const double* h_q, int64_t nq, double radius, int64_t cap, int64_t* h_offsets,
             int32_t* h_idx, double* h_dist);
/* Device buffers, on the context stream.  On trees of 2^16 nodes or more the
 * uniform grid (L1/L2 in 2..8 dimensions, knn_grid on) and the select path
 * (k > 1, knn_select on) read counts of queries left for the scans back to
 * the host, so such calls return with their work done; other calls only
 * enqueue. */
int mptg_knn_dev(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius,
                 int32_t* d_idx, double* d_dist);

/* ---- scene: the Scenario's valid() as data.  Replaces the per-worker
 *      Scenario copy (src/mpt/impl/prrt/prrt.hpp:326,343). */
int mptg_scene_boxes(mptg_ctx* ctx, const double* h_boxes /* n x [x0,y0,x1,y1] */, int n);
int mptg_scene_grid(mptg_ctx* ctx, const uint8_t* h_cells /* h x w, nonzero = obstacle */,
                    int width, int height);
int mptg_scene_arm(mptg_ctx* ctx, const double* h_lengths, int nlinks, double robot_radius,
                   const double* h_disks /* n x [cx,cy,r] */, int ndisks);
/* SE(3) rigid body among triangles (replaces SE3RigidBodyScenario::valid,
 * demo/se3_rigid_body_scenario.hpp:288-312, whose collide() is FCL's).  The
 * space must be MPTG_METRIC_SE3 with states [qx,qy,qz,qw,tx,ty,tz]; pose =
 * Translation * Quaternion (:258-261).  Triangles are 9 floats (3 vertices):
 * robot in its own frame (<= 2^16), environment in the world frame (<= 2^20).
 * valid(q) = no (robot triangle, env triangle) pair whose boxes overlap and
 * which intersect (FP32 separating-axis test).  The link form must be DMV. */
int mptg_scene_mesh(mptg_ctx* ctx, const float* h_robot, int nrobot, const float* h_env, int nenv);
/* Holonomic2DPointScenario (demo/holonomic_2d_point_scenario.hpp:95-113):
 * valid(q) = q outside every circle (squared distance to the centre > r^2,
 * Circle::pointIsValid, shape_hierarchy.hpp:222-226) and every closed
 * rectangle (Rect::pointIsValid, :177-182).  Circles are [cx, cy, r] (at most
 * 4096), rectangles [x0, y0, x1, y1] (at most 4096); the space must be 2-D. */
int mptg_scene_shapes(mptg_ctx* ctx, const double* h_circles, int ncircles, const double* h_rects, int nrects);
/* The Scenario's own valid()/link() as host callbacks (the reference calls them
 * per state from its workers, src/mpt/impl/prrt/prrt.hpp:439,454-457;
 * src/mpt/scenario.hpp:93-130): sampling, nearest neighbours, steering, the
 * tree and PRRT*'s choose-parent/rewire stay on the device, and every batch of
 * states or edges the planner (or mptg_valid / mptg_link) needs decided is
 * copied to the host and handed to one call.  valid_fn sets ok[i] = valid(q_i)
 * (0/1) for n states of `dim` doubles; link_fn sets ok[i] = link(from_i, to_i)
 * for n edges; either returns 0, or nonzero to fail the calling entry point
 * with MPTG_E_ARG.  The callbacks run on the thread that called the entry
 * point and may split the batch over threads of their own.  link() here is the
 * Scenario's whole link(): the planner calls valid(new) first, then link(near,
 * new) (prrt.hpp:439-457).  checks and valid_calls count no host calls (0).
 * Planner iterations with a host scene run step by step (mptg_prrt_graph's
 * setting is not used). */
typedef int (*mptg_valid_fn)(void* user, const double* h_q, int64_t n, uint8_t* h_ok);
typedef int (*mptg_link_fn)(void* user, const double* h_from, const double* h_to, int64_t n, uint8_t* h_ok);
int mptg_scene_host(mptg_ctx* ctx, mptg_valid_fn valid_fn, mptg_link_fn link_fn, void* user);
/* link(): DMV step or bisection tolerance */
int mptg_set_link(mptg_ctx* ctx, int mode, double param);

/* ---- motion validation: replaces Scenario::valid / Scenario::link
 *      (src/mpt/impl/prrt/prrt.hpp:439,454-457).  verdict[i] = link(from_i,
 *      to_i); checks[i] = valid() calls the given order makes before the
 *      verdict is known (reference order reproduces the reference's count).
 *      endpoints_known_valid: endpoint checks inside link() are counted but
 *      not evaluated (the planner has already validated both ends). */
#define MPTG_CHECKS_OVER_LIMIT 0xffffffffu /* edge beyond the work limits (more than
                                                 2^24 DMV waypoints or 30 bisection levels):
                                                 not checked, verdict 0; host-pointer calls
                                                 return MPTG_E_RANGE */
int mptg_valid(mptg_ctx* ctx, const double* h_q, int64_t n, uint8_t* h_ok);
int mptg_link(mptg_ctx* ctx, const double* h_from, const double* h_to, int64_t n, int order,
              int endpoints_known_valid, uint8_t* h_verdict, uint32_t* h_checks);
int mptg_link_dev(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                  int endpoints_known_valid, uint8_t* d_verdict, uint32_t* d_checks);
/* mptg_link_dev without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It covers
 * discrete-motion-validator scenes (boxes, occupancy grids, arms with DMV
 * links; not meshes or bisection links: MPTG_E_STATE) in the reference order,
 * with the same kernels, verdicts and counts as mptg_link_dev.  The buffers
 * the kernels need beyond the inputs and outputs (order tables, long-edge
 * list and count) live in one static device array of 3.1 MB per GPU,
 * allocated with the library and held until the process ends; the first
 * call on a device writes its order tables there (synchronously).  At most
 * 2^18 edges per call (MPTG_E_RANGE beyond).  Every call on a GPU shares that
 * array: calls that overlap on one GPU (several contexts or streams) race. */
int mptg_link_dev_noalloc(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                          int endpoints_known_valid, uint8_t* d_verdict, uint32_t* d_checks);

/* ---- batched PRRT: replaces PRRT::solve / Worker::solve / addSample
 *      (src/mpt/impl/prrt/prrt.hpp:194-201,358-452) with batch-synchronous
 *      iterations: the samples of one iteration extend the tree as it stood
 *      at the iteration's start and are appended in sample order.  batch = 1
 *      is the reference's single-worker loop.  Samples come from the
 *      project's Philox stream (seed, sample index). */
typedef struct {
    uint64_t seed;
    double range;        /* setRange (+inf = no steering)     prrt.hpp:168-171 */
    double goal_bias;    /* setGoalBias (default 0.01)        prrt.hpp:118,159-162 */
    double goal_radius;  /* GoalState radius                  goal_state.hpp:64-69 */
    int32_t batch;       /* samples per iteration */
    int32_t order;       /* mptg_check_order for link() */
} mptg_prrt_params;

typedef struct {
    int64_t iterations, samples, zero_dist, invalid_states;
    int64_t link_calls, link_ok, valid_calls, goals, size;
    /* PRRT* only (zero for PRRT) */
    int64_t rewire_tests;    /* neighbours considered (Stats::rewireTests, prrt_star.hpp:564) */
    int64_t rewires;         /* neighbours moved under a new node (Stats::rewireCount) */
    int64_t parent_changes;  /* new nodes whose parent is not the near node */
    int64_t cp_links;        /* choose-parent link() calls evaluated (speculatively, in parallel) */
    int64_t rw_links;        /* rewire link() calls evaluated (speculatively, in parallel) */
    int64_t deferred;        /* rewire pairs whose condition held only because the new node's
                                cost fell earlier in its own iteration's rewire pass */
    int64_t graph_launches;  /* hipGraph launches (mptg_prrt_graph), not in the reference */
    int64_t rw_late;         /* of rw_links: queued within the rewire pass, in batches, when
                                it reached a pair with no verdict (at least `deferred`) */
} mptg_prrt_stats;

int mptg_prrt_init(mptg_ctx* ctx, const mptg_prrt_params* p, const double* h_lo, const double* h_hi,
                   const double* h_goal);
/* start state(s): tree must be empty before the first call (addStart, prrt.hpp:181-187) */
int mptg_prrt_add_start(mptg_ctx* ctx, const double* h_q);
/* Runs iterations until size >= target_size or max_iters iterations. */
int mptg_prrt_run(mptg_ctx* ctx, int64_t target_size, int64_t max_iters, mptg_prrt_stats* out);
/* Plain PRRT only: run `iterations_per_launch` iterations per hipGraph launch
   (0, the default: enqueue every iteration and read the counters after each).
   The device stops itself at the size target or the iteration bound, so the
   tree, the counters and the statistics equal the step-by-step run's; only
   the host sees them every `iterations_per_launch` iterations.  Replaces the
   per-iteration loop of Worker::solve (src/mpt/impl/prrt/prrt.hpp:358-375)
   when the host has nothing to do between iterations. */
int mptg_prrt_graph(mptg_ctx* ctx, int iterations_per_launch);
int mptg_prrt_stats_get(mptg_ctx* ctx, mptg_prrt_stats* out);
/* PRRT*: the path cost and parent each goal node had when it was found, in discovery order
   (foundGoal keeps that edge as the solution, src/mpt/impl/prrt_star/prrt_star.hpp:203-216;
   solution() starts from it, :324-336).  Returns the count via *n (at most 4096 goals kept). */
int mptg_prrt_goal_edges(mptg_ctx* ctx, double* h_cost, int32_t* h_parent, int64_t cap, int64_t* n);
/* A Scenario's own sampler (src/mpt/impl/scenario_sampler.hpp:145-187: scenario.sample(rng) or
   scenario.sampler()) in place of the uniform box sampler, for PRRT and PRRT*.  Every iteration
   calls fn once, on the thread that called mptg_prrt_run, for the iteration's n = batch sample
   indices first .. first + n - 1 (first = iteration * batch): it writes n states of `dim` doubles
   to h_q and ok[i] = 0 for a slot that has no sample (the reference's std::optional<State>
   without a value: that slot is skipped, as addSample skips it, prrt.hpp:400-403) and returns 0,
   or nonzero to fail mptg_prrt_run with MPTG_E_ARG.  Goal biasing stays with the planner: when
   slot 0 takes the bias coin it samples the goal (prrt.hpp:377-384).  fn = NULL restores the box
   sampler.  Iterations with a custom sampler run step by step (no graph capture). */
typedef int (*mptg_sample_fn)(void* user, uint64_t seed, int64_t first, int64_t n, double* h_q, uint8_t* h_ok);
int mptg_prrt_sampler(mptg_ctx* ctx, mptg_sample_fn fn, void* user);
/* ids of goal nodes in discovery order (foundGoal, prrt.hpp:139-146); returns count via *n */
int mptg_prrt_goals(mptg_ctx* ctx, int32_t* h_ids, int64_t cap, int64_t* n);

/* ---- batched PRRT*: replaces PRRTStar::solve / Worker::addSample
 *      (src/mpt/impl/prrt_star/prrt_star.hpp:447-688).  Per iteration: extend
 *      every sample against the iteration's tree (as PRRT, with dNear
 *      recomputed after steering, :535); choose each survivor's parent among
 *      its k nearest in ascending cost + distance order (:552-605); append in
 *      sample order; then rewire survivors in sample order, neighbours in
 *      distance order, against the costs current at that point, pushing each
 *      cost decrease down the moved subtree (:626-688).  batch = 1 is the
 *      reference's non-concurrent single-worker loop.  After
 *      mptg_prrt_star_init, mptg_prrt_add_start / _run / _stats_get / _goals
 *      drive the PRRT* loop; the tree must be grown from mptg_prrt_add_start
 *      (nodes from mptg_tree_insert carry no costs: MPTG_E_STATE). */
enum mptg_rewire_neighbors {
    MPTG_REWIRE_K_NEAREST = 0, /* rewire_k_nearest  planner_tags.hpp:49, rrg_rewire_neighbors.hpp:46-92 */
    MPTG_REWIRE_R_NEAREST = 1  /* rewire_r_nearest  planner_tags.hpp:50, rrg_rewire_neighbors.hpp:94-152 */
};

typedef struct {
    int32_t k;             /* rewire_k_nearest: neighbours per survivor, 1..64; 0 = the
                              reference's ceil(kRRG * ln(n + 1)), kRRG = factor * e *
                              (1 + 1/d) (rrg_rewire_neighbors.hpp:46-67), n = tree size
                              at the iteration start, d = the space's dimension (6 for
                              SE(3)); MPTG_E_RANGE once that exceeds 64 (n around 10^8 at
                              d = 7, 1.6 * 10^6 at d = 2).  Ignored for r-nearest. */
    int32_t neighbors;     /* mptg_rewire_neighbors; rewire_r_nearest takes every node
                              within rRRG (ln(n + 1) / (n + 1))^(1/d) (:99-128), in lists
                              of any length (MPTG_E_RANGE only past 2^31 entries in one
                              iteration) */
    double rewire_factor;  /* setRewireFactor (default 1.1, prrt_star.hpp:176,224) */
} mptg_prrt_star_params;

int mptg_prrt_star_init(mptg_ctx* ctx, const mptg_prrt_params* p, const mptg_prrt_star_params* sp,
                        const double* h_lo, const double* h_hi, const double* h_goal);
/* path cost (Edge::cost) of nodes [first, first + n) of a PRRT* tree */
int mptg_prrt_costs(mptg_ctx* ctx, int64_t first, int64_t n, double* h_cost);

/* ---- collision surrogate (no counterpart in the reference): a
 *      64-128-128-128-1 ReLU MLP (BF16 operands, FP32 accumulation) trained
 *      on the device on exact-checker labels.  It only orders the waypoint
 *      checks of link() (MPTG_ORDER_SURROGATE, arm scene): verdicts still come
 *      from the exact checker.  Parameters are FP32, laid out
 *      W1[64][128] b1 W2[128][128] b2 W3[128][128] b3 w4[128] b4. */
int64_t mptg_surrogate_nparams(void);
int mptg_surrogate_init(mptg_ctx* ctx, uint64_t seed);  /* He-uniform weights, zero biases */
int mptg_surrogate_get(mptg_ctx* ctx, float* h_params, int64_t n);
int mptg_surrogate_set(mptg_ctx* ctx, const float* h_params, int64_t n);
/* 64 BF16 features per arm state (joint positions / total length, cos/sin of
 * the first 12 cumulative angles) */
int mptg_surrogate_features(mptg_ctx* ctx, const double* d_states, int64_t n, uint16_t* d_feat);
/* one training sample per edge: a waypoint of link(from_i, to_i), its
 * features and its exact label (1 = collides) */
int mptg_surrogate_samples(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, uint64_t seed,
                           uint16_t* d_feat, float* d_label);
int mptg_surrogate_forward(mptg_ctx* ctx, const uint16_t* d_feat, int64_t n, float* d_logits);
/* mean binary cross-entropy over n rows (multiple of 16) and its gradient */
int mptg_surrogate_loss_grad(mptg_ctx* ctx, const uint16_t* d_feat, const float* d_label, int64_t n,
                             float* h_loss, float* h_grad);
/* Adam (0.9, 0.999, 1e-8) over `epochs` passes of batches of `batch` rows */
int mptg_surrogate_train(mptg_ctx* ctx, const uint16_t* d_feat, const float* d_label, int64_t n, int epochs,
                         int batch, float lr, float* h_epoch_losses);
/* scores of every waypoint of every edge: d_scores[i * 512 + reference ordinal] */
int mptg_surrogate_score_edges(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n,
                               float* d_scores);

/* ---- diagnostics (no counterpart in the reference): evaluates the device
 *      build of the project math used by the exact path, for bit-equality
 *      tests against the host build. */
enum mptg_diag_op { MPTG_DIAG_SIN = 0, MPTG_DIAG_COS = 1, MPTG_DIAG_ACOS = 2, MPTG_DIAG_SQRT = 3 };
int mptg_diag_math(mptg_ctx* ctx, int op, const double* h_x, int64_t n, double* h_y);
/* ---- autotuner (no reference counterpart: the reference has no launch
   shapes to choose) --------------------------------------------------------
   Launch shapes the device kernels take from the context.  None of them
   changes a result: they only split the same work differently. */
typedef struct mptg_tune_params {
    int32_t knn_blocks_per_cu;      /* kNN scan blocks per CU the tree split aims at (1..256) */
    int32_t dmv_long_waves_per_cu;  /* grid of the persistent DMV long-edge kernel (1..256) */
    int32_t mesh_waves_per_cu;      /* grid of the persistent mesh kernels (1..256) */
    int32_t arm_tail_waves;         /* waves sharing one arm edge's continuation rounds (1..8) */
    int32_t arm_cull;               /* arm link disk culling: 0 one walk per wave over the wave's
                                       box, 1 each lane over its own grid cells */
    int32_t mesh_edge_waves;        /* waves sharing one mesh edge's waypoints (1..32) */
    int32_t knn_select;             /* k > 1 on trees of 2^16 nodes or more: 1 bound, collect and
                                       select (registers of k = 1 in the scan), 0 the k-list scan */
    int32_t knn_grid;               /* L1/L2 in 2..8 dimensions on trees of 2^16 nodes or more: 1 a
                                       uniform grid built per call, searched cell by cell; 0 the scans */
    int32_t dmv_edge_lanes;         /* DMV link: lanes sharing one edge's checks (1, 2, 4, .., 64) */
    int32_t knn_q8;                 /* k = 1 under L1 in 8 or more dimensions: 1 a byte prefilter
                                       (four coordinates per sum-of-absolute-differences), 0 FP32 */
} mptg_tune_params;

enum {
    MPTG_TUNE_KNN = 0,
    MPTG_TUNE_DMV_LONG = 1,
    MPTG_TUNE_MESH = 2,
    MPTG_TUNE_GRAPH = 3,
    MPTG_TUNE_ARM_TAIL = 4,
    MPTG_TUNE_ARM_CULL = 5,
    MPTG_TUNE_MESH_EDGE = 6,
    MPTG_TUNE_DMV_LANES = 7
};

/* What one tuning call tried and chose: the knob, each candidate value and
   its median time (ms per call; per planner iteration for MPTG_TUNE_GRAPH). */
typedef struct mptg_tune_report {
    int32_t knob;
    int32_t ncand;
    int32_t chosen;
    int32_t _pad;
    int32_t value[8];
    float ms[8];
} mptg_tune_report;

int mptg_tune_get(mptg_ctx* ctx, mptg_tune_params* out);
int mptg_tune_set(mptg_ctx* ctx, const mptg_tune_params* p);
/* Times kNN of the nq device queries d_q against the current tree for each
   candidate split and keeps the fastest (replaces the nigh KD-tree's fixed
   traversal, src/mpt/impl/prrt/prrt.hpp:406-409, which has nothing to tune).
   The split shapes the scans; a tree the uniform grid serves (knn_grid on)
   sends every candidate through the grid, so switch knn_grid off first to
   tune the scans themselves. */
int mptg_autotune_knn(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, mptg_tune_report* rep);
/* Times link() of n device edges (reference order) for each candidate of the
   scene's link launch shape -- for DMV scenes the lanes per edge and then
   the grid of the persistent long-edge kernel, for meshes the grid and then the waves per edge, for the arm the
   disk culling and then the waves per continued edge (rep: the last knob
   swept) -- and keeps the fastest;
   the per-edge bisection kernels have none (MPTG_E_STATE). */
int mptg_autotune_link(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n,
                       mptg_tune_report* rep);
/* Grows the tree (these iterations count, as in mptg_prrt_run) with each
   candidate of the planner's launch knob and keeps the fastest per
   iteration; stops early at target_size.  Plain PRRT: iterations per graph
   launch (0 = step by step).  PRRT*: waves per mesh edge, or per continued
   arm edge (other scenes: MPTG_E_STATE). */
int mptg_autotune_prrt(mptg_ctx* ctx, int64_t target_size, mptg_tune_report* rep);

/* one arm-scene link() call (as mptg_link_dev) timed phase by phase with HIP
 * events on the context stream: h_ms[0] surrogate scoring of every edge's
 * picks (surrogate order; one kernel), [1] packed head round, [2] the gap
 * before the continuation, [3] continuation, [4] irregular-edge walk (ms) */
int mptg_link_phases(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                     int endpoints_known_valid, uint8_t* d_verdict, uint32_t* d_checks, float* h_ms);
/* plain device-to-device copy kernel of `bytes` (multiple of 16, 16-byte
 * aligned), asynchronous: the known-size reference for hardware byte counters */
int mptg_diag_copy(mptg_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes);
/* Work of a link() batch on the arm scene in the given order, from untimed
 * instrumented copies of the kernels (used to price a launch in FLOPs):
 *   h_work[0..4] = the GPU kernel's executed work: valid() calls, forward-
 *                  kinematics links, group broad-phase tests, link tests
 *                  resolved by the far-disk bound, link tests that ran the
 *                  exact expression;
 *   h_work[5..8] = the reference algorithm's work on the same edges: valid()
 *                  calls, FK links, segment-disk tests (every disk until the
 *                  first hit), and those of the tests whose disk centre lies
 *                  within (link length + r + R) of the link's start joint,
 *                  the pairs no exact per-link distance bound clears. */
int mptg_link_work(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                   int endpoints_known_valid, uint64_t* h_work);

#ifdef __cplusplus
}
#endif
#endif
