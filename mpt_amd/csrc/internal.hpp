// Library-internal context, error plumbing and device-buffer helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mptg.h"
#include "hd.hpp"

namespace mptg {

struct Surrogate;

constexpr int kMaxDim = 32;
constexpr int kMaxK = 64;

// Device buffer owned by a context.  Grows, never shrinks.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct SceneDev {
    int kind = MPTG_SCENE_NONE;
    // boxes
    double* boxes = nullptr;
    int nboxes = 0;
    // grid: one bit per cell (1 = obstacle) in 8 x 8 tiles, tile (tx, ty) at
    // ty * grid_tx + tx, cell (x, y) at bit 8 * (y % 8) + x % 8
    uint64_t* grid = nullptr;
    int width = 0, height = 0, grid_tx = 0;
    // arm
    double* lengths = nullptr;
    int nlinks = 0;
    double robot_radius = 0;
    double* disks = nullptr;  // [cx, cy, (r+R)^2, r+R] per disk, reference order
    int ndisks = 0;
    DevBuf arm_disks, arm_lim;  // kernel layout of the disks (ascending cx), per-link bounds
    DevBuf arm_index;           // [xmin, 1/bucket width, max (r+R), 0] + int32 bucket starts
    double total_length = 0;    // sum of link lengths (feature scale)
    // mesh (SE(3) rigid body): env BVH nodes, env triangles, robot triangles
    DevBuf mesh_nodes, mesh_tris, mesh_robot;
    int mesh_nnodes = 0, mesh_nrobot = 0;
    // shapes: circles [cx, cy, r^2, r]; the rectangles are `boxes`
    double* circles = nullptr;
    int ncircles = 0;
    // host scene: the caller's valid()/link() callbacks
    mptg_valid_fn host_valid = nullptr;
    mptg_link_fn host_link = nullptr;
    void* host_user = nullptr;
};

struct PrrtState {
    bool init = false;
    bool counters_valid = false;  // device counters mirror the host state
    mptg_prrt_stats stats_base{};  // host stats when the device counters were reset
    DevBuf q, near, ndist, from, to, status, mask, verdict, checks, counters, goal_buf, blk;
    // valid() before nearest (no range limit): listed valid samples, their ids,
    // the list count, and their nearest node and distance
    DevBuf qv, vidx, vcnt, kn_i, kn_d, vcheck;
    // a Scenario's own sampler (host callback): staged samples, their flags,
    // the slots without a sample
    mptg_sample_fn sampler = nullptr;
    void* sampler_user = nullptr;
    DevBuf hs_q, hs_ok, hs_skip;
    std::vector<double> hs_host_q;
    std::vector<uint8_t> hs_host_ok;
    mptg_prrt_params p{};
    double lo[kMaxDim], hi[kMaxDim], goal[kMaxDim];
    int64_t iteration = 0;       // next iteration index
    int64_t goals = 0;           // goal nodes found so far
    // iterations per graph launch (0: one launch sequence per iteration)
    int32_t graph_iters = 0;
    // the captured iterations, kept across runs while nothing else ran:
    // valid at context epoch g_epoch for g_iters iterations, tree capacity
    // g_cap and a kNN split planned for g_plan nodes
    hipGraphExec_t g_exec = nullptr;
    uint64_t g_epoch = 0;
    int32_t g_iters = 0;
    int64_t g_cap = 0, g_plan = 0;
    mptg_prrt_stats stats{};
    std::vector<int32_t> goal_ids;
    // PRRT*: each goal node's path cost and parent when it was found
    std::vector<double> goal_cost0;
    std::vector<int32_t> goal_parent0;
    // PRRT*
    bool star = false;
    int32_t star_k = 20;          // 0: the reference's k(n) formula
    int32_t star_mode = MPTG_REWIRE_K_NEAREST;  // or MPTG_REWIRE_R_NEAREST
    double rewire_factor = 1.1;
    // neighbour lists: k-nearest rows of star_k slots, or radius lists in
    // CSR form (nbh_off per survivor, nbh_own = survivor of each entry)
    DevBuf dnear, sidx, qs, nbh_idx, nbh_d, nbh_off, nbh_own, marked, pc, cpar, cp_from, cp_to, cp_ver, cp_chk,
        pair_cp, rw_from, rw_to, rw_ver, rw_chk, pair_rw, front;
};

}  // namespace mptg

struct mptg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int cus = 0;

    // space
    int metric = -1;
    int dim = 0;
    double w_so3 = 1, w_l2 = 1;

    // tree (node id = insertion order)
    int64_t size = 0, cap = 0;
    double* x64 = nullptr;   // [cap][dim] AoS, FP64 (exact coordinates)
    float* x32 = nullptr;    // [cap][dp]  AoS, FP32 shadow for the prefilter
    int32_t* parent = nullptr;
    // PRRT* node data: path cost and child lists (first child, next and
    // previous sibling; -1 = none), valid for every node while star_ok
    double* cost = nullptr;
    int32_t *child = nullptr, *next = nullptr, *prev = nullptr;
    bool star_ok = true;
    double coord_bound = 0;  // max |coordinate| over inserted nodes and bounds
    // nodes [0, valid_upto) are known to pass the scene's valid(); some_invalid:
    // a node was found invalid (both reset when the tree is cleared or the
    // scene changes)
    int64_t valid_upto = 0;
    bool some_invalid = false;

    // scene
    mptg::SceneDev scene;
    int link_mode = MPTG_LINK_DMV;
    double link_param = 1.0;

    // scratch (grow-only)
    // kNN k > 1 over large trees (bound, collect, select): subset results,
    // bounds + counts, candidates, overflow list, gathered queries, results
    mptg::DevBuf knn_sel[6];
    // kNN over a uniform grid: cell of each node, counting-sort arrays,
    // sorted rows and ids, overflow list, gathered overflow queries +
    // results, query cells and order
    mptg::DevBuf knn_grid[6];
    mptg::DevBuf dmv_table16;  // the DMV ring order for 256..1023 interior waypoints (uint16 rows)
    // radius search: list offsets, scanned offsets, counts, scan scratch;
    // unsorted pairs and sort scratch; staged host queries; sorted results
    mptg::DevBuf rnn_buf, rnn_pairs, rnn_q, rnn_out;
    mptg::DevBuf scratch[12];  // 0-2 host staging, 3 kNN seed bounds, 4-5 kNN partials, 6 DMV table, 7 edge lists, 8 scores,
                               // 9 surrogate picks, 10 continuation list
    // pinned host staging
    void* pinned = nullptr;
    size_t pinned_bytes = 0;

    mptg::PrrtState prrt;
    bool force_generic = false;  // route link() to the per-edge reference kernels (testing)
    mptg::Surrogate* surrogate = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    // per-phase timing of one arm link() call (mptg_link_phases): events
    // recorded between its launches when `phase_timing` is set
    bool phase_timing = false;
    // a stream capture is open: no allocation, synchronisation or memset
    bool capturing = false;
    // bumped by every entry point and every (re)allocation: a captured
    // graph is replayed only while the epoch it was captured at holds
    uint64_t epoch = 0;
    // launch shapes (mptg_autotune_*): kNN blocks per CU, waves per CU of the
    // persistent DMV long-edge kernel and of the mesh kernels, waves per edge
    // of the arm continuation, arm link disk culling (0 wave box, 1 per lane),
    // waves per mesh edge, kNN select and grid paths, lanes per DMV edge, the
    // byte prefilter of the k = 1 L1 scan
    mptg_tune_params tune{8, 16, 32, 2, 0, 2, 1, 1, 8, 1};
    hipEvent_t phase_ev[6] = {};
};

namespace mptg {

extern thread_local std::string g_create_err;

// Marks phase boundary i of a timed arm link() call (no-op otherwise).
inline void phase_mark(mptg_ctx* ctx, int i) {
    if (ctx->phase_timing) (void)hipEventRecord(ctx->phase_ev[i], ctx->stream);
}

inline int fail(mptg_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    else g_create_err = msg;
    return code;
}

#define MPTG_HIP(ctx, call)                                                                \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::mptg::fail((ctx), MPTG_E_HIP,                                         \
                                std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

// FP32 shadow row pitch (floats) for a dimension.
inline int pitch32(int dim) { return dim <= 2 ? 2 : (dim <= 4 ? 4 : (dim <= 8 ? 8 : ((dim + 3) / 4) * 4)); }

int ensure(mptg_ctx* ctx, DevBuf& b, size_t bytes);
int ensure_pinned(mptg_ctx* ctx, size_t bytes);
// Entry-point prologue: selects the context's device and retires captured
// graphs (the call may change state they hold).
int check_device_ready(mptg_ctx* ctx);
// The same without retiring graphs, for entry points that change nothing a
// graph holds (their scratch growth retires graphs through ensure()).
int check_device(mptg_ctx* ctx);
// Largest |coordinate| of the initialised planner's sampling box and goal
// (0 without a planner), with a 1e-12 relative margin.
double planner_coord_bound(const mptg_ctx* ctx);

// kernels launched from other translation units
// n_dev: tree size read on the device (tree_n then only plans the split);
// skip: device flag that turns the launch into a no-op
// Allocates the kNN partial-result buffers for nq queries over a tree of
// tree_n nodes (what launch_knn would allocate).
int knn_reserve(mptg_ctx* ctx, int64_t nq, int k, int64_t tree_n);
// Grid kNN (knn_grid.hip): whether it serves a tree of n nodes in this
// context, and the search itself; the m queries it leaves (*m, listed at
// *over on the device) are the caller's to answer.
bool knn_grid_applies(const mptg_ctx* ctx, int64_t n);
int knn_grid_path(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx, double* d_dist,
                  int64_t n, int64_t** over, uint32_t* m);
// Radius search on device-resident queries (rnn.hip), over tree nodes
// [0, ctx->size): rnn_count counts every query's nodes within r and returns
// the total (one 8-byte readback); rnn_offsets is then the device array of
// nq + 1 list offsets; rnn_fill writes the lists, each in ascending
// (distance, id) order, and (if owner is not null) the query of each entry.
int rnn_count(mptg_ctx* ctx, const double* d_q, int64_t nq, double r, int64_t* total);
const int64_t* rnn_offsets(const mptg_ctx* ctx);
int rnn_fill(mptg_ctx* ctx, const double* d_q, int64_t nq, double r, int64_t total, double* d_dist, int32_t* d_idx,
             int32_t* d_owner);
bool rnn_dim_ok(int dim);
// Zeroes n 32-bit words with a kernel (graph-capturable, unlike a memset).
int launch_zero_u32(mptg_ctx* ctx, uint32_t* p, int64_t n);
// Zeroes the 32-bit words *a and *b with one small kernel.
int launch_zero_pair(mptg_ctx* ctx, uint32_t* a, uint32_t* b);
// Sets n 32-bit words to v with a kernel.
int launch_fill_u32(mptg_ctx* ctx, uint32_t* p, int64_t n, uint32_t v);
// nq_dev (k = 1, L1/L2): only the first *nq_dev of the nq query slots are
// live (nq stays the layout); the scans run, planned for half the slots.
int launch_knn(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx,
               double* d_dist, int64_t tree_n, const int64_t* n_dev = nullptr, const int32_t* skip = nullptr,
               const uint32_t* nq_dev = nullptr);
int launch_link(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                int known_valid, uint8_t* d_verdict, uint32_t* d_checks);
int launch_link_masked(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                       int known_valid, const uint8_t* d_mask, uint8_t* d_verdict, uint32_t* d_checks);
int launch_valid(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_ok);
// mptg_link_dev_noalloc's work (validate.hip): the DMV kernels on the
// library's static device scratch, no allocation
int link_dev_noalloc(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                     int known_valid, uint8_t* d_verdict, uint32_t* d_checks);
// mesh scene: one wave per state / edge
int launch_mesh_valid(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_ok);
// Host scenes (host_scene.hip): the caller's callbacks over device rows.
// host_valid_dev: d_ok[i] = valid(q_i).  host_filter: for masked rows, valid()
// false sets d_status[i] = bad and clears d_mask[i].  host_link_dev: for
// masked edges (all if d_mask is null), d_verdict[i] = link(from_i, to_i) and
// d_checks[i] = 0.  All synchronous.
int host_valid_dev(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_ok);
int host_filter(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_mask, uint8_t* d_status, uint8_t bad);
int host_link_dev(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, const uint8_t* d_mask,
                  uint8_t* d_verdict, uint32_t* d_checks);
int launch_mesh_link(mptg_ctx* ctx, const uint8_t* d_table, const double* d_from, const double* d_to, int64_t n,
                     int known_valid, const uint8_t* d_mask, uint8_t* d_verdict, uint32_t* d_checks);
// planners: states[b] with mask[b] set that collide get status[b] = invalid_code, mask[b] = 0
int launch_mesh_filter(mptg_ctx* ctx, const double* d_states, int64_t n, uint8_t* d_mask, uint8_t* d_status,
                       int invalid_code);
// Arm link kernels in `order`; with the surrogate order every edge's picks
// are scored first, and the packed round checks them.
int launch_arm_link(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int known_valid,
                    const uint8_t* mask, uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr,
                    unsigned long long* work, int order);
// surrogate scores of every waypoint of every (masked) arm edge:
// scores[e * 512 + reference ordinal]
int surrogate_score_edges(mptg_ctx* ctx, const double* from, const double* to, int64_t n, const uint8_t* mask,
                          float* scores);
// The surrogate's picks of every (masked) regular arm edge of at most 9
// levels: the reference ordinals of the 4 best-scored nodes of its top
// tree, best first, at picks[e * 4 + r].
int surrogate_picks(mptg_ctx* ctx, const double* from, const double* to, int64_t n, const uint8_t* mask,
                    uint16_t* picks);

}  // namespace mptg
