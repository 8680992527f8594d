// Batched PRRT extension loop on the device (replaces Worker::solve /
// Worker::addSample, src/mpt/impl/prrt/prrt.hpp:358-452).
//
// One iteration = B samples, each extended against the tree as it stood when
// the iteration began; accepted nodes are appended in sample order.  Per
// sample (reference line numbers in src/mpt/impl/prrt/prrt.hpp):
//   goal bias while no goal is known, unscaled bias      :377-387 (scaledBias :373 unused)
//   nearest                                              :416
//   d == 0 -> reject                                     :427-428
//   d > range -> interpolate(near, rand, range / d)      :430-434
//   valid(new)                                           :439-440
//   link(near, new)                                      :442
//   goal check d(new, goal) <= radius                    :443, src/mpt/goal_state.hpp:64-69
//   insert + foundGoal                                   :446-450
// Samples are drawn from the Philox stream (seed, it * B + b).
#include <algorithm>
#include <cmath>
#include <vector>

#include "prrt_dev.hpp"

namespace mptg {

namespace {

// Append in three launches, one sample per thread: count (per-block
// accepted and goal counts, the counters' sums), write (node ids in sample
// order: the block's base is the sum over the blocks before it) and commit
// (tree size, goals, the next iteration and the stop flag).  Every kernel is
// a no-op once the stop flag is set.
constexpr int kAppendBlock = 256;

__device__ __forceinline__ int block_rank(bool flag, int* wsum, int* total) {
    // rank of this thread's flag among the block's set flags (thread order);
    // wsum: 16 ints of LDS.  Contains __syncthreads: call from every thread.
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const uint64_t m = __ballot(flag);
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (lane == 0) wsum[wid] = __popcll(m);
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < kAppendBlock / 64; ++w) {
        if (w < wid) base += wsum[w];
        tot += wsum[w];
    }
    *total = tot;
    return base + __popcll(m & below);
}

// Sum over the block (kAppendBlock threads) of two int64 values; lds: 32
// int64 of LDS.  Contains __syncthreads: call from every thread.
__device__ __forceinline__ void block_sum2(int64_t* a, int64_t* b, int64_t* lds) {
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    int64_t x = *a, y = *b;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        x += __shfl_xor(x, o);
        y += __shfl_xor(y, o);
    }
    if (lane == 0) {
        lds[wid] = x;
        lds[16 + wid] = y;
    }
    __syncthreads();
    x = y = 0;
    for (int w = 0; w < kAppendBlock / 64; ++w) {
        x += lds[w];
        y += lds[16 + w];
    }
    *a = x;
    *b = y;
}

__global__ __launch_bounds__(kAppendBlock) void prrt_count(SpaceArgs sp, PrrtDev pd, double goal_radius, int B,
                                                           const double* __restrict__ to,
                                                           uint8_t* __restrict__ status,
                                                           const uint8_t* __restrict__ verdict,
                                                           const uint32_t* __restrict__ checks,
                                                           uint8_t* __restrict__ flags, int32_t* __restrict__ blk,
                                                           Counters* __restrict__ cnt) {
    __shared__ int wsum[16];
    __shared__ unsigned long long part[5];
    const int D = sp.dim;
    const int t = threadIdx.x;
    if (blockIdx.x == 0 && t == 0) {  // what prrt_write builds on
        cnt->size_seen = cnt->size;
        cnt->goals_seen = cnt->goals;
        cnt->done_seen = cnt->done;
    }
    if (cnt->done) return;  // block-uniform
    if (t < 5) part[t] = 0;
    const int b = blockIdx.x * kAppendBlock + t;
    bool ok = false, goal = false;
    unsigned long long lz = 0, li = 0, lc = 0, lo = 0, lv = 0;
    if (b < B) {
        uint8_t st = status[b];
        if (st == ST_LINK_FAIL) {
            lc = 1;
            lv = checks[b];
            if (verdict[b]) {
                st = ST_OK;
                status[b] = ST_OK;
            }
        }
        lz = st == ST_ZERO;
        li = st == ST_INVALID;
        if (st == ST_OK) {
            lo = 1;
            ok = true;
            goal = space_dist(sp, to + (int64_t)b * D, pd.goal) <= goal_radius;
        }
        flags[b] = (uint8_t)((ok ? 1 : 0) | (goal ? 2 : 0));
    }
    int nok, ngoal;
    (void)block_rank(ok, wsum, &nok);
    __syncthreads();
    (void)block_rank(goal, wsum, &ngoal);
    if (lz) atomicAdd(&part[0], lz);
    if (li) atomicAdd(&part[1], li);
    if (lc) atomicAdd(&part[2], lc);
    if (lo) atomicAdd(&part[3], lo);
    if (lv) atomicAdd(&part[4], lv);
    __syncthreads();
    if (t == 0) {
        blk[2 * blockIdx.x] = nok;
        blk[2 * blockIdx.x + 1] = ngoal;
        atomicAdd((unsigned long long*)&cnt->zero, part[0]);
        atomicAdd((unsigned long long*)&cnt->invalid, part[1]);
        atomicAdd((unsigned long long*)&cnt->link_calls, part[2]);
        atomicAdd((unsigned long long*)&cnt->link_ok, part[3]);
        atomicAdd((unsigned long long*)&cnt->valid_calls, part[4]);
    }
}

__global__ __launch_bounds__(kAppendBlock) void prrt_write(int D, int dp, int B, const double* __restrict__ to,
                                                           const int32_t* __restrict__ near,
                                                           const uint8_t* __restrict__ flags,
                                                           const int32_t* __restrict__ blk, double* __restrict__ x64,
                                                           float* __restrict__ x32, int32_t* __restrict__ parent,
                                                           Counters* __restrict__ cnt,
                                                           int32_t* __restrict__ goal_ids, int64_t goal_cap) {
    __shared__ int wsum[16];
    __shared__ int64_t red[32];
    const int t = threadIdx.x;
    // the *_seen values stay fixed while this kernel runs; size, goals,
    // iteration and done change only in its last block, after its writes
    if (cnt->done_seen) return;  // block-uniform
    // this block's first node id and goal slot: the counts of the blocks
    // before it on top of the tree
    int64_t before = 0, gbefore = 0;
    for (int j = t; j < (int)blockIdx.x; j += kAppendBlock) {
        before += blk[2 * j];
        gbefore += blk[2 * j + 1];
    }
    block_sum2(&before, &gbefore, red);
    const int64_t base = cnt->size_seen + before, gbase = cnt->goals_seen + gbefore;
    __syncthreads();
    const int b = blockIdx.x * kAppendBlock + t;
    const uint8_t f = b < B ? flags[b] : 0;
    const bool ok = f & 1, goal = f & 2;
    int tot_ok, tot_goal;
    const int r = block_rank(ok, wsum, &tot_ok);
    __syncthreads();
    const int gr = block_rank(goal, wsum, &tot_goal);
    if (ok) {
        const int64_t id = base + r;
        for (int i = 0; i < D; ++i) {
            const double v = to[(int64_t)b * D + i];
            x64[id * D + i] = v;
            x32[id * dp + i] = (float)v;
        }
        parent[id] = near[b];
        if (goal) {
            const int64_t g = gbase + gr;
            if (g < goal_cap) goal_ids[g] = (int32_t)id;
        }
    }
    // the last block commits: every block before it is counted in `before`
    if (blockIdx.x == gridDim.x - 1 && t == 0) {
        const int64_t size = base + tot_ok;
        cnt->size = size;
        cnt->goals = gbase + tot_goal;
        cnt->iteration += 1;
        cnt->done = (size >= cnt->target || cnt->iteration >= cnt->iter_end) ? 1 : 0;
    }
}

}  // namespace

using PrrtBufs = PrrtState;
static PrrtBufs& bufs(mptg_ctx* ctx) { return ctx->prrt; }

int prrt_reserve(mptg_ctx* ctx, int B) {
    PrrtBufs& pb = bufs(ctx);
    int D = ctx->dim, rc;
    if ((rc = ensure(ctx, pb.q, sizeof(double) * D * B))) return rc;
    if ((rc = ensure(ctx, pb.near, sizeof(int32_t) * B))) return rc;
    if ((rc = ensure(ctx, pb.ndist, sizeof(double) * B))) return rc;
    if ((rc = ensure(ctx, pb.from, sizeof(double) * D * B))) return rc;
    if ((rc = ensure(ctx, pb.to, sizeof(double) * D * B))) return rc;
    if ((rc = ensure(ctx, pb.status, B))) return rc;
    if ((rc = ensure(ctx, pb.mask, B))) return rc;
    if ((rc = ensure(ctx, pb.verdict, B))) return rc;
    if ((rc = ensure(ctx, pb.checks, sizeof(uint32_t) * B))) return rc;
    if ((rc = ensure(ctx, pb.counters, kCounterBlock))) return rc;
    if ((rc = ensure(ctx, pb.goal_buf, sizeof(int32_t) * 4096))) return rc;
    if ((rc = ensure(ctx, pb.blk, sizeof(int32_t) * 2 * ((B + kAppendBlock - 1) / kAppendBlock)))) return rc;
    if ((rc = ensure(ctx, pb.qv, sizeof(double) * D * B))) return rc;
    if ((rc = ensure(ctx, pb.vidx, sizeof(int32_t) * B))) return rc;
    if ((rc = ensure(ctx, pb.vcnt, 256))) return rc;
    if ((rc = ensure(ctx, pb.kn_i, sizeof(int32_t) * B))) return rc;
    if ((rc = ensure(ctx, pb.kn_d, sizeof(double) * B))) return rc;
    return MPTG_OK;
}

// valid() before nearest applies to plain PRRT with no range limit in an
// L1/L2 space with a per-state valid() (not the mesh scene's whole-wave
// check), in dimensions the nearest-neighbour scans serve.
static bool prevalid_eligible(const mptg_ctx* ctx) {
    const PrrtState& s = ctx->prrt;
    const int D = ctx->dim;
    const bool scan_dim = (D >= 1 && D <= 8) || D == 10 || D == 12 || D == 16 || D == 20;
    return s.init && !s.star && std::isinf(s.p.range) && s.p.range > 0 && ctx->metric != MPTG_METRIC_SE3 &&
           ctx->scene.kind != MPTG_SCENE_NONE && ctx->scene.kind != MPTG_SCENE_MESH &&
           ctx->scene.kind != MPTG_SCENE_HOST && !s.sampler && scan_dim &&
           !(ctx->scene.kind == MPTG_SCENE_ARM && ctx->scene.nlinks > D) &&
           !(ctx->scene.kind != MPTG_SCENE_ARM && D < 2);
}

// Checks valid() on the tree nodes not yet known valid ([valid_upto, size)):
// sets valid_upto = size, or some_invalid.  Synchronous.
static int tree_validity_refresh(mptg_ctx* ctx) {
    if (ctx->some_invalid || ctx->valid_upto >= ctx->size) return MPTG_OK;
    PrrtState& s = ctx->prrt;
    const int64_t first = ctx->valid_upto, n = ctx->size - first;
    int rc;
    if ((rc = ensure(ctx, s.vcheck, (size_t)n))) return rc;
    if ((rc = launch_valid(ctx, ctx->x64 + first * ctx->dim, n, (uint8_t*)s.vcheck.p))) return rc;
    std::vector<uint8_t> ok((size_t)n);
    MPTG_HIP(ctx, hipMemcpyAsync(ok.data(), s.vcheck.p, (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (uint8_t v : ok)
        if (!v) {
            ctx->some_invalid = true;
            return MPTG_OK;
        }
    ctx->valid_upto = ctx->size;
    return MPTG_OK;
}

void prrt_graph_drop(mptg_ctx* ctx) {
    PrrtState& s = ctx->prrt;
    if (s.g_exec) (void)hipGraphExecDestroy(s.g_exec);
    s.g_exec = nullptr;
}

void prrt_free(mptg_ctx* ctx) {
    prrt_graph_drop(ctx);
    PrrtBufs& pb = bufs(ctx);
    DevBuf* all[] = {&pb.q,       &pb.near,    &pb.ndist,  &pb.from,    &pb.to,       &pb.status,  &pb.mask,
                     &pb.verdict, &pb.checks,  &pb.counters, &pb.goal_buf, &pb.dnear, &pb.sidx,   &pb.qs,
                     &pb.nbh_idx, &pb.nbh_d,   &pb.marked, &pb.pc,      &pb.cpar,     &pb.cp_from, &pb.cp_to,
                     &pb.cp_ver,  &pb.cp_chk,  &pb.pair_cp, &pb.rw_from, &pb.rw_to,    &pb.rw_ver,  &pb.rw_chk,
                     &pb.pair_rw, &pb.front,   &pb.blk,     &pb.qv,      &pb.vidx,
                     &pb.vcnt,    &pb.kn_i,    &pb.kn_d,   &pb.vcheck, &pb.nbh_off, &pb.nbh_own, &pb.hs_q,
                     &pb.hs_ok,   &pb.hs_skip};
    for (DevBuf* b : all) {
        (void)hipFree(b->p);
        b->p = nullptr;
        b->bytes = 0;
    }
}

int tree_reserve(mptg_ctx* ctx, int64_t cap);
int prrt_star_run(mptg_ctx* ctx, int64_t target_size, int64_t max_iters);

double planner_coord_bound(const mptg_ctx* ctx) {
    const PrrtState& s = ctx->prrt;
    if (!s.init) return 0.0;
    double bound = 0.0;
    for (int i = 0; i < ctx->dim; ++i) {
        bound = std::max(bound, std::max(std::fabs(s.lo[i]), std::fabs(s.hi[i])));
        bound = std::max(bound, std::fabs(s.goal[i]));
    }
    if (ctx->metric == MPTG_METRIC_SE3) bound = std::max(bound, 1.0);
    return bound * (1.0 + 1e-12);
}

// Writes the stop condition of one run into the device counters: the next
// iteration index, the size target and the iteration bound; clears `done`.
int counters_arm(mptg_ctx* ctx, void* dcnt, int64_t iteration, int64_t target, int64_t max_iters) {
    struct Tail {
        int64_t iteration, target, iter_end;
        int32_t done, done_seen;
        int64_t size_seen, goals_seen;
    } t;
    static_assert(sizeof(Tail) == sizeof(Counters) - offsetof(Counters, iteration), "counter tail layout");
    t.iteration = iteration;
    t.target = target;
    t.iter_end = max_iters > INT64_MAX - iteration ? INT64_MAX : iteration + max_iters;
    t.done = 0;
    t.done_seen = 0;
    t.size_seen = 0;
    t.goals_seen = 0;
    MPTG_HIP(ctx, hipMemcpyAsync((char*)dcnt + offsetof(Counters, iteration), &t, sizeof(t),
                                 hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

namespace {

// Enqueues one PRRT iteration.  it >= 0: eager (host-known iteration index
// and tree size); it < 0: graph body (both read from the device counters,
// the kNN split planned for plan_n nodes).  pre: valid() before nearest
// (no range limit, every tree node valid; see prrt_prevalid) -- the eager
// grid search keeps its own order, so trees it serves are searched after.
int enqueue_iteration(mptg_ctx* ctx, const PrrtDev& pd, int64_t it, int64_t plan_n, bool pre) {
    PrrtState& s = ctx->prrt;
    PrrtBufs& pb = bufs(ctx);
    const int B = s.p.batch, D = ctx->dim;
    Counters* dcnt = (Counters*)pb.counters.p;
    const bool dev = it < 0;
    int rc;
    if (pre && !dev && knn_grid_applies(ctx, ctx->size)) pre = false;
    uint32_t* vcnt = (uint32_t*)pb.vcnt.p;  // cleared by prrt_sample when the pre path runs
    hipLaunchKernelGGL(prrt_sample, dim3(sample_blocks(ctx->metric, D, B)), dim3(kBlock), 0, ctx->stream, pd,
                       ctx->metric, D, s.p.seed, s.p.goal_bias, it, B, (const Counters*)dcnt, (double*)pb.q.p,
                       pre ? vcnt : nullptr);
    MPTG_HIP(ctx, hipGetLastError());
    uint8_t* skip = nullptr;  // custom sampler: slots without a sample (eager iterations only)
    if (s.sampler && (rc = take_host_samples(ctx, it, (double*)pb.q.p, dcnt, &skip))) return rc;
    if (pre) {
        const bool arm8 = ctx->scene.kind == MPTG_SCENE_ARM && ctx->scene.nlinks == D && (D == 5 || D == 10 || D == 20);
        dispatch_dim(D, [&]<int DD>() {
            if constexpr (DD == 5 || DD == 10 || DD == 20) {
                if (arm8) {
                    hipLaunchKernelGGL((prrt_prevalid<DD, kPrevalidLanes>),
                                       dim3((unsigned)(((int64_t)B * kPrevalidLanes + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                                       ctx->stream, scene_args(ctx), D, B, (const double*)pb.q.p,
                                       (uint8_t*)pb.status.p, (uint8_t*)pb.mask.p, (double*)pb.qv.p,
                                       (int32_t*)pb.vidx.p, vcnt, (const Counters*)dcnt);
                    return;
                }
            }
            hipLaunchKernelGGL((prrt_prevalid<DD, 1>), dim3((unsigned)((B + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               ctx->stream, scene_args(ctx), D, B, (const double*)pb.q.p, (uint8_t*)pb.status.p,
                               (uint8_t*)pb.mask.p, (double*)pb.qv.p, (int32_t*)pb.vidx.p, vcnt,
                               (const Counters*)dcnt);
        });
        MPTG_HIP(ctx, hipGetLastError());
        if ((rc = launch_knn(ctx, (const double*)pb.qv.p, B, 1, __builtin_inf(), (int32_t*)pb.kn_i.p,
                             (double*)pb.kn_d.p, dev ? plan_n : ctx->size, dev ? &dcnt->size : nullptr,
                             dev ? &dcnt->done : nullptr, vcnt)))
            return rc;
        dispatch_dim(D, [&]<int DD>() {
            hipLaunchKernelGGL((prrt_steer_listed<DD>), dim3((unsigned)((B + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               ctx->stream, D, B, (const double*)ctx->x64, (const double*)pb.qv.p,
                               (const int32_t*)pb.vidx.p, (const uint32_t*)vcnt, (const int32_t*)pb.kn_i.p,
                               (const double*)pb.kn_d.p, (int32_t*)pb.near.p, (double*)pb.from.p, (double*)pb.to.p,
                               (uint8_t*)pb.status.p, (uint8_t*)pb.mask.p, (const Counters*)dcnt);
        });
    } else {
        if ((rc = launch_knn(ctx, (const double*)pb.q.p, B, 1, __builtin_inf(), (int32_t*)pb.near.p,
                             (double*)pb.ndist.p, dev ? plan_n : ctx->size, dev ? &dcnt->size : nullptr,
                             dev ? &dcnt->done : nullptr)))
            return rc;
        dispatch_dim(D, [&]<int DD>() {
            launch_steer<DD>(ctx, s.p.range, B, (const double*)pb.q.p, (const int32_t*)pb.near.p,
                             (const double*)pb.ndist.p, (double*)pb.from.p, (double*)pb.to.p, (uint8_t*)pb.status.p,
                             (uint8_t*)pb.mask.p, (double*)nullptr, (const Counters*)dcnt);
        });
    }
    MPTG_HIP(ctx, hipGetLastError());
    if (skip && (rc = apply_skip(ctx, skip, (uint8_t*)pb.status.p, (uint8_t*)pb.mask.p, dcnt))) return rc;
    if (ctx->scene.kind == MPTG_SCENE_MESH &&
        (rc = launch_mesh_filter(ctx, (const double*)pb.to.p, B, (uint8_t*)pb.mask.p, (uint8_t*)pb.status.p,
                                 ST_INVALID)))
        return rc;
    // host scenes: valid(new) from the caller, then link(near, new)
    if (ctx->scene.kind == MPTG_SCENE_HOST &&
        (rc = host_filter(ctx, (const double*)pb.to.p, B, (uint8_t*)pb.mask.p, (uint8_t*)pb.status.p, ST_INVALID)))
        return rc;
    if ((rc = launch_link_masked(ctx, (const double*)pb.from.p, (const double*)pb.to.p, B, s.p.order, 1,
                                 (const uint8_t*)pb.mask.p, (uint8_t*)pb.verdict.p, (uint32_t*)pb.checks.p)))
        return rc;
    const unsigned ab = (unsigned)((B + kAppendBlock - 1) / kAppendBlock);
    hipLaunchKernelGGL(prrt_count, dim3(ab), dim3(kAppendBlock), 0, ctx->stream, space_args(ctx), pd,
                       s.p.goal_radius, B, (const double*)pb.to.p, (uint8_t*)pb.status.p, (const uint8_t*)pb.verdict.p,
                       (const uint32_t*)pb.checks.p, (uint8_t*)pb.mask.p, (int32_t*)pb.blk.p, dcnt);
    hipLaunchKernelGGL(prrt_write, dim3(ab), dim3(kAppendBlock), 0, ctx->stream, D, pitch32(D), B,
                       (const double*)pb.to.p, (const int32_t*)pb.near.p, (const uint8_t*)pb.mask.p,
                       (const int32_t*)pb.blk.p, ctx->x64, ctx->x32, ctx->parent, dcnt,
                       (int32_t*)pb.goal_buf.p, (int64_t)4096);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// Captures `iters` iterations into an executable graph.  Everything the
// body allocates must already be allocated: the capture flag turns any
// growth into an error instead of a synchronising reallocation.
int capture_iterations(mptg_ctx* ctx, const PrrtDev& pd, int iters, int64_t plan_n, bool pre, hipGraphExec_t* out) {
    *out = nullptr;
    const bool timing = ctx->phase_timing;
    ctx->phase_timing = false;
    ctx->capturing = true;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
    int rc = MPTG_OK;
    if (e != hipSuccess) {
        ctx->capturing = false;
        ctx->phase_timing = timing;
        return fail(ctx, MPTG_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    for (int k = 0; k < iters && rc == MPTG_OK; ++k) rc = enqueue_iteration(ctx, pd, -1, plan_n, pre);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    ctx->capturing = false;
    ctx->phase_timing = timing;
    if (rc == MPTG_OK && e != hipSuccess)
        rc = fail(ctx, MPTG_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (rc == MPTG_OK) {
        e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            *out = nullptr;
            rc = fail(ctx, MPTG_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    }
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

}  // namespace

}  // namespace mptg

using namespace mptg;

extern "C" {

int mptg_prrt_init(mptg_ctx* ctx, const mptg_prrt_params* p, const double* h_lo, const double* h_hi,
                   const double* h_goal) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (!p || !h_lo || !h_hi || !h_goal) return fail(ctx, MPTG_E_ARG, "null planner argument");
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called first");
    if (ctx->scene.kind == MPTG_SCENE_NONE) return fail(ctx, MPTG_E_STATE, "no scene uploaded");
    if (p->batch < 1 || p->batch > (1 << 20)) return fail(ctx, MPTG_E_RANGE, "batch must be in 1..2^20");
    if (!(p->range > 0)) return fail(ctx, MPTG_E_ARG, "range must be > 0");
    if (!(p->goal_bias >= 0 && p->goal_bias <= 1)) return fail(ctx, MPTG_E_ARG, "goal bias must be in [0,1]");
    if (p->order != MPTG_ORDER_REFERENCE && p->order != MPTG_ORDER_SURROGATE)
        return fail(ctx, MPTG_E_ARG, "unknown check order");
    for (int i = 0; i < ctx->dim; ++i)
        if (!(h_lo[i] <= h_hi[i]) || !std::isfinite(h_lo[i]) || !std::isfinite(h_hi[i]) || !std::isfinite(h_goal[i]))
            return fail(ctx, MPTG_E_ARG, "bounds and goal must be finite with lo <= hi");
    PrrtState& s = ctx->prrt;
    s.p = *p;
    s.star = false;
    for (int i = 0; i < ctx->dim; ++i) {
        s.lo[i] = h_lo[i];
        s.hi[i] = h_hi[i];
        s.goal[i] = h_goal[i];
    }
    s.init = true;
    ctx->coord_bound = std::max(ctx->coord_bound, planner_coord_bound(ctx));
    if (prevalid_eligible(ctx)) {  // the check the first run would make
        int rc2 = tree_validity_refresh(ctx);
        if (rc2) return rc2;
    }
    s.iteration = 0;
    s.goals = 0;
    s.goal_ids.clear();
    s.goal_cost0.clear();
    s.goal_parent0.clear();
    s.stats = mptg_prrt_stats{};
    s.counters_valid = false;
    return MPTG_OK;
}

int mptg_prrt_add_start(mptg_ctx* ctx, const double* h_q) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (!ctx->prrt.init) return fail(ctx, MPTG_E_STATE, "mptg_prrt_init must be called first");
    int32_t par = -1;
    const bool star_ok = ctx->star_ok;
    rc = mptg_tree_insert(ctx, h_q, &par, 1);
    if (rc) return rc;
    // a start is a root: path cost 0 (Edge() at prrt_star.hpp:264-272), no children
    const int64_t id = ctx->size - 1;
    const double zero = 0.0;
    const int32_t none[3] = {-1, -1, -1};
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->cost + id, &zero, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->child + id, &none[0], sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->next + id, &none[1], sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->prev + id, &none[2], sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->star_ok = star_ok;
    ctx->prrt.counters_valid = false;
    return MPTG_OK;
}

int mptg_prrt_run(mptg_ctx* ctx, int64_t target_size, int64_t max_iters, mptg_prrt_stats* out) {
    const uint64_t entry_epoch = ctx ? ctx->epoch : 0;
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    PrrtState& s = ctx->prrt;
    // the graph of the previous run stays valid if no entry point ran since
    if (s.g_exec && s.g_epoch == entry_epoch) s.g_epoch = ctx->epoch;
    if (!s.init) return fail(ctx, MPTG_E_STATE, "mptg_prrt_init must be called first");
    if (ctx->size == 0) return fail(ctx, MPTG_E_EMPTY, "there are no valid initial states");
    if (target_size > (int64_t)INT32_MAX) return fail(ctx, MPTG_E_RANGE, "target size must be < 2^31");
    if (s.star) {
        rc = prrt_star_run(ctx, target_size, max_iters);
        if (out) *out = s.stats;
        return rc;
    }
    const int B = s.p.batch;
    const int D = ctx->dim;
    if ((rc = prrt_reserve(ctx, B))) return rc;
    PrrtBufs& pb = bufs(ctx);
    Counters* dcnt = (Counters*)pb.counters.p;
    if (!s.counters_valid) {
        Counters c{};
        c.size = ctx->size;
        c.goals = s.goals;
        MPTG_HIP(ctx, hipMemcpyAsync(dcnt, &c, sizeof(c), hipMemcpyHostToDevice, ctx->stream));
        s.counters_valid = true;
        s.stats_base = s.stats;
    }
    if (max_iters <= 0 || ctx->size >= target_size) {
        if (out) *out = s.stats;
        return MPTG_OK;
    }
    // every node appended is valid (the planner checked it); with the whole
    // tree known valid and no range limit, samples are checked before the
    // nearest-neighbour search
    if (prevalid_eligible(ctx) && (rc = tree_validity_refresh(ctx))) return rc;
    const bool tree_valid = !ctx->some_invalid && ctx->valid_upto == ctx->size;
    const bool pre = tree_valid && prevalid_eligible(ctx);
    if ((rc = counters_arm(ctx, dcnt, s.iteration, target_size, max_iters))) return rc;
    if ((rc = ensure_pinned(ctx, sizeof(Counters)))) return rc;
    Counters* hc = (Counters*)ctx->pinned;
    PrrtDev pd;
    for (int i = 0; i < D; ++i) {
        pd.lo[i] = s.lo[i];
        pd.hi[i] = s.hi[i];
        pd.goal[i] = s.goal[i];
    }
    // graph mode: `graph_iters` iterations per launch, the device stopping
    // itself at the target or the iteration bound.  A graph holds tree and
    // scratch addresses and by-value arguments, so it is replayed only while
    // the context epoch it was captured at holds (no entry point, no
    // allocation since), for the same tree capacity and iteration count.
    // Before a capture one step-by-step iteration sizes every buffer.
    // host scenes and custom samplers call back into the host within every iteration
    const int G = (ctx->scene.kind == MPTG_SCENE_HOST || s.sampler) ? 0 : s.graph_iters;
    int64_t k = 0;
    bool warm = false;
    while (k < max_iters && ctx->size < target_size) {
        int64_t ran = 1;
        const int64_t span = (int64_t)G * B;
        bool replay = G > 0 && (warm || (s.g_exec && s.g_epoch == ctx->epoch && s.g_iters == G));
        if (replay && (rc = tree_reserve(ctx, ctx->size + span))) break;
        if (!replay) {
            if ((rc = tree_reserve(ctx, ctx->size + B))) break;
            if ((rc = enqueue_iteration(ctx, pd, s.iteration, 0, pre))) break;
            warm = G > 0;
        } else {
            if (!s.g_exec || s.g_epoch != ctx->epoch || s.g_iters != G || s.g_cap != ctx->cap ||
                ctx->size + span > s.g_plan) {
                prrt_graph_drop(ctx);
                const int64_t plan = std::min<int64_t>(ctx->cap, 2 * (ctx->size + span));
                if ((rc = knn_reserve(ctx, B, 1, plan))) break;
                if ((rc = capture_iterations(ctx, pd, G, plan, pre, &s.g_exec))) break;
                s.g_epoch = ctx->epoch;
                s.g_iters = G;
                s.g_cap = ctx->cap;
                s.g_plan = plan;
            }
            hipError_t e = hipGraphLaunch(s.g_exec, ctx->stream);
            if (e != hipSuccess) {
                rc = fail(ctx, MPTG_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
                break;
            }
            s.stats.graph_launches++;
            ran = -1;  // read back below
        }
        if (hipMemcpyAsync(hc, dcnt, sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = fail(ctx, MPTG_E_HIP, "reading the planner counters failed");
            break;
        }
        if (ran < 0) ran = hc->iteration - s.iteration;
        if (ran <= 0) {
            rc = fail(ctx, MPTG_E_STATE, "a graph launch ran no iteration");
            break;
        }
        ctx->size = hc->size;
        s.iteration += ran;
        k += ran;
        s.stats.iterations += ran;
        s.stats.samples += ran * B;
        if (hc->goals > s.goals) {
            int64_t ng = std::min<int64_t>(hc->goals, 4096);
            s.goal_ids.resize((size_t)ng);
            if (hipMemcpy(s.goal_ids.data(), pb.goal_buf.p, sizeof(int32_t) * ng, hipMemcpyDeviceToHost) !=
                hipSuccess) {
                rc = fail(ctx, MPTG_E_HIP, "reading the goal ids failed");
                break;
            }
            s.goals = hc->goals;
        }
        s.stats.zero_dist = s.stats_base.zero_dist + hc->zero;
        s.stats.invalid_states = s.stats_base.invalid_states + hc->invalid;
        s.stats.link_calls = s.stats_base.link_calls + hc->link_calls;
        s.stats.link_ok = s.stats_base.link_ok + hc->link_ok;
        s.stats.valid_calls = s.stats_base.valid_calls + hc->valid_calls;
        s.stats.goals = s.goals;
        s.stats.size = ctx->size;
    }
    if (rc) {
        prrt_graph_drop(ctx);
        return rc;
    }
    if (tree_valid) ctx->valid_upto = ctx->size;
    if (out) *out = s.stats;
    return MPTG_OK;
}

int mptg_prrt_graph(mptg_ctx* ctx, int iterations_per_launch) {
    if (!ctx) return fail(ctx, MPTG_E_ARG, "null context");
    if (iterations_per_launch < 0 || iterations_per_launch > 1024)
        return fail(ctx, MPTG_E_RANGE, "iterations per graph launch must be in 0..1024");
    ctx->prrt.graph_iters = iterations_per_launch;
    if (iterations_per_launch == 0) prrt_graph_drop(ctx);
    return MPTG_OK;
}

int mptg_prrt_stats_get(mptg_ctx* ctx, mptg_prrt_stats* out) {
    if (!ctx || !out) return fail(ctx, MPTG_E_ARG, "null argument");
    *out = ctx->prrt.stats;
    out->size = ctx->size;
    return MPTG_OK;
}

int mptg_prrt_goals(mptg_ctx* ctx, int32_t* h_ids, int64_t cap, int64_t* n) {
    if (!ctx || !n) return fail(ctx, MPTG_E_ARG, "null argument");
    const auto& g = ctx->prrt.goal_ids;
    *n = ctx->prrt.goals;
    for (int64_t i = 0; i < cap && i < (int64_t)g.size(); ++i) h_ids[i] = g[i];
    return MPTG_OK;
}

int mptg_prrt_sampler(mptg_ctx* ctx, mptg_sample_fn fn, void* user) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    ctx->prrt.sampler = fn;
    ctx->prrt.sampler_user = fn ? user : nullptr;
    return MPTG_OK;
}

int mptg_prrt_goal_edges(mptg_ctx* ctx, double* h_cost, int32_t* h_parent, int64_t cap, int64_t* n) {
    if (!ctx || !n) return fail(ctx, MPTG_E_ARG, "null argument");
    const PrrtState& s = ctx->prrt;
    if (!s.star) return fail(ctx, MPTG_E_STATE, "goal edges are kept for PRRT* trees");
    *n = (int64_t)s.goal_cost0.size();
    for (int64_t i = 0; i < cap && i < *n; ++i) {
        if (h_cost) h_cost[i] = s.goal_cost0[(size_t)i];
        if (h_parent) h_parent[i] = s.goal_parent0[(size_t)i];
    }
    return MPTG_OK;
}

}  // extern "C"
