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

#include "prrt_dev.hpp"

namespace mptg {

namespace {

// Single-block stable compaction: accepted samples get consecutive node ids
// in sample order; also accumulates the counters.
__global__ __launch_bounds__(1024) void prrt_append(SpaceArgs sp, PrrtDev pd, double goal_radius, int B, int dp,
                                                    const double* __restrict__ to, const int32_t* __restrict__ near,
                                                    uint8_t* __restrict__ status, const uint8_t* __restrict__ verdict,
                                                    const uint32_t* __restrict__ checks, double* __restrict__ x64,
                                                    float* __restrict__ x32, int32_t* __restrict__ parent,
                                                    Counters* __restrict__ cnt, int32_t* __restrict__ goal_ids,
                                                    int64_t goal_cap) {
    __shared__ int wsum[16];
    __shared__ int gsum[16];
    __shared__ int64_t base_s, gbase_s;
    __shared__ int64_t zero_s, inv_s, lc_s, ok_s, vc_s;
    const int D = sp.dim;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    if (t == 0) {
        base_s = cnt->size;
        gbase_s = cnt->goals;
        zero_s = inv_s = lc_s = ok_s = vc_s = 0;
    }
    __syncthreads();
    int64_t lzero = 0, linv = 0, llc = 0, lok = 0, lvc = 0;
    for (int c0 = 0; c0 < B; c0 += 1024) {
        int b = c0 + t;
        bool ok = false, goal = false;
        if (b < B) {
            uint8_t st = status[b];
            if (st == ST_LINK_FAIL) {
                ++llc;
                lvc += checks[b];
                if (verdict[b]) {
                    st = ST_OK;
                    status[b] = ST_OK;
                }
            }
            lzero += st == ST_ZERO;
            linv += st == ST_INVALID;
            if (st == ST_OK) {
                ++lok;
                ok = true;
                goal = space_dist(sp, to + (int64_t)b * D, pd.goal) <= goal_radius;
            }
        }
        // wave-level ballots then block prefix over 16 waves
        uint64_t m = __ballot(ok);
        uint64_t gm = __ballot(goal);
        uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
        int wrank = __popcll(m & below);
        int grank = __popcll(gm & below);
        if (lane == 0) {
            wsum[wid] = __popcll(m);
            gsum[wid] = __popcll(gm);
        }
        __syncthreads();
        int wbase = 0, gwbase = 0, tot = 0, gtot = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wid) {
                wbase += wsum[w];
                gwbase += gsum[w];
            }
            tot += wsum[w];
            gtot += gsum[w];
        }
        if (ok) {
            int64_t id = base_s + wbase + wrank;
            for (int i = 0; i < D; ++i) {
                double v = to[(int64_t)b * D + i];
                x64[id * D + i] = v;
                x32[id * dp + i] = (float)v;
            }
            parent[id] = near[b];
            if (goal) {
                int64_t g = gbase_s + gwbase + grank;
                if (g < goal_cap) goal_ids[g] = (int32_t)id;
            }
        }
        __syncthreads();
        if (t == 0) {
            base_s += tot;
            gbase_s += gtot;
        }
        __syncthreads();
    }
    atomicAdd((unsigned long long*)&zero_s, (unsigned long long)lzero);
    atomicAdd((unsigned long long*)&inv_s, (unsigned long long)linv);
    atomicAdd((unsigned long long*)&lc_s, (unsigned long long)llc);
    atomicAdd((unsigned long long*)&ok_s, (unsigned long long)lok);
    atomicAdd((unsigned long long*)&vc_s, (unsigned long long)lvc);
    __syncthreads();
    if (t == 0) {
        cnt->size = base_s;
        cnt->goals = gbase_s;
        cnt->zero += zero_s;
        cnt->invalid += inv_s;
        cnt->link_calls += lc_s;
        cnt->link_ok += ok_s;
        cnt->valid_calls += vc_s;
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
    if ((rc = ensure(ctx, pb.counters, sizeof(Counters)))) return rc;
    if ((rc = ensure(ctx, pb.goal_buf, sizeof(int32_t) * 4096))) return rc;
    return MPTG_OK;
}

void prrt_free(mptg_ctx* ctx) {
    PrrtBufs& pb = bufs(ctx);
    DevBuf* all[] = {&pb.q,       &pb.near,    &pb.ndist,  &pb.from,    &pb.to,       &pb.status,  &pb.mask,
                     &pb.verdict, &pb.checks,  &pb.counters, &pb.goal_buf, &pb.dnear, &pb.sidx,   &pb.qs,
                     &pb.nbh_idx, &pb.nbh_d,   &pb.marked, &pb.pc,      &pb.cpar,     &pb.cp_from, &pb.cp_to,
                     &pb.cp_ver,  &pb.cp_chk,  &pb.pair_cp, &pb.rw_from, &pb.rw_to,    &pb.rw_ver,  &pb.rw_chk,
                     &pb.pair_rw, &pb.front,   &pb.pend,   &pb.star_cnt};
    for (DevBuf* b : all) {
        (void)hipFree(b->p);
        b->p = nullptr;
        b->bytes = 0;
    }
}

int tree_reserve(mptg_ctx* ctx, int64_t cap);
int prrt_star_run(mptg_ctx* ctx, int64_t target_size, int64_t max_iters);

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
    PrrtState& s = ctx->prrt;
    s.p = *p;
    s.star = false;
    double bound = ctx->coord_bound;
    for (int i = 0; i < ctx->dim; ++i) {
        if (!(h_lo[i] <= h_hi[i]) || !std::isfinite(h_lo[i]) || !std::isfinite(h_hi[i]))
            return fail(ctx, MPTG_E_ARG, "bounds must be finite with lo <= hi");
        s.lo[i] = h_lo[i];
        s.hi[i] = h_hi[i];
        s.goal[i] = h_goal[i];
        bound = std::max(bound, std::max(std::fabs(h_lo[i]), std::fabs(h_hi[i])));
        bound = std::max(bound, std::fabs(h_goal[i]));
    }
    if (ctx->metric == MPTG_METRIC_SE3) bound = std::max(bound, 1.0);
    ctx->coord_bound = bound * (1.0 + 1e-12);
    s.init = true;
    s.iteration = 0;
    s.goals = 0;
    s.goal_ids.clear();
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
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    PrrtState& s = ctx->prrt;
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
    if ((rc = ensure_pinned(ctx, sizeof(Counters)))) return rc;
    Counters* hc = (Counters*)ctx->pinned;
    PrrtDev pd;
    for (int i = 0; i < D; ++i) {
        pd.lo[i] = s.lo[i];
        pd.hi[i] = s.hi[i];
        pd.goal[i] = s.goal[i];
    }
    const unsigned gb = (unsigned)((B + kBlock - 1) / kBlock);
    for (int64_t k = 0; k < max_iters && ctx->size < target_size; ++k) {
        if ((rc = tree_reserve(ctx, ctx->size + B))) return rc;
        const int64_t it = s.iteration;
        hipLaunchKernelGGL(prrt_sample, dim3(gb), dim3(kBlock), 0, ctx->stream, pd, ctx->metric, D, s.p.seed, s.p.goal_bias, it,
                           B, (const Counters*)dcnt, (double*)pb.q.p);
        MPTG_HIP(ctx, hipGetLastError());
        if ((rc = launch_knn(ctx, (const double*)pb.q.p, B, 1, __builtin_inf(), (int32_t*)pb.near.p,
                             (double*)pb.ndist.p, ctx->size)))
            return rc;
        dispatch_dim(D, [&]<int DD>() {
            hipLaunchKernelGGL((prrt_steer<DD>), dim3(gb), dim3(kBlock), 0, ctx->stream, space_args(ctx),
                               scene_args(ctx), s.p.range, B, (const double*)ctx->x64, (const double*)pb.q.p,
                               (const int32_t*)pb.near.p, (const double*)pb.ndist.p, (double*)pb.from.p,
                               (double*)pb.to.p, (uint8_t*)pb.status.p, (uint8_t*)pb.mask.p,
                               (double*)nullptr);
        });
        MPTG_HIP(ctx, hipGetLastError());
        if (ctx->scene.kind == MPTG_SCENE_MESH &&
            (rc = launch_mesh_filter(ctx, (const double*)pb.to.p, B, (uint8_t*)pb.mask.p, (uint8_t*)pb.status.p,
                                     ST_INVALID)))
            return rc;
        if ((rc = launch_link_masked(ctx, (const double*)pb.from.p, (const double*)pb.to.p, B, s.p.order, 1,
                                     (const uint8_t*)pb.mask.p, (uint8_t*)pb.verdict.p, (uint32_t*)pb.checks.p)))
            return rc;
        hipLaunchKernelGGL(prrt_append, dim3(1), dim3(1024), 0, ctx->stream, space_args(ctx), pd, s.p.goal_radius, B,
                           pitch32(D), (const double*)pb.to.p, (const int32_t*)pb.near.p, (uint8_t*)pb.status.p,
                           (const uint8_t*)pb.verdict.p, (const uint32_t*)pb.checks.p, ctx->x64, ctx->x32,
                           ctx->parent, dcnt, (int32_t*)pb.goal_buf.p, (int64_t)4096);
        MPTG_HIP(ctx, hipGetLastError());
        MPTG_HIP(ctx, hipMemcpyAsync(hc, dcnt, sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx->size = hc->size;
        s.iteration++;
        s.stats.iterations++;
        s.stats.samples += B;
        if (hc->goals > s.goals) {
            int64_t ng = std::min<int64_t>(hc->goals, 4096);
            s.goal_ids.resize((size_t)ng);
            MPTG_HIP(ctx, hipMemcpy(s.goal_ids.data(), pb.goal_buf.p, sizeof(int32_t) * ng, hipMemcpyDeviceToHost));
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
    if (out) *out = s.stats;
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

}  // extern "C"
