// Launch-shape autotuner.  Each knob only changes how the same work is split
// over the chip (kNN tree splits, grids of the persistent link kernels,
// planner iterations per graph launch), never a result, so the tuner picks
// by time alone: HIP events around device work, the host clock around whole
// planner iterations (their launch and synchronisation cost is the point).
#include <algorithm>
#include <chrono>
#include <vector>

#include "internal.hpp"

namespace mptg {

namespace {

constexpr int kReps = 3;

// Median event time (ms) of kReps calls of fn on the context stream, after
// one untimed call that also makes every allocation.
template <typename F>
int time_median(mptg_ctx* ctx, F&& fn, float* ms) {
    int rc = fn();
    if (rc) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    MPTG_HIP(ctx, hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(ctx, MPTG_E_HIP, "hipEventCreate failed");
    }
    std::vector<float> t;
    for (int r = 0; r < kReps && rc == MPTG_OK; ++r) {
        float x = 0;
        if (hipEventRecord(a, ctx->stream) != hipSuccess) rc = fail(ctx, MPTG_E_HIP, "hipEventRecord failed");
        if (rc == MPTG_OK) rc = fn();
        if (rc == MPTG_OK && (hipEventRecord(b, ctx->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                              hipEventElapsedTime(&x, a, b) != hipSuccess))
            rc = fail(ctx, MPTG_E_HIP, "event timing failed");
        t.push_back(x);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return MPTG_OK;
}

// Tries each candidate through set(value) + the timed call, keeps the
// fastest (the earliest on a tie) and leaves it set.
template <typename Set, typename Call>
int sweep(mptg_ctx* ctx, int knob, const std::vector<int>& cand, Set&& set, Call&& call, mptg_tune_report* rep) {
    mptg_tune_report r{};
    r.knob = knob;
    r.ncand = (int32_t)cand.size();
    int best = -1;
    for (size_t i = 0; i < cand.size(); ++i) {
        set(cand[i]);
        float ms = 0;
        int rc = time_median(ctx, call, &ms);
        if (rc) return rc;
        r.value[i] = cand[i];
        r.ms[i] = ms;
        if (best < 0 || ms < r.ms[best]) best = (int)i;
    }
    set(cand[best]);
    r.chosen = cand[best];
    if (rep) *rep = r;
    return MPTG_OK;
}

bool in_range(int v) { return v >= 1 && v <= 256; }

}  // namespace

}  // namespace mptg

using namespace mptg;

extern "C" {

int mptg_tune_get(mptg_ctx* ctx, mptg_tune_params* out) {
    if (!ctx || !out) return fail(ctx, MPTG_E_ARG, "null argument");
    *out = ctx->tune;
    return MPTG_OK;
}

int mptg_tune_set(mptg_ctx* ctx, const mptg_tune_params* p) {
    int rc = check_device_ready(ctx);  // also retires captured graphs, which hold the old grids
    if (rc) return rc;
    if (!p) return fail(ctx, MPTG_E_ARG, "null argument");
    if (!in_range(p->knn_blocks_per_cu) || !in_range(p->dmv_long_waves_per_cu) || !in_range(p->mesh_waves_per_cu))
        return fail(ctx, MPTG_E_RANGE, "launch shapes must be in 1..256");
    if (p->arm_tail_waves < 1 || p->arm_tail_waves > 8)
        return fail(ctx, MPTG_E_RANGE, "arm continuation waves per edge must be in 1..8");
    if (p->arm_cull != 0 && p->arm_cull != 1) return fail(ctx, MPTG_E_RANGE, "arm culling is 0 or 1");
    if (p->mesh_edge_waves < 1 || p->mesh_edge_waves > 32)
        return fail(ctx, MPTG_E_RANGE, "mesh waves per edge must be in 1..32");
    if (p->knn_select != 0 && p->knn_select != 1) return fail(ctx, MPTG_E_RANGE, "knn_select is 0 or 1");
    if (p->knn_grid != 0 && p->knn_grid != 1) return fail(ctx, MPTG_E_RANGE, "knn_grid is 0 or 1");
    if (p->dmv_edge_lanes < 1 || p->dmv_edge_lanes > 64 || (p->dmv_edge_lanes & (p->dmv_edge_lanes - 1)))
        return fail(ctx, MPTG_E_RANGE, "lanes per DMV edge must be a power of two in 1..64");
    if (p->knn_q8 != 0 && p->knn_q8 != 1) return fail(ctx, MPTG_E_RANGE, "knn_q8 is 0 or 1");
    ctx->tune = *p;
    return MPTG_OK;
}

int mptg_autotune_knn(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, mptg_tune_report* rep) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (nq <= 0 || !d_q) return fail(ctx, MPTG_E_ARG, "autotune_knn needs device queries");
    if (k < 1 || k > kMaxK) return fail(ctx, MPTG_E_RANGE, "k must be in 1..64");
    if (ctx->size <= 0) return fail(ctx, MPTG_E_EMPTY, "the tree is empty");
    DevBuf& out = ctx->scratch[11];
    if ((rc = ensure(ctx, out, (sizeof(double) + sizeof(int32_t)) * (size_t)k * nq))) return rc;
    double* dist = (double*)out.p;
    int32_t* idx = (int32_t*)(dist + (size_t)k * nq);
    return sweep(
        ctx, MPTG_TUNE_KNN, {1, 2, 4, 8, 16, 32, 64}, [&](int v) { ctx->tune.knn_blocks_per_cu = v; },
        [&] { return launch_knn(ctx, d_q, nq, k, __builtin_inf(), idx, dist, ctx->size); }, rep);
}

int mptg_autotune_link(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, mptg_tune_report* rep) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (n <= 0 || !d_from || !d_to) return fail(ctx, MPTG_E_ARG, "autotune_link needs device edges");
    const bool arm_wave = ctx->scene.kind == MPTG_SCENE_ARM && ctx->link_mode == MPTG_LINK_BISECT &&
                          ctx->scene.nlinks == ctx->dim && (ctx->dim == 5 || ctx->dim == 10 || ctx->dim == 20);
    if (ctx->force_generic || ctx->scene.kind == MPTG_SCENE_HOST || !(ctx->link_mode == MPTG_LINK_DMV || arm_wave))
        return fail(ctx, MPTG_E_STATE, "this scene's link kernels have no launch shape to tune");
    DevBuf& out = ctx->scratch[11];
    if ((rc = ensure(ctx, out, (sizeof(uint32_t) + 1) * (size_t)n))) return rc;
    uint32_t* checks = (uint32_t*)out.p;
    uint8_t* verdict = (uint8_t*)(checks + n);
    auto call = [&] {
        return launch_link_masked(ctx, d_from, d_to, n, MPTG_ORDER_REFERENCE, 0, nullptr, verdict, checks);
    };
    if (arm_wave) {
        if ((rc = sweep(ctx, MPTG_TUNE_ARM_CULL, {0, 1}, [&](int v) { ctx->tune.arm_cull = v; }, call, nullptr)))
            return rc;
        return sweep(ctx, MPTG_TUNE_ARM_TAIL, {1, 2, 4, 8}, [&](int v) { ctx->tune.arm_tail_waves = v; }, call, rep);
    }
    if (ctx->scene.kind == MPTG_SCENE_MESH) {
        if ((rc = sweep(ctx, MPTG_TUNE_MESH_EDGE, {1, 2, 4, 8, 16, 32},
                        [&](int v) { ctx->tune.mesh_edge_waves = v; }, call, nullptr)))
            return rc;
        return sweep(ctx, MPTG_TUNE_MESH, {8, 16, 32, 64, 128}, [&](int v) { ctx->tune.mesh_waves_per_cu = v; },
                     call, rep);
    }
    if ((rc = sweep(ctx, MPTG_TUNE_DMV_LANES, {1, 2, 4, 8, 16, 32, 64},
                    [&](int v) { ctx->tune.dmv_edge_lanes = v; }, call, nullptr)))
        return rc;
    return sweep(ctx, MPTG_TUNE_DMV_LONG, {2, 4, 8, 16, 32, 64},
                 [&](int v) { ctx->tune.dmv_long_waves_per_cu = v; }, call, rep);
}

int mptg_autotune_prrt(mptg_ctx* ctx, int64_t target_size, mptg_tune_report* rep) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    PrrtState& s = ctx->prrt;
    if (!s.init) return fail(ctx, MPTG_E_STATE, "mptg_prrt_init must be called first");
    if (ctx->size == 0) return fail(ctx, MPTG_E_EMPTY, "there are no valid initial states");
    if (ctx->scene.kind == MPTG_SCENE_HOST)
        return fail(ctx, MPTG_E_STATE, "host scenes run step by step: no launch shape to tune");
    // the knob: plain PRRT, iterations per graph launch; PRRT* (step by step),
    // the scene's per-edge link split (waves per mesh edge or per continued
    // arm edge), whose best value depends on the planner's own edge mix
    int knob;
    std::vector<int> cand;
    int32_t* field;
    int timed;
    if (!s.star) {
        knob = MPTG_TUNE_GRAPH;
        cand = {0, 2, 4, 8, 16};
        field = &s.graph_iters;
        timed = 16;
    } else if (ctx->scene.kind == MPTG_SCENE_MESH) {
        knob = MPTG_TUNE_MESH_EDGE;
        cand = {1, 2, 4, 8, 16};
        field = &ctx->tune.mesh_edge_waves;
        timed = 4;
    } else if (ctx->scene.kind == MPTG_SCENE_ARM && ctx->link_mode == MPTG_LINK_BISECT) {
        knob = MPTG_TUNE_ARM_TAIL;
        cand = {1, 2, 4, 8};
        field = &ctx->tune.arm_tail_waves;
        timed = 4;
    } else {
        return fail(ctx, MPTG_E_STATE, "this PRRT* scene has no launch shape to tune");
    }
    // each candidate per round: one untimed call (a graph capture, or the
    // first launches of the new shape), then `timed` timed iterations;
    // rounds alternate the candidate order so that tree growth favours none;
    // median over rounds.  Every iteration counts, as in mptg_prrt_run.
    constexpr int kRounds = 3;
    std::vector<std::vector<double>> per(cand.size());
    const int32_t keep = *field;
    bool stop = false;
    for (int round = 0; round < kRounds && !stop; ++round) {
        for (size_t j = 0; j < cand.size() && !stop; ++j) {
            const size_t i = round % 2 == 0 ? j : cand.size() - 1 - j;
            *field = cand[i];
            ++ctx->epoch;  // a changed launch shape retires captured graphs
            if ((rc = mptg_prrt_run(ctx, target_size, knob == MPTG_TUNE_GRAPH ? std::max(cand[i], 1) : 1,
                                    nullptr)))
                break;
            const int64_t it0 = s.iteration;
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = mptg_prrt_run(ctx, target_size, timed, nullptr))) break;
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const int64_t ran = s.iteration - it0;
            if (ran > 0) per[i].push_back(sec * 1e3 / (double)ran);
            stop = ctx->size >= target_size;
        }
        if (rc) break;
    }
    if (rc) {
        *field = keep;
        return rc;
    }
    mptg_tune_report r{};
    r.knob = knob;
    int pick = -1;
    for (size_t i = 0; i < cand.size(); ++i) {
        double best = -1.0;
        if (!per[i].empty()) {
            std::sort(per[i].begin(), per[i].end());
            best = per[i][per[i].size() / 2];
        }
        r.value[r.ncand] = cand[i];
        r.ms[r.ncand] = (float)best;
        ++r.ncand;
        if (best >= 0 && (pick < 0 || best < r.ms[pick])) pick = (int)i;
    }
    *field = pick >= 0 ? cand[pick] : keep;
    ++ctx->epoch;
    r.chosen = *field;
    if (rep) *rep = r;
    return MPTG_OK;
}

}  // extern "C"
