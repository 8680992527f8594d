// Device kernels shared by the batched PRRT and PRRT* loops: sampling and
// steering (src/mpt/impl/prrt/prrt.hpp:377-440,
// src/mpt/impl/prrt_star/prrt_star.hpp:470-540).
#pragma once
#include "sample_hd.hpp"
#include "space_dev.hpp"

namespace mptg {
namespace {

constexpr int kBlock = 256;

// ST_SKIP: a custom sampler gave no sample for the slot (counted nowhere)
enum : uint8_t { ST_ZERO = 0, ST_INVALID = 1, ST_LINK_FAIL = 2, ST_OK = 3, ST_SKIP = 4 };

// Planner counters, resident on the device.  `iteration` is the index of the
// next iteration; `done` turns every planner kernel of an iteration into a
// no-op once size >= target or iteration >= iter_end, so a captured graph of
// several iterations never runs past where the step-by-step loop stops.
struct Counters {
    int64_t size;
    int64_t goals;
    int64_t zero, invalid, link_calls, link_ok, valid_calls;
    int64_t iteration, target, iter_end;
    int32_t done;
    // size, goals and done as the iteration's append step found them
    // (prrt_count writes them, prrt_write reads them while its last block
    // commits the new size)
    int32_t done_seen;
    int64_t size_seen, goals_seen;
};

// The counters' device block: Counters at 0, PRRT*'s StarCounters at
// kStarCounterOffset, so that one copy brings both to the host.
constexpr size_t kStarCounterOffset = 256;
constexpr size_t kCounterBlock = 512;
static_assert(sizeof(Counters) <= kStarCounterOffset, "Counters fit below the PRRT* counters");

struct PrrtDev {
    double lo[kMaxDim], hi[kMaxDim], goal[kMaxDim];
};

// Threads per sample of prrt_sample: one per coordinate in the boxes (each
// coordinate is its own counter of the stream), one per state for SO(3) x R^3.
__host__ __device__ inline int sample_threads(int metric, int D) { return metric == kSampleMetricSE3 ? 1 : D; }
inline unsigned sample_blocks(int metric, int D, int B) {
    return (unsigned)(((int64_t)B * sample_threads(metric, D) + kBlock - 1) / kBlock);
}

// it < 0: the iteration index is the device counter's (graph replays).
// Thread t draws coordinate t % D of sample t / D (boxes; coalesced
// stores), or the whole state for SO(3) x R^3.  zero_word (may be null): a
// counter of the iteration's later kernels, cleared here so that it needs no
// launch of its own.
__global__ __launch_bounds__(kBlock) void prrt_sample(PrrtDev pd, int metric, int D, uint64_t seed, double goal_bias,
                                                      int64_t it, int B, const Counters* __restrict__ cnt,
                                                      double* __restrict__ q, uint32_t* __restrict__ zero_word) {
    const int per = sample_threads(metric, D);
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (zero_word && t == 0) *zero_word = 0u;
    if (t >= (int64_t)B * per || cnt->done) return;
    const int64_t b = t / per;
    if (it < 0) it = cnt->iteration;
    const uint64_t sidx = (uint64_t)it * (uint64_t)B + (uint64_t)b;
    // only sample slot 0 of an iteration (the reference's worker 0) takes the
    // goal-bias coin, and only while no goal is known (prrt.hpp:365-388)
    const bool bias = b == 0 && cnt->goals == 0 && goal_bias > 0;
    const bool useGoal = bias && uniform_at(seed, sidx, RNG_PURPOSE_BIAS, 0) < goal_bias;
    if (per == D) {
        const int i = (int)(t - b * per);
        q[t] = useGoal ? pd.goal[i]
                       : uniform_in(uniform_at(seed, sidx, RNG_PURPOSE_SAMPLE, (uint32_t)i), pd.lo[i], pd.hi[i]);
        return;
    }
    double v[7];
    if (useGoal) {
        for (int i = 0; i < 7; ++i) v[i] = pd.goal[i];
    } else {
        sample_state(seed, sidx, metric, 7, pd.lo, pd.hi, v);
    }
    for (int i = 0; i < 7; ++i) q[b * 7 + i] = v[i];
}

// A Scenario's own sampler (scenario_sampler.hpp:145-187): its staged
// samples replace the batch's box samples, except slot 0 when it took the
// goal-bias coin (the same draw as prrt_sample); a slot with no sample keeps
// its box sample as a placeholder and is marked in `skip`.
__global__ __launch_bounds__(kBlock) void prrt_take_samples(int D, uint64_t seed, double goal_bias, int64_t it, int B,
                                                            const Counters* __restrict__ cnt,
                                                            const double* __restrict__ staged,
                                                            const uint8_t* __restrict__ ok, double* __restrict__ q,
                                                            uint8_t* __restrict__ skip) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= B || cnt->done) return;
    const uint64_t sidx = (uint64_t)it * (uint64_t)B + (uint64_t)b;
    const bool bias = b == 0 && cnt->goals == 0 && goal_bias > 0;
    if (bias && uniform_at(seed, sidx, RNG_PURPOSE_BIAS, 0) < goal_bias) {
        skip[b] = 0;
        return;
    }
    if (!ok[b]) {
        skip[b] = 1;
        return;
    }
    skip[b] = 0;
    for (int i = 0; i < D; ++i) q[(int64_t)b * D + i] = staged[(int64_t)b * D + i];
}

// After the steer step: slots without a sample take no part (addSample on
// an empty std::optional returns at once, prrt.hpp:400-403).
__global__ __launch_bounds__(kBlock) void prrt_apply_skip(int B, const uint8_t* __restrict__ skip,
                                                          uint8_t* __restrict__ status, uint8_t* __restrict__ linkmask,
                                                          const Counters* __restrict__ cnt) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= B || cnt->done || !skip[b]) return;
    status[b] = ST_SKIP;
    linkmask[b] = 0;
}

template <int D>
__global__ __launch_bounds__(kBlock) void prrt_steer(SpaceArgs sp, SceneArgs sc, double range, int B,
                                                     const double* __restrict__ x64, const double* __restrict__ q,
                                                     const int32_t* __restrict__ near,
                                                     const double* ndist, double* __restrict__ from,
                                                     double* __restrict__ to, uint8_t* __restrict__ status,
                                                     uint8_t* __restrict__ linkmask, double* dnear,
                                                     const Counters* __restrict__ cnt) {
    constexpr int DA = D > 0 ? D : kMaxDim;
    const int dd = D > 0 ? D : sp.dim;
    int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= B) return;
    if (cnt->done) {  // nothing to link in an iteration past the stop
        linkmask[b] = 0;
        if (dnear) dnear[b] = ndist[b];
        return;
    }
    double d = ndist[b];
    int32_t nn = near[b];
    if (dnear) dnear[b] = d;  // replaced below when the step is cut to range
    linkmask[b] = 0;
    if (d == 0 || nn < 0) {
        status[b] = ST_ZERO;
        return;
    }
    double a[DA], r[DA], s[DA];
#pragma unroll
    for (int i = 0; i < DA; ++i)
        if (i < dd) {
            a[i] = x64[(int64_t)nn * dd + i];
            r[i] = q[(int64_t)b * dd + i];
        }
    if (d > range) {
        space_interp_t<D>(sp, a, r, range / d, s);
        // PRRT* carries d(near, new) on into the cost (prrt_star.hpp:535)
        if (dnear) dnear[b] = space_dist_t<D>(sp, a, s);
    } else {
#pragma unroll
        for (int i = 0; i < DA; ++i)
            if (i < dd) s[i] = r[i];
    }
#pragma unroll
    for (int i = 0; i < DA; ++i)
        if (i < dd) {
            from[(int64_t)b * dd + i] = a[i];
            to[(int64_t)b * dd + i] = s[i];
        }
    // mesh scenes: valid(new) is a whole-wave check, done by the filter
    // kernel that runs next over the masked samples; host scenes: done by the
    // caller's valid() over the masked samples
    if (sc.kind != MPTG_SCENE_MESH && sc.kind != MPTG_SCENE_HOST && !state_valid_t<D>(sc, s)) {
        status[b] = ST_INVALID;
        return;
    }
    status[b] = ST_LINK_FAIL;  // until link() says otherwise
    linkmask[b] = 1;
}

// prrt_steer for arm scenes with D = nlinks: kArmLanes lanes per sample
// share its disk tests (arm_ok_split); lane 0 of each group writes.  Same
// outputs as prrt_steer.
constexpr int kArmLanes = 8;
constexpr int kPrevalidLanes = 4;  // lanes per sample in prrt_prevalid on arm scenes

template <int D>
__global__ __launch_bounds__(kBlock) void prrt_steer_arm(SpaceArgs sp, SceneArgs sc, double range, int B,
                                                         const double* __restrict__ x64, const double* __restrict__ q,
                                                         const int32_t* __restrict__ near,
                                                         const double* ndist, double* __restrict__ from,
                                                         double* __restrict__ to, uint8_t* __restrict__ status,
                                                         uint8_t* __restrict__ linkmask, double* dnear,
                                                         const Counters* __restrict__ cnt) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int b = (int)(t / kArmLanes);
    const bool lead = (t % kArmLanes) == 0;
    const bool inb = b < B;
    bool need = false;
    double s[D];
#pragma unroll
    for (int i = 0; i < D; ++i) s[i] = 0.0;
    if (inb && cnt->done) {
        if (lead) {
            linkmask[b] = 0;
            if (dnear) dnear[b] = ndist[b];
        }
    } else if (inb) {
        const double d = ndist[b];
        const int32_t nn = near[b];
        if (dnear && lead) dnear[b] = d;  // replaced below when the step is cut to range
        if (d == 0 || nn < 0) {
            if (lead) {
                linkmask[b] = 0;
                status[b] = ST_ZERO;
            }
        } else {
            double a[D], r[D];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                a[i] = x64[(int64_t)nn * D + i];
                r[i] = q[(int64_t)b * D + i];
            }
            if (d > range) {
                space_interp_t<D>(sp, a, r, range / d, s);
                if (dnear && lead) dnear[b] = space_dist_t<D>(sp, a, s);
            } else {
#pragma unroll
                for (int i = 0; i < D; ++i) s[i] = r[i];
            }
            if (lead) {
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    from[(int64_t)b * D + i] = a[i];
                    to[(int64_t)b * D + i] = s[i];
                }
            }
            need = true;
        }
    }
    const bool ok = arm_ok_split<D, kArmLanes>(sc, [&](int i) { return s[i]; }, need);
    if (need && lead) {
        status[b] = ok ? ST_LINK_FAIL : ST_INVALID;  // ST_LINK_FAIL until link() says otherwise
        linkmask[b] = ok ? 1 : 0;
    }
}

// ---- valid() before nearest (PRRT with no range limit) ---------------------
// With an infinite range the new state is the sample itself
// (prrt.hpp:430-434 never interpolates), so valid(new) does not depend on
// the nearest node and an invalid sample's outcome is fixed before the
// search: rejected as invalid, unless it is at distance 0 from a tree node
// (:427), which needs a tree node equal to it, hence an invalid tree node.
// The host takes this path only while every tree node is known valid; then
// only the valid samples are searched.  prrt_prevalid checks each sample
// (L lanes per sample: the arm's split disk tests, else one), marks invalid
// ones and lists the valid ones compactly (queries qv, sample ids vidx,
// count *vcnt; list order is immaterial: every query is independent).
template <int D, int L>
__global__ __launch_bounds__(kBlock) void prrt_prevalid(SceneArgs sc, int dim, int B, const double* __restrict__ q,
                                                        uint8_t* __restrict__ status, uint8_t* __restrict__ linkmask,
                                                        double* __restrict__ qv, int32_t* __restrict__ vidx,
                                                        uint32_t* __restrict__ vcnt,
                                                        const Counters* __restrict__ cnt) {
    constexpr int DA = D > 0 ? D : kMaxDim;
    const int dd = D > 0 ? D : dim;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int b = (int)(t / L);
    const bool lead = (t % L) == 0;
    const bool inb = b < B;
    if (cnt->done) {  // nothing to link in an iteration past the stop
        if (inb && lead) linkmask[b] = 0;
        return;
    }
    double s[DA];
#pragma unroll
    for (int i = 0; i < DA; ++i) s[i] = (inb && i < dd) ? q[(int64_t)b * dd + i] : 0.0;
    bool ok;
    if constexpr (L > 1) ok = arm_ok_split<D, L>(sc, [&](int i) { return s[i]; }, inb);
    else ok = inb && state_valid_t<D>(sc, s);
    const bool list = inb && lead && ok;
    if (inb && lead && !ok) {
        status[b] = ST_INVALID;
        linkmask[b] = 0;
    }
    // one atomic per wave for its listed samples, ranks by ballot
    const uint64_t m = __ballot(list);
    const int lane = (int)(threadIdx.x & 63);
    uint32_t base = 0;
    if (m) {
        if (lane == __builtin_ctzll(m)) base = atomicAdd(vcnt, (uint32_t)__popcll(m));
        base = __shfl(base, __builtin_ctzll(m));
    }
    if (list) {
        const uint32_t k = base + (uint32_t)__popcll(m & (lane ? (~0ull >> (64 - lane)) : 0ull));
        vidx[k] = b;
#pragma unroll
        for (int i = 0; i < DA; ++i)
            if (i < dd) qv[(int64_t)k * dd + i] = s[i];
    }
}

// The steer step for the listed samples (k < *vcnt): sample b = vidx[k] with
// its nearest node kn_i[k] at kn_d[k]; with no range limit the new state is
// the sample, already valid.  Writes what prrt_steer writes (and near[b]).
template <int D>
__global__ __launch_bounds__(kBlock) void prrt_steer_listed(int dim, int B, const double* __restrict__ x64,
                                                            const double* __restrict__ qv,
                                                            const int32_t* __restrict__ vidx,
                                                            const uint32_t* __restrict__ vcnt,
                                                            const int32_t* __restrict__ kn_i,
                                                            const double* __restrict__ kn_d, int32_t* __restrict__ near,
                                                            double* __restrict__ from, double* __restrict__ to,
                                                            uint8_t* __restrict__ status, uint8_t* __restrict__ linkmask,
                                                            const Counters* __restrict__ cnt) {
    const int dd = D > 0 ? D : dim;
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (cnt->done || k >= (int64_t)*vcnt || k >= B) return;
    const int b = vidx[k];
    const double d = kn_d[k];
    const int32_t nn = kn_i[k];
    near[b] = nn;
    if (d == 0 || nn < 0) {  // prrt.hpp:427-428
        status[b] = ST_ZERO;
        linkmask[b] = 0;
        return;
    }
    for (int i = 0; i < dd; ++i) {
        from[(int64_t)b * dd + i] = x64[(int64_t)nn * dd + i];
        to[(int64_t)b * dd + i] = qv[k * dd + i];
    }
    status[b] = ST_LINK_FAIL;  // until link() says otherwise
    linkmask[b] = 1;
}

// A Scenario's own sampler: the host callback fills the iteration's batch,
// which replaces the box samples in d_q (prrt_take_samples); *skip is the
// device array of slots without a sample.  Synchronous.
inline int take_host_samples(mptg_ctx* ctx, int64_t it, double* d_q, const Counters* dcnt, uint8_t** skip) {
    PrrtState& s = ctx->prrt;
    const int B = s.p.batch, D = ctx->dim;
    int rc;
    if ((rc = ensure(ctx, s.hs_q, sizeof(double) * (size_t)B * D))) return rc;
    if ((rc = ensure(ctx, s.hs_ok, (size_t)B))) return rc;
    if ((rc = ensure(ctx, s.hs_skip, (size_t)B))) return rc;
    s.hs_host_q.assign((size_t)B * D, 0.0);
    s.hs_host_ok.assign((size_t)B, 1);
    const int cb = s.sampler(s.sampler_user, s.p.seed, it * (int64_t)B, B, s.hs_host_q.data(), s.hs_host_ok.data());
    if (cb) return fail(ctx, MPTG_E_ARG, "sampler callback returned " + std::to_string(cb));
    MPTG_HIP(ctx, hipMemcpy(s.hs_q.p, s.hs_host_q.data(), sizeof(double) * s.hs_host_q.size(), hipMemcpyHostToDevice));
    MPTG_HIP(ctx, hipMemcpy(s.hs_ok.p, s.hs_host_ok.data(), s.hs_host_ok.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(prrt_take_samples, dim3((unsigned)((B + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, D,
                       s.p.seed, s.p.goal_bias, it, B, dcnt, (const double*)s.hs_q.p, (const uint8_t*)s.hs_ok.p, d_q,
                       (uint8_t*)s.hs_skip.p);
    MPTG_HIP(ctx, hipGetLastError());
    *skip = (uint8_t*)s.hs_skip.p;
    return MPTG_OK;
}

// Slots without a sample leave the batch after the steer step.
inline int apply_skip(mptg_ctx* ctx, const uint8_t* skip, uint8_t* status, uint8_t* mask, const Counters* dcnt) {
    const int B = ctx->prrt.p.batch;
    hipLaunchKernelGGL(prrt_apply_skip, dim3((unsigned)((B + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, B,
                       skip, status, mask, dcnt);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// Launches the steer kernel for a batch of B samples (arm scenes with
// D = nlinks in {5, 10, 20}: the split-lane kernel).
template <int D>
void launch_steer(mptg_ctx* ctx, double range, int B, const double* q, const int32_t* near, const double* ndist,
                  double* from, double* to, uint8_t* status, uint8_t* linkmask, double* dnear, const Counters* cnt) {
    if constexpr (D == 5 || D == 10 || D == 20) {
        if (ctx->scene.kind == MPTG_SCENE_ARM && ctx->scene.nlinks == D) {
            const unsigned g = (unsigned)(((int64_t)B * kArmLanes + kBlock - 1) / kBlock);
            hipLaunchKernelGGL((prrt_steer_arm<D>), dim3(g), dim3(kBlock), 0, ctx->stream, space_args(ctx),
                               scene_args(ctx), range, B, (const double*)ctx->x64, q, near, ndist, from, to, status,
                               linkmask, dnear, cnt);
            return;
        }
    }
    const unsigned gb = (unsigned)((B + kBlock - 1) / kBlock);
    hipLaunchKernelGGL((prrt_steer<D>), dim3(gb), dim3(kBlock), 0, ctx->stream, space_args(ctx), scene_args(ctx),
                       range, B, (const double*)ctx->x64, q, near, ndist, from, to, status, linkmask, dnear, cnt);
}

}  // namespace
}  // namespace mptg
