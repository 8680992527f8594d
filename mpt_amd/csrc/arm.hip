// Planar-arm link() on the GPU: one wave per edge, 64 waypoints per round.
//
// Reference semantics (demo/link_manipulator_scenario.hpp:118-138): valid(a),
// valid(b), then the pre-order of recursive midpoints (a+b)/2 down to
// L-inf(a-b) < tol; the verdict is the AND over that fixed set, the check
// count is the position of the first failing waypoint in the visiting order.
//
// Regular trees.  Each midpoint step moves an endpoint by at most 2^-53 M
// (M = max |coordinate|), so after d levels every interval width is within
// d 2^-52 M of w_max / 2^d.  When
//     w_max / 2^(h-1) - e >= tol   and   w_max / 2^h + e < tol,
//     e = (h + 1) 2^-50 M,
// every node above depth h splits and every node at depth h is a leaf: the
// tree is perfect with 2^h - 1 waypoints and pre-order ordinal -> path is a
// closed form.  Lanes then take 64 consecutive ordinals of the chosen order,
// rebuild their midpoint from (a, b) along the path (bit-identical to the
// reference's recursion), run valid() and the wave ballots for the first
// failure.  Edges whose widths fall within e of tol at depth h-1 or h (never
// observed; kept exact anyway) are walked by one lane with the stackless
// exact walk.
//
// valid() per lane: forward kinematics with the project sincos, then every
// link segment against every disk.  Disks are read with wave-uniform (scalar)
// loads.  A disk whose centre is farther than (link length + r + R) from the
// segment start cannot satisfy dist <= r + R; the test skips it after a
// bound check with a relative margin of 2^-30 that dominates the rounding of
// the exact expression, so the skip never changes a verdict.  Disks that pass
// the bound get the reference expression bit-for-bit.
#include "bisect_dev.hpp"

namespace mptg {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Sort key of a waypoint in surrogate order: ascending key = descending
// score, ties by ascending reference ordinal (-0 counts as +0, NaN last).
__device__ __forceinline__ uint64_t score_key(float s, uint32_t ordinal) {
    uint32_t hi;
    if (s != s) {
        hi = 0xffffffffu;
    } else {
        if (s == 0.0f) s = 0.0f;
        uint32_t u = bits_of(s);
        uint32_t mono = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // ascending in s
        hi = ~mono;
    }
    return ((uint64_t)hi << 32) | ordinal;
}

// Ascending bitonic sort of the wave's 64 * NE keys; element i = e * 64 +
// lane lives in key[e] of that lane, so after the sort rank r = c * 64 + lane
// is key[c] of that lane.  Partners within a lane are registers, partners in
// other lanes come by __shfl_xor.
template <int NE>
__device__ __forceinline__ void wave_bitonic_sort(uint64_t (&key)[NE], int lane) {
    constexpr int N = 64 * NE;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint64_t nk[NE];
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const int i = e * 64 + lane;
                uint64_t other;
                if (j >= 64) {
                    other = key[e ^ (j >> 6)];
                } else {
                    uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)key[e], j);
                    uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(key[e] >> 32), j);
                    other = ((uint64_t)hi << 32) | lo;
                }
                const bool lower = (i & j) == 0, asc = (i & k) == 0;
                const uint64_t mn = key[e] < other ? key[e] : other;
                const uint64_t mx = key[e] < other ? other : key[e];
                nk[e] = (lower == asc) ? mn : mx;
            }
#pragma unroll
            for (int e = 0; e < NE; ++e) key[e] = nk[e];
        }
    }
}

// One edge per wave (the kernel body; COUNT adds the work tally; ORD
// selects the visiting order: reference pre-order, or descending surrogate
// score from scores[e * kScoreStride + ordinal]).
template <int D, bool COUNT, int ORD>
__device__ __forceinline__ void arm_link_edge(const double* __restrict__ from, const double* __restrict__ to, int64_t n,
                                              int known_valid, const uint8_t* __restrict__ mask, double tol,
                                              const SceneArgs& sc, const float* __restrict__ scores,
                                              uint8_t* __restrict__ verdict,
                                              uint32_t* __restrict__ checks, int64_t* __restrict__ irregular,
                                              uint32_t* __restrict__ n_irregular, ArmWork* wk) {
    // no block-wide barriers in this kernel: waves leave independently
    const int lane = lane_id();
    const int w = uniform((int)(threadIdx.x >> 6));
    const int64_t e = (int64_t)blockIdx.x * kWavesPerBlock + w;
    if (e >= n) return;
    if (mask && !mask[e]) return;
    const double* a = from + e * D;  // wave-uniform: scalar loads
    const double* b = to + e * D;
    uint32_t base = 2;
    if (!known_valid) {
        bool okA = true, okB = true;
        if (lane == 0) okA = arm_ok_fn<D, COUNT>(sc, [&](int i) { return a[i]; }, wk);
        if (lane == 1) okB = arm_ok_fn<D, COUNT>(sc, [&](int i) { return b[i]; }, wk);
        uint64_t badA = __ballot(!okA), badB = __ballot(!okB);
        if (badA | badB) {
            if (lane == 0) {
                verdict[e] = 0;
                checks[e] = badA ? 1u : 2u;
            }
            return;
        }
    }
    // tree height from the widest coordinate, and the regularity test
    int h = 0;
    bool regular = false;
    if (!bisect_shape<D>(a, b, tol, &h, &regular)) {
        if (lane == 0) {
            verdict[e] = 0;
            checks[e] = MPTG_CHECKS_OVER_LIMIT;
        }
        return;
    }
    if (!regular) {  // exact per-edge walk in a second launch
        if (lane == 0) irregular[atomicAdd(n_irregular, 1u)] = e;
        return;
    }
    const uint32_t W = (1u << h) - 1u;
    uint64_t key[kScoreStride / kWave];
    if constexpr (ORD == MPTG_ORDER_SURROGATE) {
        if (W > kScoreStride) {  // deeper than the scored set: reference order
            if (lane == 0) irregular[atomicAdd(n_irregular, 1u)] = e;
            return;
        }
#pragma unroll
        for (int k = 0; k < kScoreStride / kWave; ++k) {
            const uint32_t o = (uint32_t)(k * kWave + lane);
            key[k] = o < W ? score_key(scores[e * kScoreStride + o], o) : ~0ull;
        }
        wave_bitonic_sort<kScoreStride / kWave>(key, lane);
    }
    for (uint32_t c0 = 0, c = 0; c0 < W; c0 += kWave, ++c) {
        uint32_t o = c0 + (uint32_t)lane;
        if constexpr (ORD == MPTG_ORDER_SURROGATE) {
            uint64_t pick = key[0];
#pragma unroll
            for (int k = 1; k < kScoreStride / kWave; ++k)
                if (c == (uint32_t)k) pick = key[k];
            o = (uint32_t)pick;  // rank c0 + lane -> reference ordinal
        }
        bool bad = false;
        if (c0 + (uint32_t)lane < W) {
            uint32_t path;
            int depth;
            preorder_path(h, o, &path, &depth);
            bad = !arm_ok_fn<D, COUNT>(sc, [&](int i) { return coord_mid(a[i], b[i], path, depth); }, wk);
        }
        uint64_t m = __ballot(bad);
        if (m) {
            if (lane == 0) {
                verdict[e] = 0;
                checks[e] = base + c0 + (uint32_t)__builtin_ctzll(m) + 1u;
            }
            return;
        }
    }
    if (lane == 0) {
        verdict[e] = 1;
        checks[e] = base + W;
    }
}


template <int D, bool COUNT, int ORD>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void arm_link_wave(
    const double* __restrict__ from, const double* __restrict__ to, int64_t n, int known_valid,
    const uint8_t* __restrict__ mask, double tol, SceneArgs sc, const float* __restrict__ scores,
    uint8_t* __restrict__ verdict, uint32_t* __restrict__ checks, int64_t* __restrict__ irregular,
    uint32_t* __restrict__ n_irregular, unsigned long long* __restrict__ work) {
    ArmWork wk{0, 0, 0, 0, 0};
    arm_link_edge<D, COUNT, ORD>(from, to, n, known_valid, mask, tol, sc, scores, verdict, checks, irregular,
                                 n_irregular, &wk);
    if constexpr (COUNT) {
        atomicAdd(&work[0], (unsigned long long)wk.calls);
        atomicAdd(&work[1], (unsigned long long)wk.fk);
        atomicAdd(&work[2], (unsigned long long)wk.group);
        atomicAdd(&work[3], (unsigned long long)wk.far);
        atomicAdd(&work[4], (unsigned long long)wk.exact);
    }
}

}  // namespace

template <int D>
static void launch_arm_d(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int kv, const uint8_t* mask,
                         const float* scores, uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr,
                         unsigned long long* work) {
    unsigned blocks = (unsigned)((n + kWavesPerBlock - 1) / kWavesPerBlock);
    dim3 g(blocks), t(kWave * kWavesPerBlock);
    const double tol = ctx->link_param;
    const SceneArgs sa = scene_args(ctx);
    if (scores) {
        if (work)
            hipLaunchKernelGGL((arm_link_wave<D, true, MPTG_ORDER_SURROGATE>), g, t, 0, ctx->stream, from, to, n, kv,
                               mask, tol, sa, scores, verdict, checks, irr, nirr, work);
        else
            hipLaunchKernelGGL((arm_link_wave<D, false, MPTG_ORDER_SURROGATE>), g, t, 0, ctx->stream, from, to, n, kv,
                               mask, tol, sa, scores, verdict, checks, irr, nirr, work);
    } else {
        if (work)
            hipLaunchKernelGGL((arm_link_wave<D, true, MPTG_ORDER_REFERENCE>), g, t, 0, ctx->stream, from, to, n, kv,
                               mask, tol, sa, scores, verdict, checks, irr, nirr, work);
        else
            hipLaunchKernelGGL((arm_link_wave<D, false, MPTG_ORDER_REFERENCE>), g, t, 0, ctx->stream, from, to, n, kv,
                               mask, tol, sa, scores, verdict, checks, irr, nirr, work);
    }
}

// Launches the wave kernel; edges it lists as irregular are left for the
// caller's exact per-edge kernel.  Returns MPTG_E_STATE when no wave kernel
// exists for this arm.
int launch_arm_link(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int known_valid,
                    const uint8_t* mask, uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr,
                    unsigned long long* work, const float* scores) {
    if (ctx->dim != ctx->scene.nlinks) return MPTG_E_STATE;
    switch (ctx->dim) {
        case 5: launch_arm_d<5>(ctx, from, to, n, known_valid, mask, scores, verdict, checks, irr, nirr, work); break;
        case 10: launch_arm_d<10>(ctx, from, to, n, known_valid, mask, scores, verdict, checks, irr, nirr, work); break;
        case 20: launch_arm_d<20>(ctx, from, to, n, known_valid, mask, scores, verdict, checks, irr, nirr, work); break;
        default: return MPTG_E_STATE;
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

}  // namespace mptg
