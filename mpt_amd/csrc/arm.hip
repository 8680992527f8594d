// Planar-arm link() on the GPU: a packed round (16 edges per wave, ranks
// 0..3), then a continuation of 64-rank rounds over several waves per edge.
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
// closed form.  Lanes take consecutive ranks of the chosen order, rebuild
// their midpoint from (a, b) along the path (bit-identical to the
// reference's recursion), run valid() and ballot for the first failure.
// Edges whose widths fall within e of tol at depth h-1 or h (never
// observed; kept exact anyway) are walked by one lane with the stackless
// exact walk.
//
// valid() per lane: forward kinematics with the project sincos, then every
// link segment against the disks that the grid culling leaves (the wave's
// box or the lane's own cells, arm_ok_link), behind a far-disk bound with a
// relative margin of 2^-30 that dominates the rounding of the exact
// expression, so a skip never changes a verdict.  Disks that pass the bound
// get the reference expression bit-for-bit.
#include "bisect_dev.hpp"

namespace mptg {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Waypoint at `rank` of the chosen order (tree of h levels): the reference
// pre-order, or the surrogate order -- the edge's kPicks picks (best-scored
// nodes of its top tree, from the scoring kernel) at ranks 0..3, then every
// other waypoint in reference order.  `sorted` holds the picks in ascending
// ordinal order (ranks >= kPicks only).
template <int ORD>
__device__ __forceinline__ uint32_t ordinal_at(uint32_t rank, const uint32_t (&sorted)[kPicks]) {
    if constexpr (ORD == MPTG_ORDER_SURROGATE) return skip_picks(rank - kPicks, sorted);
    return rank;
}

// The 4 picks of edge e in ascending ordinal order (wave-uniform loads).
__device__ __forceinline__ void sorted_picks(const uint16_t* __restrict__ picks, int64_t e, uint32_t (&p)[kPicks]) {
#pragma unroll
    for (int i = 0; i < kPicks; ++i) p[i] = picks[e * kPicks + i];
    // 5-comparator sorting network
    auto cs = [&](int i, int j) {
        const uint32_t lo = p[i] < p[j] ? p[i] : p[j], hi = p[i] < p[j] ? p[j] : p[i];
        p[i] = lo;
        p[j] = hi;
    };
    cs(0, 1);
    cs(2, 3);
    cs(0, 2);
    cs(1, 3);
    cs(1, 2);
}

constexpr int kHeadRanks = kPicks;                // ranks per edge in the packed round

// Packed round: 4 ranks per edge, 16 edges per wave (8 when each check takes
// 2 lanes, head_lanes).  Endpoint checks (when not vouched for), the tree
// shape, then ranks 0..3 of the chosen order (reference ordinals 0..3, or the
// surrogate's picks).  Edges still undecided go to the continuation list.
// Waypoint validity in the link kernels: the wave-box culling (LANE false:
// all lanes share one disk walk) or each lane's own grid cells (LANE true).
template <int D, bool COUNT, bool LANE, typename Coord>
__device__ __forceinline__ bool arm_ok_link(const SceneArgs& sc, Coord&& coord, bool active, ArmWork* wk) {
    if constexpr (LANE) return arm_ok_lane<D, COUNT>(sc, coord, active, wk);
    return arm_ok_wave<D, COUNT>(sc, coord, active, wk);
}

// Lanes per waypoint check in the packed round: the timed per-lane-culling
// kernel splits each check's disk runs over 2 lanes (arm_ok_split: the same
// disks, so the same verdicts), which doubles its waves; the work-counting
// and wave-box variants keep one lane per check.
template <bool COUNT, bool LANE>
constexpr int head_lanes() {
    return (!COUNT && LANE) ? 2 : 1;
}
template <bool COUNT, bool LANE>
constexpr int head_edges() {
    return kWave / (kHeadRanks * head_lanes<COUNT, LANE>());
}

// The 4 rank flags of edge slot `slot` from a wave ballot (one bit per lane,
// HL lanes per rank holding the same flag).
template <int HL>
__device__ __forceinline__ uint32_t head_group(uint64_t m, int slot) {
    const uint64_t w = m >> (kHeadRanks * HL * slot);
    uint32_t g = 0;
#pragma unroll
    for (int k = 0; k < kHeadRanks; ++k) g |= (uint32_t)((w >> (k * HL)) & 1u) << k;
    return g;
}

template <int D, bool COUNT, int ORD, bool LANE>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void arm_link_head(
    const double* __restrict__ from, const double* __restrict__ to, int64_t n, int known_valid,
    const uint8_t* __restrict__ mask, double tol, SceneArgs sc, const uint16_t* __restrict__ picks,
    uint8_t* __restrict__ verdict, uint32_t* __restrict__ checks, int64_t* __restrict__ irregular,
    uint32_t* __restrict__ n_irregular, int64_t* __restrict__ cont, uint32_t* __restrict__ n_cont,
    unsigned long long* __restrict__ work) {
    constexpr int HL = head_lanes<COUNT, LANE>();
    constexpr int HE = head_edges<COUNT, LANE>();
    ArmWork wk{0, 0, 0, 0, 0};
    const int lane = lane_id();
    const int slot = lane / (kHeadRanks * HL), rk = (lane / HL) % kHeadRanks;
    const bool first = lane % HL == 0;  // the lane of its group that writes
    const int64_t e = ((int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) * HE + slot;
    bool live = e < n && (!mask || mask[e]);
    // one waypoint check per group of HL lanes
    auto check = [&](auto&& coord, bool act) {
        if constexpr (HL == 1) return arm_ok_link<D, COUNT, LANE>(sc, coord, act, &wk);
        else return arm_ok_split<D, HL>(sc, coord, act);
    };
    const int64_t ec = live ? e : 0;
    const double* a = from + ec * D;
    const double* b = to + ec * D;
    const uint32_t base = 2;
    if (!known_valid) {
        const bool act = live && rk < 2;
        const bool okv = check([&](int i) { return rk == 0 ? a[i] : b[i]; }, act);  // every lane calls it
        const bool bad = act && !okv;
        const uint32_t g = head_group<HL>(__ballot(bad), slot);
        if (live && (g & 3u)) {
            if (rk == 0 && first) {
                verdict[e] = 0;
                checks[e] = (g & 1u) ? 1u : 2u;
            }
            live = false;
        }
    }
    int h = 0;
    bool regular = false;
    if (live) {
        if (!bisect_shape<D>(a, b, tol, &h, &regular)) {
            if (rk == 0 && first) {
                verdict[e] = 0;
                checks[e] = MPTG_CHECKS_OVER_LIMIT;
            }
            live = false;
        } else if (!regular || (ORD == MPTG_ORDER_SURROGATE && h > kMaxPickDepth)) {
            if (rk == 0 && first) irregular[atomicAdd(n_irregular, 1u)] = e;  // exact per-edge walk later
            live = false;
        }
    }
    const uint32_t W = live ? (1u << h) - 1u : 0u;
    // the whole wave takes part in the disk culling; lanes without a
    // waypoint check the edge's end state and do not count
    const bool act = live && (uint32_t)rk < W;
    uint32_t path = 0;
    int depth = 0;
    if (act) {
        const uint32_t o = ORD == MPTG_ORDER_SURROGATE ? (uint32_t)picks[e * kPicks + rk] : (uint32_t)rk;
        preorder_path(h, o, &path, &depth);
    }
    // every lane calls the check (its wave-wide culling shuffles read every
    // lane), and only lanes with a waypoint count
    const bool okv = check([&](int i) { return act ? coord_mid(a[i], b[i], path, depth) : b[i]; }, act);
    const bool bad = act && !okv;
    const uint32_t g = head_group<HL>(__ballot(bad), slot);
    if (live && rk == 0 && first) {
        if (g) {
            verdict[e] = 0;
            checks[e] = base + (uint32_t)__builtin_ctz(g) + 1u;
        } else if (W <= (uint32_t)kHeadRanks) {
            verdict[e] = 1;
            checks[e] = base + W;
        } else {
            checks[e] = 0xffffffffu;  // the continuation's first-failure slot
            cont[atomicAdd(n_cont, 1u)] = e;
        }
    }
    if constexpr (COUNT) {
        atomicAdd(&work[0], (unsigned long long)wk.calls);
        atomicAdd(&work[1], (unsigned long long)wk.fk);
        atomicAdd(&work[2], (unsigned long long)wk.group);
        atomicAdd(&work[3], (unsigned long long)wk.far);
        atomicAdd(&work[4], (unsigned long long)wk.exact);
    }
}

// Continuation: P waves per listed edge (P = waves_per_edge), 64 ranks per
// round from rank 4; wave p of an edge takes rounds p, p + P, p + 2P, ...
// A wave stops at its first failing waypoint and lowers the edge's slot in
// `checks` (set to ~0 by the packed round) to that rank with atomicMin; it
// also stops when the slot already holds a rank below its next round.  The
// first failing rank of the edge lies in one wave's rounds, and that wave
// cannot stop before it (every rank the others report is larger), so the
// slot ends at exactly the first failing rank of the chosen order.  No wave
// waits for another: the slot is read and lowered, never awaited.
template <int D, bool COUNT, int ORD, bool LANE>
__global__ __launch_bounds__(kWave* kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(5))) void arm_link_tail(
    const double* __restrict__ from, const double* __restrict__ to, double tol, SceneArgs sc,
    const uint16_t* __restrict__ picks, const int64_t* __restrict__ cont, const uint32_t* __restrict__ n_cont,
    int waves_per_edge, uint32_t* __restrict__ checks, unsigned long long* __restrict__ work) {
    // no block-wide barriers: waves leave independently
    ArmWork wk{0, 0, 0, 0, 0};
    const int lane = lane_id();
    const int64_t widx = (int64_t)blockIdx.x * kWavesPerBlock + uniform((int)(threadIdx.x >> 6));
    const int64_t item = widx / waves_per_edge;
    const uint32_t P = (uint32_t)waves_per_edge, p = (uint32_t)(widx % waves_per_edge);
    if (item < (int64_t)*n_cont) {
        const int64_t e = cont[item];
        const double* a = from + e * D;  // wave-uniform: scalar loads
        const double* b = to + e * D;
        int h = 0;
        bool regular = false;
        bisect_shape<D>(a, b, tol, &h, &regular);  // listed edges are regular
        const uint32_t W = (1u << h) - 1u;
        uint32_t sp[kPicks] = {0, 0, 0, 0};
        if constexpr (ORD == MPTG_ORDER_SURROGATE) sorted_picks(picks, e, sp);  // W > 4: 4 picks
        for (uint32_t c0 = kHeadRanks + kWave * p; c0 < W; c0 += kWave * P) {
            if (__hip_atomic_load(&checks[e], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < c0) break;  // uniform
            const uint32_t r = c0 + (uint32_t)lane;
            // every lane takes part in the wave's disk culling; lanes past
            // the last rank recompute the last waypoint and do not count
            const bool active = r < W;
            uint32_t path;
            int depth;
            preorder_path(h, ordinal_at<ORD>(active ? r : W - 1u, sp), &path, &depth);
            const bool bad = !arm_ok_link<D, COUNT, LANE>(
                sc, [&](int i) { return coord_mid(a[i], b[i], path, depth); }, active, &wk);
            const uint64_t m = __ballot(bad && active);
            if (m) {
                if (lane == 0) atomicMin(&checks[e], c0 + (uint32_t)__builtin_ctzll(m));
                break;
            }
        }
    }
    if constexpr (COUNT) {
        atomicAdd(&work[0], (unsigned long long)wk.calls);
        atomicAdd(&work[1], (unsigned long long)wk.fk);
        atomicAdd(&work[2], (unsigned long long)wk.group);
        atomicAdd(&work[3], (unsigned long long)wk.far);
        atomicAdd(&work[4], (unsigned long long)wk.exact);
    }
}

// Verdicts of the listed edges from their slots: ~0 = no failing waypoint.
template <int D>
__global__ __launch_bounds__(256) void arm_link_finish(const double* __restrict__ from, const double* __restrict__ to,
                                                       double tol, const int64_t* __restrict__ cont,
                                                       const uint32_t* __restrict__ n_cont,
                                                       uint8_t* __restrict__ verdict, uint32_t* __restrict__ checks) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)*n_cont) return;
    const int64_t e = cont[i];
    const uint32_t base = 2;
    const uint32_t first = checks[e];
    if (first != 0xffffffffu) {
        verdict[e] = 0;
        checks[e] = base + first + 1u;
    } else {
        int h = 0;
        bool regular = false;
        bisect_shape<D>(from + e * D, to + e * D, tol, &h, &regular);
        verdict[e] = 1;
        checks[e] = base + ((1u << h) - 1u);
    }
}

}  // namespace

template <int D, bool COUNT, int ORD, bool LANE>
static int launch_arm_phases(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int kv,
                             const uint8_t* mask, uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr,
                             int64_t* cont, uint32_t* ncont, unsigned long long* work) {
    const double tol = ctx->link_param;
    const SceneArgs sa = scene_args(ctx);
    const uint16_t* picks = nullptr;
    if constexpr (ORD == MPTG_ORDER_SURROGATE) {
        int rc;
        if ((rc = ensure(ctx, ctx->scratch[9], sizeof(uint16_t) * kPicks * n))) return rc;
        uint16_t* pk = (uint16_t*)ctx->scratch[9].p;
        if ((rc = surrogate_picks(ctx, from, to, n, mask, pk))) return rc;
        picks = pk;
    }
    phase_mark(ctx, 1);
    const int64_t per_block = (int64_t)kWavesPerBlock * head_edges<COUNT, LANE>();
    hipLaunchKernelGGL((arm_link_head<D, COUNT, ORD, LANE>), dim3((unsigned)((n + per_block - 1) / per_block)),
                       dim3(kWave * kWavesPerBlock), 0, ctx->stream, from, to, n, kv, mask, tol, sa, picks, verdict,
                       checks, irr, nirr, cont, ncont, work);
    phase_mark(ctx, 2);
    phase_mark(ctx, 3);
    int P = ctx->tune.arm_tail_waves;
    if ((n * P + kWavesPerBlock - 1) / kWavesPerBlock > (int64_t)INT32_MAX) P = 1;  // grid limit
    hipLaunchKernelGGL((arm_link_tail<D, COUNT, ORD, LANE>), dim3((unsigned)((n * P + kWavesPerBlock - 1) / kWavesPerBlock)),
                       dim3(kWave * kWavesPerBlock), 0, ctx->stream, from, to, tol, sa, picks, cont, ncont, P, checks,
                       work);
    hipLaunchKernelGGL((arm_link_finish<D>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, from, to,
                       tol, cont, ncont, verdict, checks);
    return MPTG_OK;
}

template <int D, bool COUNT>
static int launch_arm_cull(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int kv,
                           const uint8_t* mask, int order, uint8_t* verdict, uint32_t* checks, int64_t* irr,
                           uint32_t* nirr, int64_t* cont, uint32_t* ncont, unsigned long long* work) {
    const bool lane = ctx->tune.arm_cull != 0;
    if (order == MPTG_ORDER_SURROGATE) {
        if (lane) return launch_arm_phases<D, COUNT, MPTG_ORDER_SURROGATE, true>(ctx, from, to, n, kv, mask, verdict, checks, irr, nirr, cont, ncont, work);
        return launch_arm_phases<D, COUNT, MPTG_ORDER_SURROGATE, false>(ctx, from, to, n, kv, mask, verdict, checks, irr, nirr, cont, ncont, work);
    }
    if (lane) return launch_arm_phases<D, COUNT, MPTG_ORDER_REFERENCE, true>(ctx, from, to, n, kv, mask, verdict, checks, irr, nirr, cont, ncont, work);
    return launch_arm_phases<D, COUNT, MPTG_ORDER_REFERENCE, false>(ctx, from, to, n, kv, mask, verdict, checks, irr, nirr, cont, ncont, work);
}

template <int D>
static int launch_arm_d(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int kv, const uint8_t* mask,
                        int order, uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr, int64_t* cont,
                        uint32_t* ncont, unsigned long long* work) {
    if (work) return launch_arm_cull<D, true>(ctx, from, to, n, kv, mask, order, verdict, checks, irr, nirr, cont, ncont, work);
    return launch_arm_cull<D, false>(ctx, from, to, n, kv, mask, order, verdict, checks, irr, nirr, cont, ncont, work);
}


// Launches the packed round and the continuation (for the surrogate order,
// the scoring kernel's picks first); edges the kernels
// list as irregular are left for the caller's exact per-edge kernel
// (irr/nirr; *nirr is zeroed here).  Returns MPTG_E_STATE when no kernel
// exists for this arm.
int launch_arm_link(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int known_valid,
                    const uint8_t* mask, uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr,
                    unsigned long long* work, int order) {
    if (ctx->dim != ctx->scene.nlinks) return MPTG_E_STATE;
    if (ctx->dim != 5 && ctx->dim != 10 && ctx->dim != 20) return MPTG_E_STATE;
    int rc;
    if ((rc = ensure(ctx, ctx->scratch[10], sizeof(int64_t) * n + 256))) return rc;
    uint32_t* ncont = (uint32_t*)ctx->scratch[10].p;
    int64_t* cont = (int64_t*)((char*)ctx->scratch[10].p + 256);
    if ((rc = launch_zero_pair(ctx, ncont, nirr))) return rc;
    phase_mark(ctx, 0);
    switch (ctx->dim) {
        case 5: rc = launch_arm_d<5>(ctx, from, to, n, known_valid, mask, order, verdict, checks, irr, nirr, cont, ncont, work); break;
        case 10: rc = launch_arm_d<10>(ctx, from, to, n, known_valid, mask, order, verdict, checks, irr, nirr, cont, ncont, work); break;
        default: rc = launch_arm_d<20>(ctx, from, to, n, known_valid, mask, order, verdict, checks, irr, nirr, cont, ncont, work); break;
    }
    if (rc) return rc;
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

}  // namespace mptg
