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
#include "space_dev.hpp"

namespace mptg {
namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kMaxArmDepth = 30;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Pre-order ordinal -> (path, depth) in a perfect tree with h internal levels.
__device__ __forceinline__ void preorder_path(int h, uint32_t o, uint32_t* path, int* depth) {
    uint32_t p = 0;
    int d = 0;
    uint32_t sub = (1u << h) - 1u;
    while (o != 0) {
        o -= 1;
        uint32_t half = (sub - 1u) >> 1;
        if (o >= half) {
            o -= half;
            p |= 1u << d;
        }
        ++d;
        sub = half;
    }
    *path = p;
    *depth = d;
}

// One coordinate of the midpoint of the node at `path`: the reference's
// recursion (a+b)/2 restricted to that coordinate.
__device__ __forceinline__ double coord_mid(double lo, double hi, uint32_t path, int depth) {
    for (int k = 0; k < depth; ++k) {
        double m = (lo + hi) / 2;
        bool right = (path >> k) & 1u;
        lo = right ? m : lo;
        hi = right ? hi : m;
    }
    return (lo + hi) / 2;
}

template <int D>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void arm_link_wave(
    const double* __restrict__ from, const double* __restrict__ to, int64_t n, int known_valid,
    const uint8_t* __restrict__ mask, double tol, SceneArgs sc, uint8_t* __restrict__ verdict,
    uint32_t* __restrict__ checks, int64_t* __restrict__ irregular, uint32_t* __restrict__ n_irregular) {
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
        if (lane == 0) okA = arm_ok<D>(sc, a);
        if (lane == 1) okB = arm_ok<D>(sc, b);
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
    double wmax = 0.0, M = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        wmax = __builtin_fmax(wmax, __builtin_fabs(a[j] - b[j]));
        M = __builtin_fmax(M, __builtin_fmax(__builtin_fabs(a[j]), __builtin_fabs(b[j])));
    }
    int h = 0;
    double wd = wmax;  // w_max / 2^h, exact scaling
    while (!(wd < tol) && h <= kMaxArmDepth) {
        wd *= 0.5;
        ++h;
    }
    if (h > kMaxArmDepth) {
        if (lane == 0) {
            verdict[e] = 0;
            checks[e] = MPTG_CHECKS_OVER_LIMIT;
        }
        return;
    }
    const double err = (double)(h + 1) * 0x1.0p-50 * M;
    // h == 0: the root's width is exactly wmax < tol, a leaf
    const bool regular = (h == 0) || ((2.0 * wd - err >= tol) && (wd + err < tol));
    if (!regular) {  // exact per-edge walk in a second launch
        if (lane == 0) irregular[atomicAdd(n_irregular, 1u)] = e;
        return;
    }
    const uint32_t W = (1u << h) - 1u;
    for (uint32_t c0 = 0; c0 < W; c0 += kWave) {
        const uint32_t o = c0 + (uint32_t)lane;
        bool bad = false;
        if (o < W) {
            uint32_t path;
            int depth;
            preorder_path(h, o, &path, &depth);
            bad = !arm_ok_fn<D>(sc, [&](int i) { return coord_mid(a[i], b[i], path, depth); });
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

}  // namespace

template <int D>
static void launch_arm_d(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int kv, const uint8_t* mask,
                         uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr) {
    unsigned blocks = (unsigned)((n + kWavesPerBlock - 1) / kWavesPerBlock);
    hipLaunchKernelGGL((arm_link_wave<D>), dim3(blocks), dim3(kWave * kWavesPerBlock), 0, ctx->stream, from, to, n,
                       kv, mask, ctx->link_param, scene_args(ctx), verdict, checks, irr, nirr);
}

// Launches the wave kernel; edges it lists as irregular are left for the
// caller's exact per-edge kernel.  Returns MPTG_E_STATE when no wave kernel
// exists for this arm.
int launch_arm_link(mptg_ctx* ctx, const double* from, const double* to, int64_t n, int known_valid,
                    const uint8_t* mask, uint8_t* verdict, uint32_t* checks, int64_t* irr, uint32_t* nirr) {
    if (ctx->dim != ctx->scene.nlinks) return MPTG_E_STATE;
    switch (ctx->dim) {
        case 5: launch_arm_d<5>(ctx, from, to, n, known_valid, mask, verdict, checks, irr, nirr); break;
        case 10: launch_arm_d<10>(ctx, from, to, n, known_valid, mask, verdict, checks, irr, nirr); break;
        case 20: launch_arm_d<20>(ctx, from, to, n, known_valid, mask, verdict, checks, irr, nirr); break;
        default: return MPTG_E_STATE;
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

}  // namespace mptg
