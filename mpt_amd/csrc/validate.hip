// Batched state validity and motion validation (the Scenario's valid()/link(),
// called at src/mpt/impl/prrt/prrt.hpp:439,454-457).
//
// Two link() forms:
//  * DMV — DiscreteMotionValidator, src/mpt/discrete_motion_validator.hpp:71-130:
//    valid(to) first, steps = ceil(dist * (1/step)), waypoints i * (1/steps)
//    for i in [1, steps-1] visited in the 256-slot BFS-bisection ring order,
//    ranges that find the ring full are visited sequentially.  For
//    n = steps-1 <= 255 the ring can never fill (255 disjoint non-empty
//    ranges plus a popped range of size >= 2 need n >= 256), so the order is
//    a pure BFS and is read from a 32 KiB table built once per context.  A
//    group of lanes shares each such edge and checks a run of consecutive
//    positions of that order per round (link_dmv_lanes).  Longer edges go to
//    a second kernel in which one wave walks the ring in LDS and checks 64
//    waypoints per round.
//  * BISECT — LinkManipulatorScenario::link/bisectLink,
//    demo/link_manipulator_scenario.hpp:118-138: valid(a), valid(b), then the
//    pre-order of recursive midpoints (a+b)/2 until L-inf(a-b) < tol.  The
//    walk is stackless: the current interval is named by its left/right path
//    from the root and recomputed from (a, b) when needed, so each midpoint
//    has exactly the reference's bits whatever the tree's shape.
//  * BISECT_L2 — the same walk until the squared L2 length (b-a).(b-a) is
//    below tol^2: Rect::segmentIsValid/bisectSegment
//    (demo/shape_hierarchy.hpp:184-203) and PNG2dScenario::link/validSegment
//    (demo/png_2d_scenario.hpp:112-117,152-165).  In the circles-and-
//    rectangles scene each circle first takes its closed form
//    (Circle::segmentIsValid, shape_hierarchy.hpp:228-231), as the
//    scenario's link() tests circles before rectangles
//    (holonomic_2d_point_scenario.hpp:105-113); the points of the walk are
//    then tested against the rectangles.  Every rectangle's recursion visits
//    the same midpoints, so one walk that tests all rectangles per point
//    gives the same verdict; `checks` counts its points (none when a circle
//    blocks the segment).
//
// verdict is the AND over a fixed set of valid() calls, so it does not depend
// on the visiting order; `checks` counts the calls the given order makes
// before the verdict is known, endpoint checks included.
#include <algorithm>
#include <array>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "space_dev.hpp"

namespace mptg {
namespace {

constexpr int kBlock = 256;
constexpr int kTableMax = 255;  // largest n = steps-1 served by the BFS table
// The 16-bit table holds the ring's order for 256 <= n <= kTable16Max, row n
// at n(n-1)/2 - 255*256/2 (1 MiB).
constexpr int kTable16Max = 1023;
__host__ __device__ constexpr int64_t table16_row(int64_t m) {
    return m * (m - 1) / 2 - (int64_t)kTableMax * (kTableMax + 1) / 2;
}
// Work limits of one link() call; an edge beyond them is not checked and
// reports verdict 0 with checks = MPTG_CHECKS_OVER_LIMIT.
constexpr int64_t kMaxDmvWaypoints = (int64_t)1 << 24;
constexpr int kMaxBisectDepth = 30;

// valid() of n arm states, 8 lanes per state (arm_ok_split).
template <int D>
__global__ __launch_bounds__(kBlock) void valid_kernel_arm(SceneArgs sc, const double* __restrict__ q, int64_t n,
                                                           uint8_t* __restrict__ ok) {
    constexpr int L = 8;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t i = t / L;
    const bool active = i < n;
    double s[D];
#pragma unroll
    for (int j = 0; j < D; ++j) s[j] = active ? q[i * D + j] : 0.0;
    const bool v = arm_ok_split<D, L>(sc, [&](int j) { return s[j]; }, active);
    if (active && (t % L) == 0) ok[i] = v ? 1 : 0;
}

template <int D>
__global__ __launch_bounds__(kBlock) void valid_kernel(SceneArgs sc, int dim, const double* __restrict__ q,
                                                       int64_t n, uint8_t* __restrict__ ok) {
    constexpr int DA = D > 0 ? D : kMaxDim;
    const int dd = D > 0 ? D : dim;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double s[DA];
#pragma unroll
    for (int j = 0; j < DA; ++j)
        if (j < dd) s[j] = q[i * dd + j];
    ok[i] = state_valid_t<D>(sc, s) ? 1 : 0;
}

// steps = ceil(d * inv) as the reference's size_t; returns n = steps - 1
// (number of interior waypoints), or 0 when steps < 2.
template <int D>
__device__ __forceinline__ int64_t dmv_count(const SpaceArgs& sp, const double* a, const double* b, double inv,
                                             double* delta) {
    double x = __builtin_ceil(space_dist_t<D>(sp, a, b) * inv);
    if (!(x >= 2.0)) return 0;
    if (!(x < 0x1.0p62)) x = 0x1.0p62;
    int64_t steps = (int64_t)x;
    *delta = 1.0 / (double)steps;
    return steps - 1;
}

// valid() with the scene kind fixed at compile time (K = a scene kind, or 0
// for any kind, chosen at run time): the DMV kernels inline only the scene
// they check, which keeps their code and register count small.
template <int D, int K>
__device__ __forceinline__ bool valid_k(const SceneArgs& sc, const double* q) {
    if constexpr (K == MPTG_SCENE_GRID) return grid_valid(sc, q);
    else if constexpr (K == MPTG_SCENE_BOXES) return boxes_valid(sc, q);
    else return state_valid_t<D>(sc, q);
}

// Positions a lane checks per round: independent (table, cell) load chains in
// flight together, which the per-round latency hides.  Only for states of
// up to 3 coordinates, whose waypoints stay in registers.
template <int D>
constexpr int dmv_unroll() {
    return D > 0 && D <= 3 ? 2 : 1;
}

// G = 2^lg lanes per edge (1..64; lanes of one edge sit in one wave).  An
// edge's check sequence is position 0 = valid(to) (counted, and evaluated
// unless the endpoints are known valid) followed by positions 1..m = the
// interior waypoints in the BFS table's order.  Round r of the group covers
// positions [r G U, (r + 1) G U): lane g takes r G U + u G + g for u < U.  A
// ballot per u over the group's lanes gives the first failing position f
// (blocks of smaller u hold smaller positions), and the reference's count is
// f + 1 (m + 1 when none fails).  Every loop is bounded by m <= 255.  Edges
// with more waypoints go to the long-edge list after valid(to), as the
// reference checks `to` first (discrete_motion_validator.hpp:75).
template <int D, int K>
__global__ __launch_bounds__(kBlock) void link_dmv_lanes(SpaceArgs sp, SceneArgs sc, double inv, int lg,
                                                         const uint8_t* __restrict__ table,
                                                         const double* __restrict__ from,
                                                         const double* __restrict__ to, int64_t n, int known_valid,
                                                         const uint8_t* __restrict__ mask,
                                                         uint8_t* __restrict__ verdict, uint32_t* __restrict__ checks,
                                                         int64_t* __restrict__ long_list,
                                                         uint32_t* __restrict__ long_count) {
    constexpr int DA = D > 0 ? D : kMaxDim;
    constexpr int U = dmv_unroll<D>();
    const int dd = D > 0 ? D : sp.dim;
    const int G = 1 << lg;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t e = t >> lg;
    const int g = (int)(t & (G - 1));
    // the group's lanes within the wave's 64-bit ballot
    const int gbase = (int)(threadIdx.x & 63) & ~(G - 1);
    const uint64_t gbits = G == 64 ? ~0ull : ((1ull << G) - 1) << gbase;
    if (e >= n || (mask && !mask[e])) return;  // the whole group leaves together
    double a[DA], b[DA];
#pragma unroll
    for (int j = 0; j < DA; ++j)
        if (j < dd) {
            a[j] = from[e * dd + j];
            b[j] = to[e * dd + j];
        }
    double delta = 0.0;
    const int64_t m = dmv_count<D>(sp, a, b, inv, &delta);
    if (m > kTableMax) {
        if (g == 0) {
            if (!known_valid && !valid_k<D, K>(sc, b)) {
                verdict[e] = 0;
                checks[e] = 1;
            } else if (m > kMaxDmvWaypoints) {
                verdict[e] = 0;
                checks[e] = MPTG_CHECKS_OVER_LIMIT;
            } else {
                long_list[atomicAdd(long_count, 1u)] = e;
            }
        }
        return;
    }
    const int last = (int)m;
    const uint8_t* ord = table + (last * (last - 1)) / 2;
    // 2-D grids: every position evaluated without branches (the waypoint
    // (b - a) * t + a of space_interp_t, position 0 = valid(to) decided up
    // front), so a round's U table loads and U cell loads are in flight
    // together
    [[maybe_unused]] bool to_bad = false;
    if constexpr (K == MPTG_SCENE_GRID) to_bad = !known_valid && !grid_free(sc, b[0], b[1]);
    for (int p0 = 0; p0 <= last; p0 += G * U) {
        bool bad[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = p0 + u * G + g;
            if constexpr (K == MPTG_SCENE_GRID) {
                const bool inr = p >= 1 && p <= last;
                const double tt = (double)ord[inr ? p - 1 : 0] * delta;
                const bool ok = grid_free(sc, (b[0] - a[0]) * tt + a[0], (b[1] - a[1]) * tt + a[1]);
                bad[u] = p == 0 ? to_bad : (inr && !ok);
            } else {
                bad[u] = false;
                if (p == 0) {
                    bad[u] = !known_valid && !valid_k<D, K>(sc, b);
                } else if (p <= last) {
                    double q[DA];
                    space_interp_t<D>(sp, a, b, (double)ord[p - 1] * delta, q);
                    bad[u] = !valid_k<D, K>(sc, q);
                }
            }
        }
        // every ballot before any branch on them, so no position's loads
        // wait behind another's verdict
        uint64_t hits[U], any = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            hits[u] = __ballot(bad[u]) & gbits;
            any |= hits[u];
        }
        if (any) {
            int u = 0;
            while (!hits[u]) ++u;  // ends: some hits[u] is nonzero
            if (g == 0) {
                verdict[e] = 0;
                checks[e] = (uint32_t)(p0 + u * G + (__builtin_ctzll(hits[u]) - gbase) + 1);
            }
            return;
        }
    }
    if (g == 0) {
        verdict[e] = 1;
        checks[e] = (uint32_t)(1 + last);
    }
}

// One wave per long edge (more than 255 interior waypoints).  Up to
// kTable16Max waypoints the order comes from the 16-bit table of the ring's
// order and every lane takes U positions per round; longer edges replay the
// reference's ring: lane 0 walks it in LDS and emits up to 64 waypoint
// indices per round, which all lanes check.
template <int D, int K>
__global__ __launch_bounds__(64) void link_dmv_long(SpaceArgs sp, SceneArgs sc, double inv,
                                                    const uint16_t* __restrict__ table16,
                                                    const double* __restrict__ from, const double* __restrict__ to,
                                                    const int64_t* __restrict__ long_list,
                                                    const uint32_t* __restrict__ long_count,
                                                    uint8_t* __restrict__ verdict, uint32_t* __restrict__ checks) {
    __shared__ uint64_t rmin[256], rmax[256];
    __shared__ uint64_t emitted[64];
    __shared__ int cnt;
    __shared__ int fail_lane;
    const uint32_t total = *long_count;
    const int lane = threadIdx.x;
    for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
        constexpr int DA = D > 0 ? D : kMaxDim;
        constexpr int U = dmv_unroll<D>();
        const int dd = D > 0 ? D : sp.dim;
        const int64_t e = long_list[w];
        double a[DA], b[DA], q[DA];
#pragma unroll
        for (int j = 0; j < DA; ++j)
            if (j < dd) {
                a[j] = from[e * dd + j];
                b[j] = to[e * dd + j];
            }
        double delta = 0.0;
        int64_t m = dmv_count<D>(sp, a, b, inv, &delta);
        if (m <= kTable16Max) {  // wave-uniform branch: m is the edge's
            const uint16_t* ord = table16 + table16_row(m);
            int64_t done = m;
            bool hit_any = false;
            for (int64_t p0 = 0; p0 < m && !hit_any; p0 += 64 * U) {
                bool bad[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t o = p0 + u * 64 + lane;
                    bad[u] = false;
                    if (o < m) {
                        double qq[DA];
                        space_interp_t<D>(sp, a, b, (double)ord[o] * delta, qq);
                        bad[u] = !valid_k<D, K>(sc, qq);
                    }
                }
                uint64_t hits[U];
#pragma unroll
                for (int u = 0; u < U; ++u) hits[u] = __ballot(bad[u]);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (hits[u] && !hit_any) {
                        done = p0 + u * 64 + __builtin_ctzll(hits[u]) + 1;
                        hit_any = true;
                    }
            }
            if (lane == 0) {
                verdict[e] = hit_any ? 0 : 1;
                checks[e] = (uint32_t)(1 + done);
            }
            continue;
        }
        // ring state lives in lane 0's registers
        uint64_t qs = 0, qe = 1, seq_cur = 0, seq_end = 0;
        bool seq = false;
        if (lane == 0) {
            rmin[0] = 1;
            rmax[0] = (uint64_t)m;
        }
        int64_t done = 0;
        bool failed = false;
        for (;;) {
            if (lane == 0) {
                int c = 0;
                while (c < 64 && (seq || qs != qe)) {
                    if (seq) {
                        emitted[c++] = seq_cur;
                        if (seq_cur == seq_end) seq = false;
                        else ++seq_cur;
                        continue;
                    }
                    uint64_t mn = rmin[qs % 256], mx = rmax[qs % 256];
                    ++qs;
                    if (mn == mx) {
                        emitted[c++] = mn;
                    } else if (qe + 2 < qs + 256) {
                        uint64_t mid = (mn + mx) / 2;
                        emitted[c++] = mid;
                        if (mn < mid) {
                            rmin[qe % 256] = mn;
                            rmax[qe % 256] = mid - 1;
                            ++qe;
                        }
                        if (mid < mx) {
                            rmin[qe % 256] = mid + 1;
                            rmax[qe % 256] = mx;
                            ++qe;
                        }
                    } else {
                        seq = true;
                        seq_cur = mn;
                        seq_end = mx;
                    }
                }
                cnt = c;
                fail_lane = 64;
            }
            __syncthreads();
            const int c = cnt;
            if (c == 0) break;
            bool bad = false;
            if (lane < c) {
                space_interp_t<D>(sp, a, b, (double)emitted[lane] * delta, q);
                bad = !valid_k<D, K>(sc, q);
            }
            if (bad) atomicMin(&fail_lane, lane);
            __syncthreads();
            const int fl = fail_lane;
            if (fl < 64) {
                failed = true;
                done += fl + 1;
                break;
            }
            done += c;
            __syncthreads();
        }
        if (lane == 0) {
            verdict[e] = failed ? 0 : 1;
            checks[e] = (uint32_t)(1 + done);
        }
        __syncthreads();
    }
}

// Interval of the node at `path` (bit k = direction taken at level k,
// 0 = left half [lo, mid], 1 = right half [mid, hi]) with `depth` levels.
__device__ __forceinline__ void bisect_interval(int D, const double* a, const double* b, uint64_t path, int depth,
                                                double* lo, double* hi) {
    for (int j = 0; j < D; ++j) {
        lo[j] = a[j];
        hi[j] = b[j];
    }
    for (int k = 0; k < depth; ++k) {
        bool right = (path >> k) & 1;
        for (int j = 0; j < D; ++j) {
            double mid = (lo[j] + hi[j]) / 2;
            if (right) lo[j] = mid;
            else hi[j] = mid;
        }
    }
}

// Whether the interval [lo, hi] still gets a midpoint: L-inf length >= tol
// (L2 = false) or squared L2 length >= tol^2, summed left to right (L2 = true).
template <bool L2>
__device__ __forceinline__ bool bisect_internal(int D, const double* lo, const double* hi, double tol) {
    double m = 0.0;
    if constexpr (L2) {
        for (int j = 0; j < D; ++j) {
            const double t = hi[j] - lo[j];
            m = m + t * t;
        }
        return !(m < tol * tol);
    } else {
        for (int j = 0; j < D; ++j) m = __builtin_fmax(m, __builtin_fabs(lo[j] - hi[j]));
        return !(m < tol);
    }
}

// A point of the walk: the scene's valid(), or for circles and rectangles
// the rectangles alone (the circles took their closed form first).
__device__ __forceinline__ bool bisect_point_ok(const SceneArgs& sc, const double* q) {
    return sc.kind == MPTG_SCENE_SHAPES ? boxes_valid(sc, q) : state_valid(sc, q);
}

template <bool L2>
__device__ void bisect_edge(int D, const SceneArgs& sc, double tol, const double* __restrict__ from,
                            const double* __restrict__ to, int64_t e, int known_valid, uint8_t* __restrict__ verdict,
                            uint32_t* __restrict__ checks) {
    double a[kMaxDim], b[kMaxDim], lo[kMaxDim], hi[kMaxDim];
    for (int j = 0; j < D; ++j) {
        a[j] = from[e * D + j];
        b[j] = to[e * D + j];
    }
    if (sc.kind == MPTG_SCENE_SHAPES && !circles_segment_clear(sc, a, b)) {
        verdict[e] = 0;
        checks[e] = 0;
        return;
    }
    uint32_t c = 1;
    if (!known_valid && !bisect_point_ok(sc, a)) {
        verdict[e] = 0;
        checks[e] = c;
        return;
    }
    ++c;
    if (!known_valid && !bisect_point_ok(sc, b)) {
        verdict[e] = 0;
        checks[e] = c;
        return;
    }
    uint64_t path = 0;
    int depth = 0;
    for (;;) {
        bisect_interval(D, a, b, path, depth, lo, hi);
        if (bisect_internal<L2>(D, lo, hi, tol)) {
            for (int j = 0; j < D; ++j) lo[j] = (lo[j] + hi[j]) / 2;
            ++c;
            if (!bisect_point_ok(sc, lo)) {
                verdict[e] = 0;
                checks[e] = c;
                return;
            }
            if (depth >= kMaxBisectDepth) {
                verdict[e] = 0;
                checks[e] = MPTG_CHECKS_OVER_LIMIT;
                return;
            }
            ++depth;  // descend left (bit stays 0)
        } else {
            while (depth > 0 && ((path >> (depth - 1)) & 1)) {
                path &= ~(1ull << (depth - 1));
                --depth;
            }
            if (depth == 0) break;
            path |= 1ull << (depth - 1);
        }
    }
    verdict[e] = 1;
    checks[e] = c;
}

template <bool L2>
__global__ __launch_bounds__(kBlock) void link_bisect(int D, SceneArgs sc, double tol,
                                                      const double* __restrict__ from, const double* __restrict__ to,
                                                      int64_t n, int known_valid, const uint8_t* __restrict__ mask,
                                                      uint8_t* __restrict__ verdict, uint32_t* __restrict__ checks) {
    int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n || (mask && !mask[e])) return;
    bisect_edge<L2>(D, sc, tol, from, to, e, known_valid, verdict, checks);
}

// Exact walk of the edges the wave kernel listed (grid-stride over the list).
__global__ __launch_bounds__(kBlock) void link_bisect_list(int D, SceneArgs sc, double tol,
                                                           const double* __restrict__ from,
                                                           const double* __restrict__ to, int known_valid,
                                                           const int64_t* __restrict__ list,
                                                           const uint32_t* __restrict__ list_count,
                                                           uint8_t* __restrict__ verdict,
                                                           uint32_t* __restrict__ checks) {
    const int64_t cnt = *list_count;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * kBlock)
        bisect_edge<false>(D, sc, tol, from, to, list[i], known_valid, verdict, checks);
}

// Instrumented bisection link (not timed): same walk and verdicts, and
// accumulates [waypoint valid() calls, FK links, segment-disk tests, tests
// within the per-link reach bound] for the arm scene into work[0..3].
__global__ __launch_bounds__(kBlock) void link_bisect_work(int D, SceneArgs sc, double tol,
                                                           const double* __restrict__ from,
                                                           const double* __restrict__ to, int64_t n, int known_valid,
                                                           unsigned long long* __restrict__ work) {
    int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n) return;
    double a[kMaxDim], b[kMaxDim], lo[kMaxDim], hi[kMaxDim];
    for (int j = 0; j < D; ++j) {
        a[j] = from[e * D + j];
        b[j] = to[e * D + j];
    }
    uint64_t calls = 0, links = 0, tests = 0, near = 0;
    bool ok = true;
    if (!known_valid) {
        ++calls;
        ok = arm_valid_counted(sc, a, &links, &tests, &near);
        if (ok) {
            ++calls;
            ok = arm_valid_counted(sc, b, &links, &tests, &near);
        }
    }
    uint64_t path = 0;
    int depth = 0;
    while (ok) {
        bisect_interval(D, a, b, path, depth, lo, hi);
        if (bisect_internal<false>(D, lo, hi, tol)) {
            for (int j = 0; j < D; ++j) lo[j] = (lo[j] + hi[j]) / 2;
            ++calls;
            if (!arm_valid_counted(sc, lo, &links, &tests, &near)) break;
            if (depth >= kMaxBisectDepth) break;
            ++depth;
        } else {
            while (depth > 0 && ((path >> (depth - 1)) & 1)) {
                path &= ~(1ull << (depth - 1));
                --depth;
            }
            if (depth == 0) break;
            path |= 1ull << (depth - 1);
        }
    }
    atomicAdd(&work[0], (unsigned long long)calls);
    atomicAdd(&work[1], (unsigned long long)links);
    atomicAdd(&work[2], (unsigned long long)tests);
    atomicAdd(&work[3], (unsigned long long)near);
}

// Host replay of the reference ring for n <= 255 (never overflows): the BFS
// order of the bisection ranges.  Table row n starts at n(n-1)/2; t holds
// kTable8Entries bytes.  No heap: at most n ranges are ever queued.
constexpr size_t kTable8Entries = (size_t)kTableMax * (kTableMax + 1) / 2;
void fill_dmv_table(uint8_t* t) {
    for (int m = 1; m <= kTableMax; ++m) {
        size_t off = (size_t)m * (m - 1) / 2, c = 0;
        int qmin[kTableMax], qmax[kTableMax];
        int qe = 0;
        qmin[qe] = 1;
        qmax[qe++] = m;
        for (int h = 0; h < qe; ++h) {
            const int mn = qmin[h], mx = qmax[h];
            const int mid = (mn + mx) / 2;
            t[off + c++] = (uint8_t)mid;
            if (mn < mid) {
                qmin[qe] = mn;
                qmax[qe++] = mid - 1;
            }
            if (mid < mx) {
                qmin[qe] = mid + 1;
                qmax[qe++] = mx;
            }
        }
    }
}

std::vector<uint8_t> build_dmv_table() {
    std::vector<uint8_t> t(kTable8Entries);
    fill_dmv_table(t.data());
    return t;
}

// Host replay of the reference's 256-slot ring with its sequential fallback
// (discrete_motion_validator.hpp:102-127) for 256 <= n <= kTable16Max: row n
// of the 16-bit table lists the waypoint indices 1..n in check order.
constexpr size_t kTable16Entries = (size_t)table16_row(kTable16Max + 1);
void fill_dmv_table16(uint16_t* t) {
    for (int m = kTableMax + 1; m <= kTable16Max; ++m) {
        size_t c = (size_t)table16_row(m);
        int rmin[256], rmax[256];
        uint64_t qs = 0, qe = 1;
        rmin[0] = 1;
        rmax[0] = m;
        while (qs != qe) {
            const int mn = rmin[qs % 256], mx = rmax[qs % 256];
            ++qs;
            if (mn == mx) {
                t[c++] = (uint16_t)mn;
            } else if (qe + 2 < qs + 256) {
                const int mid = (mn + mx) / 2;
                t[c++] = (uint16_t)mid;
                if (mn < mid) {
                    rmin[qe % 256] = mn;
                    rmax[qe % 256] = mid - 1;
                    ++qe;
                }
                if (mid < mx) {
                    rmin[qe % 256] = mid + 1;
                    rmax[qe % 256] = mx;
                    ++qe;
                }
            } else {
                for (int i = mn; i <= mx; ++i) t[c++] = (uint16_t)i;
            }
        }
    }
}

std::vector<uint16_t> build_dmv_table16() {
    std::vector<uint16_t> t(kTable16Entries);
    fill_dmv_table16(t.data());
    return t;
}

}  // namespace

static int ensure_tables(mptg_ctx* ctx) {
    DevBuf& tb = ctx->scratch[6];
    if (!tb.p) {
        std::vector<uint8_t> t = build_dmv_table();
        int rc = ensure(ctx, tb, t.size());
        if (rc) return rc;
        MPTG_HIP(ctx, hipMemcpyAsync(tb.p, t.data(), t.size(), hipMemcpyHostToDevice, ctx->stream));
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    DevBuf& t16 = ctx->dmv_table16;
    if (!t16.p) {
        std::vector<uint16_t> t = build_dmv_table16();
        int rc = ensure(ctx, t16, t.size() * sizeof(uint16_t));
        if (rc) return rc;
        MPTG_HIP(ctx, hipMemcpyAsync(t16.p, t.data(), t.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                                     ctx->stream));
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return MPTG_OK;
}

int launch_valid(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_ok) {
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called first");
    if (ctx->scene.kind == MPTG_SCENE_NONE) return fail(ctx, MPTG_E_STATE, "no scene uploaded");
    if (ctx->scene.kind == MPTG_SCENE_ARM && ctx->scene.nlinks > ctx->dim)
        return fail(ctx, MPTG_E_ARG, "arm has more links than the state has coordinates");
    if (ctx->scene.kind != MPTG_SCENE_ARM && ctx->dim < 2) return fail(ctx, MPTG_E_ARG, "2-D scenes need dim >= 2");
    if (ctx->scene.kind == MPTG_SCENE_MESH) return launch_mesh_valid(ctx, d_q, n, d_ok);
    if (ctx->scene.kind == MPTG_SCENE_HOST) return host_valid_dev(ctx, d_q, n, d_ok);
    if (ctx->scene.kind == MPTG_SCENE_ARM && ctx->scene.nlinks == ctx->dim &&
        (ctx->dim == 5 || ctx->dim == 10 || ctx->dim == 20)) {
        const unsigned g = (unsigned)((n * 8 + kBlock - 1) / kBlock);
        switch (ctx->dim) {
            case 5: hipLaunchKernelGGL(valid_kernel_arm<5>, dim3(g), dim3(kBlock), 0, ctx->stream, scene_args(ctx), d_q, n, d_ok); break;
            case 10: hipLaunchKernelGGL(valid_kernel_arm<10>, dim3(g), dim3(kBlock), 0, ctx->stream, scene_args(ctx), d_q, n, d_ok); break;
            default: hipLaunchKernelGGL(valid_kernel_arm<20>, dim3(g), dim3(kBlock), 0, ctx->stream, scene_args(ctx), d_q, n, d_ok); break;
        }
        MPTG_HIP(ctx, hipGetLastError());
        return MPTG_OK;
    }
    dispatch_dim(ctx->dim, [&]<int D>() {
        hipLaunchKernelGGL((valid_kernel<D>), dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, scene_args(ctx), ctx->dim, d_q, n, d_ok);
    });
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

static int launch_dmv(mptg_ctx* ctx, const uint8_t* table8, const uint16_t* table16, int64_t* long_list,
                      uint32_t* long_count, const double* d_from, const double* d_to, int64_t n, int known_valid,
                      const uint8_t* d_mask, uint8_t* d_verdict, uint32_t* d_checks);

int launch_link(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order, int known_valid,
                uint8_t* d_verdict, uint32_t* d_checks) {
    return launch_link_masked(ctx, d_from, d_to, n, order, known_valid, nullptr, d_verdict, d_checks);
}

int launch_link_masked(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                       int known_valid, const uint8_t* d_mask, uint8_t* d_verdict, uint32_t* d_checks) {
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called first");
    if (ctx->scene.kind == MPTG_SCENE_NONE) return fail(ctx, MPTG_E_STATE, "no scene uploaded");
    if (ctx->scene.kind == MPTG_SCENE_ARM && ctx->scene.nlinks > ctx->dim)
        return fail(ctx, MPTG_E_ARG, "arm has more links than the state has coordinates");
    if (ctx->scene.kind != MPTG_SCENE_ARM && ctx->dim < 2) return fail(ctx, MPTG_E_ARG, "2-D scenes need dim >= 2");
    if (order != MPTG_ORDER_REFERENCE && order != MPTG_ORDER_SURROGATE) return fail(ctx, MPTG_E_ARG, "unknown order");
    if (order == MPTG_ORDER_SURROGATE &&
        (ctx->scene.kind != MPTG_SCENE_ARM || ctx->link_mode != MPTG_LINK_BISECT || ctx->force_generic))
        return fail(ctx, MPTG_E_STATE, "surrogate order is implemented for the arm scene with bisection links");
    if (n > ((int64_t)1 << 31)) return fail(ctx, MPTG_E_RANGE, "at most 2^31 edges per call");
    unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    if (ctx->scene.kind == MPTG_SCENE_HOST) return host_link_dev(ctx, d_from, d_to, n, d_mask, d_verdict, d_checks);
    if (ctx->scene.kind == MPTG_SCENE_MESH && ctx->link_mode != MPTG_LINK_DMV)
        return fail(ctx, MPTG_E_STATE, "mesh scenes use the discrete motion validator link");
    if (ctx->link_mode == MPTG_LINK_BISECT) {
        if (ctx->scene.kind == MPTG_SCENE_ARM && !ctx->force_generic) {
            int rc;
            if ((rc = ensure(ctx, ctx->scratch[7], sizeof(int64_t) * n + 256))) return rc;
            uint32_t* nirr = (uint32_t*)ctx->scratch[7].p;
            int64_t* irr = (int64_t*)((char*)ctx->scratch[7].p + 256);
            rc = launch_arm_link(ctx, d_from, d_to, n, known_valid, d_mask, d_verdict, d_checks, irr, nirr, nullptr,
                                 order);
            if (rc == MPTG_OK) {
                phase_mark(ctx, 4);
                // listed (irregular) edges: exact per-edge walk; usually an empty launch
                hipLaunchKernelGGL(link_bisect_list, dim3(16), dim3(kBlock), 0, ctx->stream, ctx->dim,
                                   scene_args(ctx), ctx->link_param, d_from, d_to, known_valid, (const int64_t*)irr,
                                   (const uint32_t*)nirr, d_verdict, d_checks);
                MPTG_HIP(ctx, hipGetLastError());
                phase_mark(ctx, 5);
                return MPTG_OK;
            }
            if (rc != MPTG_E_STATE) return rc;  // E_STATE: no wave kernel for this arm, use the generic one
        }
        hipLaunchKernelGGL(link_bisect<false>, dim3(blocks), dim3(kBlock), 0, ctx->stream, ctx->dim,
                           scene_args(ctx), ctx->link_param, d_from, d_to, n, known_valid, d_mask, d_verdict,
                           d_checks);
        MPTG_HIP(ctx, hipGetLastError());
        return MPTG_OK;
    }
    if (ctx->link_mode == MPTG_LINK_BISECT_L2) {
        hipLaunchKernelGGL(link_bisect<true>, dim3(blocks), dim3(kBlock), 0, ctx->stream, ctx->dim, scene_args(ctx),
                           ctx->link_param, d_from, d_to, n, known_valid, d_mask, d_verdict, d_checks);
        MPTG_HIP(ctx, hipGetLastError());
        return MPTG_OK;
    }
    int rc = ensure_tables(ctx);
    if (rc) return rc;
    if (ctx->scene.kind == MPTG_SCENE_MESH)
        return launch_mesh_link(ctx, (const uint8_t*)ctx->scratch[6].p, d_from, d_to, n, known_valid, d_mask,
                                d_verdict, d_checks);
    if ((rc = ensure(ctx, ctx->scratch[7], sizeof(int64_t) * n + 256))) return rc;
    uint32_t* long_count = (uint32_t*)ctx->scratch[7].p;
    int64_t* long_list = (int64_t*)((char*)ctx->scratch[7].p + 256);
    return launch_dmv(ctx, (const uint8_t*)ctx->scratch[6].p, (const uint16_t*)ctx->dmv_table16.p, long_list,
                      long_count, d_from, d_to, n, known_valid, d_mask, d_verdict, d_checks);
}

// The DMV link kernels over n edges, with the order tables and the long-edge
// list (n entries) and count given by the caller.
static int launch_dmv(mptg_ctx* ctx, const uint8_t* table8, const uint16_t* table16, int64_t* long_list,
                      uint32_t* long_count, const double* d_from, const double* d_to, int64_t n, int known_valid,
                      const uint8_t* d_mask, uint8_t* d_verdict, uint32_t* d_checks) {
    int rc;
    if ((rc = launch_zero_u32(ctx, long_count, 1))) return rc;
    double inv = 1.0 / ctx->link_param;
    unsigned lg = (unsigned)std::min<int64_t>(n, (int64_t)ctx->cus * ctx->tune.dmv_long_waves_per_cu);
    // lanes per edge, fewer where n edges at the tuned count would pass 2^31 threads
    int lanes_log2 = __builtin_ctz((unsigned)ctx->tune.dmv_edge_lanes);
    while (lanes_log2 > 0 && (n << lanes_log2) > ((int64_t)1 << 31)) --lanes_log2;
    const unsigned sblocks = (unsigned)(((n << lanes_log2) + kBlock - 1) / kBlock);
    auto launch = [&]<int D, int K>() {
        hipLaunchKernelGGL((link_dmv_lanes<D, K>), dim3(sblocks), dim3(kBlock), 0, ctx->stream, space_args(ctx),
                           scene_args(ctx), inv, lanes_log2, table8, d_from, d_to, n, known_valid, d_mask, d_verdict,
                           d_checks, long_list, long_count);
        hipLaunchKernelGGL((link_dmv_long<D, K>), dim3(lg), dim3(64), 0, ctx->stream, space_args(ctx),
                           scene_args(ctx), inv, table16, d_from, d_to, long_list, long_count, d_verdict, d_checks);
    };
    // 2-D grids and boxes get kernels of their own scene kind
    if (ctx->dim == 2 && ctx->scene.kind == MPTG_SCENE_GRID)
        launch.template operator()<2, MPTG_SCENE_GRID>();
    else if (ctx->dim == 2 && ctx->scene.kind == MPTG_SCENE_BOXES)
        launch.template operator()<2, MPTG_SCENE_BOXES>();
    else
        dispatch_dim(ctx->dim, [&]<int D>() { launch.template operator()<D, 0>(); });
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// d_work[0..4]: the wave kernel's executed work (instrumented twin of the
// timed kernel); d_work[5..8]: the reference algorithm's work on the same
// edges (reference order, every disk tested until the first hit, and the
// tests within the per-link reach bound).
int launch_link_work(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order, int known_valid,
                     unsigned long long* d_work, uint8_t* d_verdict, uint32_t* d_checks) {
    if (ctx->scene.kind != MPTG_SCENE_ARM || ctx->link_mode != MPTG_LINK_BISECT)
        return fail(ctx, MPTG_E_STATE, "work counting is implemented for the arm scene with bisection links");
    int rc;
    if ((rc = ensure(ctx, ctx->scratch[7], sizeof(int64_t) * n + 256))) return rc;
    uint32_t* nirr = (uint32_t*)ctx->scratch[7].p;
    int64_t* irr = (int64_t*)((char*)ctx->scratch[7].p + 256);
    MPTG_HIP(ctx, hipMemsetAsync(d_work, 0, 9 * sizeof(unsigned long long), ctx->stream));
    rc = launch_arm_link(ctx, d_from, d_to, n, known_valid, nullptr, d_verdict, d_checks, irr, nirr, d_work, order);
    if (rc) return rc == MPTG_E_STATE ? fail(ctx, rc, "no wave kernel for this arm") : rc;
    hipLaunchKernelGGL(link_bisect_work, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       ctx->dim, scene_args(ctx), ctx->link_param, d_from, d_to, n, known_valid, d_work + 5);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// ---- allocation-free DMV link (mptg_link_dev_noalloc) ---------------------
// Everything the DMV kernels need beyond the caller's inputs and outputs sits
// in one static device array per GPU, `scratch`:
//   [kNaCountOff, +256)            the long-edge count
//   [kNaTableOff, +kTable8Entries) the 8-bit BFS order table (n <= 255)
//   [kNaTable16Off, ...)           the 16-bit ring-order table (256..1023)
//   [kNaListOff, ...)              the long-edge list, kNoAllocMaxEdges int64
// Host code uses the device address hipGetSymbolAddress returns for it (kept
// per device), never the array's name.  The tables are written once per
// device at the first call on it; the count is cleared by a kernel on every
// call.  Calls that overlap on one GPU share the array and race.
namespace {
constexpr int64_t kNoAllocMaxEdges = (int64_t)1 << 18;
constexpr size_t kNaCountOff = 0;
constexpr size_t kNaTableOff = 256;
constexpr size_t kNaTable16Off = kNaTableOff + ((kTable8Entries + 255) / 256) * 256;
constexpr size_t kNaListOff = kNaTable16Off + ((kTable16Entries * sizeof(uint16_t) + 255) / 256) * 256;
constexpr size_t kNaBytes = kNaListOff + sizeof(int64_t) * (size_t)kNoAllocMaxEdges;

__device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
struct NoAllocDevice {
    std::once_flag once;
    unsigned char* ptr = nullptr;  // device address of `scratch` on this device
    hipError_t err = hipSuccess;
};
NoAllocDevice g_noalloc[kNaMaxDevices];
// host images of the two tables (static storage: the set-up allocates nothing)
uint8_t g_table8[kTable8Entries];
uint16_t g_table16[kTable16Entries];
std::once_flag g_tables_once;

// Per-device set-up, once, with the context's device current: the array's
// address and the order tables.
int noalloc_scratch(mptg_ctx* ctx, unsigned char** out) {
    if (ctx->device < 0 || ctx->device >= kNaMaxDevices) return fail(ctx, MPTG_E_RANGE, "device id beyond 63");
    std::call_once(g_tables_once, [] {
        fill_dmv_table(g_table8);
        fill_dmv_table16(g_table16);
    });
    NoAllocDevice& d = g_noalloc[ctx->device];
    std::call_once(d.once, [&] {
        void* p = nullptr;
        d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (d.err != hipSuccess) return;
        unsigned char* base = static_cast<unsigned char*>(p);
        d.err = hipMemcpy(base + kNaTableOff, g_table8, sizeof(g_table8), hipMemcpyHostToDevice);
        if (d.err != hipSuccess) return;
        d.err = hipMemcpy(base + kNaTable16Off, g_table16, sizeof(g_table16), hipMemcpyHostToDevice);
        if (d.err != hipSuccess) return;
        d.ptr = base;
    });
    if (!d.ptr)
        return fail(ctx, MPTG_E_HIP, std::string("static scratch set-up failed: ") + hipGetErrorString(d.err));
    *out = d.ptr;
    return MPTG_OK;
}
}  // namespace

int link_dev_noalloc(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                     int known_valid, uint8_t* d_verdict, uint32_t* d_checks) {
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called first");
    if (ctx->scene.kind == MPTG_SCENE_NONE) return fail(ctx, MPTG_E_STATE, "no scene uploaded");
    if (ctx->link_mode != MPTG_LINK_DMV || ctx->scene.kind == MPTG_SCENE_MESH || ctx->scene.kind == MPTG_SCENE_HOST)
        return fail(ctx, MPTG_E_STATE, "the allocation-free link covers discrete-motion-validator scenes "
                                       "(boxes, grids, arms, circles and rectangles with DMV links)");
    if (ctx->scene.kind == MPTG_SCENE_ARM && ctx->scene.nlinks > ctx->dim)
        return fail(ctx, MPTG_E_ARG, "arm has more links than the state has coordinates");
    if (ctx->scene.kind != MPTG_SCENE_ARM && ctx->dim < 2) return fail(ctx, MPTG_E_ARG, "2-D scenes need dim >= 2");
    if (order != MPTG_ORDER_REFERENCE) return fail(ctx, MPTG_E_STATE, "DMV links check in the reference order");
    if (n > kNoAllocMaxEdges)
        return fail(ctx, MPTG_E_RANGE, "the allocation-free link takes at most 2^18 edges per call");
    if (n == 0) return MPTG_OK;
    unsigned char* base = nullptr;
    int rc = noalloc_scratch(ctx, &base);
    if (rc) return rc;
    return launch_dmv(ctx, base + kNaTableOff, reinterpret_cast<const uint16_t*>(base + kNaTable16Off),
                      reinterpret_cast<int64_t*>(base + kNaListOff), reinterpret_cast<uint32_t*>(base + kNaCountOff),
                      d_from, d_to, n, known_valid, nullptr, d_verdict, d_checks);
}

}  // namespace mptg
