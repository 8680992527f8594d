// Exact k-nearest-neighbour search over the tree (the role of nigh::Nigh::nearest,
// src/mpt/impl/prrt/prrt.hpp:406-409 and src/mpt/impl/rrg_rewire_neighbors.hpp:66,127).
//
// Result contract: for each query the k tree nodes with the smallest FP64
// metric distance (<= radius), ordered by (distance, node id).  Distances are
// computed with the same left-to-right FP64 expression as the CPU oracle, so
// indices and distances are bit-exact.
//
// Kernel shape (gfx950, wave64):
//   * one lane owns one query; a 256-thread block holds 4 waves = 256 queries;
//   * blockIdx.y splits the tree into chunks so that a launch has enough
//     waves to fill 256 CUs even for a few thousand queries;
//   * every lane of a wave visits the same tree node at the same time, so the
//     node's FP32 row is a wave-uniform (scalar) load broadcast to all lanes;
//   * FP32 prefilter: d32 is compared against the lane's current k-th FP64
//     distance plus a proven rounding margin; only lanes that pass evaluate
//     the FP64 distance and insert into a register-resident sorted list;
//   * partial lists of the chunks are merged by a second kernel;
//   * SE(3) queries reverse the roles: a query gets a workgroup, its lanes take
//     consecutive nodes (knn_wave_se3).
//
// Prefilter soundness.  With u = 2^-24, S = max|q| + max|p| and
// eps = 2.01 u S, each FP32 coordinate difference satisfies
// |t_i - (q_i - p_i)| <= eps.  For L2 (squared distances d):
//   |d32 - d| <= 2 eps sqrt(D d) + D eps^2 + g_D (d + 2 eps sqrt(D d) + D eps^2),
//   g_D = D u / (1 - D u)
// and for L1: |d32 - d| <= D eps + g_D (d + D eps).  A node can enter the
// list only if d <= T (T = current k-th distance, squared for L2), hence only
// if d32 <= T + eta(T).  The kernels use 2 * eta plus a relative 2^-40 for
// the sqrt rounding of list entries.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <limits>

#include "internal.hpp"
#include "wave_list.hpp"

namespace mptg {
namespace {

constexpr int kBlock = 256;
// Longest k-list of the SE(3) wave kernel (one entry per lane).
constexpr int kWaveMaxK = 64;
// Fewest tree nodes per split of the scan: small trees still spread over
// many waves, while every split's partial list stays worth its merge.
constexpr int64_t kMinSplitNodes = 64;
// Tree nodes every k = 1 query measures first, for a starting bound: every
// (n / kSeedNodes)-th node, spread over the tree rather than its first nodes
// (a planner's first nodes crowd around its start).
constexpr int64_t kSeedNodes = 256;

// Register list length of a k-list scan: k rounded up to 1, 20, 32 or 64.
inline int k_bucket(int k) { return k <= 1 ? 1 : (k <= 20 ? 20 : (k <= 32 ? 32 : 64)); }

__device__ __forceinline__ bool better(double a, int ia, double b, int ib) {
    return a < b || (a == b && ia < ib);
}

template <int K>
struct TopK {
    double d[K];
    int32_t i[K];

    __device__ __forceinline__ void init() {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            d[j] = __builtin_inf();
            i[j] = INT_MAX;
        }
    }
    // Inserts (nd, ni) if it beats the last entry; keeps ascending order.
    __device__ __forceinline__ void insert(double nd, int32_t ni) {
        if (!better(nd, ni, d[K - 1], i[K - 1])) return;
        bool placed = false;
#pragma unroll
        for (int j = K - 1; j >= 1; --j) {
            bool mv = better(nd, ni, d[j - 1], i[j - 1]);
            double dj = mv ? d[j - 1] : (placed ? d[j] : nd);
            int32_t ij = mv ? i[j - 1] : (placed ? i[j] : ni);
            placed = placed || !mv;
            d[j] = dj;
            i[j] = ij;
        }
        if (!placed) {
            d[0] = nd;
            i[0] = ni;
        }
    }
};

// FP64 metric value, left-to-right, no contraction (matches the oracle).
template <int D, int METRIC>
__device__ __forceinline__ double exact_dist(const double* __restrict__ a, const double* __restrict__ b) {
    double s = 0.0;
    if constexpr (METRIC == MPTG_METRIC_L1) {
#pragma unroll
        for (int i = 0; i < D; ++i) s = s + __builtin_fabs(a[i] - b[i]);
        return s;
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double t = a[i] - b[i];
            s = s + t * t;
        }
        return __builtin_sqrt(s);
    }
}

// Prefilter threshold in the prefilter's units (squared distance for L2,
// distance for L1) from the list's k-th distance `dk` (metric value).
template <int D, int METRIC>
__device__ __forceinline__ float prefilter_threshold(double dk, double S) {
    if (!(dk < __builtin_inf())) return __builtin_inff();
    const double u = 0x1.0p-24;
    const double eps = 2.01 * u * S;
    const double g = D * u / (1.0 - D * u);
    double T, eta;
    if constexpr (METRIC == MPTG_METRIC_L1) {
        T = dk;
        eta = D * eps + g * (T + D * eps);
    } else {
        T = dk * dk;
        double r = 2.0 * eps * __builtin_sqrt(D * T) + D * eps * eps;
        eta = r + g * (T + r);
    }
    double thr = T * (1.0 + 0x1.0p-40) + 2.0 * eta + 1e-300;
    float f = (float)thr;
    // round up (thr > 0) so that the FP32 comparison never undercuts the FP64 bound
    return (double)f < thr ? float_of(bits_of(f) + 1u) : f;
}

// An upper bound, in metric units, on the exact distance of a node whose
// prefilter value (FP32 sum, squared for L2) is `a`: the inverse of the
// error bound behind prefilter_threshold.  L1: |a - d| <= D eps + g (d +
// D eps) gives d <= (a + D eps (1 + g)) / (1 - g).  L2: |a - T| <= r + g (T +
// r) with r = 2 eps sqrt(D T) + D eps^2 gives, for x = sqrt(T), (1 - g) x^2
// - 2 (1 + g) eps sqrt(D) x - (a + (1 + g) D eps^2) <= 0, bounded by the
// quadratic's larger root.  A relative 2^-40 covers the FP64 rounding.
template <int D, int METRIC>
__device__ __forceinline__ double prefilter_upper(float a, double S) {
    const double u = 0x1.0p-24;
    const double eps = 2.01 * u * S;
    const double g = D * u / (1.0 - D * u);
    double x;
    if constexpr (METRIC == MPTG_METRIC_L1) {
        x = ((double)a + D * eps * (1.0 + g)) / (1.0 - g);
    } else {
        const double bq = (1.0 + g) * eps * __builtin_sqrt((double)D);
        const double c = (double)a + (1.0 + g) * D * eps * eps;
        x = (bq + __builtin_sqrt(bq * bq + (1.0 - g) * c)) / (1.0 - g);
    }
    return x * (1.0 + 0x1.0p-40) + 1e-300;
}

// Splits actually used for a tree of n nodes when the launch was planned
// for `planned` (a graph captured for a larger tree): at least kMinSplitNodes
// per split, as the host plan, and no empty split.
__device__ __forceinline__ int64_t knn_used_splits(int64_t n, int64_t planned, int64_t* chunk) {
    int64_t s = (n + kMinSplitNodes - 1) / kMinSplitNodes;
    s = s < planned ? s : planned;
    s = s > 1 ? s : 1;
    *chunk = n > 0 ? (n + s - 1) / s : 0;
    return n > 0 ? (n + *chunk - 1) / *chunk : 1;
}

// Tree size: n, or *n_dev when given (launches captured in a graph read the
// size the previous iteration left; splits beyond those the size needs
// return at once); *skip != 0 makes the launch a no-op.
__device__ __forceinline__ bool knn_extent(const int64_t* n_dev, const int32_t* skip, int64_t* n, int64_t* chunk) {
    if (skip && *skip) return false;
    if (n_dev) {
        *n = *n_dev;
        if ((int64_t)blockIdx.y >= knn_used_splits(*n, gridDim.y, chunk)) return false;
    }
    return true;
}

// Splits of a byte-prefilter scan (k = 1) over n nodes for `live` queries,
// chosen on the device because the live count is only known there: enough
// chunks that the live query blocks times the splits fill `target` blocks,
// no more than `planned` (the partial results' splits, the grid's y extent)
// and at least kMinSplitNodes nodes per chunk.  Returns the splits used
// (every chunk non-empty) and sets the chunk length.
__device__ __forceinline__ int64_t q8_splits(int64_t n, int64_t live, int64_t planned, int64_t target,
                                             int64_t* chunk) {
    const int64_t qb = live > 0 ? (live + kBlock - 1) / kBlock : 1;
    int64_t s = (target + qb - 1) / qb;
    const int64_t max_s = (n + kMinSplitNodes - 1) / kMinSplitNodes;
    s = s < max_s ? s : max_s;
    s = s < planned ? s : planned;
    s = s > 1 ? s : 1;
    *chunk = n > 0 ? (n + s - 1) / s : 0;
    return n > 0 ? (n + *chunk - 1) / *chunk : 1;
}

// Nearest of kSeedNodes tree nodes (every (n / kSeedNodes)-th) to each query
// (exact FP64 distance): a starting bound for the k = 1 scan.  kSeedLanes
// lanes share a query, each measuring every kSeedLanes-th of those nodes;
// the minimum is exact in any order.
constexpr int kSeedLanes = 32;
template <int D, int METRIC>
__global__ __launch_bounds__(kBlock) void knn_seed(const double* __restrict__ x64, int64_t n,
                                                   const double* __restrict__ q, int64_t nq,
                                                   double* __restrict__ seed, const int64_t* n_dev,
                                                   const int32_t* skip, const uint32_t* nq_dev) {
    static_assert(64 % kSeedLanes == 0, "lane groups within a wave");
    if (skip && *skip) return;  // uniform: every thread reads the same flag
    if (nq_dev) nq = *nq_dev;   // live queries: the first *nq_dev slots
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t qi = t / kSeedLanes;
    const int part = (int)(t % kSeedLanes);
    if (n_dev) n = *n_dev;
    const int64_t ns = n < kSeedNodes ? n : kSeedNodes;
    const int64_t stride = ns > 0 ? n / ns : 1;
    double best = __builtin_inf();
    if (qi < nq) {
#pragma unroll 8
        for (int64_t j = part; j < ns; j += kSeedLanes)
            best = __builtin_fmin(best, exact_dist<D, METRIC>(q + qi * D, x64 + j * stride * D));
    }
#pragma unroll
    for (int o = 1; o < kSeedLanes; o <<= 1) best = __builtin_fmin(best, __shfl_xor(best, o));
    if (qi < nq && part == 0) seed[qi] = best;
}

// Tree nodes a block stages in LDS at a time: FP32 rows for the prefilter
// and FP64 rows for the exact distances, so a lane that passes the prefilter
// reads its node from LDS instead of waiting on a global load while the
// rest of its wave idles.
constexpr int kScanTile = 64;

template <int D, int METRIC, int K>
__global__ __launch_bounds__(kBlock) void knn_scan(const float* __restrict__ x32, const double* __restrict__ x64,
                                                   int64_t n, int64_t chunk, const double* __restrict__ q,
                                                   int64_t nq, double radius, double pbound,
                                                   double* __restrict__ out_d, int32_t* __restrict__ out_i,
                                                   const int64_t* n_dev, const int32_t* skip,
                                                   const double* __restrict__ seed, const uint32_t* nq_dev) {
    constexpr int DP = (D <= 2) ? 2 : (D <= 4 ? 4 : (D <= 8 ? 8 : ((D + 3) / 4) * 4));
    __shared__ __align__(16) float t32[kScanTile * DP];
    __shared__ __align__(16) double t64[kScanTile * D];
    if (!knn_extent(n_dev, skip, &n, &chunk)) return;  // block-uniform
    // live queries: all nq, or the first *nq_dev (nq stays the output stride)
    const int64_t live = nq_dev ? (int64_t)*nq_dev : nq;
    if ((int64_t)blockIdx.x * kBlock >= live) return;  // block-uniform
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = qi < live;
    const int64_t qc = active ? qi : 0;
    const int64_t p0 = (int64_t)blockIdx.y * chunk;
    const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;

    float q32[D];
    double mq = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double v = q[qc * D + i];
        q32[i] = (float)v;
        mq = __builtin_fmax(mq, __builtin_fabs(v));
    }
    const double S = mq + pbound;
    TopK<K> top;
    top.init();
    // seed bound (k = 1): the nearest of a subset of the tree, measured by
    // knn_seed.  The nearest node overall is no farther, so it still passes
    // the prefilter in its own chunk; other chunks skip the nodes beyond the
    // bound (their partial lists may then hold nothing) and the merge of the
    // chunks' minima is unchanged
    if (seed) radius = __builtin_fmin(radius, seed[qc]);
    float rthr = prefilter_threshold<D, METRIC>(radius, S);
    float thr = rthr;
    if (!active) thr = -1.0f;  // inactive lanes never take the slow path
    // k = 1 candidates awaiting their exact distance.  A node whose prefilter
    // value exceeds a threshold derived from a listed candidate's upper bound
    // is farther than that candidate, so it cannot be the nearest; every
    // listed node gets its exact distance before the result is written.
    constexpr int kCand = 4;
    int32_t cand[kCand];
    int ncand = 0;
    auto flush = [&]() {
#pragma unroll
        for (int c = 0; c < kCand; ++c)
            if (c < ncand) {
                const double dq = exact_dist<D, METRIC>(q + qc * D, x64 + (int64_t)cand[c] * D);
                if (dq <= radius) top.insert(dq, cand[c]);
            }
        ncand = 0;
    };

    for (int64_t t0 = p0; t0 < p1; t0 += kScanTile) {  // block-uniform bounds
        const int cnt = (int)(p1 - t0 < kScanTile ? p1 - t0 : kScanTile);
        __syncthreads();  // every wave is done with the previous tile
        for (int i = threadIdx.x; i < cnt * DP; i += kBlock) t32[i] = x32[t0 * DP + i];
        // k = 1 candidates read their FP64 rows at the flush, from global memory
        if constexpr (K > 1)
            for (int i = threadIdx.x; i < cnt * D; i += kBlock) t64[i] = x64[t0 * D + i];
        __syncthreads();
        // four tree nodes per step, one branch for all
        constexpr int U = 4;
        int j = 0;
        for (; j + U <= cnt; j += U) {
            float acc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float* row = t32 + (j + u) * DP;
                float a = 0.0f;
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    float t = q32[i] - row[i];
                    if constexpr (METRIC == MPTG_METRIC_L1) a = a + __builtin_fabsf(t);
                    else a = __builtin_fmaf(t, t, a);
                }
                acc[u] = a;
            }
            uint32_t pass = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) pass |= (acc[u] <= thr ? 1u : 0u) << u;
            if constexpr (K == 1) {
                // k = 1: list the candidates and tighten the threshold from
                // their prefilter values; their exact distances are taken
                // together, when the list is full or at the chunk's end
                while (pass) {
                    const int u = __builtin_ctz(pass);
                    pass &= pass - 1;
                    if (ncand == kCand) flush();
                    cand[ncand++] = (int32_t)(t0 + j + u);
                    float t = prefilter_threshold<D, METRIC>(prefilter_upper<D, METRIC>(acc[u], S), S);
                    thr = t < thr ? t : thr;
                }
            } else {
                // candidates in node order; a candidate that a lower
                // threshold would now reject is still safe to evaluate
                // (insert keeps the exact order)
                while (pass) {
                    const int u = __builtin_ctz(pass);
                    pass &= pass - 1;
                    double dq = exact_dist<D, METRIC>(q + qc * D, t64 + (j + u) * D);
                    if (dq <= radius) {
                        top.insert(dq, (int32_t)(t0 + j + u));
                        float t = prefilter_threshold<D, METRIC>(top.d[K - 1], S);
                        thr = t < rthr ? t : rthr;
                    }
                }
            }
        }
        for (; j < cnt; ++j) {
            const float* row = t32 + j * DP;
            float acc = 0.0f;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                float t = q32[i] - row[i];
                if constexpr (METRIC == MPTG_METRIC_L1) acc = acc + __builtin_fabsf(t);
                else acc = __builtin_fmaf(t, t, acc);
            }
            if (acc <= thr) {
                if constexpr (K == 1) {
                    if (ncand == kCand) flush();
                    cand[ncand++] = (int32_t)(t0 + j);
                    float t = prefilter_threshold<D, METRIC>(prefilter_upper<D, METRIC>(acc, S), S);
                    thr = t < thr ? t : thr;
                } else {
                    double dq = exact_dist<D, METRIC>(q + qc * D, t64 + j * D);
                    if (dq <= radius) {
                        top.insert(dq, (int32_t)(t0 + j));
                        float t = prefilter_threshold<D, METRIC>(top.d[K - 1], S);
                        thr = t < rthr ? t : rthr;
                    }
                }
            }
        }
    }
    if constexpr (K == 1) flush();
    if (!active) return;
    const int64_t base = ((int64_t)blockIdx.y * nq + qi) * K;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        out_d[base + j] = top.d[j];
        out_i[base + j] = top.i[j];
    }
}

// ---- k = 1 under L1 with a byte prefilter ------------------------------------
// Every coordinate is quantized to a byte on one grid over [-B, B] (B = the
// context's bound on every tree and query coordinate): q(x) = round((x + B)
// / s), s = 2B / 255.  The L1 distance of two byte rows is a sum of absolute
// byte differences, four coordinates per v_sad_u8 (exact integer arithmetic).
//
// Soundness.  For |x| <= B the FP32 quantization (of the FP32 shadow row, or
// of the FP64 query) satisfies |(x + B) / s - q(x)| <= 1/2 + 6e-5 (rounding to
// the nearest byte, FP32 arithmetic on values <= 255, the shadow's own
// rounding 2^-24 B / s).  Per coordinate, | |a - b| / s - |qa - qb| | <= 1 +
// 1.2e-4, so for the sum SAD over D coordinates
//     | L1 / s - SAD | <= D (1 + 1.2e-4) < kQ8Slack D,   kQ8Slack = 1.001.
// A node with SAD > T = floor((d (1 + 2^-40)) / s + kQ8Slack D) therefore has
// an exact L1 above d (by a relative 2^-40, more than the FP64 rounding of
// either sum), and cannot beat a listed distance d.  A listed candidate with
// byte distance a has L1 <= s (a + kQ8Slack D); a node with SAD > a + C2,
// C2 = ceil(2 kQ8Slack D) + 1, is farther by at least s.  Queries with a
// coordinate beyond B take no prefilter (every node gets its exact distance).
//
// One pass with a bound shared by the query's chunks.  knn_q8_seed measures
// kSeedNodes spread tree nodes exactly and starts the query's byte bound
// gthr[q] at T(their nearest).  Every byte bound taken below is at least the
// nearest node's SAD: T(d) for the distance d of a real node (the nearest is
// no farther), and a + C2 for any node of byte distance a (the nearest has
// L1 <= that node's L1 <= s (a + kQ8Slack D), so SAD <= a + C2).  knn_scan_q8
// keeps a lane bound thr = min(T(radius), gthr[q], a + C2 over its
// candidates so far), lowers gthr[q] to it with atomicMin at the end of each
// tile and reads it back at the start of the next, so each chunk's scan
// tightens the others'; a node gets its exact FP64 distance only when its
// SAD is within thr.  Nodes are tested 8 at a time against the minimum of
// their SADs first.  (Measured alternatives at a 10^5-node tree, 2 725
// queries: a separate pass for each query's exact byte minimum first, 160 +
// 223 us; byte rows in scalar registers, 4 queries per lane, 137 + 175 us;
// that layout in one pass, slower again from its divergent candidate
// checks; this one, 16 + 223 us.  Tiles of 128 or 512 nodes instead of 256
// cost occupancy: 3 and 2 waves per SIMD against 4.)
constexpr int kQ8Tile = 256;
constexpr double kQ8Slack = 1.001;
constexpr int kQ8BlocksPerCu = 4;

__device__ __forceinline__ uint32_t q8_byte(float x, float B, float inv_s) {
    float f = (x + B) * inv_s + 0.5f;
    f = f < 0.0f ? 0.0f : (f > 255.0f ? 255.0f : f);
    return (uint32_t)f;
}

// The query's byte row (NW words); false when a coordinate lies beyond B.
template <int D>
__device__ __forceinline__ bool q8_query(const double* __restrict__ q, double bound, float Bf, float inv_sf,
                                         uint32_t* qw) {
    constexpr int NW = (D + 3) / 4;
    bool inside = true;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = 4 * w + b;
            if (i < D) {
                const double x = q[i];
                inside = inside && __builtin_fabs(x) <= bound;
                v |= q8_byte((float)x, Bf, inv_sf) << (8 * b);
            }
        }
        qw[w] = v;
    }
    return inside;
}

// Stages tree nodes [t0, t0 + cnt) as byte rows, word w of node j at
// t8[w * kQ8Tile + j].  Whole block.  For D >= 8 the FP32 row pitch is 4 NW
// floats, so word w is one aligned 16-byte load; all of a thread's loads are
// issued before any is used.  Bytes of coordinates >= D stay 0, as in the
// query's row.
template <int D>
__device__ __forceinline__ void q8_stage(const float* __restrict__ x32, int64_t t0, int cnt, float Bf, float inv_sf,
                                         uint32_t* t8) {
    constexpr int DP = (D <= 2) ? 2 : (D <= 4 ? 4 : (D <= 8 ? 8 : ((D + 3) / 4) * 4));
    constexpr int NW = (D + 3) / 4;
    static_assert(D >= 8 && DP == 4 * NW, "byte rows read whole FP32 words");
    constexpr int WPT = (NW * kQ8Tile + kBlock - 1) / kBlock;  // words per thread, at most
    constexpr uint32_t kLast = D % 4 == 0 ? 0xffffffffu : ((1u << (8 * (D % 4))) - 1u);
    const int words = cnt * NW;
    float4 v[WPT];
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        const int e = threadIdx.x + k * kBlock;
        const int j = e / NW, w = e - j * NW;
        v[k] = e < words ? *reinterpret_cast<const float4*>(x32 + (t0 + j) * DP + 4 * w) : make_float4(0, 0, 0, 0);
    }
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        const int e = threadIdx.x + k * kBlock;
        if (e < words) {
            const int j = e / NW, w = e - j * NW;
            uint32_t b = q8_byte(v[k].x, Bf, inv_sf) | (q8_byte(v[k].y, Bf, inv_sf) << 8) |
                         (q8_byte(v[k].z, Bf, inv_sf) << 16) | (q8_byte(v[k].w, Bf, inv_sf) << 24);
            if (w == NW - 1) b &= kLast;
            t8[w * kQ8Tile + j] = b;
        }
    }
}

// Byte bound of a distance d (see above): T(d) = floor(d (1 + 2^-40) / s +
// kQ8Slack D); ~0 for d = inf or beyond the word.
__device__ __forceinline__ uint32_t q8_thr(double d, double inv_s, int D) {
    if (!(d < __builtin_inf())) return 0xffffffffu;
    const double t = d * (1.0 + 0x1.0p-40) * inv_s + kQ8Slack * D;
    return t >= 4294967295.0 ? 0xffffffffu : (uint32_t)t;
}

// gthr[q] = T(exact distance to the nearest of kSeedNodes spread tree nodes)
// for every live query (knn_seed's sample).  A block serves kQ8SeedQueries
// queries: it stages the sample's coordinates in LDS once (coordinate-major,
// so the lanes of a query read consecutive words), and kQ8SeedLanes lanes
// per query measure every kQ8SeedLanes-th sample node.  The minimum is exact
// in any order, and each distance is exact_dist's left-to-right sum.  (A
// wave per query reading the sample from L2 took 17 us for ~2 700 queries.)
constexpr int kQ8SeedQueries = 16;
constexpr int kQ8SeedLanes = kBlock / kQ8SeedQueries;
template <int D>
__global__ __launch_bounds__(kBlock) void knn_q8_seed(const double* __restrict__ x64, int64_t n,
                                                      const double* __restrict__ q, int64_t nq, double bound,
                                                      uint32_t* __restrict__ gthr, const int64_t* n_dev,
                                                      const int32_t* skip, const uint32_t* nq_dev) {
    static_assert(64 % kQ8SeedLanes == 0, "lane groups within a wave");
    __shared__ double xs[D * kSeedNodes];  // coordinate i of sample node j at xs[i * kSeedNodes + j]
    if (skip && *skip) return;
    if (nq_dev) nq = *nq_dev;
    if ((int64_t)blockIdx.x * kQ8SeedQueries >= nq) return;  // block-uniform
    if (n_dev) n = *n_dev;
    const int ns = (int)(n < kSeedNodes ? n : kSeedNodes);
    const int64_t stride = ns > 0 ? n / ns : 1;
    for (int e = threadIdx.x; e < ns * D; e += kBlock) {
        const int j = e / D, i = e - j * D;
        xs[i * kSeedNodes + j] = x64[(int64_t)j * stride * D + i];
    }
    __syncthreads();
    const int64_t qi = (int64_t)blockIdx.x * kQ8SeedQueries + threadIdx.x / kQ8SeedLanes;
    const int part = (int)(threadIdx.x % kQ8SeedLanes);
    double best = __builtin_inf();
    if (qi < nq) {
        double qv[D];
#pragma unroll
        for (int i = 0; i < D; ++i) qv[i] = q[qi * D + i];
#pragma unroll 2
        for (int j = part; j < ns; j += kQ8SeedLanes) {
            double x[D];
#pragma unroll
            for (int i = 0; i < D; ++i) x[i] = xs[i * kSeedNodes + j];
            best = __builtin_fmin(best, exact_dist<D, MPTG_METRIC_L1>(qv, x));
        }
    }
#pragma unroll
    for (int o = 1; o < kQ8SeedLanes; o <<= 1) best = __builtin_fmin(best, __shfl_xor(best, o));
    if (qi < nq && part == 0) gthr[qi] = q8_thr(best, 255.0 / (2.0 * bound), D);
}

template <int D>
__global__ __launch_bounds__(kBlock) void knn_scan_q8(const float* __restrict__ x32, const double* __restrict__ x64,
                                                      int64_t n, int64_t chunk, const double* __restrict__ q,
                                                      int64_t nq, double radius, double bound,
                                                      double* __restrict__ out_d, int32_t* __restrict__ out_i,
                                                      const int64_t* n_dev, const int32_t* skip,
                                                      uint32_t* __restrict__ gthr, const uint32_t* nq_dev,
                                                      int64_t target) {
    constexpr int NW = (D + 3) / 4;  // words of a byte row
    // byte rows by word: word w of tile node j at t8[w * kQ8Tile + j]
    __shared__ __align__(16) uint32_t t8[NW * kQ8Tile];
    (void)chunk;  // re-planned below from the live query count
    if (skip && *skip) return;  // block-uniform
    if (n_dev) n = *n_dev;
    const int64_t live = nq_dev ? (int64_t)*nq_dev : nq;
    if (live <= 0) return;
    // the grid's blocks, taken in order, are (query block, split) pairs of
    // the live queries: qblocks <= gridDim.x and splits <= gridDim.y, so
    // every pair has a block
    const int64_t qblocks = (live + kBlock - 1) / kBlock;
    const int64_t splits = q8_splits(n, live, gridDim.y, target, &chunk);
    const int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    if (b >= qblocks * splits) return;  // block-uniform
    const int64_t split = b / qblocks;
    const int64_t qi = (b - split * qblocks) * kBlock + threadIdx.x;
    const bool active = qi < live;
    const int64_t qc = active ? qi : 0;
    const int64_t p0 = split * chunk;
    const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
    const float Bf = (float)bound;
    const float inv_sf = (float)(255.0 / (2.0 * bound));
    const double inv_s = 255.0 / (2.0 * bound);

    uint32_t qw[NW];
    const bool inside = q8_query<D>(q + qc * D, bound, Bf, inv_sf, qw);
    constexpr uint32_t C2 = (uint32_t)(2.0 * kQ8Slack * D) + 2;
    // a lane whose query lies outside the grid checks every node exactly;
    // inactive lanes take no node
    const bool filter = active && inside;
    const bool take_any = active;
    uint32_t thr = filter ? q8_thr(radius, inv_s, D) : 0xffffffffu;
    uint32_t shared = 0xffffffffu;  // gthr[q] as last read or published
    TopK<1> top;
    top.init();
    constexpr int kCand = 4;
    int32_t cand[kCand];
    int ncand = 0;
    auto flush = [&]() {
#pragma unroll
        for (int c = 0; c < kCand; ++c)
            if (c < ncand) {
                const double dq = exact_dist<D, MPTG_METRIC_L1>(q + qc * D, x64 + (int64_t)cand[c] * D);
                if (dq <= radius) top.insert(dq, cand[c]);
            }
        ncand = 0;
    };
    auto take = [&](uint32_t sad, int64_t node) {
        if (ncand == kCand) flush();
        cand[ncand++] = (int32_t)node;
        if (filter) {
            const uint32_t t = sad > 0xffffffffu - C2 ? 0xffffffffu : sad + C2;
            thr = t < thr ? t : thr;
        }
    };

    for (int64_t t0 = p0; t0 < p1; t0 += kQ8Tile) {  // block-uniform bounds
        const int cnt = (int)(p1 - t0 < kQ8Tile ? p1 - t0 : kQ8Tile);
        if (filter) {  // the other chunks' progress
            shared = __hip_atomic_load(gthr + qc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            thr = shared < thr ? shared : thr;
        }
        __syncthreads();  // every wave is done with the previous tile
        q8_stage<D>(x32, t0, cnt, Bf, inv_sf, t8);
        __syncthreads();
        constexpr int U = 8;
        int j = 0;
        for (; j + U <= cnt; j += U) {
            uint32_t sad[U];
#pragma unroll
            for (int u = 0; u < U; ++u) sad[u] = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint4 lo = *reinterpret_cast<const uint4*>(t8 + w * kQ8Tile + j);
                const uint4 hi = *reinterpret_cast<const uint4*>(t8 + w * kQ8Tile + j + 4);
                sad[0] = __builtin_amdgcn_sad_u8(qw[w], lo.x, sad[0]);
                sad[1] = __builtin_amdgcn_sad_u8(qw[w], lo.y, sad[1]);
                sad[2] = __builtin_amdgcn_sad_u8(qw[w], lo.z, sad[2]);
                sad[3] = __builtin_amdgcn_sad_u8(qw[w], lo.w, sad[3]);
                sad[4] = __builtin_amdgcn_sad_u8(qw[w], hi.x, sad[4]);
                sad[5] = __builtin_amdgcn_sad_u8(qw[w], hi.y, sad[5]);
                sad[6] = __builtin_amdgcn_sad_u8(qw[w], hi.z, sad[6]);
                sad[7] = __builtin_amdgcn_sad_u8(qw[w], hi.w, sad[7]);
            }
            uint32_t m8 = sad[0];
#pragma unroll
            for (int u = 1; u < U; ++u) m8 = sad[u] < m8 ? sad[u] : m8;
            if (take_any && m8 <= thr) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (sad[u] <= thr) take(sad[u], t0 + j + u);
            }
        }
        for (; j < cnt; ++j) {
            uint32_t sad = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) sad = __builtin_amdgcn_sad_u8(qw[w], t8[w * kQ8Tile + j], sad);
            if (take_any && sad <= thr) take(sad, t0 + j);
        }
        if (filter && thr < shared) {  // share this chunk's progress
            atomicMin(gthr + qc, thr);
            shared = thr;
        }
    }
    flush();
    if (!active) return;
    const int64_t base = split * nq + qi;
    out_d[base] = top.d[0];
    out_i[base] = top.i[0];
}

// Merge of the chunks' partial k-lists: each lane keeps its query's k best
// over the splits' lists.
template <int K>
__global__ __launch_bounds__(kBlock) void knn_merge(const double* __restrict__ pd, const int32_t* __restrict__ pi,
                                                    int splits, int64_t nq, int k, double* __restrict__ out_d,
                                                    int32_t* __restrict__ out_i, const int32_t* skip,
                                                    const int64_t* n_dev) {
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (qi >= nq || (skip && *skip)) return;
    if (n_dev) {
        int64_t chunk;
        splits = (int)knn_used_splits(*n_dev, splits, &chunk);
    }
    TopK<K> top;
    top.init();
    if constexpr (K == 1) {
        // the minimum over the splits, eight loads in flight at a time
        int s = 0;
        for (; s + 8 <= splits; s += 8) {
            double d[8];
            int32_t id[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                d[u] = pd[(int64_t)(s + u) * nq + qi];
                id[u] = pi[(int64_t)(s + u) * nq + qi];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (better(d[u], id[u], top.d[0], top.i[0])) {
                    top.d[0] = d[u];
                    top.i[0] = id[u];
                }
        }
        for (; s < splits; ++s) {
            const double d = pd[(int64_t)s * nq + qi];
            const int32_t id = pi[(int64_t)s * nq + qi];
            if (better(d, id, top.d[0], top.i[0])) {
                top.d[0] = d;
                top.i[0] = id;
            }
        }
    } else {
        for (int s = 0; s < splits; ++s) {
            const int64_t base = ((int64_t)s * nq + qi) * K;
            for (int j = 0; j < K; ++j) {
                double d = pd[base + j];
                int32_t id = pi[base + j];
                if (id == INT_MAX) break;
                top.insert(d, id);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j < k) {
            bool have = top.i[j] != INT_MAX;
            out_d[qi * k + j] = have ? top.d[j] : __builtin_inf();
            out_i[qi * k + j] = have ? top.i[j] : -1;
        }
    }
}

// k = 1: the nearest over the splits.  kMergeLanes lanes share a query,
// each reading every kMergeLanes-th split's partial result, then a shuffle
// minimum by (distance, node id), which is exact in any order.
constexpr int kMergeLanes = 8;
__global__ __launch_bounds__(kBlock) void knn_merge1(const double* __restrict__ pd, const int32_t* __restrict__ pi,
                                                     int splits, int64_t nq, double* __restrict__ out_d,
                                                     int32_t* __restrict__ out_i, const int32_t* skip,
                                                     const int64_t* n_dev, double inf, const uint32_t* nq_dev,
                                                     int64_t n, int64_t q8_target) {
    static_assert(64 % kMergeLanes == 0, "lane groups within a wave");
    if (skip && *skip) return;  // uniform: every thread reads the same flag
    const int64_t live = nq_dev ? (int64_t)*nq_dev : nq;  // nq stays the partial results' stride
    int64_t chunk;
    if (q8_target > 0)  // the byte-prefilter scan's own split of the live queries
        splits = (int)q8_splits(n_dev ? *n_dev : n, live, splits, q8_target, &chunk);
    else if (n_dev)
        splits = (int)knn_used_splits(*n_dev, splits, &chunk);
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t qi = t / kMergeLanes;
    const int part = (int)(t % kMergeLanes);
    double bd = inf;
    int32_t bi = INT_MAX;
    if (qi < live) {
#pragma unroll 4
        for (int s = part; s < splits; s += kMergeLanes) {
            const double d = pd[(int64_t)s * nq + qi];
            const int32_t id = pi[(int64_t)s * nq + qi];
            if (better(d, id, bd, bi)) {
                bd = d;
                bi = id;
            }
        }
    }
#pragma unroll
    for (int o = 1; o < kMergeLanes; o <<= 1) {
        const double od = __shfl_xor(bd, o);
        const int32_t oi = __shfl_xor(bi, o);
        if (better(od, oi, bd, bi)) {
            bd = od;
            bi = oi;
        }
    }
    if (qi < live && part == 0) {
        const bool have = bi != INT_MAX;
        out_d[qi] = have ? bd : inf;
        out_i[qi] = have ? bi : -1;
    }
}

template <int K>
__global__ __launch_bounds__(kBlock) void knn_finish(const double* __restrict__ pd, const int32_t* __restrict__ pi,
                                                     int64_t nq, int k, double* __restrict__ out_d,
                                                     int32_t* __restrict__ out_i, const int32_t* skip,
                                                     const uint32_t* nq_dev) {
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (qi >= nq || (skip && *skip) || (nq_dev && qi >= (int64_t)*nq_dev)) return;
    for (int j = 0; j < k; ++j) {
        int32_t id = pi[qi * K + j];
        bool have = id != INT_MAX;
        out_d[qi * k + j] = have ? pd[qi * K + j] : __builtin_inf();
        out_i[qi * k + j] = have ? id : -1;
    }
}

// ---- k > 1 over large trees: bound, collect, select -------------------------
// knn_collect: every tree node of the block's chunk that passes the FP32
// prefilter for the query's bound tau[q] (a k-th distance over a subset of
// the tree, so no smaller than the k-th over all of it; the prefilter never
// drops a node whose exact distance is within the bound) goes, as an index,
// to the (chunk, query) candidate region, cap entries at most; cnt[chunk][q]
// counts them all, so a count above cap marks an overflow.  Each lane holds
// one bound and a count, owns its region (no atomics), and issues no
// dependent load in the scan: exact distances are the select kernel's.
template <int D, int METRIC>
__global__ __launch_bounds__(kBlock) void knn_collect(const float* __restrict__ x32, int64_t n, int64_t chunk,
                                                      const double* __restrict__ q, int64_t nq,
                                                      const double* __restrict__ tau, double pbound, int cap,
                                                      int32_t* __restrict__ cand_i, uint32_t* __restrict__ cnt) {
    constexpr int DP = (D <= 2) ? 2 : (D <= 4 ? 4 : (D <= 8 ? 8 : ((D + 3) / 4) * 4));
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = qi < nq;
    const int64_t qc = active ? qi : 0;
    const int64_t p0 = (int64_t)blockIdx.y * chunk;
    const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
    float q32[D];
    double mq = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double v = q[qc * D + i];
        q32[i] = (float)v;
        mq = __builtin_fmax(mq, __builtin_fabs(v));
    }
    const float thr = active ? prefilter_threshold<D, METRIC>(tau[qc], mq + pbound) : -1.0f;
    const int64_t region = ((int64_t)blockIdx.y * nq + qc) * cap;
    uint32_t c = 0;
    auto take = [&](int64_t j) {
        if (c < (uint32_t)cap) cand_i[region + c] = (int32_t)j;
        ++c;
    };
    constexpr int U = 4;
    int64_t p = p0;
    for (; p + U <= p1; p += U) {
        float acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float* row = x32 + (p + u) * DP;
            float a = 0.0f;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                float dd = q32[i] - row[i];
                if constexpr (METRIC == MPTG_METRIC_L1) a = a + __builtin_fabsf(dd);
                else a = __builtin_fmaf(dd, dd, a);
            }
            acc[u] = a;
        }
        // the four rows accumulate independently; one branch for all
        uint32_t pass = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) pass |= (acc[u] <= thr ? 1u : 0u) << u;
        while (pass) {
            const int u = __builtin_ctz(pass);
            pass &= pass - 1;
            take(p + u);
        }
    }
    for (; p < p1; ++p) {
        const float* row = x32 + p * DP;
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            float dd = q32[i] - row[i];
            if constexpr (METRIC == MPTG_METRIC_L1) a = a + __builtin_fabsf(dd);
            else a = __builtin_fmaf(dd, dd, a);
        }
        if (a <= thr) take(p);
    }
    if (active) cnt[(int64_t)blockIdx.y * nq + qi] = c;
}

// k best (exact distance, index) of each query's candidates over the chunks,
// among those within the radius; a query with an overflowed region goes to
// `over` for the k-list scan instead.  Candidates beyond the bound cannot
// displace the k best: at least k tree nodes lie within it.
template <int D, int METRIC, int K>
__global__ __launch_bounds__(kBlock) void knn_select(const double* __restrict__ x64, const double* __restrict__ q,
                                                     const int32_t* __restrict__ cand_i,
                                                     const uint32_t* __restrict__ cnt, int cap, int splits,
                                                     int64_t nq, int k, double radius, double* __restrict__ out_d,
                                                     int32_t* __restrict__ out_i, int64_t* __restrict__ over,
                                                     uint32_t* __restrict__ n_over) {
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (qi >= nq) return;
    for (int sp = 0; sp < splits; ++sp)
        if (cnt[(int64_t)sp * nq + qi] > (uint32_t)cap) {
            over[atomicAdd(n_over, 1u)] = qi;
            return;
        }
    double qq[D];
#pragma unroll
    for (int i = 0; i < D; ++i) qq[i] = q[qi * D + i];
    TopK<K> top;
    top.init();
    for (int sp = 0; sp < splits; ++sp) {
        const int64_t region = ((int64_t)sp * nq + qi) * cap;
        const uint32_t c = cnt[(int64_t)sp * nq + qi];
        for (uint32_t j = 0; j < c; ++j) {
            const int32_t id = cand_i[region + j];
            const double dq = exact_dist<D, METRIC>(qq, x64 + (int64_t)id * D);
            if (dq <= radius) top.insert(dq, id);
        }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {  // static indices: the list stays in registers
        if (j < k) {
            const bool have = top.i[j] != INT_MAX;
            out_d[qi * k + j] = have ? top.d[j] : __builtin_inf();
            out_i[qi * k + j] = have ? top.i[j] : -1;
        }
    }
}

// Gathers the listed queries (D coordinates each) into a compact array, and
// scatters their k results back.
__global__ __launch_bounds__(kBlock) void knn_gather(const double* __restrict__ q, int D,
                                                     const int64_t* __restrict__ list, int64_t m,
                                                     double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= m * D) return;
    out[t] = q[list[t / D] * D + t % D];
}
__global__ __launch_bounds__(kBlock) void knn_scatter(const double* __restrict__ d, const int32_t* __restrict__ i,
                                                      const int64_t* __restrict__ list, int64_t m, int k,
                                                      double* __restrict__ out_d, int32_t* __restrict__ out_i) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= m * k) return;
    const int64_t row = list[t / k];
    out_d[row * k + t % k] = d[t];
    out_i[row * k + t % k] = i[t];
}

// ---- SE(3): Cartesian[Scaled(SO3, w_so3), Scaled(L2, w_l2)] ----------------
}  // namespace
}  // namespace mptg

#include "pmath.hpp"

namespace mptg {
namespace {

// SE(3) distance w_so3 * so3 + w_l2 * |dt| in two terms (the oracle's
// expression; the sum of the two rounded products is the same either way
// round).  The translation term is a lower bound of the distance: the
// rotation term is >= 0 and rounding is monotone.
__device__ __forceinline__ double se3_trans(const double* a, const double* b, double w_l2) {
    double s = 0.0;
    for (int i = 4; i < 7; ++i) {
        double t = a[i] - b[i];
        s = s + t * t;
    }
    return w_l2 * __builtin_sqrt(s);
}

__device__ __forceinline__ double se3_rot(const double* a, const double* b, double w_so3) {
    double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    double ad = __builtin_fabs(dot);
    double so3 = ad >= 1.0 ? 0.0 : pm::acos(ad);
    return w_so3 * so3;
}

// One workgroup of W waves per query.  Wave w measures tree nodes
// w * 64 + lane, stepping by 64 W, into its k-list; a node's rotation term
// (the acos) is evaluated only when its translation term does not exceed the
// current k-th distance (nor the radius), so once the list has settled most
// waves skip it.  The W lists meet in LDS, where an entry's rank among all
// of them -- (distance, id) keys are distinct -- is its output position.
template <int W>
__global__ __launch_bounds__(64 * W) void knn_wave_se3(const double* __restrict__ x64, int64_t n,
                                                       const double* __restrict__ q, int k, double radius,
                                                       double w_so3, double w_l2, double inf,
                                                       double* __restrict__ out_d, int32_t* __restrict__ out_i,
                                                       const int64_t* n_dev, const int32_t* skip) {
    __shared__ double sd[W * kWaveMaxK];
    __shared__ int32_t si[W * kWaveMaxK];
    __shared__ int real_s[W];
    if (skip && *skip) return;
    if (n_dev) n = *n_dev;
    const int64_t qi = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double qq[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) qq[i] = q[qi * 7 + i];
    WaveList wl;
    wl.init(k, inf);
    for (int64_t base = (int64_t)w * 64; base < n; base += 64 * W) {
        const int64_t p = base + lane;
        bool pass = false;
        double dq = 0.0;
        if (p < n) {
            const double* x = x64 + p * 7;
            const double tr = se3_trans(qq, x, w_l2);
            if (tr <= __builtin_fmin(wl.kd, radius)) {
                dq = se3_rot(qq, x, w_so3) + tr;
                pass = dq <= radius && wl_before(dq, (int32_t)p, wl.kd, wl.ki);
            }
        }
        wl.offer(pass, dq, (int32_t)p, lane);
    }
    const bool mine = lane < k;
    const unsigned long long real = __ballot(mine && wl.li != INT_MAX);
    if (mine) {
        sd[w * k + lane] = wl.ld;
        si[w * k + lane] = wl.li;
    }
    if (lane == 0) real_s[w] = __popcll(real);
    __syncthreads();
    int total = 0;
#pragma unroll
    for (int v = 0; v < W; ++v) total += real_s[v];
    for (int c = threadIdx.x; c < W * k; c += 64 * W) {
        const int32_t ci = si[c];
        if (ci == INT_MAX) continue;
        const double cd = sd[c];
        int rank = 0;
        for (int e = 0; e < W * k; ++e) rank += (si[e] != INT_MAX && wl_before(sd[e], si[e], cd, ci)) ? 1 : 0;
        if (rank < k) {
            out_d[qi * k + rank] = cd;
            out_i[qi * k + rank] = ci;
        }
    }
    for (int r = threadIdx.x; r < k; r += 64 * W)
        if (r >= total) {
            out_d[qi * k + r] = inf;
            out_i[qi * k + r] = -1;
        }
}

struct ScanExtent {
    int64_t n, chunk;
    const int64_t* n_dev;
    const int32_t* skip;
    double* seed;            // k = 1: per-query starting bound (nullptr: none)
    const uint32_t* nq_dev;  // live queries on the device (nullptr: all nq)
    bool q8;                 // k = 1 under L1: the byte prefilter, its shared bounds in `seed`
    int64_t q8_target;       // byte prefilter: blocks its device-side split aims at
};

template <int D, int METRIC, int K>
void launch_scan(mptg_ctx* ctx, dim3 grid, const ScanExtent& x, const double* q, int64_t nq, double radius,
                 double* pd, int32_t* pi) {
    if constexpr (K == 1 && METRIC == MPTG_METRIC_L1 && D >= 8) {
        if (x.q8) {
            // the seed buffer (nq doubles) holds the queries' shared byte
            // bounds instead
            uint32_t* gthr = reinterpret_cast<uint32_t*>(x.seed);
            hipLaunchKernelGGL((knn_q8_seed<D>), dim3((unsigned)((nq + kQ8SeedQueries - 1) / kQ8SeedQueries)),
                               dim3(kBlock), 0, ctx->stream, ctx->x64, x.n, q, nq, ctx->coord_bound, gthr, x.n_dev,
                               x.skip, x.nq_dev);
            hipLaunchKernelGGL((knn_scan_q8<D>), grid, dim3(kBlock), 0, ctx->stream, ctx->x32, ctx->x64, x.n,
                               x.chunk, q, nq, radius, ctx->coord_bound, pd, pi, x.n_dev, x.skip, gthr, x.nq_dev,
                               x.q8_target);
            return;
        }
    }
    if (K == 1 && x.seed)
        hipLaunchKernelGGL((knn_seed<D, METRIC>), dim3((unsigned)((nq * kSeedLanes + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, ctx->stream, ctx->x64, x.n, q, nq, x.seed, x.n_dev, x.skip, x.nq_dev);
    hipLaunchKernelGGL((knn_scan<D, METRIC, K>), grid, dim3(kBlock), 0, ctx->stream, ctx->x32, ctx->x64, x.n, x.chunk,
                       q, nq, radius, ctx->coord_bound, pd, pi, x.n_dev, x.skip, (const double*)(K == 1 ? x.seed : nullptr),
                       x.nq_dev);
}

template <int D, int K>
bool dispatch_metric(mptg_ctx* ctx, dim3 grid, const ScanExtent& x, const double* q, int64_t nq, double radius,
                     double* pd, int32_t* pi) {
    if (ctx->metric == MPTG_METRIC_L1) launch_scan<D, MPTG_METRIC_L1, K>(ctx, grid, x, q, nq, radius, pd, pi);
    else launch_scan<D, MPTG_METRIC_L2, K>(ctx, grid, x, q, nq, radius, pd, pi);
    return true;
}

template <int K>
bool dispatch_dim(mptg_ctx* ctx, dim3 grid, const ScanExtent& x, const double* q, int64_t nq, double radius,
                  double* pd, int32_t* pi) {
    switch (ctx->dim) {
        case 1: return dispatch_metric<1, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 2: return dispatch_metric<2, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 3: return dispatch_metric<3, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 4: return dispatch_metric<4, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 5: return dispatch_metric<5, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 6: return dispatch_metric<6, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 7: return dispatch_metric<7, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 8: return dispatch_metric<8, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 10: return dispatch_metric<10, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 12: return dispatch_metric<12, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 16: return dispatch_metric<16, K>(ctx, grid, x, q, nq, radius, pd, pi);
        case 20: return dispatch_metric<20, K>(ctx, grid, x, q, nq, radius, pd, pi);
        default: return false;
    }
}

}  // namespace

// Chooses the number of tree chunks so that a launch has ~8 waves per SIMD
// worth of work, with at least kMinSplitNodes nodes per chunk.
static int64_t choose_splits(mptg_ctx* ctx, int64_t nq, int64_t n, int K) {
    int64_t qblocks = (nq + kBlock - 1) / kBlock;
    int64_t want_blocks = (int64_t)ctx->cus * ctx->tune.knn_blocks_per_cu;
    int64_t s = (want_blocks + qblocks - 1) / qblocks;
    int64_t max_s = (n + kMinSplitNodes - 1) / kMinSplitNodes;
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    // the merge reads splits * K candidates per query, one latency each
    const int64_t cap = K == 1 ? 4096 : 64;
    if (s > cap) s = cap;
    return s;
}

static int64_t plan_splits(mptg_ctx* ctx, int64_t nq, int64_t n, int64_t* chunk, int K = 1) {
    int64_t splits = n > 0 ? choose_splits(ctx, nq, n, K) : 1;
    *chunk = n > 0 ? (n + splits - 1) / splits : 0;
    return n > 0 ? (n + *chunk - 1) / *chunk : 1;
}

int knn_reserve(mptg_ctx* ctx, int64_t nq, int k, int64_t tree_n) {
    const int K = k_bucket(k);
    int64_t chunk;
    // k = 1 launches with a device query count plan for half the slots,
    // which never takes fewer splits
    const int64_t splits = plan_splits(ctx, K == 1 ? (nq + 1) / 2 : nq, tree_n, &chunk, K);
    int rc;
    if ((rc = ensure(ctx, ctx->scratch[4], sizeof(double) * K * nq * splits))) return rc;
    if (K == 1 && (rc = ensure(ctx, ctx->scratch[3], sizeof(double) * nq))) return rc;
    return ensure(ctx, ctx->scratch[5], sizeof(int32_t) * K * nq * splits);
}

// The k-list scan: each lane keeps its query's k best in registers.
static int knn_scan_path(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx,
                         double* d_dist, int64_t tree_n, const int64_t* n_dev, const int32_t* skip,
                         const uint32_t* nq_dev = nullptr) {
    const int K = k_bucket(k);
    // with n_dev, tree_n is only the size the split count is planned for;
    // with nq_dev (live queries on the device, k = 1) the split is planned
    // for half of the nq slots
    int64_t n = tree_n, chunk;
    const int64_t splits = plan_splits(ctx, nq_dev ? (nq + 1) / 2 : nq, n, &chunk, K);
    size_t pbytes_d = sizeof(double) * K * nq * splits, pbytes_i = sizeof(int32_t) * K * nq * splits;
    int rc;
    if ((rc = ensure(ctx, ctx->scratch[4], pbytes_d))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[5], pbytes_i))) return rc;
    // a seed bound pays off when the tree is split many ways: every chunk
    // starts from it
    double* seed = nullptr;
    if (K == 1 && splits >= 8 && ctx->metric != MPTG_METRIC_SE3) {
        if ((rc = ensure(ctx, ctx->scratch[3], sizeof(double) * nq))) return rc;
        seed = (double*)ctx->scratch[3].p;
    }
    // k = 1 under L1 in 8 or more dimensions: the byte prefilter
    const bool q8 = seed && ctx->metric == MPTG_METRIC_L1 && ctx->dim >= 8 && ctx->tune.knn_q8 &&
                    ctx->coord_bound > 0 && ctx->coord_bound < 1e300;
    // the byte prefilter's blocks (4 waves, 4 resident per CU) re-split on
    // the device to one full round for the live queries
    const int64_t q8_target = q8 ? (int64_t)ctx->cus * kQ8BlocksPerCu : 0;
    const ScanExtent ext{n, chunk, n_dev, skip, seed, nq_dev, q8, q8_target};
    double* pd = (double*)ctx->scratch[4].p;
    int32_t* pi = (int32_t*)ctx->scratch[5].p;
    dim3 grid((unsigned)((nq + kBlock - 1) / kBlock), (unsigned)splits);
    bool ok;
    if (K == 1) ok = dispatch_dim<1>(ctx, grid, ext, d_q, nq, radius, pd, pi);
    else if (K == 20) ok = dispatch_dim<20>(ctx, grid, ext, d_q, nq, radius, pd, pi);
    else if (K == 32) ok = dispatch_dim<32>(ctx, grid, ext, d_q, nq, radius, pd, pi);
    else ok = dispatch_dim<64>(ctx, grid, ext, d_q, nq, radius, pd, pi);
    if (!ok) return fail(ctx, MPTG_E_ARG, "unsupported dimension for the nearest-neighbour kernel");
    MPTG_HIP(ctx, hipGetLastError());
    dim3 g1((unsigned)((nq + kBlock - 1) / kBlock));
    if (splits > 1) {
        if (K == 1) hipLaunchKernelGGL(knn_merge1, dim3((unsigned)((nq * kMergeLanes + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, pd, pi, (int)splits, nq, d_dist, d_idx, skip, n_dev, std::numeric_limits<double>::infinity(), nq_dev, n, q8_target);
        else if (K == 20) hipLaunchKernelGGL((knn_merge<20>), g1, dim3(kBlock), 0, ctx->stream, pd, pi, (int)splits, nq, k, d_dist, d_idx, skip, n_dev);
        else if (K == 32) hipLaunchKernelGGL((knn_merge<32>), g1, dim3(kBlock), 0, ctx->stream, pd, pi, (int)splits, nq, k, d_dist, d_idx, skip, n_dev);
        else hipLaunchKernelGGL((knn_merge<64>), g1, dim3(kBlock), 0, ctx->stream, pd, pi, (int)splits, nq, k, d_dist, d_idx, skip, n_dev);
    } else {
        if (K == 1) hipLaunchKernelGGL((knn_finish<1>), g1, dim3(kBlock), 0, ctx->stream, pd, pi, nq, k, d_dist, d_idx, skip, nq_dev);
        else if (K == 20) hipLaunchKernelGGL((knn_finish<20>), g1, dim3(kBlock), 0, ctx->stream, pd, pi, nq, k, d_dist, d_idx, skip, nq_dev);
        else if (K == 32) hipLaunchKernelGGL((knn_finish<32>), g1, dim3(kBlock), 0, ctx->stream, pd, pi, nq, k, d_dist, d_idx, skip, nq_dev);
        else hipLaunchKernelGGL((knn_finish<64>), g1, dim3(kBlock), 0, ctx->stream, pd, pi, nq, k, d_dist, d_idx, skip, nq_dev);
    }
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// SE(3): one workgroup of W waves per query, W a power of two up to 16 for
// about 4096 waves per launch, and no more waves than 64-node groups.
static int knn_wave_se3_path(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx,
                             double* d_dist, int64_t tree_n, const int64_t* n_dev, const int32_t* skip) {
    if (k < 1 || k > kWaveMaxK) return fail(ctx, MPTG_E_RANGE, "SE(3) nearest neighbours: k must be 1..64");
    if (nq >= ((int64_t)1 << 31)) return fail(ctx, MPTG_E_RANGE, "at most 2^31 - 1 queries");
    if (nq == 0) return MPTG_OK;
    const int64_t want = (4096 + nq - 1) / nq, groups = std::max<int64_t>(1, (tree_n + 63) / 64);
    int W = 1;
    while (W < 16 && 2 * W <= want && 2 * W <= groups) W *= 2;
    const double inf = std::numeric_limits<double>::infinity();
    const dim3 g((unsigned)nq);
#define MPTG_WAVE_SE3(WW)                                                                                         \
    hipLaunchKernelGGL((knn_wave_se3<WW>), g, dim3(64 * WW), 0, ctx->stream, (const double*)ctx->x64, tree_n, d_q, \
                       k, radius, ctx->w_so3, ctx->w_l2, inf, d_dist, d_idx, n_dev, skip)
    switch (W) {
        case 1: MPTG_WAVE_SE3(1); break;
        case 2: MPTG_WAVE_SE3(2); break;
        case 4: MPTG_WAVE_SE3(4); break;
        case 8: MPTG_WAVE_SE3(8); break;
        default: MPTG_WAVE_SE3(16); break;
    }
#undef MPTG_WAVE_SE3
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

namespace {
__global__ __launch_bounds__(kBlock) void knn_bound(const double* __restrict__ sub_d, int64_t nq, int k,
                                                    double radius, double* __restrict__ tau) {
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (qi < nq) tau[qi] = __builtin_fmin(sub_d[qi * k + k - 1], radius);
}

template <int D, int METRIC>
void launch_collect_select(mptg_ctx* ctx, dim3 grid, int64_t n, int64_t chunk, const double* q, int64_t nq,
                           const double* tau, int cap, int splits, int k, double radius, int32_t* ci, uint32_t* cnt,
                           double* out_d, int32_t* out_i, int64_t* over, uint32_t* n_over) {
    hipLaunchKernelGGL((knn_collect<D, METRIC>), grid, dim3(kBlock), 0, ctx->stream, ctx->x32, n, chunk, q, nq, tau,
                       ctx->coord_bound, cap, ci, cnt);
    if (k <= 20)
        hipLaunchKernelGGL((knn_select<D, METRIC, 20>), dim3(grid.x), dim3(kBlock), 0, ctx->stream, ctx->x64, q,
                           (const int32_t*)ci, (const uint32_t*)cnt, cap, splits, nq, k, radius, out_d, out_i, over,
                           n_over);
    else
        hipLaunchKernelGGL((knn_select<D, METRIC, 32>), dim3(grid.x), dim3(kBlock), 0, ctx->stream, ctx->x64, q,
                           (const int32_t*)ci, (const uint32_t*)cnt, cap, splits, nq, k, radius, out_d, out_i, over,
                           n_over);
}

template <int D>
void dispatch_collect(mptg_ctx* ctx, dim3 grid, int64_t n, int64_t chunk, const double* q, int64_t nq,
                      const double* tau, int cap, int splits, int k, double radius, int32_t* ci, uint32_t* cnt,
                      double* out_d, int32_t* out_i, int64_t* over, uint32_t* n_over) {
    if (ctx->metric == MPTG_METRIC_L1)
        launch_collect_select<D, MPTG_METRIC_L1>(ctx, grid, n, chunk, q, nq, tau, cap, splits, k, radius, ci, cnt,
                                                 out_d, out_i, over, n_over);
    else
        launch_collect_select<D, MPTG_METRIC_L2>(ctx, grid, n, chunk, q, nq, tau, cap, splits, k, radius, ci, cnt,
                                                 out_d, out_i, over, n_over);
}
}  // namespace

// One refinement level over the tree prefix [0, n): collect within the
// bounds tau, select the k best; queries whose candidates overflowed take
// the k-list scan over the prefix (gathered, scanned, scattered back).
static int knn_refine(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, const double* tau,
                      int64_t n, double* out_d, int32_t* out_i) {
    const int K = k <= 20 ? 20 : 32;
    const int D = ctx->dim;
    int64_t chunk;
    const int64_t splits = plan_splits(ctx, nq, n, &chunk);
    // per (chunk, query) region: about twice a chunk's share of the 16 k
    // nodes a bound from a 1/16 prefix admits on spread-out trees; 32 at least
    const int cap = (int)std::max<int64_t>(32, (32 * (int64_t)K + splits - 1) / splits);
    int rc;
    DevBuf* b = ctx->knn_sel;
    if ((rc = ensure(ctx, b[2], sizeof(int32_t) * (size_t)nq * splits * cap))) return rc;
    if ((rc = ensure(ctx, b[3], sizeof(int64_t) * (size_t)nq + sizeof(uint32_t) * (size_t)nq * splits + 256)))
        return rc;
    int32_t* ci = (int32_t*)b[2].p;
    uint32_t* n_over = (uint32_t*)b[3].p;
    int64_t* over = (int64_t*)((char*)b[3].p + 256);
    uint32_t* cnt = (uint32_t*)(over + nq);
    if ((rc = launch_zero_u32(ctx, n_over, 1))) return rc;
    const unsigned g1 = (unsigned)((nq + kBlock - 1) / kBlock);
    const dim3 grid(g1, (unsigned)splits);
    switch (D) {
        case 2: dispatch_collect<2>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 3: dispatch_collect<3>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 4: dispatch_collect<4>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 5: dispatch_collect<5>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 6: dispatch_collect<6>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 7: dispatch_collect<7>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 8: dispatch_collect<8>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 10: dispatch_collect<10>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        case 20: dispatch_collect<20>(ctx, grid, n, chunk, d_q, nq, tau, cap, (int)splits, k, radius, ci, cnt, out_d, out_i, over, n_over); break;
        default: return fail(ctx, MPTG_E_STATE, "no collect kernel for this dimension");
    }
    MPTG_HIP(ctx, hipGetLastError());
    uint32_t m = 0;
    MPTG_HIP(ctx, hipMemcpyAsync(&m, n_over, sizeof(m), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (m == 0) return MPTG_OK;
    if ((rc = ensure(ctx, b[4], sizeof(double) * (size_t)m * D))) return rc;
    if ((rc = ensure(ctx, b[5], (sizeof(double) + sizeof(int32_t)) * (size_t)m * k))) return rc;
    double* gq = (double*)b[4].p;
    double* gd = (double*)b[5].p;
    int32_t* gi = (int32_t*)(gd + (size_t)m * k);
    hipLaunchKernelGGL(knn_gather, dim3((unsigned)(((int64_t)m * D + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, d_q, D, (const int64_t*)over, (int64_t)m, gq);
    if ((rc = knn_scan_path(ctx, gq, m, k, radius, gi, gd, n, nullptr, nullptr))) return rc;
    hipLaunchKernelGGL(knn_scatter, dim3((unsigned)(((int64_t)m * k + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, (const double*)gd, (const int32_t*)gi, (const int64_t*)over, (int64_t)m, k, out_d,
                       out_i);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

// k > 1 over a large tree, refined over growing tree prefixes: the exact k
// best over the first 1/256 of the tree (k-list scan) bound every query;
// collect within the bound over the first 1/16 and select its exact k best,
// which bound the whole tree's; collect and select over the whole tree.  A
// prefix's k-th distance is no smaller than the whole tree's, the collect
// keeps every node within the bound and the select orders candidates by
// exact (distance, index), so the result equals the k-list scan's.
static int knn_select_path(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx,
                           double* d_dist, int64_t n) {
    int rc;
    DevBuf* b = ctx->knn_sel;
    if ((rc = ensure(ctx, b[0], (sizeof(double) + sizeof(int32_t)) * (size_t)nq * k))) return rc;
    if ((rc = ensure(ctx, b[1], sizeof(double) * (size_t)nq))) return rc;
    double* sub_d = (double*)b[0].p;
    int32_t* sub_i = (int32_t*)(sub_d + (size_t)nq * k);
    double* tau = (double*)b[1].p;
    const unsigned g1 = (unsigned)((nq + kBlock - 1) / kBlock);
    if ((rc = knn_scan_path(ctx, d_q, nq, k, radius, sub_i, sub_d, n / 256, nullptr, nullptr))) return rc;
    hipLaunchKernelGGL(knn_bound, dim3(g1), dim3(kBlock), 0, ctx->stream, (const double*)sub_d, nq, k, radius, tau);
    if ((rc = knn_refine(ctx, d_q, nq, k, radius, tau, n / 16, sub_d, sub_i))) return rc;
    hipLaunchKernelGGL(knn_bound, dim3(g1), dim3(kBlock), 0, ctx->stream, (const double*)sub_d, nq, k, radius, tau);
    return knn_refine(ctx, d_q, nq, k, radius, tau, n, d_dist, d_idx);
}

constexpr int64_t kSelectMinTree = 1 << 16;  // smaller trees: the k-list scan

int launch_knn(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx,
               double* d_dist, int64_t tree_n, const int64_t* n_dev, const int32_t* skip, const uint32_t* nq_dev) {
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called first");
    if (nq > ((int64_t)1 << 31)) return fail(ctx, MPTG_E_RANGE, "at most 2^31 queries");
    if (!(radius >= 0)) return fail(ctx, MPTG_E_ARG, "radius must be >= 0");
    if (nq_dev) {  // live queries known only on the device: the k = 1 scans
        if (k != 1 || ctx->metric == MPTG_METRIC_SE3)
            return fail(ctx, MPTG_E_STATE, "a device query count needs k = 1 and an L1/L2 space");
        return knn_scan_path(ctx, d_q, nq, k, radius, d_idx, d_dist, tree_n, n_dev, skip, nq_dev);
    }
    if (ctx->metric == MPTG_METRIC_SE3) return knn_wave_se3_path(ctx, d_q, nq, k, radius, d_idx, d_dist, tree_n, n_dev, skip);
    const bool sel_dim = ctx->dim <= 8 || ctx->dim == 10 || ctx->dim == 20;
    const bool eager = nq > 0 && !n_dev && !skip && !ctx->capturing;
    // (the select path keeps lists of up to 32; larger k takes the k-list scan)
    const bool select = eager && k > 1 && k <= 32 && tree_n >= kSelectMinTree && ctx->tune.knn_select &&
                        ctx->metric != MPTG_METRIC_SE3 && ctx->dim >= 2 && sel_dim;
    if (eager && knn_grid_applies(ctx, tree_n)) {
        // the grid answers most queries; the rest take the scans, gathered
        // into a compact batch and scattered back
        int64_t* over = nullptr;
        uint32_t m = 0;
        int rc = knn_grid_path(ctx, d_q, nq, k, radius, d_idx, d_dist, tree_n, &over, &m);
        if (rc || m == 0) return rc;
        const int D = ctx->dim;
        DevBuf& g = ctx->knn_grid[4];
        if ((rc = ensure(ctx, g, (sizeof(double) * D + sizeof(double) * k + sizeof(int32_t) * k) * (size_t)m)))
            return rc;
        double* gq = (double*)g.p;
        double* gd = gq + (size_t)m * D;
        int32_t* gi = (int32_t*)(gd + (size_t)m * k);
        hipLaunchKernelGGL(knn_gather, dim3((unsigned)(((int64_t)m * D + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, d_q, D, (const int64_t*)over, (int64_t)m, gq);
        // the select path streams the tree three times whatever the batch:
        // few queries take the k-list scan
        rc = select && m >= 2048 ? knn_select_path(ctx, gq, m, k, radius, gi, gd, tree_n)
                                 : knn_scan_path(ctx, gq, m, k, radius, gi, gd, tree_n, nullptr, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(knn_scatter, dim3((unsigned)(((int64_t)m * k + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, (const double*)gd, (const int32_t*)gi, (const int64_t*)over, (int64_t)m, k,
                           d_dist, d_idx);
        MPTG_HIP(ctx, hipGetLastError());
        return MPTG_OK;
    }
    if (select) return knn_select_path(ctx, d_q, nq, k, radius, d_idx, d_dist, tree_n);
    return knn_scan_path(ctx, d_q, nq, k, radius, d_idx, d_dist, tree_n, n_dev, skip);
}

}  // namespace mptg
