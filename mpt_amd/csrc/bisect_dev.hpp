// Recursive-bisection waypoint sets (the arm's link(),
// demo/link_manipulator_scenario.hpp:118-138) in closed form.
//
// The midpoints of link(a, b) form a binary tree: node = interval, midpoint
// (lo+hi)/2 checked when L-inf(lo-hi) >= tol, children [lo, mid], [mid, hi].
// When the tree is perfect with h internal levels, pre-order ordinal <->
// (path, depth) is a closed form and any waypoint can be rebuilt from (a, b)
// along its path with exactly the reference's arithmetic.
#pragma once
#include "space_dev.hpp"

namespace mptg {

constexpr int kMaxArmDepth = 30;
// Waypoint slots per edge in the surrogate score buffer (trees up to h = 9).
constexpr int kScoreStride = 512;

// Pre-order ordinal -> (path, depth) in a perfect tree with h internal levels.
__host__ __device__ __forceinline__ void preorder_path(int h, uint32_t o, uint32_t* path, int* depth) {
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

// (path, depth) -> pre-order ordinal in a perfect tree with h levels (the
// inverse of preorder_path: each step down adds one, a right step also
// skips the left subtree of 2^(h-1-k) - 1 nodes).
__host__ __device__ __forceinline__ uint32_t path_preorder(int h, uint32_t path, int depth) {
    uint32_t o = 0;
    for (int k = 0; k < depth; ++k) o += 1u + (((path >> k) & 1u) ? (1u << (h - 1 - k)) - 1u : 0u);
    return o;
}

// The surrogate order of an edge's waypoints: the kPicks waypoints with the
// highest scores among the top tree -- the 2^kPickLevels - 1 nodes of depths
// 0 .. kPickLevels-1, spread evenly along the edge (every node when the tree
// is smaller) -- ties by reference ordinal, then every other waypoint in
// reference order.
constexpr int kPickLevels = 4;
constexpr int kPicks = 4;
// Surrogate-ordered edges have trees of at most this many levels (at most
// kScoreStride waypoints); the link kernels walk deeper ones in reference
// order.
constexpr int kMaxPickDepth = 9;

// Node i of the top tree in breadth-first order (root 0, then each level
// left to right) -> (path, depth); path bit k is step k (1 = right).
__host__ __device__ __forceinline__ void bfs_path(uint32_t i, uint32_t* path, int* depth) {
    const int d = 31 - __builtin_clz(i + 1u);
    const uint32_t idx = i + 1u - (1u << d);  // first step in the top bit
    uint32_t p = 0;
    for (int k = 0; k < d; ++k) p |= ((idx >> (d - 1 - k)) & 1u) << k;
    *path = p;
    *depth = d;
}

// The k-th reference ordinal that is not one of the picks p[0..3] (sorted
// ascending): each pick at or below the running ordinal moves it up by one.
__host__ __device__ __forceinline__ uint32_t skip_picks(uint32_t k, const uint32_t (&p)[kPicks]) {
    uint32_t o = k;
#pragma unroll
    for (int i = 0; i < kPicks; ++i) o += p[i] <= o ? 1u : 0u;
    return o;
}

// One coordinate of the midpoint of the node at `path`: the reference's
// recursion (a+b)/2 restricted to that coordinate.  (A form with one select
// per level -- the child's interval is {m_k, the end the step keeps}, and
// that end changes only when the path turns -- gives the same bits, but the
// continuation kernel spilled with it at 5 waves per SIMD and ran 3% slower.)
__host__ __device__ __forceinline__ double coord_mid(double lo, double hi, uint32_t path, int depth) {
    for (int k = 0; k < depth; ++k) {
        double m = (lo + hi) / 2;
        bool right = (path >> k) & 1u;
        lo = right ? m : lo;
        hi = right ? hi : m;
    }
    return (lo + hi) / 2;
}


// Height h and perfect-tree test of link(a, b) (see arm.hip for the bound).
// Returns false when the edge needs more than kMaxArmDepth levels.
template <int D>
__device__ __forceinline__ bool bisect_shape(const double* a, const double* b, double tol, int* h_out, bool* regular) {
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
    if (h > kMaxArmDepth) return false;
    const double err = (double)(h + 1) * 0x1.0p-50 * M;
    // h == 0: the root's width is exactly wmax < tol, a leaf
    *regular = (h == 0) || ((2.0 * wd - err >= tol) && (wd + err < tol));
    *h_out = h;
    return true;
}

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


}  // namespace mptg
