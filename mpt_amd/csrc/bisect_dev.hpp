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

}  // namespace mptg
