// A wave's sorted k-list of (distance, node id) candidates, one entry per
// lane: lane j < k holds the j-th best in ascending (distance, id) order.
// Shared by the kernels that give one query a whole wave (grid queries, the
// SE(3) scan).
#pragma once
#include <climits>

#include <hip/hip_runtime.h>

namespace mptg {

__device__ __forceinline__ bool wl_before(double a, int ia, double b, int ib) {
    return a < b || (a == b && ia < ib);
}

__device__ __forceinline__ double wl_readlane(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// (kd, ki) mirror lane k - 1's entry (the k-th best) as wave-uniform values.
// +inf comes in as a value (a kernel argument), not a constant: this
// toolchain encodes a 64-bit scalar move of the constant 0x7ff0000000000000
// with a 32-bit literal holding its low half, so a uniform +inf held in
// scalar registers would read as 0.0.
struct WaveList {
    double ld;
    int32_t li;
    double kd;
    int32_t ki;
    int k;
    __device__ __forceinline__ void init(int k_, double inf) {
        ld = inf;
        li = INT_MAX;
        kd = inf;
        ki = INT_MAX;
        k = k_;
    }
    // Inserts a wave-uniform candidate that beats (kd, ki): every entry
    // after it moves one lane up.
    __device__ __forceinline__ void insert(double cd, int32_t ci, int lane) {
        const double pd = __shfl_up(ld, 1);
        const int32_t pi = __shfl_up(li, 1);
        const bool shift = lane != 0 && wl_before(cd, ci, pd, pi);
        const bool here = !shift && wl_before(cd, ci, ld, li);
        ld = shift ? pd : (here ? cd : ld);
        li = shift ? pi : (here ? ci : li);
        kd = wl_readlane(ld, k - 1);
        ki = __builtin_amdgcn_readlane(li, k - 1);
    }
    // Offers every lane's (d, id) where `pass` holds, in lane order; only
    // those that still beat the k-th best go in.
    __device__ __forceinline__ void offer(bool pass, double d, int32_t id, int lane) {
        for (unsigned long long m = __ballot(pass); m; m &= m - 1) {
            const int b = __builtin_ctzll(m);
            const double cd = wl_readlane(d, b);
            const int32_t ci = __builtin_amdgcn_readlane(id, b);
            if (wl_before(cd, ci, kd, ki)) insert(cd, ci, lane);
        }
    }
};

}  // namespace mptg
