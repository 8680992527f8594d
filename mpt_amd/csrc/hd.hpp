// Host/device shared primitives of the project's bit-exact contract.
//
// Everything in this header is compiled twice: by hipcc for gfx950 kernels and
// by the host compiler for the CPU side, always with FP contraction disabled.
// Each routine is a fixed sequence of IEEE-754 binary64/binary32 operations
// (+, -, *, /, sqrt, explicit fma only where written), so host and device
// produce identical bits for identical inputs.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MPTG_HD __host__ __device__ inline
#else
#define MPTG_HD inline
#endif

namespace mptg {

MPTG_HD uint64_t bits_of(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}

MPTG_HD double double_of(uint64_t u) {
    double x;
    __builtin_memcpy(&x, &u, 8);
    return x;
}

MPTG_HD uint32_t bits_of(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    return u;
}

MPTG_HD float float_of(uint32_t u) {
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

// BF16 <-> FP32, round to nearest even (NaN stays NaN).
MPTG_HD uint16_t float_to_bf16(float f) {
    uint32_t u = bits_of(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

MPTG_HD float bf16_to_float(uint16_t h) { return float_of((uint32_t)h << 16); }

// ---------------------------------------------------------------------------
// Philox-4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw: "Parallel
// random numbers: as easy as 1, 2, 3", SC'11).  Multipliers and Weyl key
// increments are the published ones.  One call maps a 128-bit counter and a
// 64-bit key to 128 random bits; samples are addressed by counter, so any
// thread can draw any sample without shared state.
// ---------------------------------------------------------------------------
struct Philox4 {
    uint32_t v[4];
};

MPTG_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                              uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    Philox4 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// 53-bit uniform in [0, 1): top 53 bits of a 64-bit word times 2^-53.
MPTG_HD double u64_to_unit(uint32_t hi, uint32_t lo) {
    uint64_t u = ((uint64_t)hi << 32) | lo;
    return (double)(u >> 11) * 0x1.0p-53;
}

// The sample stream used by the planner and the synthetic scene builders.
// Counter = (index lo, index hi, draw, purpose), key = seed.  Each call yields
// two doubles in [0,1).
struct Uniform2 {
    double a, b;
};

MPTG_HD Uniform2 uniform2(uint64_t seed, uint64_t index, uint32_t draw, uint32_t purpose) {
    Philox4 r = philox4x32_10((uint32_t)index, (uint32_t)(index >> 32), draw, purpose,
                              (uint32_t)seed, (uint32_t)(seed >> 32));
    Uniform2 u;
    u.a = u64_to_unit(r.v[0], r.v[1]);
    u.b = u64_to_unit(r.v[2], r.v[3]);
    return u;
}

// j-th uniform double of a (seed, index, purpose) stream.
MPTG_HD double uniform_at(uint64_t seed, uint64_t index, uint32_t purpose, uint32_t j) {
    Uniform2 u = uniform2(seed, index, j >> 1, purpose);
    return (j & 1) ? u.b : u.a;
}

// std::uniform_real_distribution<double>(lo, hi) mapping: u * (hi - lo) + lo.
MPTG_HD double uniform_in(double u, double lo, double hi) {
    return u * (hi - lo) + lo;
}

// Purpose tags for the planner's streams (fourth Philox counter word).
enum : uint32_t {
    RNG_PURPOSE_BIAS = 0x42494153u,    // goal-bias coin
    RNG_PURPOSE_SAMPLE = 0x53414d50u,  // state coordinates
};

}  // namespace mptg
