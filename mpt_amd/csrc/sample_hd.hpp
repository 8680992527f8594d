// State sampling shared by host and device: one Philox-addressed draw per
// sample index, so any thread (or the CPU) draws any sample.
//
// Box spaces (src/mpt/uniform_box_sampler.hpp:60-68): coordinate i from
//   uniform(seed, sample, SAMPLE, i) mapped to [lo_i, hi_i).
// SE(3) = SO(3) x R^3 (src/mpt/impl/uniform_sampler_so3.hpp:55-68 for the
// rotation, the box sampler for the translation):
//   a = u0 in [0,1), b = u1 * 2 pi, c = u2 * 2 pi,
//   (w, x, y, z) = (sqrt(1-a) sin b, sqrt(1-a) cos b, sqrt(a) sin c, sqrt(a) cos c),
//   stored [x, y, z, w]; translation i from u(3+i) mapped to [lo_{4+i}, hi_{4+i}).
// sin/cos are the project's (pmath.hpp); sqrt is correctly rounded on both.
#pragma once
#include "hd.hpp"
#include "pmath.hpp"

namespace mptg {

constexpr int kSampleMetricSE3 = 2;  // same value as MPTG_METRIC_SE3

MPTG_HD void sample_state(uint64_t seed, uint64_t sidx, int metric, int dim, const double* lo, const double* hi,
                          double* q) {
    if (metric == kSampleMetricSE3) {
        const double a = uniform_at(seed, sidx, RNG_PURPOSE_SAMPLE, 0u);
        const double b = uniform_at(seed, sidx, RNG_PURPOSE_SAMPLE, 1u) * 6.283185307179586;
        const double c = uniform_at(seed, sidx, RNG_PURPOSE_SAMPLE, 2u) * 6.283185307179586;
        double sb, cb, sc, cc;
        pm::sincos(b, &sb, &cb);
        pm::sincos(c, &sc, &cc);
        const double ra = __builtin_sqrt(1.0 - a), rb = __builtin_sqrt(a);
        q[0] = ra * cb;
        q[1] = rb * sc;
        q[2] = rb * cc;
        q[3] = ra * sb;
        for (int i = 0; i < 3; ++i)
            q[4 + i] = uniform_in(uniform_at(seed, sidx, RNG_PURPOSE_SAMPLE, 3u + (uint32_t)i), lo[4 + i], hi[4 + i]);
        return;
    }
    for (int i = 0; i < dim; ++i) q[i] = uniform_in(uniform_at(seed, sidx, RNG_PURPOSE_SAMPLE, (uint32_t)i), lo[i], hi[i]);
}

}  // namespace mptg
