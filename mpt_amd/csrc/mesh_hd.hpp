// Rigid-body mesh collision primitives shared by host and device (built with
// FP contraction off on both, so every result is the same bits on both).
//
// Used for the SE(3) scene whose valid() the reference asks FCL for
// (demo/se3_rigid_body_scenario.hpp:288-312: collide(robot at pose, env)).
// This project's predicate, in FP32:
//   pose      R, t from the state's unit quaternion (x, y, z, w) and
//             translation, R by the standard quaternion-to-matrix expansion
//             (as Eigen's toRotationMatrix), evaluated in FP64 and rounded
//             once to FP32; vertex v -> ((R0 v.x + R1 v.y) + R2 v.z) + t
//             (Eigen::Translation * Quaternion, :258-261)
//   collide   exists a robot triangle a and an environment triangle b whose
//             axis-aligned boxes overlap (closed) and which intersect (closed)
//   intersect separating-axis test over the 17 candidate axes of two
//             triangles: both normals, the 9 edge-edge cross products and the
//             6 in-plane edge normals (exact for coplanar pairs too, in exact
//             arithmetic).  A triangle pair is disjoint iff some axis has the
//             two projection intervals strictly apart.
// The box test is part of the predicate, so any conservative culling of
// pairs by boxes (the device BVH, the oracle's grid) gives the same verdict.
#pragma once
#include <stdint.h>

#include "hd.hpp"

namespace mptg {

// One environment triangle as the device reads it: 16 floats (64 bytes):
// v0 v1 v2 (9), box lo (3), box hi (3), pad.
constexpr int kTriFloats = 16;
// BVH node (32 bytes): box, then either children (a, b >= 1) or a leaf
// (a = first triangle, b = -count).
struct MeshNode {
    float lo[3], hi[3];
    int32_t a, b;
};
constexpr int kLeafTris = 4;
constexpr int kMaxBvhDepth = 40;

struct Pose32 {
    float r[9];  // row-major
    float t[3];
};

// Pose of SE(3) state [qx, qy, qz, qw, tx, ty, tz].
MPTG_HD Pose32 pose_from_state(const double* s) {
    const double x = s[0], y = s[1], z = s[2], w = s[3];
    const double tx = 2.0 * x, ty = 2.0 * y, tz = 2.0 * z;
    const double twx = tx * w, twy = ty * w, twz = tz * w;
    const double txx = tx * x, txy = ty * x, txz = tz * x;
    const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
    Pose32 p;
    p.r[0] = (float)(1.0 - (tyy + tzz));
    p.r[1] = (float)(txy - twz);
    p.r[2] = (float)(txz + twy);
    p.r[3] = (float)(txy + twz);
    p.r[4] = (float)(1.0 - (txx + tzz));
    p.r[5] = (float)(tyz - twx);
    p.r[6] = (float)(txz - twy);
    p.r[7] = (float)(tyz + twx);
    p.r[8] = (float)(1.0 - (txx + tyy));
    p.t[0] = (float)s[4];
    p.t[1] = (float)s[5];
    p.t[2] = (float)s[6];
    return p;
}

MPTG_HD void xform(const Pose32& p, const float* v, float* o) {
    o[0] = ((p.r[0] * v[0] + p.r[1] * v[1]) + p.r[2] * v[2]) + p.t[0];
    o[1] = ((p.r[3] * v[0] + p.r[4] * v[1]) + p.r[5] * v[2]) + p.t[1];
    o[2] = ((p.r[6] * v[0] + p.r[7] * v[1]) + p.r[8] * v[2]) + p.t[2];
}

MPTG_HD float fmin3(float a, float b, float c) {
    float m = a < b ? a : b;
    return m < c ? m : c;
}
MPTG_HD float fmax3(float a, float b, float c) {
    float m = a > b ? a : b;
    return m > c ? m : c;
}

// box of a triangle t[9] -> lo[3], hi[3]
MPTG_HD void tri_box(const float* t, float* lo, float* hi) {
    for (int i = 0; i < 3; ++i) {
        lo[i] = fmin3(t[i], t[3 + i], t[6 + i]);
        hi[i] = fmax3(t[i], t[3 + i], t[6 + i]);
    }
}

MPTG_HD bool boxes_overlap(const float* alo, const float* ahi, const float* blo, const float* bhi) {
    return alo[0] <= bhi[0] && blo[0] <= ahi[0] && alo[1] <= bhi[1] && blo[1] <= ahi[1] && alo[2] <= bhi[2] &&
           blo[2] <= ahi[2];
}

MPTG_HD void sub3(const float* a, const float* b, float* o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

MPTG_HD void cross3(const float* a, const float* b, float* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

MPTG_HD float dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// Are triangles p and q (9 floats each) strictly apart along axis ax?
MPTG_HD bool apart_on(const float* ax, const float* p, const float* q) {
    const float p0 = dot3(ax, p), p1 = dot3(ax, p + 3), p2 = dot3(ax, p + 6);
    const float q0 = dot3(ax, q), q1 = dot3(ax, q + 3), q2 = dot3(ax, q + 6);
    return fmax3(p0, p1, p2) < fmin3(q0, q1, q2) || fmax3(q0, q1, q2) < fmin3(p0, p1, p2);
}

// Closed triangle-triangle intersection by separating axes.
MPTG_HD bool tri_tri_overlap(const float* p, const float* q) {
    float ep[3][3], eq[3][3], np[3], nq[3], ax[3];
    sub3(p + 3, p, ep[0]);
    sub3(p + 6, p + 3, ep[1]);
    sub3(p, p + 6, ep[2]);
    sub3(q + 3, q, eq[0]);
    sub3(q + 6, q + 3, eq[1]);
    sub3(q, q + 6, eq[2]);
    cross3(ep[0], ep[1], np);
    if (apart_on(np, p, q)) return false;
    cross3(eq[0], eq[1], nq);
    if (apart_on(nq, p, q)) return false;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cross3(ep[i], eq[j], ax);
            if (apart_on(ax, p, q)) return false;
        }
    for (int i = 0; i < 3; ++i) {
        cross3(np, ep[i], ax);
        if (apart_on(ax, p, q)) return false;
        cross3(nq, eq[i], ax);
        if (apart_on(ax, p, q)) return false;
    }
    return true;
}

// Collision of one robot triangle (already posed) with one environment
// triangle, boxes included: the pair predicate above.
MPTG_HD bool pair_collides(const float* a, const float* alo, const float* ahi, const float* b, const float* blo,
                           const float* bhi) {
    return boxes_overlap(alo, ahi, blo, bhi) && tri_tri_overlap(a, b);
}

// ---- SO(3) x R^3 sampling (src/mpt/impl/uniform_sampler_so3.hpp:55-68) ------
// u0 -> a in [0,1), u1, u2 -> b, c in [0, 2 pi); quaternion
// (w, x, y, z) = (sqrt(1-a) sin b, sqrt(1-a) cos b, sqrt(a) sin c, sqrt(a) cos c)
// stored as [x, y, z, w].
constexpr double kTwoPi = 6.283185307179586;

}  // namespace mptg
