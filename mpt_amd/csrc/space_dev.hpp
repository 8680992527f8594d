// Device-side metrics, interpolation and scene validity.
//
// Every expression is the FP64 left-to-right form the project fixes for
// bit-exact results (built with FP contraction off):
//   L2   sqrt(sum_i (a_i - b_i)^2)        L1  sum_i |a_i - b_i|
//   SE3  w_so3 * acos(|<qa,qb>|) + w_l2 * ||ta - tb||      (src/mpt/se3_space.hpp)
//   interpolate L_p: (b_i - a_i) * t + a_i                 (src/mpt/lp_space.hpp:44-54)
//   interpolate SO3: slerp with lerp branch and sign flip  (src/mpt/so3_space.hpp:43-83)
#pragma once
#include "internal.hpp"
#include "pmath.hpp"

namespace mptg {

struct SpaceArgs {
    int metric;
    int dim;
    double w_so3, w_l2;
};

// Disk record of the arm scene: centre, (r+R)^2 exactly as the reference
// computes it, and r+R for the far-disk bound.
struct ArmDisk {
    double cx, cy, rr2, rt;
};

struct SceneArgs {
    int kind;
    const double* boxes;
    int nboxes;
    const uint64_t* grid;  // 8 x 8 bit tiles, see grid_valid
    int width, height, grid_tx;
    const double* lengths;
    int nlinks;
    const double* disks;  // [cx, cy, (r + R)^2, r + R], reference order
    int ndisks;
    const ArmDisk* arm_disks;  // same disks, kernel layout, ordered by grid cell (row-major)
    const double* arm_lim;     // per-link segment-length bound
    const double* arm_index;   // disk grid, see ArmGrid
    const double* circles;     // shapes: [cx, cy, r^2, r] per circle
    int ncircles;
};

// Uniform grid over the disk centres.  Disks are stored by cell, row-major
// (cell y, then cell x), so the cells x0..x1 of one row hold one contiguous
// run of disks: starts[row * nx + x0] .. starts[row * nx + x1 + 1].  Layout
// of arm_index (doubles): [x0, y0, 1/cell width, 1/cell height, max (r+R),
// nx, ny, 0], then int32 starts[nx * ny + 1].
constexpr int kArmGridMax = 32;  // cells per axis at most
constexpr int kDiskBatch = 4;    // disks loaded together in the culled loop

// Cell of a coordinate along one axis; the host build and the device query
// use this same expression, so it is monotone in v on both.
MPTG_HD int arm_cell(double v, double v0, double inv, int n) {
    const double t = (v - v0) * inv;
    return t < 0.0 ? 0 : (t >= (double)n ? n - 1 : (int)t);
}

struct ArmGrid {
    double x0, y0, ix, iy, rtmax;
    int nx, ny;
    const int32_t* starts;
};

__device__ __forceinline__ ArmGrid arm_grid(const double* g) {
    return ArmGrid{g[0], g[1], g[2], g[3], g[4], (int)g[5], (int)g[6], reinterpret_cast<const int32_t*>(g + 8)};
}

inline SpaceArgs space_args(const mptg_ctx* ctx) { return SpaceArgs{ctx->metric, ctx->dim, ctx->w_so3, ctx->w_l2}; }

inline SceneArgs scene_args(const mptg_ctx* ctx) {
    const SceneDev& s = ctx->scene;
    return SceneArgs{s.kind,    s.boxes,  s.nboxes, s.grid,   s.width,
                     s.height,  s.grid_tx, s.lengths, s.nlinks, s.disks, s.ndisks,
                     (const ArmDisk*)s.arm_disks.p, (const double*)s.arm_lim.p, (const double*)s.arm_index.p,
                     s.circles, s.ncircles};
}

__device__ __forceinline__ double so3_dist(const double* a, const double* b) {
    double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    double ad = __builtin_fabs(dot);
    return ad >= 1.0 ? 0.0 : pm::acos(ad);
}

__device__ __forceinline__ double space_dist(const SpaceArgs& sp, const double* a, const double* b) {
    double s = 0.0;
    if (sp.metric == MPTG_METRIC_L1) {
        for (int i = 0; i < sp.dim; ++i) s = s + __builtin_fabs(a[i] - b[i]);
        return s;
    }
    if (sp.metric == MPTG_METRIC_SE3) {
        for (int i = 4; i < 7; ++i) {
            double t = a[i] - b[i];
            s = s + t * t;
        }
        return sp.w_so3 * so3_dist(a, b) + sp.w_l2 * __builtin_sqrt(s);
    }
    for (int i = 0; i < sp.dim; ++i) {
        double t = a[i] - b[i];
        s = s + t * t;
    }
    return __builtin_sqrt(s);
}

__device__ __forceinline__ void space_interp(const SpaceArgs& sp, const double* a, const double* b, double t,
                                             double* q) {
    int i0 = 0;
    if (sp.metric == MPTG_METRIC_SE3) {
        double d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
        double ad = __builtin_fabs(d);
        double s0, s1;
        if (d >= 1.0 - 0x1.0p-52) {
            s0 = 1.0 - t;
            s1 = t;
        } else {
            double theta = pm::acos(ad);
            double sinTheta = pm::sin(theta);
            s0 = pm::sin((1.0 - t) * theta) / sinTheta;
            s1 = pm::sin(t * theta) / sinTheta;
        }
        if (d < 0) s1 = -s1;
        for (int i = 0; i < 4; ++i) q[i] = s0 * a[i] + s1 * b[i];
        i0 = 4;
    }
    for (int i = i0; i < sp.dim; ++i) q[i] = (b[i] - a[i]) * t + a[i];
}

// distPointSegmentSquared, demo/shape_hierarchy.hpp:259-270
__device__ __forceinline__ double seg_dist_sq(double px, double py, double s0x, double s0y, double s1x, double s1y) {
    double vx = s1x - s0x, vy = s1y - s0y;
    double wx = px - s0x, wy = py - s0y;
    double c1 = vx * wx + vy * wy;
    if (c1 <= 0) return wx * wx + wy * wy;
    double c2 = vx * vx + vy * vy;
    if (c2 <= c1) {
        double ex = px - s1x, ey = py - s1y;
        return ex * ex + ey * ey;
    }
    double b = c1 / c2;
    double ex = (s0x - px) + vx * b;
    double ey = (s0y - py) + vy * b;
    return ex * ex + ey * ey;
}

// LinkManipulatorScenario::valid, demo/link_manipulator_scenario.hpp:99-116
__device__ __forceinline__ bool arm_valid(const SceneArgs& sc, const double* q) {
    double fx = 0.0, fy = 0.0, angle = 0.0;
    for (int i = 0; i < sc.nlinks; ++i) {
        angle += q[i];
        double s, c;
        pm::sincos(angle, &s, &c);
        double L = sc.lengths[i];
        double tx = fx + L * c;
        double ty = fy + L * s;
        for (int j = 0; j < sc.ndisks; ++j) {
            const double* dk = sc.disks + 4 * j;
            if (!(seg_dist_sq(dk[0], dk[1], fx, fy, tx, ty) > dk[2])) return false;
        }
        fx = tx;
        fy = ty;
    }
    return true;
}

// arm_valid with work counters: FK links evaluated, segment-disk tests
// evaluated (same early exits), and of those the tests whose disk centre lies
// within (link length + r + R) of the link's start joint, which no exact
// per-link distance bound can clear; used to price a launch in FLOPs.
__device__ __forceinline__ bool arm_valid_counted(const SceneArgs& sc, const double* q, uint64_t* links,
                                                  uint64_t* tests, uint64_t* near) {
    double fx = 0.0, fy = 0.0, angle = 0.0;
    for (int i = 0; i < sc.nlinks; ++i) {
        angle += q[i];
        double s, c;
        pm::sincos(angle, &s, &c);
        *links += 1;
        double L = sc.lengths[i];
        double tx = fx + L * c;
        double ty = fy + L * s;
        for (int j = 0; j < sc.ndisks; ++j) {
            const double* dk = sc.disks + 4 * j;
            *tests += 1;
            const double dx = dk[0] - fx, dy = dk[1] - fy, reach = L + dk[3];
            if (dx * dx + dy * dy <= reach * reach) *near += 1;
            if (!(seg_dist_sq(dk[0], dk[1], fx, fy, tx, ty) > dk[2])) return false;
        }
        fx = tx;
        fy = ty;
    }
    return true;
}

// Rect::pointIsValid over all boxes, demo/shape_hierarchy.hpp:177-182
__device__ __forceinline__ bool boxes_valid(const SceneArgs& sc, const double* q) {
    for (int i = 0; i < sc.nboxes; ++i) {
        const double* b = sc.boxes + 4 * i;
        if (q[0] >= b[0] && q[0] <= b[2] && q[1] >= b[1] && q[1] <= b[3]) return false;
    }
    return true;
}

// PNG2dScenario::valid, demo/png_2d_scenario.hpp:104-110, plus a bounds check.
// The grid is one bit per cell in 8 x 8 tiles (a 4096^2 grid is 2 MiB, so it
// stays in one XCD's L2), tile-row-major, cell (x, y) at bit 8 * (y % 8) +
// x % 8 of tile (y / 8) * grid_tx + x / 8: a waypoint's 8 x 8 neighbourhood
// is one 8-byte word, and 16 tiles of a tile row share a 128-byte line.
// The cell is x = (int)(q0 + 0.5) (truncation), and x lies in [0, width)
// exactly when q0 + 0.5 lies in (-1, width); likewise y.  Written without
// branches: every lane loads a cell word (cell (0, 0) for points outside the
// grid) and the bounds decide afterwards, so unrolled callers keep their
// loads in flight together.
__device__ __forceinline__ bool grid_free(const SceneArgs& sc, double qx, double qy) {
    const double fx = qx + 0.5, fy = qy + 0.5;
    const bool in = (fx > -1.0) & (fx < (double)sc.width) & (fy > -1.0) & (fy < (double)sc.height);
    const int x = (int)(in ? fx : 0.0), y = (int)(in ? fy : 0.0);
    const uint64_t w = sc.grid[(int64_t)(y >> 3) * sc.grid_tx + (x >> 3)];
    return in & (((w >> (((y & 7) << 3) | (x & 7))) & 1ull) == 0);
}

__device__ __forceinline__ bool grid_valid(const SceneArgs& sc, const double* q) { return grid_free(sc, q[0], q[1]); }

// Circle::pointIsValid (demo/shape_hierarchy.hpp:222-226): the squared
// distance to the centre, x^2 + y^2 in that order, exceeds r^2.
__device__ __forceinline__ bool circles_valid(const SceneArgs& sc, const double* q) {
    for (int i = 0; i < sc.ncircles; ++i) {
        const double* c = sc.circles + 4 * i;
        const double dx = q[0] - c[0], dy = q[1] - c[1];
        if (!(dx * dx + dy * dy > c[2])) return false;
    }
    return true;
}

// Holonomic2DPointScenario::valid (demo/holonomic_2d_point_scenario.hpp:95-103):
// circles, then rectangles (the AND does not depend on the order).
__device__ __forceinline__ bool shapes_valid(const SceneArgs& sc, const double* q) {
    return circles_valid(sc, q) && boxes_valid(sc, q);
}

// Circle::segmentIsValid's closed form (demo/shape_hierarchy.hpp:228-231,
// distPointSegmentSquared :259-270) for every circle: the squared distance
// from the centre p to the segment s0 s1, with v = s1 - s0, w = p - s0,
// c1 = v.w; c1 <= 0: |w|^2; c2 = |v|^2 <= c1: |p - s1|^2; else
// |(s0 - p) + v (c1 / c2)|^2; it must exceed r^2.
__device__ __forceinline__ bool circles_segment_clear(const SceneArgs& sc, const double* a, const double* b) {
    const double vx = b[0] - a[0], vy = b[1] - a[1];
    for (int i = 0; i < sc.ncircles; ++i) {
        const double* c = sc.circles + 4 * i;
        const double wx = c[0] - a[0], wy = c[1] - a[1];
        const double c1 = vx * wx + vy * wy;
        double d2;
        if (c1 <= 0) {
            d2 = wx * wx + wy * wy;
        } else {
            const double c2 = vx * vx + vy * vy;
            if (c2 <= c1) {
                const double ex = c[0] - b[0], ey = c[1] - b[1];
                d2 = ex * ex + ey * ey;
            } else {
                const double t = c1 / c2;
                const double ex = (a[0] - c[0]) + vx * t, ey = (a[1] - c[1]) + vy * t;
                d2 = ex * ex + ey * ey;
            }
        }
        if (!(d2 > c[2])) return false;
    }
    return true;
}

// distPointSegmentSquared(c, s0, s1) > (r+R)^2, skipping disks whose centre
// is farther than lim + r + R from the segment start.  lim bounds the FP
// segment length with a relative margin of 2^-30 that dominates the rounding
// of the exact expression, so a skipped disk would have passed the exact
// test; every other disk gets the reference expression bit-for-bit.
__device__ __forceinline__ bool disk_clear(const ArmDisk& dk, double fx, double fy, double tx, double ty, double vx,
                                           double vy, double lim) {
    double wx = dk.cx - fx, wy = dk.cy - fy;
    double w2 = wx * wx + wy * wy;
    double far = lim + dk.rt;
    if (w2 > far * far) return true;
    double c1 = vx * wx + vy * wy;
    double d;
    if (c1 <= 0) {
        d = w2;
    } else {
        double c2 = vx * vx + vy * vy;
        if (c2 <= c1) {
            double ex = dk.cx - tx, ey = dk.cy - ty;
            d = ex * ex + ey * ey;
        } else {
            double b = c1 / c2;
            double ex = (fx - dk.cx) + vx * b;
            double ey = (fy - dk.cy) + vy * b;
            d = ex * ex + ey * ey;
        }
    }
    return d > dk.rr2;
}

// Arm validity for a state of D = nlinks coordinates: same verdict as
// arm_valid (the reference's FK and segment-vs-disk predicate), computed
// disk-major with a sound broad phase:
//   * forward kinematics with the reference's expression
//     p_{i+1} = p_i + L_i (cos, sin)(sum of angles), one group of kArmGroup
//     links at a time (group-major keeps only that group's joints live);
//   * a group's points all lie within its
//     half-length S_g/2 of m_g = (p_first + p_last)/2 (triangle inequality);
//   * for every disk, a group whose circle (m_g, rho_g + r + R) misses the
//     disk centre is skipped; rho_g = S_g/2 (1 + 2^-30) + 1e-12 dominates the
//     FK rounding, so every skipped (disk, link) pair would have passed the
//     exact test;
//   * the remaining pairs run disk_clear (the reference expression).
// coord(i) yields the state's i-th coordinate, so callers can generate the
// coordinates on the fly instead of holding the state in registers.
constexpr int kArmGroup = 5;

// Work tally of arm_ok_fn (instrumented builds only): FK links, group tests,
// link tests resolved by the far bound, link tests that ran the exact
// expression, valid() calls.
struct ArmWork {
    uint32_t fk, group, far, exact, calls;
};

template <int D, bool COUNT = false, typename Coord>
__device__ __forceinline__ bool arm_ok_fn(const SceneArgs& sc, Coord&& coord, ArmWork* wk = nullptr) {
    constexpr int G = (D + kArmGroup - 1) / kArmGroup;
    if constexpr (COUNT) wk->calls += 1;
    double px = 0.0, py = 0.0, angle = 0.0;  // first joint of the current group
#pragma unroll
    for (int g = 0; g < G; ++g) {
        constexpr int GS = kArmGroup;
        const int i0 = g * GS;
        const int n = (i0 + GS <= D) ? GS : D - i0;
        double jx[GS + 1], jy[GS + 1];
        jx[0] = px;
        jy[0] = py;
#pragma unroll
        for (int k = 0; k < GS; ++k) {
            if (k < n) {
                angle += coord(i0 + k);
                double s, c;
                pm::sincos(angle, &s, &c);
                double L = sc.lengths[i0 + k];
                jx[k + 1] = jx[k] + L * c;
                jy[k + 1] = jy[k] + L * s;
                if constexpr (COUNT) wk->fk += 1;
            }
        }
        double last_x = jx[0], last_y = jy[0];
#pragma unroll
        for (int k = 1; k <= GS; ++k)
            if (k <= n) {
                last_x = jx[k];
                last_y = jy[k];
            }
        const double mx = (jx[0] + last_x) / 2, my = (jy[0] + last_y) / 2;
        const double rho = sc.arm_lim[D + g];
        for (int j = 0; j < sc.ndisks; ++j) {
            const ArmDisk dk = sc.arm_disks[j];
            double dx = dk.cx - mx, dy = dk.cy - my;
            double R = rho + dk.rt;
            if constexpr (COUNT) wk->group += 1;
            if (dx * dx + dy * dy > R * R) continue;
#pragma unroll
            for (int k = 0; k < GS; ++k) {
                if (k < n) {
                    double vx = jx[k + 1] - jx[k], vy = jy[k + 1] - jy[k];
                    if constexpr (COUNT) {
                        double wx = dk.cx - jx[k], wy = dk.cy - jy[k];
                        double far = sc.arm_lim[i0 + k] + dk.rt;
                        if (wx * wx + wy * wy > far * far) wk->far += 1;
                        else wk->exact += 1;
                    }
                    if (!disk_clear(dk, jx[k], jy[k], jx[k + 1], jy[k + 1], vx, vy, sc.arm_lim[i0 + k]))
                        return false;
                }
            }
        }
        px = last_x;
        py = last_y;
    }
    return true;
}

__device__ __forceinline__ double wave_min_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// arm_ok_fn for a whole wave whose lanes check nearby states (one edge's
// waypoints): every lane must call it (lanes with active == false compute on
// any valid coordinates and return true).  Per group, the lanes' group-circle
// centres span a box; a disk farther than (group half-length + r + R)
// (1 + 2^-40) + 1e-12 outside that box fails every lane's own group test
// (the margin dominates the rounding of dx*dx + dy*dy > R*R for coordinates
// below 10^3), so the wave skips it with one uniform branch.  Other disks go
// through the per-lane tests of arm_ok_fn unchanged.  Lanes that already
// found a collision leave the box; the verdicts equal arm_ok_fn's.
template <int D, bool COUNT = false, typename Coord>
__device__ __forceinline__ bool arm_ok_wave(const SceneArgs& sc, Coord&& coord, bool active, ArmWork* wk = nullptr) {
    constexpr int G = (D + kArmGroup - 1) / kArmGroup;
    if constexpr (COUNT) wk->calls += active ? 1 : 0;
    bool ok = true;
    double px = 0.0, py = 0.0, angle = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        constexpr int GS = kArmGroup;
        const int i0 = g * GS;
        const int n = (i0 + GS <= D) ? GS : D - i0;
        double jx[GS + 1], jy[GS + 1];
        jx[0] = px;
        jy[0] = py;
#pragma unroll
        for (int k = 0; k < GS; ++k) {
            if (k < n) {
                angle += coord(i0 + k);
                double s, c;
                pm::sincos(angle, &s, &c);
                double L = sc.lengths[i0 + k];
                jx[k + 1] = jx[k] + L * c;
                jy[k + 1] = jy[k] + L * s;
                if constexpr (COUNT) wk->fk += (active && ok) ? 1 : 0;
            }
        }
        double last_x = jx[0], last_y = jy[0];
#pragma unroll
        for (int k = 1; k <= GS; ++k)
            if (k <= n) {
                last_x = jx[k];
                last_y = jy[k];
            }
        const double mx = (jx[0] + last_x) / 2, my = (jy[0] + last_y) / 2;
        const bool live = active && ok;
        if (__ballot(live) == 0) break;  // wave-uniform
        const double inf = __builtin_inf();
        const double bx0 = wave_min_f64(live ? mx : inf), bx1 = wave_max_f64(live ? mx : -inf);
        const double by0 = wave_min_f64(live ? my : inf), by1 = wave_max_f64(live ? my : -inf);
        const double rho = sc.arm_lim[D + g];
        // disks whose centre lies within the box widened by the largest
        // (group half-length + r + R) sit in the grid cells that box covers,
        // one contiguous run per cell row; every other disk is farther than
        // its own bound below
        const ArmGrid ag = arm_grid(sc.arm_index);
        const double Rmax = (rho + ag.rtmax) * (1.0 + 0x1.0p-40) + 1e-12;
        const int gx0 = arm_cell(bx0 - Rmax, ag.x0, ag.ix, ag.nx), gx1 = arm_cell(bx1 + Rmax, ag.x0, ag.ix, ag.nx);
        const int gy0 = arm_cell(by0 - Rmax, ag.y0, ag.iy, ag.ny), gy1 = arm_cell(by1 + Rmax, ag.y0, ag.iy, ag.ny);
        for (int gy = gy0; gy <= gy1; ++gy) {
        const int j0 = ag.starts[gy * ag.nx + gx0];
        const int j1 = ag.starts[gy * ag.nx + gx1 + 1];
        for (int jb = j0; jb < j1; jb += kDiskBatch) {
            ArmDisk batch[kDiskBatch];
#pragma unroll
            for (int u = 0; u < kDiskBatch; ++u) batch[u] = sc.arm_disks[jb + u < j1 ? jb + u : j1 - 1];
#pragma unroll
            for (int u = 0; u < kDiskBatch; ++u) {
                if (jb + u >= j1) break;  // uniform
                const ArmDisk dk = batch[u];
                const double Rw = (rho + dk.rt) * (1.0 + 0x1.0p-40) + 1e-12;
                if (dk.cx < bx0 - Rw || dk.cx > bx1 + Rw || dk.cy < by0 - Rw || dk.cy > by1 + Rw) continue;  // uniform
                if (!(active && ok)) continue;
                double dx = dk.cx - mx, dy = dk.cy - my;
                double R = rho + dk.rt;
                if constexpr (COUNT) wk->group += 1;
                if (dx * dx + dy * dy > R * R) continue;
#pragma unroll
                for (int k = 0; k < GS; ++k) {
                    if (k < n && ok) {
                        double vx = jx[k + 1] - jx[k], vy = jy[k + 1] - jy[k];
                        if constexpr (COUNT) {
                            double wx = dk.cx - jx[k], wy = dk.cy - jy[k];
                            double far = sc.arm_lim[i0 + k] + dk.rt;
                            if (wx * wx + wy * wy > far * far) wk->far += 1;
                            else wk->exact += 1;
                        }
                        if (!disk_clear(dk, jx[k], jy[k], jx[k + 1], jy[k + 1], vx, vy, sc.arm_lim[i0 + k]))
                            ok = false;
                    }
                }
            }
        }
        }  // grid rows
        px = last_x;
        py = last_y;
    }
    return ok;
}

// arm_ok_fn for one lane on its own (states unrelated to their neighbours'):
// per group, only the disks whose centre lies within (group half-length +
// the largest r + R) (1 + 2^-40) + 1e-12 of the lane's group-circle centre,
// per axis, can pass the group test; they sit in the grid cells that box
// covers, one contiguous run per cell row.  Every other disk fails
// arm_ok_fn's group test (the margin dominates the rounding of
// dx*dx + dy*dy > R*R for coordinates below 10^3), so the verdict equals
// arm_ok_fn's.  No cross-lane operation: lanes may call it divergently;
// active == false returns true at once.
template <int D, bool COUNT = false, typename Coord>
__device__ __forceinline__ bool arm_ok_lane(const SceneArgs& sc, Coord&& coord, bool active = true,
                                            ArmWork* wk = nullptr) {
    if (!active) return true;
    constexpr int G = (D + kArmGroup - 1) / kArmGroup;
    if constexpr (COUNT) wk->calls += 1;
    const ArmGrid ag = arm_grid(sc.arm_index);
    double px = 0.0, py = 0.0, angle = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        constexpr int GS = kArmGroup;
        const int i0 = g * GS;
        const int n = (i0 + GS <= D) ? GS : D - i0;
        double jx[GS + 1], jy[GS + 1];
        jx[0] = px;
        jy[0] = py;
#pragma unroll
        for (int k = 0; k < GS; ++k) {
            if (k < n) {
                angle += coord(i0 + k);
                double s, c;
                pm::sincos(angle, &s, &c);
                double L = sc.lengths[i0 + k];
                jx[k + 1] = jx[k] + L * c;
                jy[k + 1] = jy[k] + L * s;
                if constexpr (COUNT) wk->fk += 1;
            }
        }
        double last_x = jx[0], last_y = jy[0];
#pragma unroll
        for (int k = 1; k <= GS; ++k)
            if (k <= n) {
                last_x = jx[k];
                last_y = jy[k];
            }
        const double mx = (jx[0] + last_x) / 2, my = (jy[0] + last_y) / 2;
        const double rho = sc.arm_lim[D + g];
        const double Rmax = (rho + ag.rtmax) * (1.0 + 0x1.0p-40) + 1e-12;
        const int gx0 = arm_cell(mx - Rmax, ag.x0, ag.ix, ag.nx), gx1 = arm_cell(mx + Rmax, ag.x0, ag.ix, ag.nx);
        const int gy0 = arm_cell(my - Rmax, ag.y0, ag.iy, ag.ny), gy1 = arm_cell(my + Rmax, ag.y0, ag.iy, ag.ny);
        for (int gy = gy0; gy <= gy1; ++gy) {
            const int j1 = ag.starts[gy * ag.nx + gx1 + 1];
            for (int j = ag.starts[gy * ag.nx + gx0]; j < j1; ++j) {
                const ArmDisk dk = sc.arm_disks[j];
                double dx = dk.cx - mx, dy = dk.cy - my;
                double R = rho + dk.rt;
                if constexpr (COUNT) wk->group += 1;
                if (dx * dx + dy * dy > R * R) continue;
#pragma unroll
                for (int k = 0; k < GS; ++k) {
                    if (k < n) {
                        double vx = jx[k + 1] - jx[k], vy = jy[k + 1] - jy[k];
                        if constexpr (COUNT) {
                            double wx = dk.cx - jx[k], wy = dk.cy - jy[k];
                            double far = sc.arm_lim[i0 + k] + dk.rt;
                            if (wx * wx + wy * wy > far * far) wk->far += 1;
                            else wk->exact += 1;
                        }
                        if (!disk_clear(dk, jx[k], jy[k], jx[k + 1], jy[k + 1], vx, vy, sc.arm_lim[i0 + k]))
                            return false;
                    }
                }
            }
        }
        px = last_x;
        py = last_y;
    }
    return true;
}

// arm_ok_lane with L lanes per state: the L lanes of an aligned group of the
// wave check the same state, each a 1/L share of every group's disk run
// (disks j0 + u, j0 + u + L, ...), and OR their hits after each group of
// links.  Same disks tested as arm_ok_lane, so the same verdict.  Every lane
// of the wave must call it (the OR is a shuffle); lanes with active == false
// test no disk and return true.
template <int D, int L, typename Coord>
__device__ __forceinline__ bool arm_ok_split(const SceneArgs& sc, Coord&& coord, bool active) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L must be a power of two up to 64");
    constexpr int G = (D + kArmGroup - 1) / kArmGroup;
    const int u = (int)(threadIdx.x & 63) & (L - 1);
    const ArmGrid ag = arm_grid(sc.arm_index);
    bool ok = true;
    double px = 0.0, py = 0.0, angle = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        constexpr int GS = kArmGroup;
        const int i0 = g * GS;
        const int n = (i0 + GS <= D) ? GS : D - i0;
        double jx[GS + 1], jy[GS + 1];
        jx[0] = px;
        jy[0] = py;
#pragma unroll
        for (int k = 0; k < GS; ++k) {
            if (k < n) {
                angle += coord(i0 + k);
                double s, c;
                pm::sincos(angle, &s, &c);
                double L_ = sc.lengths[i0 + k];
                jx[k + 1] = jx[k] + L_ * c;
                jy[k + 1] = jy[k] + L_ * s;
            }
        }
        double last_x = jx[0], last_y = jy[0];
#pragma unroll
        for (int k = 1; k <= GS; ++k)
            if (k <= n) {
                last_x = jx[k];
                last_y = jy[k];
            }
        const double mx = (jx[0] + last_x) / 2, my = (jy[0] + last_y) / 2;
        const double rho = sc.arm_lim[D + g];
        const double Rmax = (rho + ag.rtmax) * (1.0 + 0x1.0p-40) + 1e-12;
        int gx0 = 0, gx1 = -1, gy0 = 0, gy1 = -1;
        if (active && ok) {
            gx0 = arm_cell(mx - Rmax, ag.x0, ag.ix, ag.nx);
            gx1 = arm_cell(mx + Rmax, ag.x0, ag.ix, ag.nx);
            gy0 = arm_cell(my - Rmax, ag.y0, ag.iy, ag.ny);
            gy1 = arm_cell(my + Rmax, ag.y0, ag.iy, ag.ny);
        }
        bool hit = false;
        // the group's disk runs, row after row, dealt to its L lanes in turn
        for (int gy = gy0; gy <= gy1 && !hit; ++gy) {
        const int j0 = ag.starts[gy * ag.nx + gx0], j1 = ag.starts[gy * ag.nx + gx1 + 1];
        for (int j = j0 + u; j < j1 && !hit; j += L) {
            const ArmDisk dk = sc.arm_disks[j];
            double dx = dk.cx - mx, dy = dk.cy - my;
            double R = rho + dk.rt;
            if (dx * dx + dy * dy > R * R) continue;
#pragma unroll
            for (int k = 0; k < GS; ++k) {
                if (k < n && !hit) {
                    double vx = jx[k + 1] - jx[k], vy = jy[k + 1] - jy[k];
                    if (!disk_clear(dk, jx[k], jy[k], jx[k + 1], jy[k + 1], vx, vy, sc.arm_lim[i0 + k])) hit = true;
                }
            }
        }
        }  // grid rows
#pragma unroll
        for (int o = 1; o < L; o <<= 1) hit = (__shfl_xor(hit ? 1 : 0, o) != 0) || hit;
        ok = ok && !hit;
        px = last_x;
        py = last_y;
    }
    return ok;
}

template <int D>
__device__ __forceinline__ bool arm_ok(const SceneArgs& sc, const double* q) {
    return arm_ok_lane<D>(sc, [&](int i) { return q[i]; });
}

// Validity with the state size known at compile time (D > 0) or at run time
// (D == 0, generic fallback).
template <int D>
__device__ __forceinline__ bool state_valid_t(const SceneArgs& sc, const double* q) {
    if constexpr (D > 0) {
        if (sc.kind == MPTG_SCENE_ARM && sc.nlinks == D) return arm_ok<D>(sc, q);
    }
    switch (sc.kind) {
        case MPTG_SCENE_BOXES: return boxes_valid(sc, q);
        case MPTG_SCENE_GRID: return grid_valid(sc, q);
        case MPTG_SCENE_ARM: return arm_valid(sc, q);
        case MPTG_SCENE_SHAPES: return shapes_valid(sc, q);
        default: return true;
    }
}

template <int D>
__device__ __forceinline__ double space_dist_t(const SpaceArgs& sp, const double* a, const double* b) {
    // SE(3) states have 7 coordinates (mptg_space enforces it): any other
    // compile-time size is L1 or L2, and the arrays never leave registers
    if constexpr (D > 0) {
        if (D != 7 || sp.metric != MPTG_METRIC_SE3) {
            double s = 0.0;
            if (sp.metric == MPTG_METRIC_L1) {
#pragma unroll
                for (int i = 0; i < D; ++i) s = s + __builtin_fabs(a[i] - b[i]);
                return s;
            }
#pragma unroll
            for (int i = 0; i < D; ++i) {
                double t = a[i] - b[i];
                s = s + t * t;
            }
            return __builtin_sqrt(s);
        }
    }
    return space_dist(sp, a, b);
}

template <int D>
__device__ __forceinline__ void space_interp_t(const SpaceArgs& sp, const double* a, const double* b, double t,
                                               double* q) {
    if constexpr (D > 0) {
        if (D != 7 || sp.metric != MPTG_METRIC_SE3) {
#pragma unroll
            for (int i = 0; i < D; ++i) q[i] = (b[i] - a[i]) * t + a[i];
            return;
        }
    }
    if constexpr (D == 0 || D == 7) space_interp(sp, a, b, t, q);
}

// Runs f.template operator()<D>() with D the compile-time state size, or
// D = 0 (run-time size) for dimensions without an instantiation.
template <typename F>
inline void dispatch_dim(int dim, F&& f) {
    switch (dim) {
        case 2: f.template operator()<2>(); break;
        case 3: f.template operator()<3>(); break;
        case 5: f.template operator()<5>(); break;
        case 7: f.template operator()<7>(); break;
        case 10: f.template operator()<10>(); break;
        case 20: f.template operator()<20>(); break;
        default: f.template operator()<0>(); break;
    }
}

__device__ __forceinline__ bool state_valid(const SceneArgs& sc, const double* q) {
    switch (sc.kind) {
        case MPTG_SCENE_BOXES: return boxes_valid(sc, q);
        case MPTG_SCENE_GRID: return grid_valid(sc, q);
        case MPTG_SCENE_ARM: return arm_valid(sc, q);
        case MPTG_SCENE_SHAPES: return shapes_valid(sc, q);
        default: return true;
    }
}

}  // namespace mptg
