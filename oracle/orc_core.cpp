// CPU oracle: spaces, exact nearest neighbours, state validity and motion
// validation — TEST INFRASTRUCTURE ONLY (loaded by tests, the smoke check and
// the benchmark's CPU-baseline leg, never by the product library).
//
// Restated reference behaviour (paths relative to the reference root):
//   * L_p interpolation  (b - a) * t + a per coordinate   src/mpt/lp_space.hpp:44-54
//   * SO(3) slerp with the lerp branch and sign flip       src/mpt/so3_space.hpp:43-83
//   * Cartesian/scaled interpolation per component         src/mpt/cartesian_space.hpp:43-68,
//                                                          src/mpt/scaled_space.hpp:43-51
//   * metrics (nigh, absent): L2/L1 sums, SO(3) acos(|<a,b>|), weighted sum;
//     values pinned by test/lp_space_test.cpp:39-50, test/so3_space_test.cpp:39-56,
//     test/se3_space_test.cpp:45-89 (restated in tests/test_oracle_kat.py)
//   * DiscreteMotionValidator: valid(to) first, steps = ceil(d * (1/step)),
//     waypoints i*(1/steps), 256-slot BFS bisection ring with sequential
//     fallback                                              src/mpt/discrete_motion_validator.hpp:71-130
//   * boxes: Rect::pointIsValid (closed box)               demo/shape_hierarchy.hpp:177-182
//   * grid: PNG2dScenario::valid ((int)(q + 0.5) lookup)   demo/png_2d_scenario.hpp:104-110
//           (+ a bounds check the reference omits)
//   * arm: LinkManipulatorScenario::valid (FK, segment vs every disk, first
//     hit exits) and bisectLink (L-inf < tol, (a+b)/2, pre-order)
//                                                          demo/link_manipulator_scenario.hpp:99-138
//     with Circle::segmentIsValid / distPointSegmentSquared demo/shape_hierarchy.hpp:234-237,259-270
//   * shapes: Holonomic2DPointScenario::valid/link (circles, then rectangles;
//     Circle::pointIsValid, the closed-form Circle::segmentIsValid, and per
//     rectangle the endpoints and the midpoint recursion bisectSegment to
//     squared length < tol^2)                              demo/holonomic_2d_point_scenario.hpp:95-113,
//                                                          demo/shape_hierarchy.hpp:184-231,259-270
//   * BISECT_L2 link: valid(a), valid(b), then recursive midpoints until the
//     squared L2 length is < tol^2                         demo/png_2d_scenario.hpp:112-117,152-165
//   * mesh: SE3RigidBodyScenario::valid (robot vs environment triangles)
//                                                          demo/se3_rigid_body_scenario.hpp:288-312
//     (orc_mesh.cpp)
//
// Parity pinning: see tests/test_oracle_*.py.  The nearest-neighbour order
// "ascending (FP64 distance, index)" and the left-to-right summation order are
// this project's definition (nigh's KD-tree tie order and summation order
// cannot be observed without nigh): parity of the GPU path is exact against
// this definition, and the definition itself is pinned by brute-force
// invariants and hand-checked cases.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <utility>
#include <vector>

#include "../mpt_amd/csrc/pmath.hpp"
#include "../mpt_amd/csrc/sample_hd.hpp"

namespace orc_internal {
bool valid_mesh(const orc_scene* sc, const double* q);
}  // namespace orc_internal

namespace {

const double kInf = std::numeric_limits<double>::infinity();

double so3_dot(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

double dist_l2(const double* a, const double* b, int d) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) {
        double t = a[i] - b[i];
        s = s + t * t;
    }
    return std::sqrt(s);
}

double dist_l1(const double* a, const double* b, int d) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) s = s + std::fabs(a[i] - b[i]);
    return s;
}

double dist_so3(const double* a, const double* b) {
    double d = std::fabs(so3_dot(a, b));
    if (d >= 1.0) return 0.0;
    return mptg::pm::acos(d);
}

double distance(const orc_scene* sc, const double* a, const double* b) {
    switch (sc->metric) {
        case ORC_METRIC_L1: return dist_l1(a, b, sc->dim);
        case ORC_METRIC_SE3:
            return sc->w_so3 * dist_so3(a, b) + sc->w_l2 * dist_l2(a + 4, b + 4, 3);
        default: return dist_l2(a, b, sc->dim);
    }
}

void lerp(const double* a, const double* b, double t, double* q, int d) {
    for (int i = 0; i < d; ++i) q[i] = (b[i] - a[i]) * t + a[i];
}

void slerp(const double* a, const double* b, double t, double* q) {
    double d = so3_dot(a, b);
    double ad = std::fabs(d);
    double s0, s1;
    if (d >= 1.0 - std::numeric_limits<double>::epsilon()) {
        s0 = 1.0 - t;
        s1 = t;
    } else {
        double theta = mptg::pm::acos(ad);
        double sinTheta = mptg::pm::sin(theta);
        s0 = mptg::pm::sin((1.0 - t) * theta) / sinTheta;
        s1 = mptg::pm::sin(t * theta) / sinTheta;
    }
    if (d < 0) s1 = -s1;
    for (int i = 0; i < 4; ++i) q[i] = s0 * a[i] + s1 * b[i];
}

void interpolate(const orc_scene* sc, const double* a, const double* b, double t, double* q) {
    if (sc->metric == ORC_METRIC_SE3) {
        slerp(a, b, t, q);
        lerp(a + 4, b + 4, t, q + 4, 3);
    } else {
        lerp(a, b, t, q, sc->dim);
    }
}

// distPointSegmentSquared(pt, s0, s1), demo/shape_hierarchy.hpp:259-270
double dist_point_segment_sq(double px, double py, double s0x, double s0y, double s1x, double s1y) {
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

bool valid_arm(const orc_scene* sc, const double* q) {
    double fx = 0.0, fy = 0.0, angle = 0.0;
    double R = sc->robot_radius;
    for (int i = 0; i < sc->nlinks; ++i) {
        angle += q[i];
        double s, c;
        mptg::pm::sincos(angle, &s, &c);
        double L = sc->lengths[i];
        double tx = fx + L * c;
        double ty = fy + L * s;
        for (int j = 0; j < sc->ndisks; ++j) {
            const double* dk = sc->disks + 3 * j;
            double rr = dk[2] + R;
            if (!(dist_point_segment_sq(dk[0], dk[1], fx, fy, tx, ty) > rr * rr)) return false;
        }
        fx = tx;
        fy = ty;
    }
    return true;
}

// valid_arm with its work counted, in the reference's loop order: links whose
// forward kinematics ran, segment-disk tests made (every disk until the first
// hit), and of those the tests whose disk centre lies within (link length +
// r + R) of the link's start joint -- the pairs that no exact per-link
// distance bound can clear, since every point of the link lies within its
// length of that joint.
bool valid_arm_work(const orc_scene* sc, const double* q, uint64_t* fk, uint64_t* tests, uint64_t* near) {
    double fx = 0.0, fy = 0.0, angle = 0.0;
    double R = sc->robot_radius;
    for (int i = 0; i < sc->nlinks; ++i) {
        angle += q[i];
        double s, c;
        mptg::pm::sincos(angle, &s, &c);
        double L = sc->lengths[i];
        double tx = fx + L * c;
        double ty = fy + L * s;
        ++*fk;
        for (int j = 0; j < sc->ndisks; ++j) {
            const double* dk = sc->disks + 3 * j;
            double rr = dk[2] + R;
            ++*tests;
            double dx = dk[0] - fx, dy = dk[1] - fy, reach = L + rr;
            if (dx * dx + dy * dy <= reach * reach) ++*near;
            if (!(dist_point_segment_sq(dk[0], dk[1], fx, fy, tx, ty) > rr * rr)) return false;
        }
        fx = tx;
        fy = ty;
    }
    return true;
}

bool valid_boxes(const orc_scene* sc, const double* q) {
    for (int i = 0; i < sc->nboxes; ++i) {
        const double* b = sc->boxes + 4 * i;
        if (q[0] >= b[0] && q[0] <= b[2] && q[1] >= b[1] && q[1] <= b[3]) return false;
    }
    return true;
}

bool valid_grid(const orc_scene* sc, const double* q) {
    double fx = q[0] + 0.5, fy = q[1] + 0.5;
    // (int) truncation as the reference; states outside the grid are invalid
    // (the reference indexes without a check).
    if (!(fx > -1.0 && fx < (double)sc->width + 1.0 && fy > -1.0 && fy < (double)sc->height + 1.0))
        return false;
    int x = (int)fx, y = (int)fy;
    if (x < 0 || y < 0 || x >= sc->width || y >= sc->height) return false;
    return sc->grid[(int64_t)sc->width * y + x] == 0;
}

// Circle::pointIsValid: (q - c).squaredNorm() > r * r, demo/shape_hierarchy.hpp:222-226
bool valid_circles(const orc_scene* sc, const double* q) {
    for (int i = 0; i < sc->ncircles; ++i) {
        const double* c = sc->circles + 3 * i;
        double dx = q[0] - c[0], dy = q[1] - c[1];
        if (!(dx * dx + dy * dy > c[2] * c[2])) return false;
    }
    return true;
}

bool valid(const orc_scene* sc, const double* q) {
    switch (sc->kind) {
        case ORC_SCENE_SHAPES: return valid_circles(sc, q) && valid_boxes(sc, q);
        case ORC_SCENE_BOXES: return valid_boxes(sc, q);
        case ORC_SCENE_GRID: return valid_grid(sc, q);
        case ORC_SCENE_ARM: return valid_arm(sc, q);
        case ORC_SCENE_MESH: return orc_internal::valid_mesh(sc, q);
        default: return true;
    }
}

// Visits the DiscreteMotionValidator waypoint indices in check order; the
// visitor returns false to stop.  Returns false iff stopped.
template <typename Visit>
bool dmv_visit(std::size_t steps, Visit&& visit) {
    constexpr std::size_t Q = 256;
    std::array<std::pair<std::size_t, std::size_t>, Q> queue;
    queue[0] = std::make_pair(std::size_t(1), steps - 1);
    std::size_t qStart = 0, qEnd = 1;
    while (qStart != qEnd) {
        auto [mn, mx] = queue[qStart++ % Q];
        if (mn == mx) {
            if (!visit(mn)) return false;
        } else if (qEnd + 2 < qStart + Q) {
            std::size_t mid = (mn + mx) / 2;
            if (!visit(mid)) return false;
            if (mn < mid) queue[qEnd++ % Q] = std::make_pair(mn, mid - 1);
            if (mid < mx) queue[qEnd++ % Q] = std::make_pair(mid + 1, mx);
        } else {
            for (std::size_t i = mn; i <= mx; ++i)
                if (!visit(i)) return false;
        }
    }
    return true;
}

std::size_t dmv_steps(const orc_scene* sc, const double* a, const double* b) {
    double inv = 1.0 / sc->link_param;
    double s = std::ceil(distance(sc, a, b) * inv);
    return (std::size_t)s;
}

// Pre-order visit of the recursive-midpoint waypoints (bisectLink).  The
// visitor receives the midpoint state and returns false to stop.
// BISECT stops below L-inf length tol (bisectLink); BISECT_L2 below squared L2
// length tol^2 (bisectSegment / validSegment: (b - a).squaredNorm() < tol * tol).
template <typename Visit>
bool bisect_visit(const orc_scene* sc, const double* a, const double* b, Visit& visit, int depth = 0) {
    int d = sc->dim;
    if (sc->link_mode == ORC_LINK_BISECT_L2) {
        double s2 = 0.0;
        for (int i = 0; i < d; ++i) {
            double t = b[i] - a[i];
            s2 = s2 + t * t;
        }
        if (s2 < sc->link_param * sc->link_param) return true;
    } else {
        double m = 0.0;
        for (int i = 0; i < d; ++i) m = std::max(m, std::fabs(a[i] - b[i]));
        if (m < sc->link_param) return true;
    }
    double mid[64];
    for (int i = 0; i < d; ++i) mid[i] = (a[i] + b[i]) / 2;
    if (!visit(static_cast<const double*>(mid))) return false;
    if (!bisect_visit(sc, a, mid, visit, depth + 1)) return false;
    return bisect_visit(sc, mid, b, visit, depth + 1);
}

struct LinkResult {
    bool ok;
    uint32_t checks;
};

LinkResult link(const orc_scene* sc, const double* a, const double* b, bool known_valid) {
    uint32_t checks = 0;
    if (sc->link_mode == ORC_LINK_BISECT || sc->link_mode == ORC_LINK_BISECT_L2) {
        // circles and rectangles: every circle's closed form first, then the
        // endpoints and midpoints against the rectangles (every rectangle's
        // recursion visits the same midpoints, so testing them all per point
        // decides the same AND; `checks` counts those points)
        const bool shapes = sc->kind == ORC_SCENE_SHAPES;
        if (shapes)
            for (int i = 0; i < sc->ncircles; ++i) {
                const double* c = sc->circles + 3 * i;
                if (!(dist_point_segment_sq(c[0], c[1], a[0], a[1], b[0], b[1]) > c[2] * c[2])) return {false, 0};
            }
        auto point_ok = [&](const double* q) { return shapes ? valid_boxes(sc, q) : valid(sc, q); };
        ++checks;
        if (!known_valid && !point_ok(a)) return {false, checks};
        ++checks;
        if (!known_valid && !point_ok(b)) return {false, checks};
        auto visit = [&](const double* q) {
            ++checks;
            return point_ok(q);
        };
        bool ok = bisect_visit(sc, a, b, visit);
        return {ok, checks};
    }
    ++checks;
    if (!known_valid && !valid(sc, b)) return {false, checks};
    std::size_t steps = dmv_steps(sc, a, b);
    if (steps < 2) return {true, checks};
    double delta = 1.0 / (double)steps;
    double q[64];
    bool ok = dmv_visit(steps, [&](std::size_t i) {
        ++checks;
        interpolate(sc, a, b, (double)i * delta, q);
        return valid(sc, q);
    });
    return {ok, checks};
}

template <typename Fn>
void parallel_for(int64_t n, int nthreads, Fn&& fn) {
    if (nthreads <= 1 || n < 64) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> ts;
    int64_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        ts.emplace_back([lo, hi, &fn] {
            for (int64_t i = lo; i < hi; ++i) fn(i);
        });
    }
    for (auto& t : ts) t.join();
}

}  // namespace

// ---------------------------------------------------------------------------
// exported API
// ---------------------------------------------------------------------------
namespace orc_internal {
// The sample of slot `sidx`: the goal (the bias coin taken), a row of the
// custom sampler's table, or the Philox box sample.  false: the custom
// sampler gave no sample (the slot is skipped, as addSample(std::optional)
// skips an empty one, prrt.hpp:400-403); -1 via *err when the table is short.
bool sample_slot(const orc_scene* sc, const orc_prrt_params* p, uint64_t sidx, bool use_goal, double* q, int* err) {
    const int d = sc->dim;
    if (use_goal) {
        std::memcpy(q, p->goal, sizeof(double) * d);
        return true;
    }
    if (p->samples) {
        if ((int64_t)sidx >= p->nsamples) {
            *err = -1;
            return false;
        }
        if (p->samples_ok && !p->samples_ok[sidx]) return false;
        std::memcpy(q, p->samples + (int64_t)sidx * d, sizeof(double) * d);
        return true;
    }
    mptg::sample_state(p->seed, sidx, sc->metric, d, p->lo, p->hi, q);
    return true;
}

double distance(const orc_scene* sc, const double* a, const double* b) { return ::distance(sc, a, b); }
void interpolate(const orc_scene* sc, const double* a, const double* b, double t, double* q) {
    ::interpolate(sc, a, b, t, q);
}
bool valid(const orc_scene* sc, const double* q) { return ::valid(sc, q); }
}  // namespace orc_internal

extern "C" {

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    mptg::Philox4 r = mptg::philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void orc_sincos(const double* x, int64_t n, double* s, double* c) {
    for (int64_t i = 0; i < n; ++i) mptg::pm::sincos(x[i], &s[i], &c[i]);
}

void orc_acos(const double* x, int64_t n, double* y) {
    for (int64_t i = 0; i < n; ++i) y[i] = mptg::pm::acos(x[i]);
}

double orc_uniform_at(uint64_t seed, uint64_t index, uint32_t purpose, uint32_t j) {
    return mptg::uniform_at(seed, index, purpose, j);
}

double orc_distance(const orc_scene* sc, const double* a, const double* b) { return distance(sc, a, b); }

void orc_interpolate(const orc_scene* sc, const double* a, const double* b, double t, double* out) {
    interpolate(sc, a, b, t, out);
}

void orc_knn(const orc_scene* sc, const double* pts, int64_t n, const double* q, int64_t nq, int32_t k,
             double r, int32_t* idx, double* dist, int32_t nthreads) {
    int d = sc->dim;
    parallel_for(nq, nthreads, [&](int64_t qi) {
        std::vector<std::pair<double, int32_t>> best;  // sorted ascending, size <= k
        best.reserve(k + 1);
        const double* qq = q + qi * d;
        for (int64_t i = 0; i < n; ++i) {
            double di = distance(sc, qq, pts + i * d);
            if (!(di <= r)) continue;
            std::pair<double, int32_t> e(di, (int32_t)i);
            if ((int)best.size() == k && !(e < best.back())) continue;
            auto it = std::upper_bound(best.begin(), best.end(), e);
            best.insert(it, e);
            if ((int)best.size() > k) best.pop_back();
        }
        for (int j = 0; j < k; ++j) {
            if (j < (int)best.size()) {
                idx[qi * k + j] = best[j].second;
                dist[qi * k + j] = best[j].first;
            } else {
                idx[qi * k + j] = -1;
                dist[qi * k + j] = kInf;
            }
        }
    });
}

int32_t orc_valid(const orc_scene* sc, const double* q) { return valid(sc, q) ? 1 : 0; }

int64_t orc_dmv_order(int64_t steps, int64_t* out, int64_t cap) {
    if (steps < 2) return 0;
    int64_t cnt = 0;
    dmv_visit((std::size_t)steps, [&](std::size_t i) {
        if (cnt < cap) out[cnt] = (int64_t)i;
        ++cnt;
        return true;
    });
    return cnt;
}

int64_t orc_link_waypoints(const orc_scene* sc, const double* a, const double* b, double* out, int64_t cap) {
    int d = sc->dim;
    int64_t cnt = 0;
    if (sc->link_mode == ORC_LINK_BISECT || sc->link_mode == ORC_LINK_BISECT_L2) {
        auto visit = [&](const double* q) {
            if (cnt < cap) std::memcpy(out + cnt * d, q, sizeof(double) * d);
            ++cnt;
            return true;
        };
        bisect_visit(sc, a, b, visit);
        return cnt;
    }
    std::size_t steps = dmv_steps(sc, a, b);
    if (steps < 2) return 0;
    double delta = 1.0 / (double)steps;
    dmv_visit(steps, [&](std::size_t i) {
        if (cnt < cap) interpolate(sc, a, b, (double)i * delta, out + cnt * d);
        ++cnt;
        return true;
    });
    return cnt;
}

int32_t orc_link(const orc_scene* sc, const double* a, const double* b, int32_t known_valid,
                 uint32_t* checks_done) {
    LinkResult r = link(sc, a, b, known_valid != 0);
    if (checks_done) *checks_done = r.checks;
    return r.ok ? 1 : 0;
}

void orc_link_batch(const orc_scene* sc, const double* from, const double* to, int64_t n,
                    int32_t known_valid, uint8_t* verdict, uint32_t* checks_done, int32_t nthreads) {
    int d = sc->dim;
    parallel_for(n, nthreads, [&](int64_t i) {
        LinkResult r = link(sc, from + i * d, to + i * d, known_valid != 0);
        verdict[i] = r.ok ? 1 : 0;
        checks_done[i] = r.checks;
    });
}

int32_t orc_arm_link_work(const orc_scene* sc, const double* from, const double* to, int64_t n,
                          int32_t known_valid, int32_t nthreads, uint64_t* out) {
    if (sc->kind != ORC_SCENE_ARM || sc->link_mode != ORC_LINK_BISECT) return -1;
    const int d = sc->dim;
    std::vector<std::array<uint64_t, 4>> w((size_t)n);
    parallel_for(n, nthreads, [&](int64_t e) {
        std::array<uint64_t, 4> c{};
        auto check = [&](const double* q) {
            ++c[0];
            return valid_arm_work(sc, q, &c[1], &c[2], &c[3]);
        };
        const double* a = from + e * d;
        const double* b = to + e * d;
        bool ok = known_valid || (check(a) && check(b));
        if (ok) bisect_visit(sc, a, b, check);
        w[(size_t)e] = c;
    });
    for (int i = 0; i < 4; ++i) out[i] = 0;
    for (const auto& c : w)
        for (int i = 0; i < 4; ++i) out[i] += c[i];
    return 0;
}

uint32_t orc_checks_in_score_order(const orc_scene* sc, const double* a, const double* b,
                                   int32_t known_valid, const float* scores, int64_t nscores, int32_t lead,
                                   int32_t top_levels, int32_t picks) {
    int d = sc->dim;
    uint32_t checks = 0;
    int nend = sc->link_mode == ORC_LINK_BISECT ? 2 : 1;
    for (int e = 0; e < nend; ++e) {
        const double* q = (sc->link_mode == ORC_LINK_BISECT && e == 0) ? a : b;
        ++checks;
        if (!known_valid && !valid(sc, q)) return checks;
    }
    std::vector<double> wp;
    int64_t n = orc_link_waypoints(sc, a, b, nullptr, 0);
    if (n != nscores) return 0xffffffffu;
    wp.resize((size_t)n * d);
    orc_link_waypoints(sc, a, b, wp.data(), n);
    // bisection depth of each pre-order ordinal: a subtree of 2^k - 1 nodes
    // is its root, then the left and the right subtree of 2^(k-1) - 1 each
    std::vector<int> depth((size_t)n, 0);
    {
        std::vector<std::pair<int64_t, int64_t>> stack{{0, n}};  // (first ordinal, size)
        std::vector<int> sd{0};
        while (!stack.empty()) {
            auto [o, sz] = stack.back();
            int d = sd.back();
            stack.pop_back();
            sd.pop_back();
            if (sz <= 0) continue;
            depth[(size_t)o] = d;
            const int64_t half = (sz - 1) / 2;
            stack.push_back({o + 1 + half, half});
            sd.push_back(d + 1);
            stack.push_back({o + 1, half});
            sd.push_back(d + 1);
        }
    }
    // the first `lead` ordinals keep the reference order; the other top
    // waypoints follow by descending score (ties: ascending ordinal); the
    // deeper ones keep the reference order
    const int64_t l = std::min<int64_t>(std::max<int32_t>(lead, 0), n);
    std::vector<int64_t> order, deep;
    for (int64_t i = 0; i < l; ++i) order.push_back(i);
    const size_t top0 = order.size();
    for (int64_t i = l; i < n; ++i) (top_levels < 0 || depth[(size_t)i] < top_levels ? order : deep).push_back(i);
    std::stable_sort(order.begin() + (ptrdiff_t)top0, order.end(),
                     [&](int64_t x, int64_t y) { return scores[x] > scores[y]; });
    if (picks >= 0 && (size_t)picks < order.size() - top0) {
        // only the first `picks` of the scored waypoints lead; every other
        // waypoint follows in reference order
        deep.insert(deep.end(), order.begin() + (ptrdiff_t)(top0 + (size_t)picks), order.end());
        order.resize(top0 + (size_t)picks);
        std::sort(deep.begin(), deep.end());
    }
    order.insert(order.end(), deep.begin(), deep.end());
    for (int64_t o : order) {
        ++checks;
        if (!valid(sc, wp.data() + o * d)) return checks;
    }
    return checks;
}

}  // extern "C"
