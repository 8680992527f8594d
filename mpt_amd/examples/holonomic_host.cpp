// The reference's holonomic 2-D point demo (demo/holonomic_2d_point_planning.cpp:
// 1024 x 512 canvas, three circles, three rectangles, start (30, 30), goal
// (994, 482)) with its Scenario written as ordinary C++: valid() and link()
// are member functions over circles and rectangles
// (demo/holonomic_2d_point_scenario.hpp:95-113, demo/shape_hierarchy.hpp:
// 168-271), and there is no gpu_scene().  mpt_amd::Planner then keeps
// sampling, nearest neighbours, steering and the tree on the GPU and calls
// these functions for each iteration's batch on a host thread pool.
//
//   holonomic_host.bin MODE SEED TARGET RANGE GOAL_BIAS OUT [GOAL_RADIUS]
//     MODE: prrt1 (PRRT, batch 1), prrt256 (PRRT, batch 256),
//           star64 (PRRT*, batch 64, k = 20),
//           sampled256 (PRRT, batch 256, with the Scenario's own sampler:
//           the left half of the canvas, one draw in ten without a sample)
//   writes OUT: int64 n, n x 2 doubles (states), n int32 (parents), for star64
//   n doubles (path costs), then the solution: int64 m, m x 2 doubles
//   (waypoints, start first) and int64 t (edges the trajectory callback
//   saw); prints the planner's counters.
#include <array>
#include <cmath>
#include <optional>
#include <random>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <variant>
#include <vector>

#include "mpt_amd/planner.hpp"

namespace {

struct Circle {
    double cx, cy, r;
    // Circle::pointIsValid
    bool pointIsValid(const std::array<double, 2>& p) const {
        const double dx = p[0] - cx, dy = p[1] - cy;
        return dx * dx + dy * dy > r * r;
    }
    // Circle::segmentIsValid: squared distance from the centre to the segment
    bool segmentIsValid(const std::array<double, 2>& a, const std::array<double, 2>& b) const {
        const double vx = b[0] - a[0], vy = b[1] - a[1];
        const double wx = cx - a[0], wy = cy - a[1];
        const double c1 = vx * wx + vy * wy;
        double d2;
        if (c1 <= 0) {
            d2 = wx * wx + wy * wy;
        } else {
            const double c2 = vx * vx + vy * vy;
            if (c2 <= c1) {
                const double ex = cx - b[0], ey = cy - b[1];
                d2 = ex * ex + ey * ey;
            } else {
                const double t = c1 / c2;
                const double ex = (a[0] - cx) + vx * t, ey = (a[1] - cy) + vy * t;
                d2 = ex * ex + ey * ey;
            }
        }
        return d2 > r * r;
    }
};

struct Rect {
    double x0, y0, x1, y1;
    bool pointIsValid(const std::array<double, 2>& p) const {
        return !(p[0] >= x0 && p[0] <= x1 && p[1] >= y0 && p[1] <= y1);
    }
    // Rect::bisectSegment: midpoints until the squared length is below 1
    bool bisect(const std::array<double, 2>& a, const std::array<double, 2>& b) const {
        const std::array<double, 2> mid = {(a[0] + b[0]) / 2, (a[1] + b[1]) / 2};
        const double dx = b[0] - a[0], dy = b[1] - a[1];
        if (dx * dx + dy * dy < 1.0) return true;
        if (!pointIsValid(mid)) return false;
        if (!bisect(a, mid)) return false;
        return bisect(mid, b);
    }
    bool segmentIsValid(const std::array<double, 2>& a, const std::array<double, 2>& b) const {
        if (!pointIsValid(a) || !pointIsValid(b)) return false;
        return bisect(a, b);
    }
};

class Holonomic2DPoint {
   public:
    static constexpr int kDimensions = 2;
    static constexpr int kMetric = MPTG_METRIC_L2;
    using State = std::array<double, 2>;

    Holonomic2DPoint(int width, int height, std::vector<Circle> circles, std::vector<Rect> rects, const State& goal,
                     double goalRadius = 1e-6)
        : w_(width), h_(height), circles_(std::move(circles)), rects_(std::move(rects)), goal_(goal),
          goalRadius_(goalRadius) {}

    bool valid(const State& q) const {
        for (const Circle& c : circles_)
            if (!c.pointIsValid(q)) return false;
        for (const Rect& r : rects_)
            if (!r.pointIsValid(q)) return false;
        return true;
    }
    bool link(const State& a, const State& b) const {
        for (const Circle& c : circles_)
            if (!c.segmentIsValid(a, b)) return false;
        for (const Rect& r : rects_)
            if (!r.segmentIsValid(a, b)) return false;
        return true;
    }

    void bounds_to(double* lo, double* hi) const {
        lo[0] = 0;
        lo[1] = 0;
        hi[0] = w_;
        hi[1] = h_;
    }
    const State& goal_state() const { return goal_; }
    double goal_radius() const { return goalRadius_; }
    double distance(const double* a, const double* b) const {
        const double dx = a[0] - b[0], dy = a[1] - b[1];
        return std::sqrt(dx * dx + dy * dy);
    }

   private:
    int w_, h_;
    std::vector<Circle> circles_;
    std::vector<Rect> rects_;
    State goal_;
    double goalRadius_;
};

// The same Scenario with a sampler of its own (scenario_sampler.hpp:145-167):
// uniform over the left half of the canvas, and no sample (std::nullopt) for
// one draw in ten.
class Holonomic2DPointSampled : public Holonomic2DPoint {
   public:
    using Holonomic2DPoint::Holonomic2DPoint;
    template <typename RNG>
    std::optional<State> sample(RNG& rng) const {
        std::uniform_real_distribution<double> u01(0.0, 1.0);
        if (u01(rng) < 0.1) return std::nullopt;
        return State{u01(rng) * 512.0, u01(rng) * 512.0};
    }
};

template <typename Algorithm, typename Scenario = Holonomic2DPoint>
int run(const Scenario& scenario, uint64_t seed, size_t target, double range, double bias, const char* out,
        bool costs) {
    mpt_amd::Planner<Scenario, Algorithm> planner(scenario, seed);
    static_assert(decltype(planner)::kHostScene, "the Scenario has no gpu_scene()");
    planner.setRange(range);
    planner.setGoalBias(bias);
    planner.addStart({30.0, 30.0});
    planner.solveUntilSize(target);
    const mptg_prrt_stats st = planner.stats();
    std::printf("size=%zu iterations=%lld samples=%lld zero=%lld link_calls=%lld link_ok=%lld invalid=%lld goals=%lld "
                "rewires=%lld solved=%d\n",
                planner.size(), (long long)st.iterations, (long long)st.samples, (long long)st.zero_dist,
                (long long)st.link_calls, (long long)st.link_ok, (long long)st.invalid_states, (long long)st.goals,
                (long long)st.rewires, planner.solved() ? 1 : 0);
    const int64_t n = mptg_tree_size(planner.context());
    std::vector<double> states((size_t)n * 2), cost((size_t)n);
    std::vector<int32_t> parents((size_t)n);
    if (mptg_tree_read(planner.context(), 0, n, states.data(), parents.data()) != MPTG_OK) return 3;
    if (costs && mptg_prrt_costs(planner.context(), 0, n, cost.data()) != MPTG_OK) return 3;
    FILE* f = std::fopen(out, "wb");
    if (!f) return 4;
    std::fwrite(&n, sizeof(n), 1, f);
    std::fwrite(states.data(), sizeof(double), states.size(), f);
    std::fwrite(parents.data(), sizeof(int32_t), parents.size(), f);
    if (costs) std::fwrite(cost.data(), sizeof(double), cost.size(), f);
    // the solution, as waypoints and through both callback shapes
    const std::vector<Holonomic2DPoint::State> path = planner.solution();
    int64_t edges = 0, waypoints = 0;
    bool same = true;
    planner.solution([&](const Holonomic2DPoint::State& a, const std::monostate&, const Holonomic2DPoint::State& b,
                         bool) {
        same = same && a == path[(size_t)edges] && b == path[(size_t)edges + 1];
        ++edges;
    });
    planner.solution([&](const Holonomic2DPoint::State& q) { same = same && q == path[(size_t)waypoints++]; });
    const int64_t np = (int64_t)path.size();
    std::fwrite(&np, sizeof(np), 1, f);
    for (const auto& q : path) std::fwrite(q.data(), sizeof(double), 2, f);
    std::fwrite(&edges, sizeof(edges), 1, f);
    std::fclose(f);
    std::printf("solution waypoints=%lld callback_edges=%lld callbacks_agree=%d\n", (long long)np, (long long)edges,
                same && waypoints == np && edges == (np > 0 ? np - 1 : 0) ? 1 : 0);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 7 && argc != 8) {
        std::fprintf(stderr, "usage: %s prrt1|prrt256|star64 SEED TARGET RANGE GOAL_BIAS OUT [GOAL_RADIUS]\n",
                     argv[0]);
        return 1;
    }
    const std::string mode = argv[1];
    const uint64_t seed = std::strtoull(argv[2], nullptr, 10);
    const size_t target = std::strtoull(argv[3], nullptr, 10);
    const double range = std::strtod(argv[4], nullptr), bias = std::strtod(argv[5], nullptr);
    std::vector<Circle> circles = {{170.0, 140.0, 80.0}, {800.0, 70.0, 50.0}, {900.0, 380.0, 70.0}};
    std::vector<Rect> rects = {{375, 140, 520, 220}, {200, 320, 390, 390}, {600, 200, 680, 450}};
    // the demo's goal radius is 1e-6 (GoalState(1e-6, goal)); a larger one finds many goals
    const double goalRadius = argc == 8 ? std::strtod(argv[7], nullptr) : 1e-6;
    const Holonomic2DPoint scenario(1024, 512, circles, rects, {1024 - 30, 512 - 30}, goalRadius);
    try {
        if (mode == "prrt1")
            return run<mpt_amd::PRRT<mpt_amd::gpu_batch<1>, mpt_amd::max_threads<4>>>(scenario, seed, target, range,
                                                                                      bias, argv[6], false);
        if (mode == "prrt256")
            return run<mpt_amd::PRRT<mpt_amd::gpu_batch<256>>>(scenario, seed, target, range, bias, argv[6], false);
        if (mode == "star64")
            return run<mpt_amd::PRRTStar<mpt_amd::gpu_batch<64>, mpt_amd::rewire_k<20>>>(scenario, seed, target,
                                                                                         range, bias, argv[6], true);
        if (mode == "sampled256") {
            const Holonomic2DPointSampled sampled(1024, 512, circles, rects, {1024 - 30, 512 - 30}, goalRadius);
            return run<mpt_amd::PRRT<mpt_amd::gpu_batch<256>>>(sampled, seed, target, range, bias, argv[6], false);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 2;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 1;
}
