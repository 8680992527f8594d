// SE(3) rigid-body PRRT* (the shape of demo/se3_rigid_body_planning.cpp) on
// the GPU planner: a box-cluster robot among random boxes, driven through
// mpt_amd::Planner<Scenario, PRRTStar<gpu_batch<B>, rewire_k<20>>> until a
// solution is found.  Checks
// that every solution edge passes link() again and that path costs add up.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "mpt_amd/planner.hpp"
#include "mpt_amd/scenarios.hpp"

using Scenario = mpt_amd::RigidBodyMeshes;
using Algorithm = mpt_amd::PRRTStar<mpt_amd::gpu_batch<2048>, mpt_amd::rewire_k<20>, mpt_amd::report_stats<true>>;

static void box(std::vector<float>& out, const float* lo, const float* hi) {
    const float c[8][3] = {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]}, {hi[0], hi[1], lo[2]}, {lo[0], hi[1], lo[2]},
                           {lo[0], lo[1], hi[2]}, {hi[0], lo[1], hi[2]}, {hi[0], hi[1], hi[2]}, {lo[0], hi[1], hi[2]}};
    const int f[12][3] = {{0, 1, 2}, {0, 2, 3}, {4, 6, 5}, {4, 7, 6}, {0, 5, 1}, {0, 4, 5},
                          {3, 2, 6}, {3, 6, 7}, {0, 3, 7}, {0, 7, 4}, {1, 5, 6}, {1, 6, 2}};
    for (const auto& t : f)
        for (int v : t)
            for (int i = 0; i < 3; ++i) out.push_back(c[v][i]);
}

int main(int argc, char** argv) {
    uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<float> u01(0, 1);
    std::vector<float> robot, env;
    for (int i = 0; i < 4; ++i) {  // a small L-shaped robot
        float lo[3] = {-1.0f + i, -0.4f, -0.4f}, hi[3] = {-0.2f + i, 0.4f, 0.4f};
        box(robot, lo, hi);
    }
    for (int i = 0; i < 600; ++i) {  // 600 boxes in [0, 50]^3, kept off the start and goal corners
        float c[3], h[3];
        for (int k = 0; k < 3; ++k) {
            c[k] = 50 * u01(rng);
            h[k] = 0.5f + 1.5f * u01(rng);
        }
        if ((c[0] < 10 && c[1] < 10 && c[2] < 10) || (c[0] > 40 && c[1] > 40 && c[2] > 40)) continue;
        float lo[3] = {c[0] - h[0], c[1] - h[1], c[2] - h[2]}, hi[3] = {c[0] + h[0], c[1] + h[1], c[2] + h[2]};
        box(env, lo, hi);
    }
    Scenario::State start = {0, 0, 0, 1, 4, 4, 4};
    Scenario::State goal = {0, 0, 0.7071067811865476, 0.7071067811865476, 46, 46, 46};
    Scenario scenario(robot, env, goal, {0, 0, 0}, {50, 50, 50});

    mpt_amd::Planner<Scenario, Algorithm> planner(scenario, seed);
    // one goal-biased draw per iteration at most (the reference's worker 0),
    // so a larger bias than the default 0.01 for a batch of 2048
    planner.setGoalBias(0.2);
    planner.addStart(start);
    auto t0 = std::chrono::steady_clock::now();
    // until solved, as demo/se3_rigid_body_planning.cpp:195-198 with a node cap
    planner.solve([&] { return planner.solved() || planner.size() >= 200000; });
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    planner.printStats();
    std::vector<Scenario::State> path = planner.solution();
    const double cost = planner.solutionCost();
    std::printf("solved=%d size=%zu waypoints=%zu cost=%.4f ms=%.2f\n", planner.solved() ? 1 : 0, planner.size(),
                path.size(), cost, ms);
    if (path.empty()) return 2;
    std::vector<double> from, to;
    double sum = 0;
    for (size_t i = 0; i + 1 < path.size(); ++i) {
        for (int j = 0; j < 7; ++j) {
            from.push_back(path[i][j]);
            to.push_back(path[i + 1][j]);
        }
        sum += scenario.distance(path[i].data(), path[i + 1].data());
    }
    std::vector<uint8_t> verdict(path.size() - 1);
    std::vector<uint32_t> checks(path.size() - 1);
    if (mptg_link(planner.context(), from.data(), to.data(), (int64_t)verdict.size(), MPTG_ORDER_REFERENCE, 0,
                  verdict.data(), checks.data()) != MPTG_OK)
        return 3;
    for (uint8_t v : verdict)
        if (!v) return 4;
    std::printf("path_valid=1 cost_sum=%.4f\n", sum);
    return std::fabs(sum - cost) <= 1e-9 * (1 + cost) ? 0 : 5;
}
