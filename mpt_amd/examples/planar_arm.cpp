// The reference's planar-arm demo (demo/link_manipulator_planning.cpp:57-96)
// on the GPU planner: same arm, disks, start and goal, driven through
// mpt_amd::Planner<Scenario, PRRT<gpu_batch<B>>>.  Prints the solution and
// re-validates every solution edge through the C ABI; then grows the same
// tree with iterations captured in hipGraphs and checks that it is the same.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "mpt_amd/planner.hpp"
#include "mpt_amd/scenarios.hpp"

using Scenario = mpt_amd::LinkManipulator<5>;
using Algorithm = mpt_amd::PRRT<mpt_amd::gpu_batch<1024>, mpt_amd::report_stats<true>>;

int main(int argc, char** argv) {
    std::vector<double> lengths = {10.0, 12.0, 8.0, 6.0, 4.0};
    std::vector<mpt_amd::Disk> disks = {{20, -20, 8}, {-20, -30, 5}, {0, 25, 10}, {30, 10, 10}, {-30, 10, 8}};
    const double pi = M_PI;
    Scenario::State start = {-pi * 5 / 6, 0, 0, 0, 0};
    Scenario::State goal = {pi - 1e-10, 0, 0, 0, 0};
    Scenario scenario(goal, disks, lengths, 0.5);
    uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;

    mpt_amd::Planner<Scenario, Algorithm> planner(scenario, seed);
    planner.addStart(start);
    auto t0 = std::chrono::steady_clock::now();
    planner.solve([&] { return planner.solved() || planner.size() > 200000; });
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    planner.printStats();
    std::vector<Scenario::State> path = planner.solution();
    std::printf("solved=%d size=%zu waypoints=%zu ms=%.2f\n", planner.solved() ? 1 : 0, planner.size(),
                path.size(), ms);
    if (path.empty()) return 2;
    // every solution edge must pass link() again
    std::vector<double> from, to;
    for (size_t i = 0; i + 1 < path.size(); ++i)
        for (int j = 0; j < 5; ++j) {
            from.push_back(path[i][j]);
            to.push_back(path[i + 1][j]);
        }
    std::vector<uint8_t> verdict(path.size() - 1);
    std::vector<uint32_t> checks(path.size() - 1);
    if (mptg_link(planner.context(), from.data(), to.data(), (int64_t)verdict.size(), MPTG_ORDER_REFERENCE, 0,
                  verdict.data(), checks.data()) != MPTG_OK)
        return 3;
    for (uint8_t v : verdict)
        if (!v) return 4;
    const bool at_start = path.front() == start;
    double gd = 0;
    for (int j = 0; j < 5; ++j) gd += std::fabs(path.back()[j] - goal[j]);
    std::printf("path_valid=1 starts_at_start=%d goal_distance=%.3g\n", at_start ? 1 : 0, gd);
    if (!(at_start && gd <= 1e-6)) return 5;

    // the same planner with 4 iterations per graph launch, grown to the same
    // size: same iterations, same tree, same solution
    using Graphed = mpt_amd::PRRT<mpt_amd::gpu_batch<1024>, mpt_amd::graph_iterations<4>>;
    mpt_amd::Planner<Scenario, Graphed> graphed(scenario, seed);
    graphed.addStart(start);
    graphed.solveUntilSize(planner.size());
    const mptg_prrt_stats gs = graphed.stats();
    const bool same = graphed.size() == planner.size() && graphed.solution() == path &&
                      gs.iterations == planner.stats().iterations;
    std::printf("graph_equal=%d graph_launches=%lld\n", same ? 1 : 0, (long long)gs.graph_launches);
    return same && gs.graph_launches > 0 ? 0 : 6;
}
