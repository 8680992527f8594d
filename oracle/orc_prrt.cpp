// CPU oracle: batch-synchronous PRRT — TEST INFRASTRUCTURE ONLY.
//
// Restates the PRRT worker loop of the reference
//   Worker::solve           src/mpt/impl/prrt/prrt.hpp:358-399
//   Worker::addSample       src/mpt/impl/prrt/prrt.hpp:411-452
//   foundGoal               src/mpt/impl/prrt/prrt.hpp:139-146
//   GoalState::operator()   src/mpt/goal_state.hpp:64-69
// with batch-synchronous semantics: in iteration `it` the B samples with
// global indices it*B .. it*B+B-1 are each extended against the tree as it
// stood at the start of the iteration; accepted nodes are appended in sample
// order.  Only sample slot 0 of an iteration takes the goal-bias coin, as
// only the reference's worker 0 biases (prrt.hpp:365-388); it does so while no
// goal has been found at the start of the iteration (worker 0 re-checks
// goalCount_ each iteration, prrt.hpp:379) and with the unscaled bias
// (prrt.hpp:381; the scaled value at :373 is never used).  With B = 1 this is exactly the reference's
// single-worker sequential loop, with the project's Philox stream in place of
// mt19937_64:
//   coin   = uniform(seed, sample, BIAS, 0) < goalBias  -> sample = goal
//   coords = uniform(seed, sample, SAMPLE, i) mapped to [lo_i, hi_i)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "../mpt_amd/csrc/hd.hpp"
#include "../mpt_amd/csrc/sample_hd.hpp"
#include "oracle.h"

namespace orc_internal {
bool sample_slot(const orc_scene* sc, const orc_prrt_params* p, uint64_t sidx, bool use_goal, double* q, int* err);
double distance(const orc_scene* sc, const double* a, const double* b);
void interpolate(const orc_scene* sc, const double* a, const double* b, double t, double* q);
bool valid(const orc_scene* sc, const double* q);
}  // namespace orc_internal

namespace {

struct Ext {
    bool accepted = false;
    bool goal = false;
    int32_t parent = -1;
    bool zero = false, invalid = false, linked = false, skipped = false;
    uint32_t checks = 0;
    double state[64];
};

}  // namespace

extern "C" int32_t orc_link(const orc_scene* sc, const double* a, const double* b, int32_t known_valid,
                            uint32_t* checks_done);

extern "C" int64_t orc_prrt_run(const orc_scene* sc, const orc_prrt_params* p, const double* start,
                                int64_t target, int64_t max_iters, double* nodes, int32_t* parents,
                                int64_t cap, int32_t* goal_nodes, int64_t goal_cap, orc_prrt_stats* st) {
    using namespace orc_internal;
    const int d = sc->dim;
    const int B = std::max(1, p->batch);
    orc_prrt_stats s{};
    if (cap < 1) return -1;
    std::memcpy(nodes, start, sizeof(double) * d);
    parents[0] = -1;
    int64_t n = 1;
    int64_t goals = 0;
    std::vector<Ext> ext(B);
    unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    int table_err = 0;  // set only if a table row is missing (checked per iteration first)
    for (int64_t it = 0; it < max_iters && n < target; ++it) {
        // a custom sampler's table must hold the iteration's rows
        if (p->samples && (it + 1) * (int64_t)B > p->nsamples) return -4;
        const bool bias = goals == 0 && p->goal_bias > 0;
        const int64_t n0 = n;
        auto work = [&](int b) {
            Ext& e = ext[b];
            e = Ext();
            uint64_t sidx = (uint64_t)it * (uint64_t)B + (uint64_t)b;
            double q[64];
            // slot 0 of the iteration plays the reference's worker 0, the only
            // worker that biases (prrt.hpp:365-388)
            bool useGoal = bias && b == 0 && mptg::uniform_at(p->seed, sidx, mptg::RNG_PURPOSE_BIAS, 0) < p->goal_bias;
            if (!sample_slot(sc, p, sidx, useGoal, q, &table_err)) {
                e.skipped = true;
                return;
            }
            // nearest: ascending (distance, index) over the iteration's tree
            int32_t best = -1;
            double bd = std::numeric_limits<double>::infinity();
            for (int64_t j = 0; j < n0; ++j) {
                double dj = distance(sc, q, nodes + j * d);
                if (dj < bd) {
                    bd = dj;
                    best = (int32_t)j;
                }
            }
            if (bd == 0) {
                e.zero = true;
                return;
            }
            const double* near = nodes + (int64_t)best * d;
            if (bd > p->range) {
                double t = p->range / bd;
                interpolate(sc, near, q, t, e.state);
            } else {
                std::memcpy(e.state, q, sizeof(double) * d);
            }
            if (!valid(sc, e.state)) {
                e.invalid = true;
                return;
            }
            e.linked = true;
            uint32_t checks = 0;
            bool ok = orc_link(sc, near, e.state, 1, &checks) != 0;
            e.checks = checks;
            if (!ok) return;
            e.accepted = true;
            e.parent = best;
            e.goal = distance(sc, e.state, p->goal) <= p->goal_radius;
        };
        if (B >= 64 && hw > 1) {
            std::vector<std::thread> ts;
            for (unsigned t = 0; t < hw; ++t)
                ts.emplace_back([&, t] {
                    for (int b = (int)t; b < B; b += (int)hw) work(b);
                });
            for (auto& t : ts) t.join();
        } else {
            for (int b = 0; b < B; ++b) work(b);
        }
        s.iterations++;
        for (int b = 0; b < B; ++b) {
            const Ext& e = ext[b];
            s.samples++;
            s.nearest_calls += !e.skipped;
            s.zero_dist += e.zero;
            s.invalid_states += e.invalid;
            if (e.linked) {
                s.link_calls++;
                s.valid_calls += e.checks;
            }
            if (!e.accepted) continue;
            s.link_ok++;
            if (n >= cap) return -1;
            std::memcpy(nodes + n * d, e.state, sizeof(double) * d);
            parents[n] = e.parent;
            if (e.goal) {
                if (goals < goal_cap) goal_nodes[goals] = (int32_t)n;
                ++goals;
            }
            ++n;
        }
    }
    s.goals = goals;
    if (st) *st = s;
    return n;
}
