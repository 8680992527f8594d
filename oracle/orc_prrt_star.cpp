// CPU oracle: batch-synchronous PRRT* — TEST INFRASTRUCTURE ONLY.
//
// Restates the PRRT* worker of the reference
//   Worker::solve           src/mpt/impl/prrt_star/prrt_star.hpp:447-492
//   Worker::addSample       src/mpt/impl/prrt_star/prrt_star.hpp:510-657
//   nonConcurrentPushUpdate src/mpt/impl/prrt_star/prrt_star.hpp:664-688
//   k-nearest rewiring      src/mpt/impl/rrg_rewire_neighbors.hpp:46-67
// with batch-synchronous semantics.  Iteration `it` draws the B samples
// it*B .. it*B+B-1 (Philox, as the PRRT oracle) and runs four phases:
//
//  1. extend (per sample, against the tree of the iteration start):
//     nearest, d == 0 -> drop, steer to `range` and recompute dNear
//     (:526-536), valid(new) (:539), link(near, new) (:545), goal check.
//  2. choose parent (per survivor, against the same tree and its costs):
//     nbh = k nearest of new (ascending (distance, id)); visit nbh in
//     ascending (cost(nbr) + d, position) order (:559-574); stop once the
//     cost exceeds the running parent cost, at the near node, or at the
//     first nbr whose link(nbr, new) passes; every visited entry is marked
//     checked (:577-605).  The reference's "bad distance" abort (:583-585)
//     is an error here too (tested for the near node wherever it sits in
//     nbh; the distances it compares are bit-identical by construction).
//  3. append survivors in sample order with the chosen parent and cost
//     (:610-622).
//  4. rewire, sequentially over survivors in sample order and over each
//     nbh in distance order (:626-656): skip marked entries; with
//     newCost = cost(new) + d, rewire when newCost < cost(nbr) and
//     link(new, nbr) passes: nbr gets cost newCost and parent new, and
//     every node below nbr has (cost - delta) applied (:664-688).
//     cost(new) is read when its turn comes, so a survivor whose ancestor an
//     earlier survivor rewired uses its lowered cost.
//
// With B = 1 this is the reference's single-worker non-concurrent loop.
// k is either fixed or the reference's ceil(kRRG * ln(n + 1)) with
// kRRG = factor * e * (1 + 1/dim) evaluated at the tree size of the iteration
// start (:559 calls it with nn_.size()).  Sorting ties (equal cost keys) are
// broken by nbh position; the reference's std::sort leaves their order
// unspecified (parity unpinned for exact ties only).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
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

extern "C" int32_t orc_link(const orc_scene* sc, const double* a, const double* b, int32_t known_valid,
                            uint32_t* checks_done);

namespace {

constexpr int kMaxNbh = 64;  // largest k of rewire_k_nearest

struct StarExt {
    bool accepted = false, goal = false, zero = false, invalid = false, linked = false;
    uint32_t checks = 0;
    int32_t near = -1;
    double dnear = 0;
    double state[64];
    // phase 2: the neighbour list (ascending (distance, id); -1 = empty slot)
    std::vector<int32_t> nbr;
    std::vector<double> nd;
    std::vector<uint8_t> marked;
    int32_t parent = -1;
    double pc = 0;
    int64_t cp_links = 0;
    int64_t spec_cp = 0;  // walk entries before its first cost-bound or near-node stop
    bool bad_distance = false;
};

template <typename F>
void parallel_for(int n, unsigned hw, F&& f) {
    if (n >= 64 && hw > 1) {
        std::vector<std::thread> ts;
        for (unsigned t = 0; t < hw; ++t)
            ts.emplace_back([&, t] {
                for (int i = (int)t; i < n; i += (int)hw) f(i);
            });
        for (auto& t : ts) t.join();
    } else {
        for (int i = 0; i < n; ++i) f(i);
    }
}

}  // namespace

extern "C" int32_t orc_star_k(int32_t k_fixed, double factor, int32_t dim, int64_t n) {
    if (k_fixed > 0) return k_fixed;
    const double kRRG = factor * 2.718281828459045 * (1.0 + 1.0 / (double)dim);
    return (int32_t)std::ceil(kRRG * std::log((double)(n + 1)));
}

// The space's dimensions() as the reference's formulas use it: the state
// size for L_p spaces, 3 + 3 for SE(3) (the SO(3) part is a 3-manifold).
static int manifold_dim(const orc_scene* sc) { return sc->metric == ORC_METRIC_SE3 ? 6 : sc->dim; }

// RRGRewireNeighbors<rewire_r_nearest>::rewireRadius
// (src/mpt/impl/rrg_rewire_neighbors.hpp:99-128) with the sampler measures of
// uniform_box_sampler.hpp:76-78 (box: product of extents),
// uniform_sampler_so3.hpp:79-83 (pi^2), uniform_sampler_scaled.hpp:59-65
// (times the weight) and uniform_sampler_cartesian.hpp:93-98 (product over
// the components, left to right).
extern "C" double orc_star_radius(const orc_scene* sc, const double* lo, const double* hi, double factor,
                                  int64_t n) {
    const int d = manifold_dim(sc);
    const double pi = 3.14159265358979323846;
    double mu;
    if (sc->metric == ORC_METRIC_SE3) {
        double box = 1.0;
        for (int i = 4; i < 7; ++i) box = box * (hi[i] - lo[i]);
        mu = (pi * pi * sc->w_so3) * (sc->w_l2 == 1.0 ? box : box * sc->w_l2);
    } else {
        mu = 1.0;
        for (int i = 0; i < sc->dim; ++i) mu = mu * (hi[i] - lo[i]);
    }
    const double invDim = 1 / (double)d;
    const double unitBall = std::pow(std::sqrt(pi), (double)d) / std::tgamma((double)d / 2 + 1);
    const double rRRG = factor * std::pow(2 * (1 + invDim) * mu / unitBall, invDim);
    const double m = (double)(n + 1);
    return rRRG * std::pow(std::log(m) / m, invDim);
}

extern "C" int64_t orc_prrt_star_run(const orc_scene* sc, const orc_prrt_params* p, int32_t k_fixed,
                                     double rewire_factor, const double* start, int64_t target, int64_t max_iters,
                                     double* nodes, int32_t* parents, double* costs, int64_t cap,
                                     int32_t* goal_nodes, int64_t goal_cap, orc_prrt_stats* st,
                                     orc_star_stats* sst, double* goal_cost0, int32_t* goal_parent0) {
    using namespace orc_internal;
    const int d = sc->dim;
    const int B = std::max(1, p->batch);
    orc_prrt_stats s{};
    orc_star_stats ss{};
    if (cap < 1) return -1;
    std::memcpy(nodes, start, sizeof(double) * d);
    parents[0] = -1;
    costs[0] = 0.0;
    std::vector<std::vector<int32_t>> children(1);
    int64_t n = 1;
    int64_t goals = 0;
    std::vector<StarExt> ext(B);
    std::vector<int> surv;
    unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    int table_err = 0;  // set only if a table row is missing (checked per iteration first)
    for (int64_t it = 0; it < max_iters && n < target; ++it) {
        // a custom sampler's table must hold the iteration's rows
        if (p->samples && (it + 1) * (int64_t)B > p->nsamples) return -4;
        const bool bias = goals == 0 && p->goal_bias > 0;
        const int64_t n0 = n;
        // ---- phase 1: extend against the iteration's tree
        parallel_for(B, hw, [&](int b) {
            StarExt& e = ext[b];
            e = StarExt();
            uint64_t sidx = (uint64_t)it * (uint64_t)B + (uint64_t)b;
            double q[64];
            // slot 0 plays worker 0, the only biasing worker (prrt_star.hpp:457-487)
            bool useGoal = bias && b == 0 && mptg::uniform_at(p->seed, sidx, mptg::RNG_PURPOSE_BIAS, 0) < p->goal_bias;
            if (!sample_slot(sc, p, sidx, useGoal, q, &table_err)) return;  // no sample: skipped
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
                interpolate(sc, near, q, p->range / bd, e.state);
                bd = distance(sc, near, e.state);
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
            e.near = best;
            e.dnear = bd;
            e.goal = distance(sc, e.state, p->goal) <= p->goal_radius;
        });
        surv.clear();
        for (int b = 0; b < B; ++b) {
            const StarExt& e = ext[b];
            s.samples++;
            s.nearest_calls++;
            s.zero_dist += e.zero;
            s.invalid_states += e.invalid;
            s.link_calls += e.linked;
            s.valid_calls += e.checks;
            s.link_ok += e.accepted;
            if (e.accepted) surv.push_back(b);
        }
        const int m = (int)surv.size();
        // rewire_k_nearest: the k nearest (k_fixed > 0, or the formula for 0);
        // rewire_r_nearest (k_fixed = -1): every node within the radius
        const bool by_radius = k_fixed == -1;
        const int k = by_radius ? 0 : orc_star_k(k_fixed, rewire_factor, manifold_dim(sc), n0);
        const double radius = by_radius ? orc_star_radius(sc, p->lo, p->hi, rewire_factor, n0) : 0.0;
        if (!by_radius && (k < 1 || k > kMaxNbh)) return -3;
        // ---- phase 2: choose parent against the iteration's tree and costs
        parallel_for(m, hw, [&](int j) {
            StarExt& e = ext[surv[j]];
            if (by_radius) {
                std::vector<std::pair<double, int32_t>> all;
                for (int64_t i = 0; i < n0; ++i) {
                    const double di = distance(sc, e.state, nodes + i * d);
                    if (di <= radius) all.emplace_back(di, (int32_t)i);
                }
                std::sort(all.begin(), all.end());
                e.nbr.resize(all.size());
                e.nd.resize(all.size());
                for (size_t i = 0; i < all.size(); ++i) {
                    e.nd[i] = all[i].first;
                    e.nbr[i] = all[i].second;
                }
            } else {
                e.nbr.assign(k, -1);
                e.nd.assign(k, 0.0);
                orc_knn(sc, nodes, n0, e.state, 1, k, std::numeric_limits<double>::infinity(), e.nbr.data(),
                        e.nd.data(), 1);
            }
            const int nn = (int)e.nbr.size();
            e.marked.assign(nn, 0);
            std::vector<int> ord;
            std::vector<double> key(nn);
            int cnt = 0;
            for (int i = 0; i < nn; ++i)
                if (e.nbr[i] >= 0) {
                    key[i] = costs[e.nbr[i]] + e.nd[i];
                    ord.push_back(i);
                    ++cnt;
                }
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b2) { return key[a] < key[b2]; });
            e.parent = e.near;
            e.pc = costs[e.near] + e.dnear;
            for (int r = 0; r < cnt; ++r)
                if (e.nbr[ord[r]] == e.near && key[ord[r]] != e.pc) {
                    e.bad_distance = true;
                    return;
                }
            for (int r = 0; r < cnt && key[ord[r]] <= e.pc && e.nbr[ord[r]] != e.near; ++r) ++e.spec_cp;
            for (int r = 0; r < cnt; ++r) {
                const int i = ord[r];
                const double newCost = key[i];
                if (newCost > e.pc) break;
                e.marked[i] = 1;
                if (e.nbr[i] == e.near) {
                    e.pc = newCost;
                    break;
                }
                uint32_t checks = 0;
                ++e.cp_links;
                if (orc_link(sc, nodes + (int64_t)e.nbr[i] * d, e.state, 1, &checks)) {
                    e.parent = e.nbr[i];
                    e.pc = newCost;
                    break;
                }
            }
        });
        // ---- phase 3: append in sample order
        if (n + m > cap) return -1;
        for (int j = 0; j < m; ++j) {
            const StarExt& e = ext[surv[j]];
            if (e.bad_distance) return -2;
            std::memcpy(nodes + n * d, e.state, sizeof(double) * d);
            parents[n] = e.parent;
            costs[n] = e.pc;
            children.emplace_back();
            children[e.parent].push_back((int32_t)n);
            ss.cp_links += e.cp_links;
            ss.spec_cp += e.spec_cp;
            ss.parent_changes += e.parent != e.near;
            if (e.goal) {
                // the goal edge as found (foundGoal, prrt_star.hpp:203-216): its cost and parent then
                if (goals < goal_cap) {
                    goal_nodes[goals] = (int32_t)n;
                    if (goal_cost0) goal_cost0[goals] = e.pc;
                    if (goal_parent0) goal_parent0[goals] = e.parent;
                }
                ++goals;
            }
            ++n;
        }
        // ---- phase 4: rewire, survivors in sample order, nbh in distance order
        // (snap: the costs the phase starts from, to count the pairs whose
        // rewire condition only holds after an earlier rewire lowered the new
        // node's own cost)
        std::vector<int32_t> stack;
        const std::vector<double> snap(costs, costs + n);
        for (int j = 0; j < m; ++j) {
            const StarExt& e = ext[surv[j]];
            const int32_t id = (int32_t)(n0 + j);
            for (int i = 0; i < (int)e.nbr.size(); ++i) {
                const int32_t nb = e.nbr[i];
                if (nb < 0) continue;
                ss.rewire_tests++;
                if (e.marked[i]) continue;
                const bool queued = e.pc + e.nd[i] < snap[nb];
                ss.spec_rw += queued;
                const double newCost = costs[id] + e.nd[i];
                if (!(newCost < costs[nb])) continue;
                ss.deferred += !queued;
                uint32_t checks = 0;
                ss.rw_links++;
                if (!orc_link(sc, nodes + (int64_t)id * d, nodes + (int64_t)nb * d, 1, &checks)) continue;
                const double delta = costs[nb] - newCost;
                auto& sib = children[parents[nb]];
                sib.erase(std::find(sib.begin(), sib.end(), nb));
                parents[nb] = id;
                children[id].push_back(nb);
                costs[nb] = newCost;
                ss.rewires++;
                stack.assign(children[nb].begin(), children[nb].end());
                while (!stack.empty()) {
                    int32_t c = stack.back();
                    stack.pop_back();
                    costs[c] = costs[c] - delta;
                    stack.insert(stack.end(), children[c].begin(), children[c].end());
                }
            }
        }
        s.iterations++;
    }
    s.goals = goals;
    if (st) *st = s;
    if (sst) *sst = ss;
    return n;
}
