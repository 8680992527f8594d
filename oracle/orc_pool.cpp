// CPU baseline: the reference's shared-memory worker pool, restated —
// TEST/BENCHMARK INFRASTRUCTURE ONLY (timed by the benchmark's CPU leg).
//
// Follows
//   WorkerPool<...>::solve (std::thread backend)  src/mpt/impl/worker_pool_std_thread.hpp:89-120
//   ThreadPool sizing (OMP_NUM_THREADS / hardware) src/mpt/impl/thread_pool.hpp:117-130
//   PRRT Worker::solve / addSample                 src/mpt/impl/prrt/prrt.hpp:358-452
// Each worker owns an mt19937_64 and draws states with
// std::uniform_real_distribution per coordinate (src/mpt/uniform_box_sampler.hpp:60-68);
// worker 0 alone does goal biasing until a goal is known (prrt.hpp:365-388).
// All workers share one tree.  The reference's nigh KDTreeBatch is absent, so
// the shared nearest-neighbour structure here is a stand-in: a balanced
// KD-tree over a published prefix of the nodes (rebuilt when the unindexed
// tail grows past 1/8 of the indexed part) plus a linear scan of the tail.
// Readers never block; appends take a short mutex.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "oracle.h"

namespace orc_internal {
double distance(const orc_scene* sc, const double* a, const double* b);
void interpolate(const orc_scene* sc, const double* a, const double* b, double t, double* q);
bool valid(const orc_scene* sc, const double* q);
}  // namespace orc_internal

extern "C" int32_t orc_link(const orc_scene* sc, const double* a, const double* b, int32_t known_valid,
                            uint32_t* checks_done);

namespace {

using namespace orc_internal;

struct KdIndex {
    int dim = 0;
    int64_t count = 0;               // nodes [0, count) are indexed
    std::vector<int32_t> perm;       // node ids in tree order
    struct Cell {
        int32_t lo, hi;              // range in perm
        int32_t left, right;         // children (-1 for leaf)
        int32_t axis;
        double split;
    };
    std::vector<Cell> cells;

    int32_t build(const double* pts, int32_t lo, int32_t hi) {
        int32_t id = (int32_t)cells.size();
        cells.push_back({lo, hi, -1, -1, 0, 0.0});
        if (hi - lo <= 8) return id;
        int best = 0;
        double bw = -1;
        for (int a = 0; a < dim; ++a) {
            double mn = INFINITY, mx = -INFINITY;
            for (int32_t i = lo; i < hi; ++i) {
                double v = pts[(int64_t)perm[i] * dim + a];
                mn = std::min(mn, v);
                mx = std::max(mx, v);
            }
            if (mx - mn > bw) {
                bw = mx - mn;
                best = a;
            }
        }
        int32_t mid = lo + (hi - lo) / 2;
        std::nth_element(perm.begin() + lo, perm.begin() + mid, perm.begin() + hi, [&](int32_t x, int32_t y) {
            return pts[(int64_t)x * dim + best] < pts[(int64_t)y * dim + best];
        });
        double split = pts[(int64_t)perm[mid] * dim + best];
        int32_t l = build(pts, lo, mid);
        int32_t r = build(pts, mid, hi);
        cells[id].left = l;
        cells[id].right = r;
        cells[id].axis = best;
        cells[id].split = split;
        return id;
    }

    // exact nearest by (distance, id); the coordinate gap lower-bounds the
    // L1 and L2 distances
    void query(const orc_scene* sc, const double* pts, const double* q, int32_t c, double& bd, int32_t& bi) const {
        const Cell& cell = cells[c];
        if (cell.left < 0) {
            for (int32_t i = cell.lo; i < cell.hi; ++i) {
                int32_t id = perm[i];
                double d = distance(sc, q, pts + (int64_t)id * dim);
                if (d < bd || (d == bd && id < bi)) {
                    bd = d;
                    bi = id;
                }
            }
            return;
        }
        double diff = q[cell.axis] - cell.split;
        int32_t first = diff < 0 ? cell.left : cell.right;
        int32_t second = diff < 0 ? cell.right : cell.left;
        query(sc, pts, q, first, bd, bi);
        if (std::fabs(diff) <= bd) query(sc, pts, q, second, bd, bi);
    }
};

struct SharedTree {
    const orc_scene* sc;
    int dim;
    std::vector<double> pts;        // capacity rows, written once
    std::vector<int32_t> parents;
    std::atomic<int64_t> published{0};
    std::mutex append_mu, build_mu;
    std::shared_ptr<const KdIndex> index;
    bool use_kd;

    SharedTree(const orc_scene* s, int64_t cap) : sc(s), dim(s->dim), pts((size_t)cap * s->dim), parents(cap) {
        use_kd = s->metric != ORC_METRIC_SE3;
        index = std::make_shared<KdIndex>();
    }

    int64_t append(const double* q, int32_t parent) {
        std::lock_guard<std::mutex> lk(append_mu);
        int64_t n = published.load(std::memory_order_relaxed);
        if ((size_t)(n + 1) * dim > pts.size()) return -1;
        std::memcpy(&pts[(size_t)n * dim], q, sizeof(double) * dim);
        parents[n] = parent;
        published.store(n + 1, std::memory_order_release);
        return n;
    }

    void maybe_rebuild(int64_t n) {
        std::shared_ptr<const KdIndex> cur = std::atomic_load(&index);
        if (!use_kd || n - cur->count <= std::max<int64_t>(64, cur->count / 8)) return;
        std::unique_lock<std::mutex> lk(build_mu, std::try_to_lock);
        if (!lk.owns_lock()) return;
        auto nx = std::make_shared<KdIndex>();
        nx->dim = dim;
        nx->count = n;
        nx->perm.resize(n);
        for (int64_t i = 0; i < n; ++i) nx->perm[i] = (int32_t)i;
        nx->cells.reserve((size_t)(n / 4 + 16));
        nx->build(pts.data(), 0, (int32_t)n);
        std::atomic_store(&index, std::shared_ptr<const KdIndex>(nx));
    }

    std::pair<int32_t, double> nearest(const double* q) {
        int64_t n = published.load(std::memory_order_acquire);
        std::shared_ptr<const KdIndex> idx = std::atomic_load(&index);
        double bd = INFINITY;
        int32_t bi = -1;
        int64_t start = 0;
        if (use_kd && idx->count > 0) {
            idx->query(sc, pts.data(), q, 0, bd, bi);
            start = idx->count;
        }
        for (int64_t i = start; i < n; ++i) {
            double d = distance(sc, q, &pts[(size_t)i * dim]);
            if (d < bd || (d == bd && i < bi)) {
                bd = d;
                bi = (int32_t)i;
            }
        }
        maybe_rebuild(n);
        return {bi, bd};
    }
};

}  // namespace

// Runs the pool from a tree of n0 nodes (parents[i] < i, parents[0] = -1)
// until the tree holds `target` nodes or max_seconds have passed.  The
// nearest-neighbour index over the given nodes is built before the clock
// starts.
extern "C" int32_t orc_pool_run_tree(const orc_scene* sc, const orc_prrt_params* p, const double* nodes,
                                     const int32_t* parents, int64_t n0, int64_t target, double max_seconds,
                                     int32_t nthreads, orc_pool_result* out) {
    if (nthreads < 1) nthreads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    if (n0 < 1 || target < n0) return -1;
    const int d = sc->dim;
    SharedTree tree(sc, target + (int64_t)nthreads + 16);
    for (int64_t i = 0; i < n0; ++i) tree.append(nodes + i * d, parents ? parents[i] : (i ? 0 : -1));
    tree.maybe_rebuild(n0);
    std::atomic<bool> done{false};
    std::atomic<int64_t> samples{0}, links{0}, valids{0}, goals{0};
    auto t0 = std::chrono::steady_clock::now();
    auto is_done = [&]() {
        if (done.load(std::memory_order_relaxed)) return true;
        if (tree.published.load(std::memory_order_relaxed) >= target ||
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= max_seconds) {
            done.store(true, std::memory_order_relaxed);
            return true;
        }
        return false;
    };
    auto worker = [&](int no) {
        std::seed_seq seq{(uint32_t)p->seed, (uint32_t)(p->seed >> 32), (uint32_t)no};
        std::mt19937_64 rng(seq);
        std::uniform_real_distribution<double> uniform01;
        std::vector<double> q(d), s(d);
        int64_t my_samples = 0, my_links = 0, my_valids = 0;
        while (!is_done()) {
            ++my_samples;
            bool useGoal = false;
            if (no == 0 && p->goal_bias > 0 && goals.load(std::memory_order_relaxed) == 0)
                useGoal = uniform01(rng) < p->goal_bias;
            if (useGoal) {
                std::memcpy(q.data(), p->goal, sizeof(double) * d);
            } else {
                for (int i = 0; i < d; ++i) {
                    std::uniform_real_distribution<double> u(p->lo[i], p->hi[i]);
                    q[i] = u(rng);
                }
            }
            auto [near, dist] = tree.nearest(q.data());
            if (near < 0 || dist == 0) continue;
            const double* a = &tree.pts[(size_t)near * d];
            if (dist > p->range) interpolate(sc, a, q.data(), p->range / dist, s.data());
            else s = q;
            if (!valid(sc, s.data())) continue;
            uint32_t checks = 0;
            ++my_links;
            bool ok = orc_link(sc, a, s.data(), 1, &checks) != 0;
            my_valids += checks;
            if (!ok) continue;
            bool isGoal = distance(sc, s.data(), p->goal) <= p->goal_radius;
            if (tree.append(s.data(), near) < 0) break;
            if (isGoal) goals.fetch_add(1);
        }
        samples.fetch_add(my_samples);
        links.fetch_add(my_links);
        valids.fetch_add(my_valids);
    };
    std::vector<std::thread> ts;
    for (int i = 1; i < nthreads; ++i) ts.emplace_back(worker, i);
    worker(0);
    for (auto& t : ts) t.join();
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (out) {
        out->seconds = secs;
        out->size = tree.published.load();
        out->samples = samples.load();
        out->link_calls = links.load();
        out->valid_calls = valids.load();
        out->goals = goals.load();
    }
    return 0;
}

extern "C" int32_t orc_pool_run(const orc_scene* sc, const orc_prrt_params* p, const double* start, int64_t target,
                                double max_seconds, int32_t nthreads, orc_pool_result* out) {
    return orc_pool_run_tree(sc, p, start, nullptr, 1, std::max<int64_t>(target, 1), max_seconds, nthreads, out);
}
