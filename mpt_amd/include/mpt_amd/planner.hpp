// C++ host layer of the MI355X batched PRRT path: the reference planner's
// template API over the C ABI of include/mptg.h.
//
// The reference resolves planners at compile time:
//     unc::robotics::mpt::Planner<Scenario, PRRT<options...>>
// (src/mpt/planner.hpp:46-47, src/mpt/prrt.hpp:56-82).  This header keeps
// that shape:
//     mpt_amd::Planner<Scenario, mpt_amd::PRRT<mpt_amd::gpu_batch<8192>, ...>>
// and the public methods of the reference class
// (src/mpt/impl/prrt/prrt.hpp:148-316): setGoalBias/getGoalBias,
// setRange/getRange, size, addStart, solve(doneFn), solveFor, solveUntil,
// solved, solution(), solution(fn), printStats, visitGraph.
//
// Scenario contract (src/mpt/scenario.hpp:93-130): bounds, goal and
// distance as below, and valid()/link() in one of two forms:
//   * as data for the device:
//         void gpu_scene(mptg_ctx*) const;  // uploads space, obstacles, link form
//   * as the Scenario's own C++ code, when it has no gpu_scene():
//         bool valid(const State&) const;
//         bool link(const State&, const State&) const;
//         static constexpr int kMetric = MPTG_METRIC_L2;  // or _L1, or _SE3 (then kSO3Weight)
//     Sampling, nearest neighbours, steering and the tree stay on the GPU;
//     every batch of valid()/link() calls an iteration needs is run on a
//     host thread pool of max_threads<N> threads (default: the hardware's
//     thread count), as the reference's workers call them (prrt.hpp:439-457).
// Both forms provide State, kDimensions, bounds_to(lo, hi), goal_state(),
// goal_radius() and distance(a, b).  States are any type with operator[].
// A Scenario with a sample(RNG&) method (returning a State or a
// std::optional<State>; scenario_sampler.hpp:145-167) samples in place of the
// uniform box sampler, in either form (mptg_prrt_sampler).
// Errors from the C ABI are rethrown as std::runtime_error (as the reference
// throws at prrt.hpp:198); an exception thrown by the Scenario's valid() or
// link() is rethrown as itself.
#pragma once

#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <mutex>
#include <optional>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

#include "mptg.h"

namespace mpt_amd {

// ---- option tags (src/mpt/planner_tags.hpp:45-64) ---------------------------
template <int B>
struct gpu_batch {
    static_assert(B >= 1, "batch must be positive");
    static constexpr int value = B;
};
template <bool b>
struct report_stats {
    static constexpr bool value = b;
};
// Host threads for a Scenario's own valid()/link() (0: the hardware's
// thread count).  One host thread drives the GPU context either way.
template <int N>
struct max_threads {
    static_assert(N >= 0, "max_threads must be >= 0");
    static constexpr int value = N;
};
// Visiting order of link()'s waypoints: reference, or surrogate-ranked.
template <int O>
struct check_order {
    static constexpr int value = O;
};
// PRRT only: iterations per hipGraph launch; solve() then checks its done
// condition every G iterations instead of after each one.
template <int G>
struct graph_iterations {
    static_assert(G >= 0 && G <= 1024, "graph_iterations must be in 0..1024");
    static constexpr int value = G;
};
// PRRT* rewiring (src/mpt/planner_tags.hpp:49-51).  k-nearest is the
// default, with the reference's k = ceil(kRRG ln(n + 1)) (up to 64);
// rewire_k<K> fixes k (1..64).  r-nearest takes every node within the
// reference's radius (rrg_rewire_neighbors.hpp:94-152), however many.
struct rewire_k_nearest {};
struct rewire_r_nearest {};
template <int K>
struct rewire_k {
    static_assert(K >= 1 && K <= 64, "rewire_k must be in 1..64");
    static constexpr int value = K;
};

namespace detail {
inline void check(mptg_ctx* ctx, int rc) {
    if (rc != MPTG_OK) throw std::runtime_error(std::string("mptg: ") + mptg_last_error(ctx));
}

// Fork-join pool for a host Scenario's valid()/link() batches: the calling
// thread and n - 1 workers take indices from a shared counter; run() returns
// when every index is done.  The first exception any task throws is kept.
class HostPool {
   public:
    explicit HostPool(int n) {
        if (n <= 0) n = static_cast<int>(std::thread::hardware_concurrency());
        if (n <= 0) n = 1;
        for (int i = 1; i < n; ++i) workers_.emplace_back([this] { loop(); });
    }
    HostPool(const HostPool&) = delete;
    HostPool& operator=(const HostPool&) = delete;
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    int threads() const { return static_cast<int>(workers_.size()) + 1; }

    // fn(i) for i in [0, n); rethrows the first exception of any call
    void run(int64_t n, const std::function<void(int64_t)>& fn) {
        if (n <= 0) return;
        error_ = nullptr;
        if (workers_.empty() || n < 2) {
            for (int64_t i = 0; i < n; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &fn;
            n_ = n;
            next_.store(0);
            busy_ = static_cast<int>(workers_.size());
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [this] { return busy_ == 0; });
        job_ = nullptr;
        if (error_) std::rethrow_exception(error_);
    }

   private:
    void work() {
        for (int64_t i; (i = next_.fetch_add(1)) < n_;) {
            try {
                (*job_)(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(err_m_);
                if (!error_) error_ = std::current_exception();
                next_.store(n_);  // the rest is not needed
            }
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work();
            {
                std::lock_guard<std::mutex> g(m_);
                if (--busy_ == 0) done_.notify_one();
            }
        }
    }

    std::vector<std::thread> workers_;
    std::mutex m_, err_m_;
    std::condition_variable cv_, done_;
    const std::function<void(int64_t)>* job_ = nullptr;
    int64_t n_ = 0;
    std::atomic<int64_t> next_{0};
    int busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr error_;
};

template <typename S, typename = void>
struct has_gpu_scene : std::false_type {};
template <typename S>
struct has_gpu_scene<S, std::void_t<decltype(std::declval<const S&>().gpu_scene(static_cast<mptg_ctx*>(nullptr)))>>
    : std::true_type {};

// Scenario::sample(RNG&) (scenario_sampler.hpp:145-167): returns a State or
// a std::optional<State>
template <typename S, typename = void>
struct has_sample : std::false_type {};
template <typename S>
struct has_sample<S, std::void_t<decltype(std::declval<S&>().sample(std::declval<std::mt19937_64&>()))>>
    : std::true_type {};
template <typename T>
struct is_optional : std::false_type {};
template <typename T>
struct is_optional<std::optional<T>> : std::true_type {};

// splitmix64: one well-mixed 64-bit seed per (planner seed, sample index)
inline uint64_t sample_seed(uint64_t seed, uint64_t index) {
    uint64_t z = seed ^ (index * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

template <typename S, typename = void>
struct so3_weight {
    static constexpr double value = 1.0;
};
template <typename S>
struct so3_weight<S, std::void_t<decltype(S::kSO3Weight)>> {
    static constexpr double value = S::kSO3Weight;
};
}  // namespace detail

// Strategy type produced by PRRT<options...>; option order is free, as in
// the reference's option parser (src/mpt/impl/packs.hpp:105-125).
template <int Batch, bool Stats, int Order, int Graph = 0, int Threads = 0>
struct PRRTStrategy {
    static constexpr int batch = Batch;
    static constexpr bool stats = Stats;
    static constexpr int order = Order;
    static constexpr int graph = Graph;
    static constexpr int threads = Threads;
};

template <typename... Opts>
struct PRRTOptions;
template <>
struct PRRTOptions<> {
    static constexpr int batch = 8192;
    static constexpr bool stats = false;
    static constexpr int order = MPTG_ORDER_REFERENCE;
    static constexpr int graph = 0;
    static constexpr int threads = 0;
};
template <int B, typename... Rest>
struct PRRTOptions<gpu_batch<B>, Rest...> : PRRTOptions<Rest...> {
    static constexpr int batch = B;
};
template <bool S, typename... Rest>
struct PRRTOptions<report_stats<S>, Rest...> : PRRTOptions<Rest...> {
    static constexpr bool stats = S;
};
template <int N, typename... Rest>
struct PRRTOptions<max_threads<N>, Rest...> : PRRTOptions<Rest...> {
    static constexpr int threads = N;
};
template <int O, typename... Rest>
struct PRRTOptions<check_order<O>, Rest...> : PRRTOptions<Rest...> {
    static constexpr int order = O;
};
template <int G, typename... Rest>
struct PRRTOptions<graph_iterations<G>, Rest...> : PRRTOptions<Rest...> {
    static constexpr int graph = G;
};

template <typename... Opts>
using PRRT = PRRTStrategy<PRRTOptions<Opts...>::batch, PRRTOptions<Opts...>::stats, PRRTOptions<Opts...>::order,
                          PRRTOptions<Opts...>::graph, PRRTOptions<Opts...>::threads>;

// PRRT* (src/mpt/prrt_star.hpp:55-110): the PRRT options plus rewiring.
template <int Batch, bool Stats, int Order, int K, int Neighbors = MPTG_REWIRE_K_NEAREST, int Threads = 0>
struct PRRTStarStrategy {
    static constexpr int batch = Batch;
    static constexpr bool stats = Stats;
    static constexpr int order = Order;
    static constexpr int k = K;  // 0: the reference's k-nearest formula
    static constexpr int neighbors = Neighbors;
    static constexpr int threads = Threads;
};

template <typename... Opts>
struct PRRTStarOptions;
template <>
struct PRRTStarOptions<> : PRRTOptions<> {
    static constexpr int k = 0;
    static constexpr int neighbors = MPTG_REWIRE_K_NEAREST;
};
template <typename Opt, typename... Rest>
struct PRRTStarOptions<Opt, Rest...> : PRRTStarOptions<Rest...> {
    using Base = PRRTStarOptions<Rest...>;
    static constexpr int batch = PRRTOptions<Opt>::batch != PRRTOptions<>::batch ? PRRTOptions<Opt>::batch : Base::batch;
    static constexpr bool stats = PRRTOptions<Opt>::stats || Base::stats;
    static constexpr int order = PRRTOptions<Opt>::order != PRRTOptions<>::order ? PRRTOptions<Opt>::order : Base::order;
    static constexpr int threads =
        PRRTOptions<Opt>::threads != PRRTOptions<>::threads ? PRRTOptions<Opt>::threads : Base::threads;
};
template <typename... Rest>
struct PRRTStarOptions<rewire_k_nearest, Rest...> : PRRTStarOptions<Rest...> {};
template <int G, typename... Rest>
struct PRRTStarOptions<graph_iterations<G>, Rest...> : PRRTStarOptions<Rest...> {
    static_assert(G < 0, "graph_iterations applies to PRRT; PRRT* enqueues its iterations one by one");
};
template <typename... Rest>
struct PRRTStarOptions<rewire_r_nearest, Rest...> : PRRTStarOptions<Rest...> {
    static constexpr int neighbors = MPTG_REWIRE_R_NEAREST;
};
template <int K, typename... Rest>
struct PRRTStarOptions<rewire_k<K>, Rest...> : PRRTStarOptions<Rest...> {
    static constexpr int k = K;
};

template <typename... Opts>
using PRRTStar = PRRTStarStrategy<PRRTStarOptions<Opts...>::batch, PRRTStarOptions<Opts...>::stats,
                                  PRRTStarOptions<Opts...>::order, PRRTStarOptions<Opts...>::k,
                                  PRRTStarOptions<Opts...>::neighbors, PRRTStarOptions<Opts...>::threads>;

template <typename Scenario, typename Strategy>
class GpuPRRT;

template <typename Scenario, typename Algorithm>
struct PlannerResolver;
template <typename Scenario, int B, bool S, int O, int G, int T>
struct PlannerResolver<Scenario, PRRTStrategy<B, S, O, G, T>> {
    using type = GpuPRRT<Scenario, PRRTStrategy<B, S, O, G, T>>;
};
template <typename Scenario, int B, bool S, int O, int K, int N, int T>
struct PlannerResolver<Scenario, PRRTStarStrategy<B, S, O, K, N, T>> {
    using type = GpuPRRT<Scenario, PRRTStarStrategy<B, S, O, K, N, T>>;
};

template <typename Scenario, typename Algorithm>
using Planner = typename PlannerResolver<Scenario, Algorithm>::type;

// ---- the planner ---------------------------------------------------------------
template <typename S>
struct is_star : std::false_type {};
template <int B, bool S, int O, int K, int N, int T>
struct is_star<PRRTStarStrategy<B, S, O, K, N, T>> : std::true_type {};

template <typename Scenario, typename Strategy>
class GpuPRRT {
   public:
    using State = typename Scenario::State;
    static constexpr int kDim = Scenario::kDimensions;
    static constexpr bool kStar = is_star<Strategy>::value;
    // the Scenario's valid()/link() run on the host (it has no gpu_scene())
    static constexpr bool kHostScene = !detail::has_gpu_scene<Scenario>::value;

   private:
    Scenario scenario_;
    mptg_ctx* ctx_ = nullptr;
    std::unique_ptr<detail::HostPool> pool_;
    std::exception_ptr hostError_;
    uint64_t seed_;
    double maxDistance_ = std::numeric_limits<double>::infinity();
    double goalBias_ = 0.01;
    double rewireFactor_ = 1.1;
    bool configured_ = false;

    static void toDoubles(const State& s, double* out) {
        for (int i = 0; i < kDim; ++i) out[i] = static_cast<double>(s[i]);
    }
    static State fromDoubles(const double* in) {
        State s{};
        for (int i = 0; i < kDim; ++i) s[i] = in[i];
        return s;
    }

    // mptg_valid_fn / mptg_link_fn over the Scenario's own valid()/link()
    static int hostValid(void* user, const double* q, int64_t n, uint8_t* ok) {
        auto* self = static_cast<GpuPRRT*>(user);
        try {
            self->pool_->run(n, [&](int64_t i) { ok[i] = self->scenario_.valid(fromDoubles(q + i * kDim)) ? 1 : 0; });
            return 0;
        } catch (...) {
            self->hostError_ = std::current_exception();
            return 1;
        }
    }
    static int hostLink(void* user, const double* a, const double* b, int64_t n, uint8_t* ok) {
        auto* self = static_cast<GpuPRRT*>(user);
        try {
            self->pool_->run(n, [&](int64_t i) {
                ok[i] = self->scenario_.link(fromDoubles(a + i * kDim), fromDoubles(b + i * kDim)) ? 1 : 0;
            });
            return 0;
        } catch (...) {
            self->hostError_ = std::current_exception();
            return 1;
        }
    }

    // mptg_sample_fn over Scenario::sample(rng): one std::mt19937_64 per sample
    // index, seeded from (planner seed, index), so a run is reproducible; the
    // samples are drawn on the calling thread, in index order (the reference
    // gives each worker a Scenario copy of its own, prrt.hpp:326,343)
    static int hostSample(void* user, uint64_t seed, int64_t first, int64_t n, double* q, uint8_t* ok) {
        auto* self = static_cast<GpuPRRT*>(user);
        try {
            if constexpr (detail::has_sample<Scenario>::value) {
                for (int64_t i = 0; i < n; ++i) {
                    std::mt19937_64 rng(detail::sample_seed(seed, (uint64_t)(first + i)));
                    auto r = self->scenario_.sample(rng);
                    if constexpr (detail::is_optional<decltype(r)>::value) {
                        ok[i] = r ? 1 : 0;
                        if (r) toDoubles(*r, q + i * kDim);
                    } else {
                        ok[i] = 1;
                        toDoubles(r, q + i * kDim);
                    }
                }
            }
            return 0;
        } catch (...) {
            self->hostError_ = std::current_exception();
            return 1;
        }
    }

    // a failing entry point rethrows the Scenario's own exception if it caused it
    void check(int rc) {
        if (rc != MPTG_OK && hostError_) {
            std::exception_ptr e = hostError_;
            hostError_ = nullptr;
            std::rethrow_exception(e);
        }
        detail::check(ctx_, rc);
    }

    void uploadScene() {
        if constexpr (kHostScene) {
            static_assert(Scenario::kMetric == MPTG_METRIC_L2 || Scenario::kMetric == MPTG_METRIC_L1 ||
                              Scenario::kMetric == MPTG_METRIC_SE3,
                          "a Scenario without gpu_scene() names its space by kMetric");
            pool_ = std::make_unique<detail::HostPool>(Strategy::threads);
            detail::check(ctx_, mptg_space(ctx_, Scenario::kMetric, kDim, detail::so3_weight<Scenario>::value, 1.0));
            detail::check(ctx_, mptg_scene_host(ctx_, &GpuPRRT::hostValid, &GpuPRRT::hostLink, this));
        } else {
            scenario_.gpu_scene(ctx_);
        }
    }

    void configure() {
        double lo[kDim], hi[kDim], goal[kDim];
        scenario_.bounds_to(lo, hi);
        toDoubles(scenario_.goal_state(), goal);
        mptg_prrt_params p{};
        p.seed = seed_;
        p.range = maxDistance_;
        p.goal_bias = goalBias_;
        p.goal_radius = scenario_.goal_radius();
        p.batch = Strategy::batch;
        p.order = Strategy::order;
        if constexpr (kStar) {
            mptg_prrt_star_params sp{};
            sp.k = Strategy::k;
            sp.neighbors = Strategy::neighbors;
            sp.rewire_factor = rewireFactor_;
            detail::check(ctx_, mptg_prrt_star_init(ctx_, &p, &sp, lo, hi, goal));
        } else {
            detail::check(ctx_, mptg_prrt_init(ctx_, &p, lo, hi, goal));
            detail::check(ctx_, mptg_prrt_graph(ctx_, Strategy::graph));
        }
        configured_ = true;
    }

   public:
    explicit GpuPRRT(const Scenario& scenario = Scenario(), uint64_t seed = 0, int device = 0)
        : scenario_(scenario), seed_(seed) {
        if (mptg_create(device, &ctx_) != MPTG_OK)
            throw std::runtime_error(std::string("mptg: ") + mptg_last_error(nullptr));
        try {
            uploadScene();
            if constexpr (detail::has_sample<Scenario>::value)
                detail::check(ctx_, mptg_prrt_sampler(ctx_, &GpuPRRT::hostSample, this));
            configure();
        } catch (...) {
            mptg_destroy(ctx_);
            throw;
        }
    }
    GpuPRRT(const GpuPRRT&) = delete;
    GpuPRRT& operator=(const GpuPRRT&) = delete;
    ~GpuPRRT() { mptg_destroy(ctx_); }

    mptg_ctx* context() { return ctx_; }

    void setGoalBias(double bias) {
        if (!(0 <= bias && bias <= 1)) throw std::invalid_argument("goal bias must be in [0,1]");
        goalBias_ = bias;
        configure();
    }
    double getGoalBias() const { return goalBias_; }
    void setRange(double range) {
        if (!(range > 0)) throw std::invalid_argument("range must be > 0");
        maxDistance_ = range;
        configure();
    }
    double getRange() const { return maxDistance_; }

    // PRRT* only (prrt_star.hpp:225-231)
    void setRewireFactor(double factor) {
        static_assert(kStar, "setRewireFactor is a PRRT* method");
        if (!(factor > 0)) throw std::invalid_argument("rewire factor must be > 0");
        rewireFactor_ = factor;
        configure();
    }

    std::size_t size() const { return static_cast<std::size_t>(mptg_tree_size(ctx_)); }

    void addStart(const State& q) {
        double d[kDim];
        toDoubles(q, d);
        check(mptg_prrt_add_start(ctx_, d));
    }

    // One batched iteration per doneFn() check (prrt.hpp:194-201); with
    // graph_iterations<G>, one graph launch of G iterations per check.
    template <typename DoneFn>
    std::enable_if_t<std::is_same_v<bool, std::invoke_result_t<DoneFn>>> solve(DoneFn doneFn) {
        if (size() == 0) throw std::runtime_error("there are no valid initial states");
        int64_t per_check = 1;
        if constexpr (!kStar) per_check = Strategy::graph > 0 ? Strategy::graph : 1;
        while (!doneFn()) check(mptg_prrt_run(ctx_, std::numeric_limits<int32_t>::max(), per_check, nullptr));
    }

    template <typename Duration>
    void solveFor(const Duration& d) {
        auto end = std::chrono::steady_clock::now() + d;
        solve([&] { return std::chrono::steady_clock::now() >= end; });
    }

    template <typename Clock, typename Duration>
    void solveUntil(const std::chrono::time_point<Clock, Duration>& t) {
        solve([&] { return Clock::now() >= t; });
    }

    // Grow until the tree holds at least n nodes (the demos' done condition).
    void solveUntilSize(std::size_t n) {
        if (size() == 0) throw std::runtime_error("there are no valid initial states");
        check(mptg_prrt_run(ctx_, static_cast<int64_t>(n), std::numeric_limits<int64_t>::max(), nullptr));
    }

    bool solved() const {
        mptg_prrt_stats st{};
        mptg_prrt_stats_get(ctx_, &st);
        return st.goals > 0;
    }

    mptg_prrt_stats stats() const {
        mptg_prrt_stats st{};
        detail::check(ctx_, mptg_prrt_stats_get(ctx_, &st));
        return st;
    }

   private:
    void readTree(std::vector<double>& states, std::vector<int32_t>& parents) const {
        const int64_t n = mptg_tree_size(ctx_);
        states.resize(static_cast<size_t>(n) * kDim);
        parents.resize(static_cast<size_t>(n));
        if (n) detail::check(ctx_, mptg_tree_read(ctx_, 0, n, states.data(), parents.data()));
    }

    // The solution's goal node (-1 if none) and the parent its path leaves
    // it by.  PRRT (prrt.hpp:208-234): the goal of least path cost, summed
    // from the goal up, the first found on ties.  PRRT* (foundGoal,
    // prrt_star.hpp:203-216): the goal edge found with the least path cost
    // at the time it was found, the first on ties; its path leaves the goal
    // by the parent it had then and follows the current tree above it
    // (solution(), :324-336).
    int64_t bestGoal(const std::vector<double>& states, const std::vector<int32_t>& parents,
                     double* bestCost = nullptr, int64_t* firstParent = nullptr) const {
        int64_t ng = 0;
        detail::check(ctx_, mptg_prrt_goals(ctx_, nullptr, 0, &ng));
        ng = std::min<int64_t>(ng, 4096);
        std::vector<int32_t> goals(static_cast<size_t>(ng));
        if (ng) detail::check(ctx_, mptg_prrt_goals(ctx_, goals.data(), ng, &ng));
        goals.resize(static_cast<size_t>(std::min<int64_t>(ng, 4096)));
        std::vector<double> cost0;
        std::vector<int32_t> parent0;
        if constexpr (kStar) {
            int64_t ne = 0;
            cost0.resize(goals.size());
            parent0.resize(goals.size());
            if (!goals.empty())
                detail::check(ctx_, mptg_prrt_goal_edges(ctx_, cost0.data(), parent0.data(), (int64_t)goals.size(), &ne));
        }
        double best = std::numeric_limits<double>::infinity();
        int64_t bestId = -1, bestParent = -1;
        for (size_t i = 0; i < goals.size(); ++i) {
            const int32_t g = goals[i];
            double cost = 0;
            int64_t par = parents[g];
            if constexpr (kStar) {
                cost = cost0[i];
                par = parent0[i];
            } else {
                for (int64_t n = g; parents[n] >= 0; n = parents[n])
                    cost += scenario_.distance(&states[n * kDim], &states[parents[n] * kDim]);
            }
            if (bestId < 0 || cost < best) {
                best = cost;
                bestId = g;
                bestParent = par;
            }
        }
        if (bestCost) *bestCost = bestId < 0 ? std::numeric_limits<double>::quiet_NaN() : best;
        if (firstParent) *firstParent = bestParent;
        return bestId;
    }

   public:
    // link()'s trajectory type for Scenarios whose link() returns bool
    // (link_trajectory<bool>, src/mpt/impl/link_trajectory.hpp:49-56)
    using Trajectory = std::monostate;

    std::vector<State> solution() const {
        std::vector<double> states;
        std::vector<int32_t> parents;
        readTree(states, parents);
        std::vector<State> path;
        int64_t first = -1;
        const int64_t goal = bestGoal(states, parents, nullptr, &first);
        if (goal >= 0) {
            path.push_back(fromDoubles(&states[goal * kDim]));
            for (int64_t n = first; n >= 0; n = parents[n]) path.push_back(fromDoubles(&states[n * kDim]));
        }
        return std::vector<State>(path.rbegin(), path.rend());
    }

    // cost of the best solution, NaN if none (prrt_star.hpp:336-341)
    double solutionCost() const {
        std::vector<double> states;
        std::vector<int32_t> parents;
        readTree(states, parents);
        double c;
        bestGoal(states, parents, &c);
        return c;
    }

    // The best path through a callback (prrt.hpp:236-285): fn(from, trajectory,
    // to, true) once per edge, start first, when fn takes that shape
    // (is_trajectory_callback, link_trajectory.hpp:76-84); otherwise fn(state)
    // once per waypoint.
    template <typename Fn>
    void solution(Fn fn) const {
        const std::vector<State> path = solution();
        if constexpr (std::is_invocable_v<Fn&, const State&, const Trajectory&, const State&, bool>) {
            const Trajectory traj{};
            for (size_t i = 1; i < path.size(); ++i) fn(path[i - 1], traj, path[i], true);
        } else {
            for (const State& s : path) fn(s);
        }
    }

    // visitor.vertex(state) / visitor.edge(parentState) per node (prrt.hpp:300-316)
    template <typename Visitor>
    void visitGraph(Visitor&& visitor) const {
        std::vector<double> states;
        std::vector<int32_t> parents;
        readTree(states, parents);
        for (size_t n = 0; n < parents.size(); ++n) {
            visitor.vertex(fromDoubles(&states[n * kDim]));
            if (parents[n] >= 0) visitor.edge(fromDoubles(&states[parents[n] * kDim]));
        }
    }

    void printStats() const {
        mptg_prrt_stats st = stats();
        std::cout << "nodes in graph: " << size() << "\n"
                  << "solutions: " << st.goals << "\n";
        if constexpr (Strategy::stats) {
            std::cout << "iterations: " << st.iterations << "\nsamples: " << st.samples
                      << "\nlink calls: " << st.link_calls << " (" << st.link_ok << " ok, " << st.valid_calls
                      << " valid() calls)\n";
            if constexpr (kStar)
                std::cout << "rewire tests: " << st.rewire_tests << "\nrewires: " << st.rewires
                          << "\nparent changes: " << st.parent_changes << "\n";
        }
    }
};

}  // namespace mpt_amd
