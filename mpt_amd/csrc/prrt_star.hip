// Batched PRRT* on the device (replaces PRRTStar::solve / Worker::addSample,
// src/mpt/impl/prrt_star/prrt_star.hpp:447-688).
//
// One iteration = B samples against the tree (and path costs) as they stood
// when the iteration began.  Phases and kernels:
//   extend         prrt_sample, kNN(k=1), prrt_steer (dNear recomputed after
//                  steering, :535), link(near, new)          :510-545
//   survivors      star_select      compaction in sample order
//   neighbours     kNN(new, k) over the iteration's tree     :559-562
//   choose parent  star_cp_prepare  one wave per survivor, lanes = neighbours:
//                                   ranks by (cost + d, position); every
//                                   neighbour the sequential walk could reach
//                                   before the near node or the cost bound
//                                   gets its link(nbr, new) queued
//                  link()           all queued edges at once (speculative)
//                  star_cp_resolve  the sequential walk's outcome from those
//                                   verdicts: parent, cost, checked marks
//                                   (:569-605); rewire candidates under the
//                                   iteration's costs queued
//   append         star_append      ids in sample order, child lists (:610-622)
//                  link()           queued rewire edges (speculative)
//   rewire         star_rewire      one workgroup walks (survivor, neighbour)
//                                   pairs in sequential order against the
//                                   current costs (:626-656); per rewire it
//                                   relinks the node and pushes the cost
//                                   delta down the subtree level by level
//                                   (:664-688).  A pair whose condition holds
//                                   only because the new node's own cost fell
//                                   earlier in the pass has no queued verdict:
//                                   the pass stops there and queues, as one
//                                   batch, every later pair whose condition
//                                   holds under the current costs and has no
//                                   verdict; the host links the batch and the
//                                   pass resumes at the same pair.
// Link verdicts do not depend on who asks or in which order, so the
// speculative batches change the work done, never the tree.
#include <algorithm>
#include <climits>
#include <cmath>

#include "prrt_dev.hpp"

namespace mptg {

namespace {

constexpr int kStarMaxK = 64;

struct StarCounters {
    int64_t m;          // survivors of the current iteration
    int64_t ncp, nrw;   // queued choose-parent / rewire edges
    int64_t pos;        // rewire pass: next (survivor, neighbour) pair
    int64_t late_from;  // first row of the rewire pass's latest batch
    int32_t pending, bad_distance;  // pending: a batch waits for verdicts; bad_distance: 1 near-node
                                    // cost mismatch, 2 corrupt child lists
    int64_t rewire_tests, rewires, parent_changes, cp_links, rw_links, deferred, rw_late;
};
static_assert(kStarCounterOffset + sizeof(StarCounters) <= kCounterBlock, "PRRT* counters fit the block");

__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int w = __shfl_xor(v, o);
        v = w < v ? w : v;
    }
    return v;
}

// ---- survivors of the extend phase, in sample order (single block) --------
__global__ __launch_bounds__(1024) void star_select(int B, int D, const uint8_t* __restrict__ status,
                                                    const uint8_t* __restrict__ verdict,
                                                    const uint32_t* __restrict__ checks,
                                                    const double* __restrict__ to, int32_t* __restrict__ sidx,
                                                    double* __restrict__ qs, Counters* __restrict__ cnt,
                                                    StarCounters* __restrict__ sc) {
    __shared__ int wsum[16];
    __shared__ int64_t base_s;
    __shared__ unsigned long long zero_s, inv_s, lc_s, ok_s, vc_s;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    if (t == 0) {
        base_s = 0;
        zero_s = inv_s = lc_s = ok_s = vc_s = 0;
    }
    __syncthreads();
    unsigned long long lzero = 0, linv = 0, llc = 0, lok = 0, lvc = 0;
    for (int c0 = 0; c0 < B; c0 += 1024) {
        const int b = c0 + t;
        bool ok = false;
        if (b < B) {
            const uint8_t st = status[b];
            if (st == ST_LINK_FAIL) {
                ++llc;
                lvc += checks[b];
                ok = verdict[b] != 0;
            }
            lzero += st == ST_ZERO;
            linv += st == ST_INVALID;
            lok += ok;
        }
        const uint64_t msk = __ballot(ok);
        if (lane == 0) wsum[wid] = __popcll(msk);
        __syncthreads();
        int wbase = 0, tot = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wid) wbase += wsum[w];
            tot += wsum[w];
        }
        if (ok) {
            const int64_t j = base_s + wbase + __popcll(msk & lanes_below(lane));
            sidx[j] = b;
            for (int i = 0; i < D; ++i) qs[j * D + i] = to[(int64_t)b * D + i];
        }
        __syncthreads();
        if (t == 0) base_s += tot;
        __syncthreads();
    }
    atomicAdd(&zero_s, lzero);
    atomicAdd(&inv_s, linv);
    atomicAdd(&lc_s, llc);
    atomicAdd(&ok_s, lok);
    atomicAdd(&vc_s, lvc);
    __syncthreads();
    if (t == 0) {
        cnt->zero += (int64_t)zero_s;
        cnt->invalid += (int64_t)inv_s;
        cnt->link_calls += (int64_t)lc_s;
        cnt->link_ok += (int64_t)ok_s;
        cnt->valid_calls += (int64_t)vc_s;
        sc->m = base_s;
        sc->ncp = 0;
        sc->nrw = 0;
        sc->pos = 0;
        sc->pending = 0;
    }
}

// Survivor j's neighbour list: entries [lo, hi) of nbh_idx / nbh_d, in
// ascending (distance, id) order; rows of k slots (unused slots hold -1) or,
// with offsets, the CSR lists of a radius search (any length).
struct NbhList {
    int64_t lo, hi;
};

__device__ __forceinline__ NbhList nbh_list(const int64_t* __restrict__ off, int k, int64_t j) {
    if (off) return {off[j], off[j + 1]};
    return {j * k, j * k + k};
}

__device__ __forceinline__ int64_t nbh_owner(const int32_t* __restrict__ own, int k, int64_t e) {
    return own ? (int64_t)own[e] : e / k;
}

// Lexicographic (key, entry) minimum over the wave; entries with no
// candidate carry (+inf, INT64_MAX).
__device__ __forceinline__ void wave_min_key(double& key, int64_t& e) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ko = __shfl_xor(key, o);
        const int64_t eo = __shfl_xor(e, o);
        if (ko < key || (ko == key && eo < e)) {
            key = ko;
            e = eo;
        }
    }
}

__device__ __forceinline__ bool key_before(double k1, int64_t e1, double k2, int64_t e2) {
    return k1 < k2 || (k1 == k2 && e1 < e2);
}

// The walk (:580-605) visits the list in ascending (cost(nb) + d, position)
// order: keys are compared with the list position (here the entry index,
// which grows with it) breaking ties, exactly as a stable sort on the key.

// ---- choose parent, pass 1: queue the walk's link checks -----------------
// The sequential walk stops at the first neighbour whose key exceeds the
// parent cost, at the near node, or at the first passing link.  Every
// neighbour before the first of the first two stops may be reached, so all
// of their links are queued.  One wave per survivor, 64 entries per page.
__global__ __launch_bounds__(256) void star_cp_prepare(int D, int k, int64_t m, const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ sidx,
                                                       const int32_t* __restrict__ near,
                                                       const double* __restrict__ dnear,
                                                       const double* __restrict__ qs,
                                                       const int32_t* __restrict__ nbh_idx,
                                                       const double* __restrict__ nbh_d,
                                                       const double* __restrict__ x64, const double* cost,
                                                       double* __restrict__ cp_from, double* __restrict__ cp_to,
                                                       int32_t* __restrict__ pair_cp, StarCounters* sc) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m) return;  // wave-uniform
    const int b = sidx[j];
    const int32_t nr = near[b];
    const double pc0 = cost[nr] + dnear[b];
    const NbhList L = nbh_list(off, k, j);
    double sk = __builtin_inf();
    int64_t se = INT64_MAX;
    bool bad = false;
    for (int64_t e0 = L.lo; e0 < L.hi; e0 += 64) {
        const int64_t e = e0 + lane;
        const int32_t nb = e < L.hi ? nbh_idx[e] : -1;
        if (nb >= 0) {
            const double key = cost[nb] + nbh_d[e];
            bad |= nb == nr && key != pc0;
            if ((key > pc0 || nb == nr) && key_before(key, e, sk, se)) {
                sk = key;
                se = e;
            }
        }
    }
    wave_min_key(sk, se);
    if (__ballot(bad) && lane == 0) atomicOr(&sc->bad_distance, 1);
    for (int64_t e0 = L.lo; e0 < L.hi; e0 += 64) {
        const int64_t e = e0 + lane;
        const int32_t nb = e < L.hi ? nbh_idx[e] : -1;
        const bool cand = nb >= 0 && key_before(cost[nb] + nbh_d[e], e, sk, se);
        const uint64_t cm = __ballot(cand);
        unsigned long long base = 0;
        if (lane == 0 && cm) base = atomicAdd((unsigned long long*)&sc->ncp, (unsigned long long)__popcll(cm));
        base = __shfl(base, 0);
        if (e < L.hi) pair_cp[e] = -1;
        if (cand) {
            const int64_t q = (int64_t)base + __popcll(cm & lanes_below(lane));
            pair_cp[e] = (int32_t)q;
            for (int i = 0; i < D; ++i) {
                cp_from[q * D + i] = x64[(int64_t)nb * D + i];
                cp_to[q * D + i] = qs[j * D + i];
            }
        }
    }
}

// ---- choose parent, pass 2: the walk's outcome; queue rewire checks --------
__global__ __launch_bounds__(256) void star_cp_resolve(int D, int k, int64_t m, const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ sidx,
                                                       const int32_t* __restrict__ near,
                                                       const double* __restrict__ dnear,
                                                       const double* __restrict__ qs,
                                                       const int32_t* __restrict__ nbh_idx,
                                                       const double* __restrict__ nbh_d,
                                                       const double* __restrict__ x64, const double* cost,
                                                       const int32_t* __restrict__ pair_cp,
                                                       const uint8_t* __restrict__ cp_ver, int32_t* __restrict__ cpar,
                                                       double* __restrict__ pc_out, uint8_t* __restrict__ marked,
                                                       double* __restrict__ rw_from, double* __restrict__ rw_to,
                                                       int32_t* __restrict__ pair_rw, StarCounters* sc) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m) return;  // wave-uniform
    const int b = sidx[j];
    const int32_t nr = near[b];
    const double pc0 = cost[nr] + dnear[b];
    const NbhList L = nbh_list(off, k, j);
    // the first stop in walk order
    double sk = __builtin_inf();
    int64_t se = INT64_MAX;
    for (int64_t e0 = L.lo; e0 < L.hi; e0 += 64) {
        const int64_t e = e0 + lane;
        const int32_t nb = e < L.hi ? nbh_idx[e] : -1;
        if (nb >= 0) {
            const double key = cost[nb] + nbh_d[e];
            const int32_t q = pair_cp[e];
            const bool pass = q >= 0 && cp_ver[q] != 0;
            if ((key > pc0 || nb == nr || pass) && key_before(key, e, sk, se)) {
                sk = key;
                se = e;
            }
        }
    }
    wave_min_key(sk, se);
    // why it stops: 1 = cost bound (not visited), 2 = near node, 3 = passing link
    int skind = 0;
    int32_t par = nr;
    double pc = pc0;
    if (se != INT64_MAX) {
        const int32_t snb = nbh_idx[se];
        skind = sk > pc0 ? 1 : (snb == nr ? 2 : 3);
        if (skind == 2) {
            pc = sk;
        } else if (skind == 3) {
            par = snb;
            pc = sk;
        }
    }
    unsigned long long nv = 0, ncp = 0, nrw = 0;
    for (int64_t e0 = L.lo; e0 < L.hi; e0 += 64) {
        const int64_t e = e0 + lane;
        const int32_t nb = e < L.hi ? nbh_idx[e] : -1;
        const bool valid = nb >= 0;
        const double d = valid ? nbh_d[e] : 0.0;
        const double cnb = valid ? cost[nb] : 0.0;
        const bool mk = valid && (key_before(cnb + d, e, sk, se) || (e == se && skind != 1));
        // rewire candidates under the iteration's costs; costs only fall
        // during the rewire pass, so every pair it rewires while
        // cost(new) = pc is here
        const bool rw = valid && !mk && (pc + d < cnb);
        const uint64_t rm = __ballot(rw);
        unsigned long long base = 0;
        if (lane == 0 && rm) base = atomicAdd((unsigned long long*)&sc->nrw, (unsigned long long)__popcll(rm));
        base = __shfl(base, 0);
        if (e < L.hi) {
            marked[e] = mk ? 1 : 0;
            pair_rw[e] = -1;
        }
        if (rw) {
            const int64_t r = (int64_t)base + __popcll(rm & lanes_below(lane));
            pair_rw[e] = (int32_t)r;
            for (int i = 0; i < D; ++i) {
                rw_from[r * D + i] = qs[j * D + i];
                rw_to[r * D + i] = x64[(int64_t)nb * D + i];
            }
        }
        nv += __popcll(__ballot(valid));
        ncp += __popcll(__ballot(valid && pair_cp[e] >= 0));
        nrw += __popcll(rm);
    }
    if (lane == 0) {
        cpar[j] = par;
        pc_out[j] = pc;
        atomicAdd((unsigned long long*)&sc->rewire_tests, nv);
        atomicAdd((unsigned long long*)&sc->cp_links, ncp);
        atomicAdd((unsigned long long*)&sc->rw_links, nrw);
        if (par != nr) atomicAdd((unsigned long long*)&sc->parent_changes, 1ull);
    }
}

// ---- append survivors (single block) ----------------------------------------
__global__ __launch_bounds__(1024) void star_append(SpaceArgs sp, PrrtDev pd, double goal_radius, int64_t m,
                                                    int64_t n0, int dp, const double* __restrict__ qs,
                                                    const int32_t* __restrict__ cpar,
                                                    const double* __restrict__ pc, double* __restrict__ x64,
                                                    float* __restrict__ x32, int32_t* __restrict__ parent,
                                                    double* cost, int32_t* child, int32_t* next, int32_t* prev,
                                                    Counters* __restrict__ cnt, int32_t* __restrict__ goal_ids,
                                                    int64_t goal_cap) {
    __shared__ int gsum[16];
    __shared__ int64_t gbase_s;
    const int D = sp.dim;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    if (t == 0) gbase_s = cnt->goals;
    __syncthreads();
    for (int64_t c0 = 0; c0 < m; c0 += 1024) {
        const int64_t j = c0 + t;
        bool goal = false;
        if (j < m) {
            const int64_t id = n0 + j;
            for (int i = 0; i < D; ++i) {
                const double v = qs[j * D + i];
                x64[id * D + i] = v;
                x32[id * dp + i] = (float)v;
            }
            parent[id] = cpar[j];
            cost[id] = pc[j];
            child[id] = -1;
            next[id] = -1;
            prev[id] = -1;
            goal = space_dist(sp, qs + j * D, pd.goal) <= goal_radius;
        }
        const uint64_t gm = __ballot(goal);
        if (lane == 0) gsum[wid] = __popcll(gm);
        __syncthreads();
        int gw = 0, gt = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wid) gw += gsum[w];
            gt += gsum[w];
        }
        if (goal) {
            const int64_t g = gbase_s + gw + __popcll(gm & lanes_below(lane));
            if (g < goal_cap) goal_ids[g] = (int32_t)(n0 + j);
        }
        __syncthreads();
        if (t == 0) gbase_s += gt;
        __syncthreads();
    }
    // every new row is initialised: push each new node onto its parent's
    // child list (list order does not affect any result)
    for (int64_t j = t; j < m; j += 1024) {
        const int32_t id = (int32_t)(n0 + j);
        const int32_t old = atomicExch(&child[cpar[j]], id);
        next[id] = old;
        if (old >= 0) prev[old] = id;
    }
    if (t == 0) {
        cnt->size = n0 + m;
        cnt->goals = gbase_s;
    }
}

// ---- rewire pass (single block) ---------------------------------------------
struct RewireShared {
    int64_t pos;
    int first, unknown, fn, gn, bad, queued;
    int64_t seen, late;
    double delta;
};

// Moves neighbour `pair` under its survivor and pushes the cost decrease down
// its subtree (nonConcurrentPushUpdate, :664-688).  Whole block.
// The walk is bounded by the tree size `fcap` (also the capacity of each
// frontier array): a subtree can hold no more nodes, so exceeding it means
// the child lists are corrupt; the pass then flags the tree and stops.
__device__ void rewire_apply(RewireShared& s, int64_t pair, int k, const int32_t* __restrict__ own, int64_t n0,
                             int64_t fcap, const int32_t* __restrict__ nbh_idx, const double* __restrict__ nbh_d,
                             double* cost, int32_t* parent, int32_t* child, int32_t* next, int32_t* prev, int32_t* fa,
                             int32_t* fb, StarCounters* sc) {
    const int t = threadIdx.x;
    if (t == 0) {
        const int32_t id = (int32_t)(n0 + nbh_owner(own, k, pair));
        const int32_t nb = nbh_idx[pair];
        const double nc = cost[id] + nbh_d[pair];
        const double delta = cost[nb] - nc;
        const int32_t p = parent[nb], pv = prev[nb], nx = next[nb];
        if (pv >= 0) next[pv] = nx;
        else child[p] = nx;
        if (nx >= 0) prev[nx] = pv;
        const int32_t h = child[id];
        next[nb] = h;
        prev[nb] = -1;
        if (h >= 0) prev[h] = nb;
        child[id] = nb;
        parent[nb] = id;
        cost[nb] = nc;
        s.delta = delta;
        fa[0] = nb;
        s.fn = 1;
        s.gn = 0;
        s.seen = 1;
        sc->rewires += 1;
    }
    __syncthreads();
    int32_t* cur = fa;
    int32_t* nxt = fb;
    for (;;) {
        const int n = s.fn;
        if (n == 0 || s.bad) break;
        const double delta = s.delta;
        for (int x = t; x < n; x += 1024) {
            int64_t len = 0;
            for (int32_t c = child[cur[x]]; c >= 0 && len < fcap; c = next[c], ++len) {
                cost[c] = cost[c] - delta;
                const int slot = atomicAdd(&s.gn, 1);
                if (slot < fcap) nxt[slot] = c;
            }
            if (len >= fcap) s.bad = 1;
        }
        __syncthreads();
        if (t == 0) {
            s.seen += s.gn;
            if (s.seen > fcap) s.bad = 1;
            s.fn = s.bad ? 0 : s.gn;
            s.gn = 0;
        }
        __syncthreads();
        int32_t* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
}

// Pairs are the list entries in order (survivor by survivor, each list in
// order): `total` of them, survivor of entry e = own[e] (or e / k).  Verdicts
// queued before the pass have rows below nrw0; a pair the walk reaches with
// its condition holding and a row at or above nrw0 held only because its new
// node's cost fell earlier in the pass (`deferred`).  A pair reached with its
// condition holding and no row stops the pass: every pair from it on whose
// condition holds under the current costs and has no row gets one (rows
// [late_from, nrw), pair_rw set), `pending` asks the host to link them, and
// the next launch resumes at the same pair with the same costs.
__global__ __launch_bounds__(1024) void star_rewire(int D, int k, int64_t total, int64_t n0, int64_t fcap,
                                                    int64_t nrw0, const int32_t* __restrict__ own,
                                                    const int32_t* __restrict__ nbh_idx,
                                                    const double* __restrict__ nbh_d,
                                                    const uint8_t* __restrict__ marked, int32_t* pair_rw,
                                                    const uint8_t* __restrict__ rw_ver, const double* __restrict__ x64,
                                                    double* cost, int32_t* parent, int32_t* child, int32_t* next,
                                                    int32_t* prev, int32_t* fa, int32_t* fb,
                                                    double* __restrict__ rw_from, double* __restrict__ rw_to,
                                                    StarCounters* sc) {
    __shared__ RewireShared s;
    const int t = threadIdx.x;
    if (t == 0) {
        s.pos = sc->pos;
        s.first = INT_MAX;
        s.bad = 0;
        s.late = 0;
        s.queued = 0;
        sc->pending = 0;
    }
    __syncthreads();
    for (;;) {
        const int64_t p0 = s.pos;
        if (p0 >= total || s.bad) break;
        const int64_t pair = p0 + t;
        bool act = false, unk = false, late = false;
        if (pair < total) {
            const int32_t nb = nbh_idx[pair];
            const int64_t j = nbh_owner(own, k, pair);
            if (nb >= 0 && !marked[pair]) {
                const double nc = cost[n0 + j] + nbh_d[pair];
                if (nc < cost[nb]) {
                    const int32_t r = pair_rw[pair];
                    if (r >= 0) {
                        act = rw_ver[r] != 0;
                        late = r >= nrw0;
                    } else {
                        act = true;
                        unk = true;
                    }
                }
            }
        }
        if (act) atomicMin(&s.first, t);
        __syncthreads();
        const int f = s.first;
        // the walk reaches the pairs up to the first that acts (all of them
        // when none does) under these costs
        if (late && t <= f) atomicAdd((unsigned long long*)&s.late, 1ull);
        if (f == INT_MAX) {
            __syncthreads();
            if (t == 0) s.pos = p0 + 1024;
            __syncthreads();
            continue;
        }
        if (t == f) s.unknown = unk ? 1 : 0;
        __syncthreads();
        if (s.unknown) {
            const int64_t stop = p0 + f;
            const int64_t base = sc->nrw;
            for (int64_t e = stop + t; e < total; e += 1024) {
                const int32_t nb = nbh_idx[e];
                if (nb < 0 || marked[e] || pair_rw[e] >= 0) continue;
                const int64_t id = n0 + nbh_owner(own, k, e);
                if (!(cost[id] + nbh_d[e] < cost[nb])) continue;
                const int64_t r = base + atomicAdd(&s.queued, 1);
                pair_rw[e] = (int32_t)r;
                for (int i = 0; i < D; ++i) {
                    rw_from[r * D + i] = x64[id * D + i];
                    rw_to[r * D + i] = x64[(int64_t)nb * D + i];
                }
            }
            __syncthreads();
            if (t == 0) {
                sc->late_from = base;
                sc->nrw = base + s.queued;
                sc->rw_links += s.queued;
                sc->rw_late += s.queued;
                sc->deferred += s.late;
                sc->pending = 1;
                sc->pos = stop;
            }
            return;
        }
        rewire_apply(s, p0 + f, k, own, n0, fcap, nbh_idx, nbh_d, cost, parent, child, next, prev, fa, fb, sc);
        if (t == 0) {
            s.pos = p0 + f + 1;
            s.first = INT_MAX;
        }
        __syncthreads();
    }
    if (t == 0) {
        sc->pos = total;
        sc->deferred += s.late;
        if (s.bad) sc->bad_distance = 2;
    }
}

}  // namespace

int tree_reserve(mptg_ctx* ctx, int64_t cap);
int prrt_reserve(mptg_ctx* ctx, int B);

int star_k(const PrrtState& s, int dim, int64_t n) {
    if (s.star_k > 0) return s.star_k;
    const double kRRG = s.rewire_factor * 2.718281828459045 * (1.0 + 1.0 / (double)dim);
    return (int)std::ceil(kRRG * std::log((double)(n + 1)));
}

// The space's dimension as the reference's formulas take it
// (space.dimensions()): the state size for L1/L2, 6 for SE(3) (SO(3) is a
// 3-manifold, its quaternion has 4 coordinates).
static int manifold_dim(const mptg_ctx* ctx) { return ctx->metric == MPTG_METRIC_SE3 ? 6 : ctx->dim; }

// rewire_r_nearest's radius at tree size n (rrg_rewire_neighbors.hpp:99-128):
// r = rRRG (ln(n + 1) / (n + 1))^(1/d), rRRG = f (2 (1 + 1/d) mu / zeta_d)^(1/d),
// zeta_d = sqrt(pi)^d / Gamma(d/2 + 1) (the unit L2 ball), mu = the sampler's
// measure (uniform_box_sampler.hpp:76-78, uniform_sampler_scaled.hpp:59-65,
// uniform_sampler_so3.hpp:79-83, uniform_sampler_cartesian.hpp:93-98): the
// product of the box extents; SE(3): pi^2 w_so3 times the translation
// box's product (times w_l2).  Extents multiply left to right.
double star_radius(const mptg_ctx* ctx, int64_t n) {
    const PrrtState& s = ctx->prrt;
    const int d = manifold_dim(ctx);
    const double pi = 3.14159265358979323846;
    double mu;
    if (ctx->metric == MPTG_METRIC_SE3) {
        double box = 1.0;
        for (int i = 4; i < 7; ++i) box = box * (s.hi[i] - s.lo[i]);
        mu = (pi * pi * ctx->w_so3) * (ctx->w_l2 == 1.0 ? box : box * ctx->w_l2);
    } else {
        mu = 1.0;
        for (int i = 0; i < ctx->dim; ++i) mu = mu * (s.hi[i] - s.lo[i]);
    }
    const double inv = 1 / (double)d;
    const double ball = std::pow(std::sqrt(pi), (double)d) / std::tgamma((double)d / 2 + 1);
    const double rrg = s.rewire_factor * std::pow(2 * (1 + inv) * mu / ball, inv);
    const double m = (double)(n + 1);
    return rrg * std::pow(std::log(m) / m, inv);
}

// Per-survivor buffers for B samples per iteration.
static int star_reserve(mptg_ctx* ctx, int B) {
    PrrtState& pb = ctx->prrt;
    const int D = ctx->dim;
    int rc;
    if ((rc = ensure(ctx, pb.dnear, sizeof(double) * B))) return rc;
    if ((rc = ensure(ctx, pb.sidx, sizeof(int32_t) * B))) return rc;
    if ((rc = ensure(ctx, pb.qs, sizeof(double) * D * B))) return rc;
    if ((rc = ensure(ctx, pb.pc, sizeof(double) * B))) return rc;
    if ((rc = ensure(ctx, pb.cpar, sizeof(int32_t) * B))) return rc;
    return MPTG_OK;
}

// Per-entry buffers for E neighbour-list entries (each may queue one
// choose-parent and one rewire edge); owner: the CSR lists' survivor index.
static int star_reserve_entries(mptg_ctx* ctx, size_t E, bool owner) {
    PrrtState& pb = ctx->prrt;
    const int D = ctx->dim;
    E = std::max<size_t>(E, 1);
    int rc;
    if ((rc = ensure(ctx, pb.nbh_idx, sizeof(int32_t) * E))) return rc;
    if ((rc = ensure(ctx, pb.nbh_d, sizeof(double) * E))) return rc;
    if (owner && (rc = ensure(ctx, pb.nbh_own, sizeof(int32_t) * E))) return rc;
    if ((rc = ensure(ctx, pb.marked, E))) return rc;
    if ((rc = ensure(ctx, pb.cp_from, sizeof(double) * D * E))) return rc;
    if ((rc = ensure(ctx, pb.cp_to, sizeof(double) * D * E))) return rc;
    if ((rc = ensure(ctx, pb.cp_ver, E))) return rc;
    if ((rc = ensure(ctx, pb.cp_chk, sizeof(uint32_t) * E))) return rc;
    if ((rc = ensure(ctx, pb.pair_cp, sizeof(int32_t) * E))) return rc;
    if ((rc = ensure(ctx, pb.rw_from, sizeof(double) * D * E))) return rc;
    if ((rc = ensure(ctx, pb.rw_to, sizeof(double) * D * E))) return rc;
    if ((rc = ensure(ctx, pb.rw_ver, E))) return rc;
    if ((rc = ensure(ctx, pb.rw_chk, sizeof(uint32_t) * E))) return rc;
    if ((rc = ensure(ctx, pb.pair_rw, sizeof(int32_t) * E))) return rc;
    return MPTG_OK;
}

int prrt_star_run(mptg_ctx* ctx, int64_t target_size, int64_t max_iters) {
    PrrtState& s = ctx->prrt;
    if (!ctx->star_ok)
        return fail(ctx, MPTG_E_STATE, "PRRT* trees must be grown from mptg_prrt_add_start (no path costs)");
    const int B = s.p.batch;
    const int D = ctx->dim;
    int rc;
    if ((rc = prrt_reserve(ctx, B))) return rc;
    if ((rc = star_reserve(ctx, B))) return rc;
    Counters* dcnt = (Counters*)s.counters.p;
    StarCounters* dsc = (StarCounters*)((char*)s.counters.p + kStarCounterOffset);
    if (!s.counters_valid) {
        Counters c{};
        c.size = ctx->size;
        c.goals = s.goals;
        StarCounters z{};
        MPTG_HIP(ctx, hipMemcpyAsync(dcnt, &c, sizeof(c), hipMemcpyHostToDevice, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(dsc, &z, sizeof(z), hipMemcpyHostToDevice, ctx->stream));
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        s.counters_valid = true;
        s.stats_base = s.stats;
    }
    if ((rc = ensure_pinned(ctx, kStarCounterOffset + sizeof(StarCounters)))) return rc;
    Counters* hc = (Counters*)ctx->pinned;
    StarCounters* hs = (StarCounters*)((char*)ctx->pinned + kStarCounterOffset);
    auto read_counters = [&]() -> int {  // both counter sets in one copy
        MPTG_HIP(ctx, hipMemcpyAsync(hc, dcnt, kStarCounterOffset + sizeof(StarCounters), hipMemcpyDeviceToHost,
                                     ctx->stream));
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return MPTG_OK;
    };
    PrrtDev pd;
    for (int i = 0; i < D; ++i) {
        pd.lo[i] = s.lo[i];
        pd.hi[i] = s.hi[i];
        pd.goal[i] = s.goal[i];
    }
    const int dp = pitch32(D);
    for (int64_t iter = 0; iter < max_iters && ctx->size < target_size; ++iter) {
        const int64_t n0 = ctx->size;
        // rewire_k_nearest: rows of the k nearest; rewire_r_nearest: every node
        // within the radius, in CSR lists
        const bool by_radius = s.star_mode == MPTG_REWIRE_R_NEAREST;
        const int k = by_radius ? 1 : star_k(s, manifold_dim(ctx), n0);
        if (k < 1 || k > kStarMaxK)
            return fail(ctx, MPTG_E_RANGE,
                        "PRRT* neighbour count " + std::to_string(k) + " is outside 1..64 at tree size " +
                            std::to_string(n0) + " (set a fixed k)");
        if ((rc = tree_reserve(ctx, n0 + B))) return rc;
        if ((rc = ensure(ctx, s.front, sizeof(int32_t) * 2 * (size_t)(n0 + B)))) return rc;
        const int64_t it = s.iteration;
        // ---- extend
        hipLaunchKernelGGL(prrt_sample, dim3(sample_blocks(ctx->metric, D, B)), dim3(kBlock), 0, ctx->stream, pd,
                           ctx->metric, D, s.p.seed, s.p.goal_bias, it, B, (const Counters*)dcnt, (double*)s.q.p,
                           nullptr);
        MPTG_HIP(ctx, hipGetLastError());
        uint8_t* skip = nullptr;  // custom sampler: slots without a sample
        if (s.sampler && (rc = take_host_samples(ctx, it, (double*)s.q.p, dcnt, &skip))) return rc;
        if ((rc = launch_knn(ctx, (const double*)s.q.p, B, 1, __builtin_inf(), (int32_t*)s.near.p,
                             (double*)s.ndist.p, n0)))
            return rc;
        // the steer kernel writes dnear: d(near, new), the kNN distance unless
        // the step was cut to the range
        dispatch_dim(D, [&]<int DD>() {
            launch_steer<DD>(ctx, s.p.range, B, (const double*)s.q.p, (const int32_t*)s.near.p,
                             (const double*)s.ndist.p, (double*)s.from.p, (double*)s.to.p, (uint8_t*)s.status.p,
                             (uint8_t*)s.mask.p, (double*)s.dnear.p, (const Counters*)dcnt);
        });
        MPTG_HIP(ctx, hipGetLastError());
        if (skip && (rc = apply_skip(ctx, skip, (uint8_t*)s.status.p, (uint8_t*)s.mask.p, dcnt))) return rc;
        if (ctx->scene.kind == MPTG_SCENE_MESH &&
            (rc = launch_mesh_filter(ctx, (const double*)s.to.p, B, (uint8_t*)s.mask.p, (uint8_t*)s.status.p,
                                     ST_INVALID)))
            return rc;
        if (ctx->scene.kind == MPTG_SCENE_HOST &&
            (rc = host_filter(ctx, (const double*)s.to.p, B, (uint8_t*)s.mask.p, (uint8_t*)s.status.p, ST_INVALID)))
            return rc;
        if ((rc = launch_link_masked(ctx, (const double*)s.from.p, (const double*)s.to.p, B, s.p.order, 1,
                                     (const uint8_t*)s.mask.p, (uint8_t*)s.verdict.p, (uint32_t*)s.checks.p)))
            return rc;
        hipLaunchKernelGGL(star_select, dim3(1), dim3(1024), 0, ctx->stream, B, D, (const uint8_t*)s.status.p,
                           (const uint8_t*)s.verdict.p, (const uint32_t*)s.checks.p, (const double*)s.to.p,
                           (int32_t*)s.sidx.p, (double*)s.qs.p, dcnt, dsc);
        MPTG_HIP(ctx, hipGetLastError());
        if ((rc = read_counters())) return rc;
        const int64_t m = hs->m;
        if (m > 0) {
            // ---- neighbours and choose parent
            int64_t E;
            const int64_t* off = nullptr;
            const int32_t* own = nullptr;
            if (by_radius) {
                const double r = star_radius(ctx, n0);
                if ((rc = rnn_count(ctx, (const double*)s.qs.p, m, r, &E))) return rc;
                if (E > (int64_t)INT32_MAX)
                    return fail(ctx, MPTG_E_RANGE, "more than 2^31 rewire neighbours in one iteration");
                if ((rc = star_reserve_entries(ctx, (size_t)E, true))) return rc;
                if ((rc = rnn_fill(ctx, (const double*)s.qs.p, m, r, E, (double*)s.nbh_d.p, (int32_t*)s.nbh_idx.p,
                                   (int32_t*)s.nbh_own.p)))
                    return rc;
                off = rnn_offsets(ctx);
                own = (const int32_t*)s.nbh_own.p;
            } else {
                E = m * (int64_t)k;
                if ((rc = star_reserve_entries(ctx, (size_t)E, false))) return rc;
                if ((rc = launch_knn(ctx, (const double*)s.qs.p, m, k, __builtin_inf(), (int32_t*)s.nbh_idx.p,
                                     (double*)s.nbh_d.p, n0)))
                    return rc;
            }
            const unsigned wb = (unsigned)((m + 3) / 4);
            hipLaunchKernelGGL(star_cp_prepare, dim3(wb), dim3(256), 0, ctx->stream, D, k, m, off,
                               (const int32_t*)s.sidx.p, (const int32_t*)s.near.p, (const double*)s.dnear.p,
                               (const double*)s.qs.p, (const int32_t*)s.nbh_idx.p, (const double*)s.nbh_d.p,
                               (const double*)ctx->x64, (const double*)ctx->cost, (double*)s.cp_from.p,
                               (double*)s.cp_to.p, (int32_t*)s.pair_cp.p, dsc);
            MPTG_HIP(ctx, hipGetLastError());
            if ((rc = read_counters())) return rc;
            if (hs->bad_distance) return fail(ctx, MPTG_E_STATE, "bad distance: near node cost mismatch");
            if (hs->ncp > 0 &&
                (rc = launch_link(ctx, (const double*)s.cp_from.p, (const double*)s.cp_to.p, hs->ncp, s.p.order,
                                  1, (uint8_t*)s.cp_ver.p, (uint32_t*)s.cp_chk.p)))
                return rc;
            hipLaunchKernelGGL(star_cp_resolve, dim3(wb), dim3(256), 0, ctx->stream, D, k, m, off,
                               (const int32_t*)s.sidx.p, (const int32_t*)s.near.p, (const double*)s.dnear.p,
                               (const double*)s.qs.p, (const int32_t*)s.nbh_idx.p, (const double*)s.nbh_d.p,
                               (const double*)ctx->x64, (const double*)ctx->cost, (const int32_t*)s.pair_cp.p,
                               (const uint8_t*)s.cp_ver.p, (int32_t*)s.cpar.p, (double*)s.pc.p,
                               (uint8_t*)s.marked.p, (double*)s.rw_from.p, (double*)s.rw_to.p,
                               (int32_t*)s.pair_rw.p, dsc);
            MPTG_HIP(ctx, hipGetLastError());
            // ---- append
            hipLaunchKernelGGL(star_append, dim3(1), dim3(1024), 0, ctx->stream, space_args(ctx), pd,
                               s.p.goal_radius, m, n0, dp, (const double*)s.qs.p, (const int32_t*)s.cpar.p,
                               (const double*)s.pc.p, ctx->x64, ctx->x32, ctx->parent, ctx->cost, ctx->child,
                               ctx->next, ctx->prev, dcnt, (int32_t*)s.goal_buf.p, (int64_t)4096);
            MPTG_HIP(ctx, hipGetLastError());
            if ((rc = read_counters())) return rc;
            // goals found in this iteration: the path cost and parent they
            // have now, before the rewire pass (foundGoal keeps the goal edge as
            // found, prrt_star.hpp:203-216; solution() starts from it, :324-336)
            const int64_t g0 = (int64_t)s.goal_cost0.size(), g1 = std::min<int64_t>(hc->goals, 4096);
            for (int64_t g = g0; g < g1; ++g) {
                int32_t id = -1, par = -1;
                double c = 0.0;
                MPTG_HIP(ctx, hipMemcpy(&id, (int32_t*)s.goal_buf.p + g, sizeof(id), hipMemcpyDeviceToHost));
                const int64_t j = (int64_t)id - n0;
                if (j < 0 || j >= m) return fail(ctx, MPTG_E_STATE, "goal list out of step with the append");
                MPTG_HIP(ctx, hipMemcpy(&c, (double*)s.pc.p + j, sizeof(c), hipMemcpyDeviceToHost));
                MPTG_HIP(ctx, hipMemcpy(&par, (int32_t*)s.cpar.p + j, sizeof(par), hipMemcpyDeviceToHost));
                s.goal_cost0.push_back(c);
                s.goal_parent0.push_back(par);
            }
            // ---- rewire
            const int64_t nrw0 = hs->nrw;
            if (nrw0 > 0 &&
                (rc = launch_link(ctx, (const double*)s.rw_from.p, (const double*)s.rw_to.p, nrw0, s.p.order, 1,
                                  (uint8_t*)s.rw_ver.p, (uint32_t*)s.rw_chk.p)))
                return rc;
            int32_t* front = (int32_t*)s.front.p;
            for (;;) {
                hipLaunchKernelGGL(star_rewire, dim3(1), dim3(1024), 0, ctx->stream, D, k, E, n0, n0 + m, nrw0, own,
                                   (const int32_t*)s.nbh_idx.p, (const double*)s.nbh_d.p,
                                   (const uint8_t*)s.marked.p, (int32_t*)s.pair_rw.p, (const uint8_t*)s.rw_ver.p,
                                   (const double*)ctx->x64, ctx->cost, ctx->parent, ctx->child, ctx->next,
                                   ctx->prev, front, front + (n0 + B), (double*)s.rw_from.p, (double*)s.rw_to.p,
                                   dsc);
                MPTG_HIP(ctx, hipGetLastError());
                if ((rc = read_counters())) return rc;
                if (hs->bad_distance == 2) return fail(ctx, MPTG_E_STATE, "PRRT* child lists are corrupt");
                if (!hs->pending) break;
                // the batch the pass queued when it stopped
                const int64_t a = hs->late_from, b = hs->nrw;
                if (b <= a || b > E) return fail(ctx, MPTG_E_STATE, "PRRT* rewire batch out of range");
                if ((rc = launch_link(ctx, (const double*)s.rw_from.p + a * D, (const double*)s.rw_to.p + a * D,
                                      b - a, s.p.order, 1, (uint8_t*)s.rw_ver.p + a, (uint32_t*)s.rw_chk.p + a)))
                    return rc;
            }
        } else {
            if ((rc = read_counters())) return rc;
        }
        ctx->size = hc->size;
        s.iteration++;
        s.stats.iterations++;
        s.stats.samples += B;
        if (hc->goals > s.goals) {
            const int64_t ng = std::min<int64_t>(hc->goals, 4096);
            s.goal_ids.resize((size_t)ng);
            MPTG_HIP(ctx, hipMemcpy(s.goal_ids.data(), s.goal_buf.p, sizeof(int32_t) * ng, hipMemcpyDeviceToHost));
            s.goals = hc->goals;
        }
        const mptg_prrt_stats& b0 = s.stats_base;
        s.stats.zero_dist = b0.zero_dist + hc->zero;
        s.stats.invalid_states = b0.invalid_states + hc->invalid;
        s.stats.link_calls = b0.link_calls + hc->link_calls;
        s.stats.link_ok = b0.link_ok + hc->link_ok;
        s.stats.valid_calls = b0.valid_calls + hc->valid_calls;
        s.stats.rewire_tests = b0.rewire_tests + hs->rewire_tests;
        s.stats.rewires = b0.rewires + hs->rewires;
        s.stats.parent_changes = b0.parent_changes + hs->parent_changes;
        s.stats.cp_links = b0.cp_links + hs->cp_links;
        s.stats.rw_links = b0.rw_links + hs->rw_links;
        s.stats.deferred = b0.deferred + hs->deferred;
        s.stats.rw_late = b0.rw_late + hs->rw_late;
        s.stats.goals = s.goals;
        s.stats.size = ctx->size;
    }
    return MPTG_OK;
}

}  // namespace mptg

using namespace mptg;

extern "C" {

int mptg_prrt_star_init(mptg_ctx* ctx, const mptg_prrt_params* p, const mptg_prrt_star_params* sp,
                        const double* h_lo, const double* h_hi, const double* h_goal) {
    if (!ctx) return fail(nullptr, MPTG_E_ARG, "null context");
    if (!sp) return fail(ctx, MPTG_E_ARG, "null PRRT* parameters");
    if (sp->k < 0 || sp->k > kStarMaxK) return fail(ctx, MPTG_E_RANGE, "PRRT* k must be in 0..64");
    if (sp->neighbors != MPTG_REWIRE_K_NEAREST && sp->neighbors != MPTG_REWIRE_R_NEAREST)
        return fail(ctx, MPTG_E_ARG, "unknown rewire neighbour strategy");
    if (!(sp->rewire_factor > 0) || !std::isfinite(sp->rewire_factor))
        return fail(ctx, MPTG_E_ARG, "rewire factor must be a positive finite number");
    int rc = mptg_prrt_init(ctx, p, h_lo, h_hi, h_goal);
    if (rc) return rc;
    PrrtState& s = ctx->prrt;
    s.star = true;
    s.star_k = sp->k;
    s.star_mode = sp->neighbors;
    s.rewire_factor = sp->rewire_factor;
    return MPTG_OK;
}

int mptg_prrt_costs(mptg_ctx* ctx, int64_t first, int64_t n, double* h_cost) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (!ctx->prrt.star || !ctx->star_ok) return fail(ctx, MPTG_E_STATE, "the tree has no PRRT* path costs");
    if (first < 0 || n < 0 || first + n > ctx->size) return fail(ctx, MPTG_E_RANGE, "read beyond tree size");
    if (n == 0) return MPTG_OK;
    if (!h_cost) return fail(ctx, MPTG_E_ARG, "null output");
    MPTG_HIP(ctx, hipMemcpyAsync(h_cost, ctx->cost + first, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

}  // extern "C"
