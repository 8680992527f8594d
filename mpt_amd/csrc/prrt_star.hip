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
//                                   the pass stops, the host checks that one
//                                   edge, and the pass resumes.
// Link verdicts do not depend on who asks or in which order, so the
// speculative batches change the work done, never the tree.
#include <algorithm>
#include <climits>
#include <cmath>

#include "prrt_dev.hpp"

namespace mptg {

namespace {

constexpr int kStarMaxK = 32;

struct StarCounters {
    int64_t m;          // survivors of the current iteration
    int64_t ncp, nrw;   // queued choose-parent / rewire edges
    int64_t pos;        // rewire pass: next (survivor, neighbour) pair
    int64_t pend_pair;  // pair waiting for a verdict
    int32_t pending, bad_distance;  // bad_distance: 1 near-node cost mismatch, 2 corrupt child lists
    int64_t rewire_tests, rewires, parent_changes, cp_links, rw_links, deferred;
};

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

// Per-lane view of one survivor's neighbour list (lane i = i-th nearest).
struct NbhLane {
    bool valid;
    int32_t nb;
    double d, key;  // distance to the new state, cost(nb) + d
    int rank;       // position in ascending (key, lane) order among valid lanes
};

__device__ __forceinline__ NbhLane nbh_lane(int k, int64_t j, int lane, const int32_t* __restrict__ nbh_idx,
                                            const double* __restrict__ nbh_d, const double* cost) {
    NbhLane l;
    l.nb = lane < k ? nbh_idx[j * k + lane] : -1;
    l.valid = l.nb >= 0;
    l.d = l.valid ? nbh_d[j * k + lane] : 0.0;
    l.key = l.valid ? cost[l.nb] + l.d : 0.0;
    l.rank = 0;
    for (int o = 0; o < k; ++o) {
        const double ko = __shfl(l.key, o);
        const int vo = __shfl((int)l.valid, o);
        l.rank += (vo && (ko < l.key || (ko == l.key && o < lane))) ? 1 : 0;
    }
    return l;
}

// ---- choose parent, pass 1: queue the walk's link checks -----------------
// The sequential walk (:580-605) visits ranks 0, 1, ... and stops at the first
// neighbour whose key exceeds the parent cost, at the near node, or at the
// first passing link.  Every neighbour ranked before the first of the first
// two stops may be reached, so all of their links are queued.
__global__ __launch_bounds__(256) void star_cp_prepare(int D, int k, int64_t m, const int32_t* __restrict__ sidx,
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
    const NbhLane l = nbh_lane(k, j, lane, nbh_idx, nbh_d, cost);
    if (__ballot(l.valid && l.nb == nr && l.key != pc0) && lane == 0) atomicOr(&sc->bad_distance, 1);
    const int rs = wave_min(l.valid && (l.key > pc0 || l.nb == nr) ? l.rank : 64);
    const bool cand = l.valid && l.rank < rs;
    const uint64_t cm = __ballot(cand);
    unsigned long long base = 0;
    if (lane == 0 && cm) base = atomicAdd((unsigned long long*)&sc->ncp, (unsigned long long)__popcll(cm));
    base = __shfl(base, 0);
    if (lane < k) pair_cp[j * k + lane] = -1;
    if (cand) {
        const int64_t e = (int64_t)base + __popcll(cm & lanes_below(lane));
        pair_cp[j * k + lane] = (int32_t)e;
        for (int i = 0; i < D; ++i) {
            cp_from[e * D + i] = x64[(int64_t)l.nb * D + i];
            cp_to[e * D + i] = qs[j * D + i];
        }
    }
}

// ---- choose parent, pass 2: the walk's outcome; queue rewire checks --------
__global__ __launch_bounds__(256) void star_cp_resolve(int D, int k, int64_t m, const int32_t* __restrict__ sidx,
                                                       const int32_t* __restrict__ near,
                                                       const double* __restrict__ dnear,
                                                       const double* __restrict__ qs,
                                                       const int32_t* __restrict__ nbh_idx,
                                                       const double* __restrict__ nbh_d,
                                                       const double* __restrict__ x64, const double* cost,
                                                       const int32_t* __restrict__ pair_cp,
                                                       const uint8_t* __restrict__ cp_ver, int32_t* __restrict__ cpar,
                                                       double* __restrict__ pc_out, uint32_t* __restrict__ marked,
                                                       double* __restrict__ rw_from, double* __restrict__ rw_to,
                                                       int32_t* __restrict__ pair_rw, StarCounters* sc) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m) return;  // wave-uniform
    const int b = sidx[j];
    const int32_t nr = near[b];
    const double pc0 = cost[nr] + dnear[b];
    const NbhLane l = nbh_lane(k, j, lane, nbh_idx, nbh_d, cost);
    const int32_t e = l.valid ? pair_cp[j * k + lane] : -1;
    const bool pass = e >= 0 && cp_ver[e] != 0;
    const bool stop = l.valid && (l.key > pc0 || l.nb == nr || pass);
    const int rs = wave_min(stop ? l.rank : 64);
    // the stopping lane (ranks are distinct) and why it stops:
    // 1 = cost bound (not visited), 2 = near node, 3 = passing link
    const uint64_t sm = __ballot(l.valid && stop && l.rank == rs);
    const int src = sm ? __ffsll((unsigned long long)sm) - 1 : 0;
    const int kind = l.key > pc0 ? 1 : (l.nb == nr ? 2 : 3);
    const int skind = sm ? __shfl(kind, src) : 0;
    const int32_t snb = __shfl(l.nb, src);
    const double skey = __shfl(l.key, src);
    int32_t par = nr;
    double pc = pc0;
    if (skind == 2) {
        pc = skey;
    } else if (skind == 3) {
        par = snb;
        pc = skey;
    }
    const bool mk = l.valid && (l.rank < rs || (l.rank == rs && skind != 1));
    const uint64_t mm = __ballot(mk);
    // rewire candidates under the iteration's costs; costs only fall during
    // the rewire pass, so every pair it rewires while cost(new) = pc is here
    const bool rw = l.valid && !mk && (pc + l.d < cost[l.nb]);
    const uint64_t rm = __ballot(rw);
    unsigned long long base = 0;
    if (lane == 0 && rm) base = atomicAdd((unsigned long long*)&sc->nrw, (unsigned long long)__popcll(rm));
    base = __shfl(base, 0);
    if (lane < k) pair_rw[j * k + lane] = -1;
    if (rw) {
        const int64_t r = (int64_t)base + __popcll(rm & lanes_below(lane));
        pair_rw[j * k + lane] = (int32_t)r;
        for (int i = 0; i < D; ++i) {
            rw_from[r * D + i] = qs[j * D + i];
            rw_to[r * D + i] = x64[(int64_t)l.nb * D + i];
        }
    }
    const uint64_t vm = __ballot(l.valid);
    const uint64_t em = __ballot(e >= 0);
    if (lane == 0) {
        cpar[j] = par;
        pc_out[j] = pc;
        marked[j] = (uint32_t)mm;
        atomicAdd((unsigned long long*)&sc->rewire_tests, (unsigned long long)__popcll(vm));
        atomicAdd((unsigned long long*)&sc->cp_links, (unsigned long long)__popcll(em));
        atomicAdd((unsigned long long*)&sc->rw_links, (unsigned long long)__popcll(rm));
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
    int64_t pair;
    int first, unknown, pend, fn, gn, bad;
    int64_t seen;
    double delta;
};

// Moves neighbour `pair` under its survivor and pushes the cost decrease down
// its subtree (nonConcurrentPushUpdate, :664-688).  Whole block.
// The walk is bounded by the tree size `fcap` (also the capacity of each
// frontier array): a subtree can hold no more nodes, so exceeding it means
// the child lists are corrupt; the pass then flags the tree and stops.
__device__ void rewire_apply(RewireShared& s, int64_t pair, int k, int64_t n0, int64_t fcap,
                             const int32_t* __restrict__ nbh_idx, const double* __restrict__ nbh_d, double* cost,
                             int32_t* parent, int32_t* child, int32_t* next, int32_t* prev, int32_t* fa, int32_t* fb,
                             StarCounters* sc) {
    const int t = threadIdx.x;
    if (t == 0) {
        const int32_t id = (int32_t)(n0 + pair / k);
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

__global__ __launch_bounds__(1024) void star_rewire(int D, int k, int64_t m, int64_t n0,
                                                    const int32_t* __restrict__ nbh_idx,
                                                    const double* __restrict__ nbh_d,
                                                    const uint32_t* __restrict__ marked,
                                                    const int32_t* __restrict__ pair_rw,
                                                    const uint8_t* __restrict__ rw_ver,
                                                    const uint8_t* __restrict__ pend_ver,
                                                    const double* __restrict__ x64, double* cost, int32_t* parent,
                                                    int32_t* child, int32_t* next, int32_t* prev, int32_t* fa,
                                                    int32_t* fb, double* __restrict__ pend_rows, StarCounters* sc) {
    __shared__ RewireShared s;
    const int t = threadIdx.x;
    const int64_t total = m * (int64_t)k;
    const int64_t fcap = n0 + m;
    if (t == 0) {
        s.pos = sc->pos;
        s.pend = sc->pending;
        s.pair = sc->pend_pair;
        s.first = INT_MAX;
        s.bad = 0;
    }
    __syncthreads();
    if (s.pend) {
        // costs are as they were when the pass stopped at this pair, so its
        // condition still holds; the host has checked its link since
        if (pend_ver[0])
            rewire_apply(s, s.pair, k, n0, fcap, nbh_idx, nbh_d, cost, parent, child, next, prev, fa, fb, sc);
        if (t == 0) {
            sc->pending = 0;
            s.pos = s.pair + 1;
        }
        __syncthreads();
    }
    for (;;) {
        const int64_t p0 = s.pos;
        if (p0 >= total || s.bad) break;
        const int64_t pair = p0 + t;
        bool act = false, unk = false;
        if (pair < total) {
            const int32_t nb = nbh_idx[pair];
            const int64_t j = pair / k;
            const int i = (int)(pair - j * k);
            if (nb >= 0 && !((marked[j] >> i) & 1u)) {
                const double nc = cost[n0 + j] + nbh_d[pair];
                if (nc < cost[nb]) {
                    const int32_t r = pair_rw[pair];
                    if (r >= 0) {
                        act = rw_ver[r] != 0;
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
        if (f == INT_MAX) {
            __syncthreads();
            if (t == 0) s.pos = p0 + 1024;
            __syncthreads();
            continue;
        }
        if (t == f) s.unknown = unk ? 1 : 0;
        __syncthreads();
        if (s.unknown) {
            if (t == 0) {
                const int64_t pr = p0 + f;
                const int64_t id = n0 + pr / k;
                const int32_t nb = nbh_idx[pr];
                for (int i = 0; i < D; ++i) {
                    pend_rows[i] = x64[id * D + i];
                    pend_rows[D + i] = x64[(int64_t)nb * D + i];
                }
                sc->pending = 1;
                sc->pend_pair = pr;
                sc->pos = pr;
                sc->deferred += 1;
                sc->rw_links += 1;
            }
            return;
        }
        rewire_apply(s, p0 + f, k, n0, fcap, nbh_idx, nbh_d, cost, parent, child, next, prev, fa, fb, sc);
        if (t == 0) {
            s.pos = p0 + f + 1;
            s.first = INT_MAX;
        }
        __syncthreads();
    }
    if (t == 0) {
        sc->pos = total;
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

static int star_reserve(mptg_ctx* ctx, int B) {
    PrrtState& pb = ctx->prrt;
    const int D = ctx->dim;
    const size_t E = (size_t)B * kStarMaxK;
    int rc;
    if ((rc = ensure(ctx, pb.dnear, sizeof(double) * B))) return rc;
    if ((rc = ensure(ctx, pb.sidx, sizeof(int32_t) * B))) return rc;
    if ((rc = ensure(ctx, pb.qs, sizeof(double) * D * B))) return rc;
    if ((rc = ensure(ctx, pb.nbh_idx, sizeof(int32_t) * E))) return rc;
    if ((rc = ensure(ctx, pb.nbh_d, sizeof(double) * E))) return rc;
    if ((rc = ensure(ctx, pb.marked, sizeof(uint32_t) * B))) return rc;
    if ((rc = ensure(ctx, pb.pc, sizeof(double) * B))) return rc;
    if ((rc = ensure(ctx, pb.cpar, sizeof(int32_t) * B))) return rc;
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
    // two pending rows, then verdict and checks of the pending edge
    if ((rc = ensure(ctx, pb.pend, sizeof(double) * 2 * D + 16))) return rc;
    if ((rc = ensure(ctx, pb.star_cnt, sizeof(StarCounters)))) return rc;
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
    StarCounters* dsc = (StarCounters*)s.star_cnt.p;
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
    if ((rc = ensure_pinned(ctx, sizeof(Counters) + sizeof(StarCounters)))) return rc;
    Counters* hc = (Counters*)ctx->pinned;
    StarCounters* hs = (StarCounters*)((char*)ctx->pinned + sizeof(Counters));
    auto read_counters = [&]() -> int {
        MPTG_HIP(ctx, hipMemcpyAsync(hc, dcnt, sizeof(Counters), hipMemcpyDeviceToHost, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(hs, dsc, sizeof(StarCounters), hipMemcpyDeviceToHost, ctx->stream));
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
        const int k = star_k(s, D, n0);
        if (k < 1 || k > kStarMaxK)
            return fail(ctx, MPTG_E_RANGE,
                        "PRRT* neighbour count " + std::to_string(k) + " is outside 1..32 at tree size " +
                            std::to_string(n0) + " (set a fixed k)");
        if ((rc = tree_reserve(ctx, n0 + B))) return rc;
        if ((rc = ensure(ctx, s.front, sizeof(int32_t) * 2 * (size_t)(n0 + B)))) return rc;
        const int64_t it = s.iteration;
        // ---- extend
        hipLaunchKernelGGL(prrt_sample, dim3(sample_blocks(ctx->metric, D, B)), dim3(kBlock), 0, ctx->stream, pd,
                           ctx->metric, D, s.p.seed, s.p.goal_bias, it, B, (const Counters*)dcnt, (double*)s.q.p);
        MPTG_HIP(ctx, hipGetLastError());
        if ((rc = launch_knn(ctx, (const double*)s.q.p, B, 1, __builtin_inf(), (int32_t*)s.near.p,
                             (double*)s.ndist.p, n0)))
            return rc;
        MPTG_HIP(ctx, hipMemcpyAsync(s.dnear.p, s.ndist.p, sizeof(double) * B, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        dispatch_dim(D, [&]<int DD>() {
            launch_steer<DD>(ctx, s.p.range, B, (const double*)s.q.p, (const int32_t*)s.near.p,
                             (const double*)s.ndist.p, (double*)s.from.p, (double*)s.to.p, (uint8_t*)s.status.p,
                             (uint8_t*)s.mask.p, (double*)s.dnear.p, (const Counters*)dcnt);
        });
        MPTG_HIP(ctx, hipGetLastError());
        if (ctx->scene.kind == MPTG_SCENE_MESH &&
            (rc = launch_mesh_filter(ctx, (const double*)s.to.p, B, (uint8_t*)s.mask.p, (uint8_t*)s.status.p,
                                     ST_INVALID)))
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
            if ((rc = launch_knn(ctx, (const double*)s.qs.p, m, k, __builtin_inf(), (int32_t*)s.nbh_idx.p,
                                 (double*)s.nbh_d.p, n0)))
                return rc;
            const unsigned wb = (unsigned)((m + 3) / 4);
            hipLaunchKernelGGL(star_cp_prepare, dim3(wb), dim3(256), 0, ctx->stream, D, k, m,
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
            hipLaunchKernelGGL(star_cp_resolve, dim3(wb), dim3(256), 0, ctx->stream, D, k, m,
                               (const int32_t*)s.sidx.p, (const int32_t*)s.near.p, (const double*)s.dnear.p,
                               (const double*)s.qs.p, (const int32_t*)s.nbh_idx.p, (const double*)s.nbh_d.p,
                               (const double*)ctx->x64, (const double*)ctx->cost, (const int32_t*)s.pair_cp.p,
                               (const uint8_t*)s.cp_ver.p, (int32_t*)s.cpar.p, (double*)s.pc.p,
                               (uint32_t*)s.marked.p, (double*)s.rw_from.p, (double*)s.rw_to.p,
                               (int32_t*)s.pair_rw.p, dsc);
            MPTG_HIP(ctx, hipGetLastError());
            // ---- append
            hipLaunchKernelGGL(star_append, dim3(1), dim3(1024), 0, ctx->stream, space_args(ctx), pd,
                               s.p.goal_radius, m, n0, dp, (const double*)s.qs.p, (const int32_t*)s.cpar.p,
                               (const double*)s.pc.p, ctx->x64, ctx->x32, ctx->parent, ctx->cost, ctx->child,
                               ctx->next, ctx->prev, dcnt, (int32_t*)s.goal_buf.p, (int64_t)4096);
            MPTG_HIP(ctx, hipGetLastError());
            if ((rc = read_counters())) return rc;
            // ---- rewire
            if (hs->nrw > 0 &&
                (rc = launch_link(ctx, (const double*)s.rw_from.p, (const double*)s.rw_to.p, hs->nrw, s.p.order,
                                  1, (uint8_t*)s.rw_ver.p, (uint32_t*)s.rw_chk.p)))
                return rc;
            int32_t* front = (int32_t*)s.front.p;
            double* pend_rows = (double*)s.pend.p;
            uint8_t* pend_ver = (uint8_t*)(pend_rows + 2 * D);
            uint32_t* pend_chk = (uint32_t*)(pend_ver + 8);
            for (;;) {
                hipLaunchKernelGGL(star_rewire, dim3(1), dim3(1024), 0, ctx->stream, D, k, m, n0,
                                   (const int32_t*)s.nbh_idx.p, (const double*)s.nbh_d.p,
                                   (const uint32_t*)s.marked.p, (const int32_t*)s.pair_rw.p,
                                   (const uint8_t*)s.rw_ver.p, (const uint8_t*)pend_ver, (const double*)ctx->x64,
                                   ctx->cost, ctx->parent, ctx->child, ctx->next, ctx->prev, front,
                                   front + (n0 + B), pend_rows, dsc);
                MPTG_HIP(ctx, hipGetLastError());
                if ((rc = read_counters())) return rc;
                if (hs->bad_distance == 2) return fail(ctx, MPTG_E_STATE, "PRRT* child lists are corrupt");
                if (!hs->pending) break;
                if ((rc = launch_link(ctx, pend_rows, pend_rows + D, 1, s.p.order, 1, pend_ver, pend_chk)))
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
    if (sp->k < 0 || sp->k > kStarMaxK) return fail(ctx, MPTG_E_RANGE, "PRRT* k must be in 0..32");
    if (!(sp->rewire_factor > 0) || !std::isfinite(sp->rewire_factor))
        return fail(ctx, MPTG_E_ARG, "rewire factor must be a positive finite number");
    int rc = mptg_prrt_init(ctx, p, h_lo, h_hi, h_goal);
    if (rc) return rc;
    PrrtState& s = ctx->prrt;
    s.star = true;
    s.star_k = sp->k;
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
