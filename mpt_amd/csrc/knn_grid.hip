// Uniform-grid index for k-nearest-neighbour queries over large trees in up
// to 8 dimensions (sub-linear in the tree size; the reference's nigh KD-tree
// plays this role, src/mpt/impl/prrt/prrt.hpp:78-80,406-409).  Built per
// call: a handful of O(n) kernels.
//
// Grid: g cells per axis over [-B, B]^D, B = the context's bound on every
// tree coordinate, g chosen for about 32 nodes per cell.  A node's cell is
// clamp(floor((x_i + B) * g / 2B), 0, g - 1) per axis, from its FP64
// coordinates; nodes are counting-sorted by cell into one FP64 column per
// axis plus node ids, and queries are counting-sorted by cell too.
//
// Query (one wave per query, exact throughout):
//   1. the query's own cell: every node's exact FP64 distance (lanes take
//      consecutive nodes), into the wave's k-list (one entry per lane);
//   2. with k entries, the k-th distance r bounds the answer: every node of
//      the answer lies within r of the query on each axis, so in the cells
//      covering [q - r, q + r]; each of those cells is scanned unless the
//      exact distance from the query to the cell's box exceeds the current
//      k-th distance (the box widened by 1e-12 of the grid's extent, which
//      dominates the rounding of the cell assignment; the bound lowered by a
//      relative 1e-12, which dominates its own rounding);
//   3. queries whose own cell holds fewer than k nodes within the radius
//      (and no finite radius bounds them) are listed for the scans.
// Visiting order does not matter: the k-list keeps (distance, node id) order,
// so results equal the brute-force scan's bit for bit.
#include <algorithm>
#include <cmath>
#include <limits>

#include "internal.hpp"
#include "space_dev.hpp"
#include "wave_list.hpp"

namespace mptg {
namespace {

constexpr int kB = 256;
constexpr int kGridMaxDim = 8;
constexpr int kScanBlock = 1024;

struct GridArgs {
    double x0;     // -B
    double inv;    // g / 2B
    double width;  // 2B / g
    double margin; // 1e-12 * 2B
    int g;
    int64_t ncells;
};

__device__ __forceinline__ int grid_cell(double x, const GridArgs& ga) {
    // monotone in x; NaN goes to cell 0
    const double t = (x - ga.x0) * ga.inv;
    return !(t >= 0.0) ? 0 : (t >= (double)ga.g ? ga.g - 1 : (int)t);
}

template <int D>
__global__ __launch_bounds__(kB) void grid_assign(const double* __restrict__ x64, int64_t n, GridArgs ga,
                                                  uint32_t* __restrict__ cell, uint32_t* __restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    int64_t c = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        c += (int64_t)grid_cell(x64[i * D + d], ga) * stride;
        stride *= ga.g;
    }
    cell[i] = (uint32_t)c;
    atomicAdd(&count[c], 1u);
}

// Exclusive prefix sums of count[0..m) in three launches: per-block sums and
// local prefixes, a single-block scan of the block sums, the block offsets.
__global__ __launch_bounds__(kScanBlock) void scan_local(const uint32_t* __restrict__ in, int64_t m,
                                                         uint32_t* __restrict__ out, uint32_t* __restrict__ sums) {
    __shared__ uint32_t ws[kScanBlock / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + t;
    const uint32_t v = i < m ? in[i] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int k = 0; k < kScanBlock / 64; ++k) {
        if (k < w) base += ws[k];
        tot += ws[k];
    }
    if (i < m) out[i] = base + x - v;
    if (t == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void scan_sums(uint32_t* __restrict__ sums, int nb) {
    // nb <= kScanBlock * 4: each thread owns up to 4 consecutive sums
    __shared__ uint32_t part[kScanBlock];
    const int t = threadIdx.x;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = t * 4 + j;
        v[j] = i < nb ? sums[i] : 0u;
        s += v[j];
    }
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < kScanBlock; o <<= 1) {  // Hillis-Steele over the thread totals
        const uint32_t y = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int i = t * 4 + j;
        if (i < nb) sums[i] = run;
        run += v[j];
    }
}

__global__ __launch_bounds__(kScanBlock) void scan_add(uint32_t* __restrict__ out, int64_t m,
                                                       const uint32_t* __restrict__ sums) {
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    if (i < m) out[i] += sums[blockIdx.x];
}

template <int D>
__global__ __launch_bounds__(kB) void grid_scatter(const double* __restrict__ x64, int64_t n,
                                                   const uint32_t* __restrict__ cell,
                                                   const uint32_t* __restrict__ start, uint32_t* __restrict__ fill,
                                                   double* __restrict__ sx, int32_t* __restrict__ sid) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cell[i];
    const int64_t pos = (int64_t)start[c] + atomicAdd(&fill[c], 1u);
#pragma unroll
    for (int d = 0; d < D; ++d) sx[d * n + pos] = x64[i * D + d];  // one column per axis
    sid[pos] = (int32_t)i;
}

// Every node of cell c (wave-uniform) within the radius that beats the k-th
// best, into the list; lanes take consecutive nodes (contiguous rows).  For
// L2 the square root is taken only for sums that can reach the k-th
// distance: sqrt(s) <= t implies s <= t^2 (1 + 2^-51); the test keeps 2^-40.
template <int D, int METRIC>
__device__ __forceinline__ void wave_scan_cell(const double* __restrict__ sx, int64_t n,
                                               const int32_t* __restrict__ sid, const uint32_t* __restrict__ start,
                                               uint32_t c, const double* qq, double radius, WaveList& wl, int lane) {
    const uint32_t j0 = start[c], j1 = start[c + 1];
    for (uint32_t base = j0; base < j1; base += 64) {
        const uint32_t j = base + (uint32_t)lane;
        bool pass = false;
        double dq = 0.0;
        int32_t id = 0;
        if (j < j1) {
            const double* col = sx + j;
            double s = 0.0;
            if constexpr (METRIC == MPTG_METRIC_L1) {
#pragma unroll
                for (int i = 0; i < D; ++i) s = s + __builtin_fabs(qq[i] - col[i * n]);
                dq = s;
                pass = dq <= radius && dq <= wl.kd;
            } else {
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    const double t = qq[i] - col[i * n];
                    s = s + t * t;
                }
                const double thr = __builtin_fmin(wl.kd, radius);
                if (thr < 1e-140 || s <= thr * thr * (1.0 + 0x1.0p-40)) {
                    dq = __builtin_sqrt(s);
                    pass = dq <= radius && dq <= wl.kd;
                }
            }
            if (pass) {
                id = sid[j];
                pass = wl_before(dq, id, wl.kd, wl.ki);
            }
        }
        wl.offer(pass, dq, id, lane);
    }
}

// One wave per query, queries in cell order (qorder: neighbouring waves
// share cells in cache).  The wave scans the query's own cell, which bounds
// the answer by the k-th distance r (or the radius); then it tests the
// cells of the box [q - r, q + r], 64 at a time, each lane one cell: a cell
// is scanned when the exact lower bound from the query to it does not exceed
// the current k-th distance.  A query left without k nodes within the radius
// after its own cell (infinite radius) is listed in `over` and writes
// nothing.
template <int D, int METRIC>
__global__ __launch_bounds__(kB) void grid_query(const double* __restrict__ sx, int64_t n,
                                                 const int32_t* __restrict__ sid,
                                                 const uint32_t* __restrict__ start, GridArgs ga,
                                                 const double* __restrict__ q, const uint32_t* __restrict__ qorder,
                                                 int64_t nq, int k, double radius, double inf,
                                                 double* __restrict__ out_d, int32_t* __restrict__ out_i,
                                                 int64_t* __restrict__ over, uint32_t* __restrict__ n_over) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * (kB / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (t >= nq) return;
    const int64_t qi = (int64_t)qorder[t];
    double qq[D];
    uint32_t own = 0, gstride[D];
    uint32_t stride = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        qq[d] = q[qi * D + d];
        own += (uint32_t)grid_cell(qq[d], ga) * stride;
        gstride[d] = stride;
        stride *= (uint32_t)ga.g;
    }
    own = (uint32_t)__builtin_amdgcn_readfirstlane((int)own);
    WaveList wl;
    wl.init(k, inf);
    wave_scan_cell<D, METRIC>(sx, n, sid, start, own, qq, radius, wl, lane);

    const double r = __builtin_fmin(wl.kd, radius);
    int lo[D], ext[D];
    int64_t cells = 1;
    if (r < inf) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            // a node within r has |x - q| <= r (1 + 2^-50) per axis; the
            // margin covers that and the rounding of q -+ r
            const double m = ga.margin + (__builtin_fabs(qq[d]) + r) * 4e-15;
            lo[d] = __builtin_amdgcn_readfirstlane(grid_cell(qq[d] - r - m, ga));
            ext[d] = __builtin_amdgcn_readfirstlane(grid_cell(qq[d] + r + m, ga)) - lo[d] + 1;
            cells *= ext[d];
        }
    }
    if (!(r < inf)) {
        if (lane == 0) over[atomicAdd(n_over, 1u)] = qi;
        return;
    }
    // lane l tests box cell it = base + l; cur holds its mixed-radix digits
    // and steps by 64 (digits inc) per round
    int cur0[D], inc[D];
    {
        int a = lane, b = 64;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            cur0[d] = a % ext[d];
            a /= ext[d];
            inc[d] = b % ext[d];
            b /= ext[d];
        }
    }
    // Two passes over the box: the first scans the cells of the first 64
    // chunks within half the current bound (the near ones, which tighten it
    // fastest), the second every other cell the bound does not exclude.
    // Lane ch keeps the mask of chunk ch's cells the first pass scanned.
    unsigned long long done = 0;
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        // L1: lb against f t; L2: lb (a squared sum) against f^2 t^2, with
        // relative slack 1e-12 either side for rounding; bounds below 1e-140
        // prune nothing
        const double f = pass == 0 ? 0.5 : 1.0;
        auto need = [&](double b) {
            const double thr = f * __builtin_fmin(wl.kd, radius);
            if constexpr (METRIC == MPTG_METRIC_L1) return b * (1.0 - 1e-12) <= thr;
            else return thr < 1e-140 || b * (1.0 - 1e-12) <= thr * thr * (1.0 + 1e-12);
        };
        int cur[D];
#pragma unroll
        for (int d = 0; d < D; ++d) cur[d] = cur0[d];
        const int64_t end = pass == 0 ? (cells < 64 * 64 ? cells : 64 * 64) : cells;
        int64_t ch = 0;
        for (int64_t base = 0; base < end; base += 64, ++ch) {
            uint32_t cc = 0;
            double lb = 0.0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int x = lo[d] + cur[d];
                cc += (uint32_t)x * gstride[d];
                // gap from the query to the cell's box widened by the margin
                const double blo = ga.x0 + (double)x * ga.width - ga.margin;
                const double bhi = ga.x0 + (double)(x + 1) * ga.width + ga.margin;
                const double gap = qq[d] < blo ? blo - qq[d] : (qq[d] > bhi ? qq[d] - bhi : 0.0);
                if constexpr (METRIC == MPTG_METRIC_L1) lb = lb + gap;
                else lb = lb + gap * gap;
            }
            const int cl = (int)ch;
            const unsigned long long skip =
                pass == 0 || ch >= 64
                    ? 0ull
                    : ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(done >> 32), cl) << 32) |
                          (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)done, cl);
            const bool want = base + lane < cells && cc != own && !((skip >> lane) & 1ull) && need(lb);
            unsigned long long scanned = 0;
            for (unsigned long long m = __ballot(want); m; m &= m - 1) {
                const int b = __builtin_ctzll(m);
                if (need(wl_readlane(lb, b))) {
                    wave_scan_cell<D, METRIC>(sx, n, sid, start, (uint32_t)__builtin_amdgcn_readlane((int)cc, b),
                                              qq, radius, wl, lane);
                    scanned |= 1ull << b;
                }
            }
            if (pass == 0 && lane == ch) done = scanned;
            int carry = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int v = cur[d] + inc[d] + carry;
                carry = v >= ext[d] ? 1 : 0;
                cur[d] = carry ? v - ext[d] : v;
            }
        }
    }
    if (lane < k) {
        const bool have = wl.li != INT_MAX;
        out_d[qi * k + lane] = have ? wl.ld : inf;
        out_i[qi * k + lane] = have ? wl.li : -1;
    }
}

template <int D>
__global__ __launch_bounds__(kB) void grid_assign_queries(const double* __restrict__ q, int64_t nq, GridArgs ga,
                                                          uint32_t* __restrict__ qcell, uint32_t* __restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= nq) return;
    uint32_t c = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        c += (uint32_t)grid_cell(q[i * D + d], ga) * stride;
        stride *= (uint32_t)ga.g;
    }
    qcell[i] = c;
    atomicAdd(&count[c], 1u);
}

__global__ __launch_bounds__(kB) void grid_order_queries(const uint32_t* __restrict__ qcell, int64_t nq,
                                                         const uint32_t* __restrict__ start,
                                                         uint32_t* __restrict__ fill, uint32_t* __restrict__ qorder) {
    const int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (i >= nq) return;
    const uint32_t c = qcell[i];
    qorder[start[c] + atomicAdd(&fill[c], 1u)] = (uint32_t)i;
}

// Exclusive prefix sums of count[0..m) into start (sums: nb block totals).
void launch_scan(mptg_ctx* ctx, const uint32_t* count, int64_t m, uint32_t* start, uint32_t* sums) {
    const unsigned nb = (unsigned)((m + kScanBlock - 1) / kScanBlock);
    hipLaunchKernelGGL(scan_local, dim3(nb), dim3(kScanBlock), 0, ctx->stream, count, m, start, sums);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanBlock), 0, ctx->stream, sums, (int)nb);
    hipLaunchKernelGGL(scan_add, dim3(nb), dim3(kScanBlock), 0, ctx->stream, start, m, (const uint32_t*)sums);
}

// Per-call arrays: node cells, node and query counting-sort arrays (count,
// start, fill of m = cells + 1 entries each, block sums), sorted node rows
// and ids, query cells and order.
struct GridBufs {
    uint32_t *cell, *count, *start, *fill, *sums, *qcount, *qstart, *qfill, *qsums, *qcell, *qorder;
    double* sx;
    int32_t* sid;
};

template <int D, int METRIC>
void launch_query(mptg_ctx* ctx, int64_t n, const GridBufs& gb, const GridArgs& ga, const double* q, int64_t nq,
                  int k, double radius, double* od, int32_t* oi, int64_t* over, uint32_t* n_over) {
    const unsigned g = (unsigned)((nq + kB / 64 - 1) / (kB / 64));
    hipLaunchKernelGGL((grid_query<D, METRIC>), dim3(g), dim3(kB), 0, ctx->stream, (const double*)gb.sx, n,
                       (const int32_t*)gb.sid, (const uint32_t*)gb.start, ga, q, (const uint32_t*)gb.qorder, nq, k,
                       radius, std::numeric_limits<double>::infinity(), od, oi, over, n_over);
}

template <int D>
void launch_build_query(mptg_ctx* ctx, int64_t n, const GridArgs& ga, const GridBufs& gb, const double* q,
                        int64_t nq, int k, double radius, double* od, int32_t* oi, int64_t* over, uint32_t* n_over) {
    const unsigned gn = (unsigned)((n + kB - 1) / kB), gq = (unsigned)((nq + kB - 1) / kB);
    const int64_t m = ga.ncells + 1;
    // nodes: cell, counting sort into rows ordered by cell
    hipLaunchKernelGGL((grid_assign<D>), dim3(gn), dim3(kB), 0, ctx->stream, (const double*)ctx->x64, n, ga, gb.cell,
                       gb.count);
    launch_scan(ctx, gb.count, m, gb.start, gb.sums);
    hipLaunchKernelGGL((grid_scatter<D>), dim3(gn), dim3(kB), 0, ctx->stream, (const double*)ctx->x64, n,
                       (const uint32_t*)gb.cell, (const uint32_t*)gb.start, gb.fill, gb.sx, gb.sid);
    // queries: the same cells, in cell order
    hipLaunchKernelGGL((grid_assign_queries<D>), dim3(gq), dim3(kB), 0, ctx->stream, q, nq, ga, gb.qcell, gb.qcount);
    launch_scan(ctx, gb.qcount, m, gb.qstart, gb.qsums);
    hipLaunchKernelGGL(grid_order_queries, dim3(gq), dim3(kB), 0, ctx->stream, (const uint32_t*)gb.qcell, nq,
                       (const uint32_t*)gb.qstart, gb.qfill, gb.qorder);
    if (ctx->metric == MPTG_METRIC_L1)
        launch_query<D, MPTG_METRIC_L1>(ctx, n, gb, ga, q, nq, k, radius, od, oi, over, n_over);
    else
        launch_query<D, MPTG_METRIC_L2>(ctx, n, gb, ga, q, nq, k, radius, od, oi, over, n_over);
}

}  // namespace

bool knn_grid_applies(const mptg_ctx* ctx, int64_t n) {
    return ctx->tune.knn_grid && ctx->dim >= 2 && ctx->dim <= kGridMaxDim && n >= ((int64_t)1 << 16) &&
           (ctx->metric == MPTG_METRIC_L1 || ctx->metric == MPTG_METRIC_L2) && ctx->coord_bound > 0 &&
           std::isfinite(ctx->coord_bound);
}

int knn_grid_path(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx,
                  double* d_dist, int64_t n, int64_t** over_out, uint32_t* m_over) {
    const int D = ctx->dim;
    // cells per axis: about 32 nodes per cell, fewer than 2^22 cells (the
    // prefix sum covers 2^22 entries)
    int g = std::max(2, (int)std::floor(std::pow((double)n / 32.0, 1.0 / D)));
    while (g > 2 && std::pow((double)g, (double)D) >= (double)(1 << 22)) --g;
    int64_t ncells = 1;
    for (int d = 0; d < D; ++d) ncells *= g;
    const double B = ctx->coord_bound;
    GridArgs ga{-B, g / (2.0 * B), 2.0 * B / g, 1e-12 * 2.0 * B, g, ncells};
    int rc;
    DevBuf* b = ctx->knn_grid;
    const int64_t m = ncells + 1;
    const int64_t nb = (m + kScanBlock - 1) / kScanBlock;
    if (nb > 4 * kScanBlock) return fail(ctx, MPTG_E_RANGE, "grid too fine");
    if (nq >= ((int64_t)1 << 32) || n >= ((int64_t)1 << 31)) return fail(ctx, MPTG_E_RANGE, "grid: too many points");
    if ((rc = ensure(ctx, b[0], sizeof(uint32_t) * (size_t)n))) return rc;
    if ((rc = ensure(ctx, b[1], sizeof(uint32_t) * (size_t)(6 * m + 2 * nb)))) return rc;
    if ((rc = ensure(ctx, b[2], (sizeof(double) * D + sizeof(int32_t)) * (size_t)n))) return rc;
    if ((rc = ensure(ctx, b[3], 256 + sizeof(int64_t) * (size_t)nq))) return rc;
    if ((rc = ensure(ctx, b[5], 2 * sizeof(uint32_t) * (size_t)nq))) return rc;
    uint32_t* n_over = (uint32_t*)b[3].p;
    int64_t* over = (int64_t*)((char*)b[3].p + 256);
    GridBufs gb;
    gb.cell = (uint32_t*)b[0].p;
    gb.count = (uint32_t*)b[1].p;
    gb.start = gb.count + m;
    gb.fill = gb.start + m;
    gb.qcount = gb.fill + m;
    gb.qstart = gb.qcount + m;
    gb.qfill = gb.qstart + m;
    gb.sums = gb.qfill + m;
    gb.qsums = gb.sums + nb;
    gb.sx = (double*)b[2].p;
    gb.sid = (int32_t*)(gb.sx + (size_t)n * D);
    gb.qcell = (uint32_t*)b[5].p;
    gb.qorder = gb.qcell + nq;
    // count and fill of nodes, then of queries: 3 m words each, contiguous
    if ((rc = launch_zero_u32(ctx, gb.count, m))) return rc;
    if ((rc = launch_zero_u32(ctx, gb.fill, 4 * m))) return rc;
    if ((rc = launch_zero_u32(ctx, n_over, 1))) return rc;
    switch (D) {
        case 2: launch_build_query<2>(ctx, n, ga, gb, d_q, nq, k, radius, d_dist, d_idx, over, n_over); break;
        case 3: launch_build_query<3>(ctx, n, ga, gb, d_q, nq, k, radius, d_dist, d_idx, over, n_over); break;
        case 4: launch_build_query<4>(ctx, n, ga, gb, d_q, nq, k, radius, d_dist, d_idx, over, n_over); break;
        case 5: launch_build_query<5>(ctx, n, ga, gb, d_q, nq, k, radius, d_dist, d_idx, over, n_over); break;
        case 6: launch_build_query<6>(ctx, n, ga, gb, d_q, nq, k, radius, d_dist, d_idx, over, n_over); break;
        case 7: launch_build_query<7>(ctx, n, ga, gb, d_q, nq, k, radius, d_dist, d_idx, over, n_over); break;
        default: launch_build_query<8>(ctx, n, ga, gb, d_q, nq, k, radius, d_dist, d_idx, over, n_over); break;
    }
    MPTG_HIP(ctx, hipGetLastError());
    MPTG_HIP(ctx, hipMemcpyAsync(m_over, n_over, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *over_out = over;
    return MPTG_OK;
}

}  // namespace mptg
