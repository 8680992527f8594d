// Every tree node within a radius of each query, nearest first: nigh's
// nearest(nbh, q, SIZE_MAX, r), the call behind rewire_r_nearest
// (src/mpt/impl/rrg_rewire_neighbors.hpp:124-128).  The lists have no length
// limit, so they are built in two passes over the tree:
//   1. rnn_pass<count>: per (query, tree chunk) the number of nodes within r;
//   2. an exclusive scan of those counts (query-major, chunks in order
//      within a query) gives every (query, chunk) its first slot and every
//      query its list offset; only the total comes back to the host;
//   3. rnn_pass<fill> writes (distance, id) pairs in node order, and
//   4. a stable segmented radix sort on the FP64 distance orders each
//      query's pairs by (distance, id): ties keep the fill's node order.
// Distances are the metric's FP64 value (L1/L2 sums left to right, SE(3) as
// the other kernels); for L1/L2 an FP32 prefilter with the proven margin of
// the nearest-neighbour scans skips the exact evaluation of far nodes.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <string>

#include "space_dev.hpp"

namespace mptg {
namespace {

constexpr int kBlock = 256;
constexpr int64_t kRnnChunk = 4096;       // tree nodes per block row of a pass (at least)
constexpr int64_t kRnnSlots = 1ll << 25;  // (query, chunk) counters per call (at most)
constexpr int kRnnMaxChunks = 65535;      // grid rows

// |d32 - d| bound of the FP32 prefilter (as the nearest-neighbour scans):
// u = 2^-24, S = max|q| + max|p|, eps = 2.01 u S, g = D u / (1 - D u); L1:
// D eps + g (d + D eps); L2 (squared units): 2 eps sqrt(D T) + D eps^2 plus
// g of that and T.  Returns the FP32 threshold for the exact radius r.
template <int D, int METRIC>
__device__ __forceinline__ float rnn_threshold(double r, double S) {
    if (!(r < __builtin_inf())) return __builtin_inff();
    const double u = 0x1.0p-24;
    const double eps = 2.01 * u * S;
    const double g = D * u / (1.0 - D * u);
    double T, eta;
    if constexpr (METRIC == MPTG_METRIC_L1) {
        T = r;
        eta = D * eps + g * (T + D * eps);
    } else {
        T = r * r;
        const double e2 = 2.0 * eps * __builtin_sqrt(D * T) + D * eps * eps;
        eta = e2 + g * (T + e2);
    }
    const double thr = T * (1.0 + 0x1.0p-40) + 2.0 * eta + 1e-300;
    float f = (float)thr;
    // round up (thr > 0): the FP32 comparison never undercuts the FP64 bound
    if ((double)f < thr) f = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, f) + 1u);
    return f;
}

template <int D, int METRIC>
__device__ __forceinline__ double rnn_exact(const double* a, const double* b) {
    double s = 0.0;
    if constexpr (METRIC == MPTG_METRIC_L1) {
#pragma unroll
        for (int i = 0; i < D; ++i) s = s + __builtin_fabs(a[i] - b[i]);
        return s;
    } else {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double t = a[i] - b[i];
            s = s + t * t;
        }
        return __builtin_sqrt(s);
    }
}

// Pass over the block row's chunk of the tree (chunk nodes from
// blockIdx.y * chunk): FILL = false counts into cnt[q * chunks + row], true
// writes pairs (and the query as owner) at off[q * chunks + row] onwards.
// One lane per query; every lane of a wave visits the same node (scalar row
// loads).
template <int D, int METRIC, bool FILL>
__global__ __launch_bounds__(kBlock) void rnn_pass(const float* __restrict__ x32, const double* __restrict__ x64,
                                                   int64_t n, int64_t chunk, const double* __restrict__ q, int64_t nq,
                                                   double r, double pbound, int64_t* __restrict__ cnt,
                                                   const int64_t* __restrict__ off, double* __restrict__ out_d,
                                                   int32_t* __restrict__ out_i, int32_t* __restrict__ owner) {
    constexpr int DP = (D <= 2) ? 2 : (D <= 4 ? 4 : (D <= 8 ? 8 : ((D + 3) / 4) * 4));
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = qi < nq;
    const int64_t qc = active ? qi : 0;
    const int64_t slot = qc * gridDim.y + blockIdx.y;
    const int64_t p0 = (int64_t)blockIdx.y * chunk;
    const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
    const double* qq = q + qc * D;
    float q32[D];
    double mq = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        q32[i] = (float)qq[i];
        mq = __builtin_fmax(mq, __builtin_fabs(qq[i]));
    }
    const float thr = active ? rnn_threshold<D, METRIC>(r, mq + pbound) : -1.0f;
    int64_t c = 0;
    int64_t w = FILL && active ? off[slot] : 0;
    for (int64_t p = p0; p < p1; ++p) {
        const float* row = x32 + p * DP;
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const float t = q32[i] - row[i];
            if constexpr (METRIC == MPTG_METRIC_L1) a = a + __builtin_fabsf(t);
            else a = __builtin_fmaf(t, t, a);
        }
        if (a <= thr) {
            const double d = rnn_exact<D, METRIC>(qq, x64 + p * D);
            if (d <= r) {
                if constexpr (FILL) {
                    out_d[w] = d;
                    out_i[w] = (int32_t)p;
                    if (owner) owner[w] = (int32_t)qi;
                    ++w;
                } else {
                    ++c;
                }
            }
        }
    }
    if (!FILL && active) cnt[slot] = c;
}

__device__ __forceinline__ double rnn_se3(const double* a, const double* b, double w_so3, double w_l2) {
    const double dot = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    const double ad = __builtin_fabs(dot);
    const double so3 = ad >= 1.0 ? 0.0 : pm::acos(ad);
    double s = 0.0;
    for (int i = 4; i < 7; ++i) {
        const double t = a[i] - b[i];
        s = s + t * t;
    }
    return w_so3 * so3 + w_l2 * __builtin_sqrt(s);
}

// SE(3): exact distances for every pair (no prefilter).
template <bool FILL>
__global__ __launch_bounds__(kBlock) void rnn_pass_se3(const double* __restrict__ x64, int64_t n, int64_t chunk,
                                                       const double* __restrict__ q, int64_t nq, double r,
                                                       double w_so3, double w_l2, int64_t* __restrict__ cnt,
                                                       const int64_t* __restrict__ off, double* __restrict__ out_d,
                                                       int32_t* __restrict__ out_i, int32_t* __restrict__ owner) {
    const int64_t qi = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (qi >= nq) return;
    const int64_t slot = qi * gridDim.y + blockIdx.y;
    const int64_t p0 = (int64_t)blockIdx.y * chunk;
    const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
    int64_t c = 0;
    int64_t w = FILL ? off[slot] : 0;
    for (int64_t p = p0; p < p1; ++p) {
        const double d = rnn_se3(q + qi * 7, x64 + p * 7, w_so3, w_l2);
        if (d <= r) {
            if constexpr (FILL) {
                out_d[w] = d;
                out_i[w] = (int32_t)p;
                if (owner) owner[w] = (int32_t)qi;
                ++w;
            } else {
                ++c;
            }
        }
    }
    if (!FILL) cnt[slot] = c;
}

// List offsets from the scanned (query, chunk) slots: seg[q] = off[q * chunks],
// seg[nq] = the total.
__global__ __launch_bounds__(kBlock) void rnn_segments(const int64_t* __restrict__ off,
                                                       const int64_t* __restrict__ cnt, int64_t nq, int64_t chunks,
                                                       int64_t* __restrict__ seg) {
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q < nq) seg[q] = off[q * chunks];
    else if (q == nq) seg[q] = off[nq * chunks - 1] + cnt[nq * chunks - 1];
}

template <bool FILL>
void rnn_launch(mptg_ctx* ctx, dim3 grid, int64_t chunk, const double* q, int64_t nq, double r, int64_t* cnt,
                const int64_t* off, double* od, int32_t* oi, int32_t* own) {
    if (ctx->metric == MPTG_METRIC_SE3) {
        hipLaunchKernelGGL((rnn_pass_se3<FILL>), grid, dim3(kBlock), 0, ctx->stream, ctx->x64, ctx->size, chunk, q,
                           nq, r, ctx->w_so3, ctx->w_l2, cnt, off, od, oi, own);
        return;
    }
    auto go = [&]<int D>() {
        if (ctx->metric == MPTG_METRIC_L1)
            hipLaunchKernelGGL((rnn_pass<D, MPTG_METRIC_L1, FILL>), grid, dim3(kBlock), 0, ctx->stream, ctx->x32,
                               ctx->x64, ctx->size, chunk, q, nq, r, ctx->coord_bound, cnt, off, od, oi, own);
        else
            hipLaunchKernelGGL((rnn_pass<D, MPTG_METRIC_L2, FILL>), grid, dim3(kBlock), 0, ctx->stream, ctx->x32,
                               ctx->x64, ctx->size, chunk, q, nq, r, ctx->coord_bound, cnt, off, od, oi, own);
    };
    switch (ctx->dim) {
        case 1: go.template operator()<1>(); break;
        case 2: go.template operator()<2>(); break;
        case 3: go.template operator()<3>(); break;
        case 4: go.template operator()<4>(); break;
        case 5: go.template operator()<5>(); break;
        case 6: go.template operator()<6>(); break;
        case 7: go.template operator()<7>(); break;
        case 8: go.template operator()<8>(); break;
        case 10: go.template operator()<10>(); break;
        case 12: go.template operator()<12>(); break;
        case 16: go.template operator()<16>(); break;
        case 20: go.template operator()<20>(); break;
        default: break;  // rejected by the caller
    }
}

// Tree chunking of one call: chunks (grid rows) and nodes per chunk.
struct RnnPlan {
    int64_t chunks, chunk;
};

RnnPlan rnn_plan(int64_t n, int64_t nq) {
    int64_t chunks = std::max<int64_t>(1, (n + kRnnChunk - 1) / kRnnChunk);
    chunks = std::min<int64_t>(chunks, std::max<int64_t>(1, kRnnSlots / std::max<int64_t>(nq, 1)));
    chunks = std::min<int64_t>(chunks, kRnnMaxChunks);
    const int64_t chunk = std::max<int64_t>(1, (n + chunks - 1) / chunks);
    return {chunks, chunk};
}

// rnn_buf layout: list offsets [nq + 1], scanned slots [nq * chunks],
// counts [nq * chunks], scan scratch.
struct RnnLayout {
    int64_t *seg, *off, *cnt;
    void* tmp;
    size_t tmp_bytes;
};

RnnLayout rnn_layout(void* base, int64_t nq, int64_t slots, size_t tmp_bytes) {
    RnnLayout l;
    l.seg = (int64_t*)base;
    l.off = l.seg + nq + 1;
    l.cnt = l.off + slots;
    l.tmp = (void*)(((uintptr_t)(l.cnt + slots) + 255) & ~(uintptr_t)255);
    l.tmp_bytes = tmp_bytes;
    return l;
}

}  // namespace

bool rnn_dim_ok(int d) { return (d >= 1 && d <= 8) || d == 10 || d == 12 || d == 16 || d == 20; }

int rnn_count(mptg_ctx* ctx, const double* d_q, int64_t nq, double r, int64_t* total) {
    int rc;
    const RnnPlan pl = rnn_plan(ctx->size, nq);
    const int64_t slots = nq * pl.chunks;
    size_t tmp_bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (const int64_t*)nullptr, (int64_t*)nullptr, (int)slots,
                                         ctx->stream) != hipSuccess)
        return fail(ctx, MPTG_E_HIP, "radius search: scan sizing failed");
    const size_t bytes = sizeof(int64_t) * (size_t)(nq + 1 + 2 * slots) + 256 + tmp_bytes;
    if ((rc = ensure(ctx, ctx->rnn_buf, bytes))) return rc;
    const RnnLayout l = rnn_layout(ctx->rnn_buf.p, nq, slots, tmp_bytes);
    const dim3 grid((unsigned)((nq + kBlock - 1) / kBlock), (unsigned)pl.chunks);
    rnn_launch<false>(ctx, grid, pl.chunk, d_q, nq, r, l.cnt, nullptr, nullptr, nullptr, nullptr);
    MPTG_HIP(ctx, hipGetLastError());
    size_t tb = l.tmp_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(l.tmp, tb, (const int64_t*)l.cnt, l.off, (int)slots, ctx->stream) !=
        hipSuccess)
        return fail(ctx, MPTG_E_HIP, "radius search: scan failed");
    hipLaunchKernelGGL(rnn_segments, dim3((unsigned)((nq + 1 + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       (const int64_t*)l.off, (const int64_t*)l.cnt, nq, pl.chunks, l.seg);
    MPTG_HIP(ctx, hipGetLastError());
    MPTG_HIP(ctx, hipMemcpyAsync(total, l.seg + nq, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

const int64_t* rnn_offsets(const mptg_ctx* ctx) { return (const int64_t*)ctx->rnn_buf.p; }

int rnn_fill(mptg_ctx* ctx, const double* d_q, int64_t nq, double r, int64_t total, double* d_dist, int32_t* d_idx,
             int32_t* d_owner) {
    if (total <= 0) return MPTG_OK;
    if (total > (int64_t)INT32_MAX) return fail(ctx, MPTG_E_RANGE, "more than 2^31 neighbours in one call");
    int rc;
    const RnnPlan pl = rnn_plan(ctx->size, nq);
    const RnnLayout l = rnn_layout(ctx->rnn_buf.p, nq, nq * pl.chunks, 0);
    // unsorted pairs, sort scratch
    size_t tmp_bytes = 0;
    if (hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tmp_bytes, (const double*)nullptr, (double*)nullptr,
                                                    (const int32_t*)nullptr, (int32_t*)nullptr, (int)total, (int)nq,
                                                    (const int64_t*)nullptr, (const int64_t*)nullptr, 0, 64,
                                                    ctx->stream) != hipSuccess)
        return fail(ctx, MPTG_E_HIP, "radius search: sort sizing failed");
    const size_t pairs = (size_t)total * (sizeof(double) + sizeof(int32_t));
    if ((rc = ensure(ctx, ctx->rnn_pairs, pairs + tmp_bytes + 512))) return rc;
    double* d0 = (double*)ctx->rnn_pairs.p;
    int32_t* i0 = (int32_t*)(d0 + total);
    void* tmp = (void*)(((uintptr_t)(i0 + total) + 255) & ~(uintptr_t)255);
    const dim3 grid((unsigned)((nq + kBlock - 1) / kBlock), (unsigned)pl.chunks);
    rnn_launch<true>(ctx, grid, pl.chunk, d_q, nq, r, nullptr, l.off, d0, i0, d_owner);
    MPTG_HIP(ctx, hipGetLastError());
    // stable: equal distances keep ascending ids
    if (hipcub::DeviceSegmentedRadixSort::SortPairs(tmp, tmp_bytes, d0, d_dist, i0, d_idx, (int)total, (int)nq,
                                                    l.seg, l.seg + 1, 0, 64, ctx->stream) != hipSuccess)
        return fail(ctx, MPTG_E_HIP, "radius search: sort failed");
    return MPTG_OK;
}

}  // namespace mptg

using namespace mptg;

extern "C" int mptg_rnn(mptg_ctx* ctx, const double* h_q, int64_t nq, double radius, int64_t cap,
                        int64_t* h_offsets, int32_t* h_idx, double* h_dist) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (nq < 0 || cap < 0 || !h_offsets || (nq > 0 && !h_q) || (cap > 0 && (!h_idx || !h_dist)))
        return fail(ctx, MPTG_E_ARG, "bad radius-search buffers");
    if (!(radius >= 0)) return fail(ctx, MPTG_E_ARG, "radius must be >= 0");
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called first");
    if (ctx->metric != MPTG_METRIC_SE3 && !rnn_dim_ok(ctx->dim))
        return fail(ctx, MPTG_E_ARG, "unsupported dimension for the radius search");
    if (nq > ((int64_t)1 << 24)) return fail(ctx, MPTG_E_RANGE, "at most 2^24 queries per call");
    h_offsets[0] = 0;
    if (nq == 0) return MPTG_OK;
    const int D = ctx->dim;
    const size_t qb = sizeof(double) * D * nq;
    if ((rc = ensure(ctx, ctx->rnn_q, qb))) return rc;
    double* dq = (double*)ctx->rnn_q.p;
    MPTG_HIP(ctx, hipMemcpyAsync(dq, h_q, qb, hipMemcpyHostToDevice, ctx->stream));
    int64_t total = 0;
    if ((rc = rnn_count(ctx, dq, nq, radius, &total))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(h_offsets, rnn_offsets(ctx), sizeof(int64_t) * (nq + 1), hipMemcpyDeviceToHost,
                                 ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (total > cap) return fail(ctx, MPTG_E_RANGE, "the neighbour lists need " + std::to_string(total) +
                                                        " entries (offsets are filled)");
    if (total == 0) return MPTG_OK;
    if (total > (int64_t)INT32_MAX) return fail(ctx, MPTG_E_RANGE, "more than 2^31 neighbours in one call");
    if ((rc = ensure(ctx, ctx->rnn_out, (size_t)total * (sizeof(double) + sizeof(int32_t))))) return rc;
    double* od = (double*)ctx->rnn_out.p;
    int32_t* oi = (int32_t*)(od + total);
    if ((rc = rnn_fill(ctx, dq, nq, radius, total, od, oi, nullptr))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(h_dist, od, sizeof(double) * total, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(h_idx, oi, sizeof(int32_t) * total, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}
