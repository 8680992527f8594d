// Diagnostic entry point: evaluates the device's project math on a batch of
// arguments so tests can compare it bit-for-bit with the host build.
#include "internal.hpp"
#include <algorithm>

#include "pmath.hpp"

namespace mptg {
namespace {

__global__ void diag_math_kernel(int op, const double* __restrict__ x, int64_t n, double* __restrict__ y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = x[i];
    double r;
    switch (op) {
        case MPTG_DIAG_SIN: r = pm::sin(v); break;
        case MPTG_DIAG_COS: r = pm::cos(v); break;
        case MPTG_DIAG_ACOS: r = pm::acos(v); break;
        default: r = __builtin_sqrt(v); break;
    }
    y[i] = r;
}

// Plain device copy (16-byte vector loads and stores, grid-stride): the
// reference point for byte counters (read n bytes, write n bytes).
__global__ __launch_bounds__(256) void diag_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                        int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

}  // namespace
}  // namespace mptg

using namespace mptg;

extern "C" int mptg_diag_copy(mptg_ctx* ctx, void* d_dst, const void* d_src, uint64_t bytes) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (bytes % 16 || ((uintptr_t)d_dst | (uintptr_t)d_src) % 16 || (bytes && (!d_dst || !d_src)))
        return fail(ctx, MPTG_E_ARG, "copy needs 16-byte aligned buffers and a multiple of 16 bytes");
    const int64_t n16 = (int64_t)(bytes / 16);
    if (n16 == 0) return MPTG_OK;
    const unsigned g = (unsigned)std::min<int64_t>((n16 + 255) / 256, (int64_t)ctx->cus * 32);
    hipLaunchKernelGGL(diag_copy_kernel, dim3(g), dim3(256), 0, ctx->stream, (const uint4*)d_src, (uint4*)d_dst, n16);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

namespace mptg {
int launch_link_work(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order, int known_valid,
                     unsigned long long* d_work, uint8_t* d_verdict, uint32_t* d_checks);
}

extern "C" int mptg_link_work(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                              int endpoints_known_valid, uint64_t* h_work) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (n < 0 || !h_work || (n > 0 && (!d_from || !d_to))) return fail(ctx, MPTG_E_ARG, "bad buffers");
    if ((rc = ensure(ctx, ctx->scratch[3], 9 * sizeof(uint64_t)))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[2], (sizeof(uint32_t) + 1) * (size_t)std::max<int64_t>(n, 1)))) return rc;
    unsigned long long* w = (unsigned long long*)ctx->scratch[3].p;
    uint32_t* chk = (uint32_t*)ctx->scratch[2].p;
    uint8_t* ver = (uint8_t*)(chk + std::max<int64_t>(n, 1));
    if ((rc = launch_link_work(ctx, d_from, d_to, n, order, endpoints_known_valid, w, ver, chk))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(h_work, w, 9 * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

extern "C" int mptg_link_phases(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                                int endpoints_known_valid, uint8_t* d_verdict, uint32_t* d_checks, float* h_ms) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (!h_ms) return fail(ctx, MPTG_E_ARG, "null output");
    if (ctx->scene.kind != MPTG_SCENE_ARM || ctx->link_mode != MPTG_LINK_BISECT || ctx->force_generic)
        return fail(ctx, MPTG_E_STATE, "phase timing covers the arm scene's bisection link");
    for (hipEvent_t& e : ctx->phase_ev)
        if (!e) MPTG_HIP(ctx, hipEventCreate(&e));
    ctx->phase_timing = true;
    rc = launch_link(ctx, d_from, d_to, n, order, endpoints_known_valid, d_verdict, d_checks);
    ctx->phase_timing = false;
    if (rc) return rc;
    MPTG_HIP(ctx, hipEventSynchronize(ctx->phase_ev[5]));
    for (int i = 0; i < 5; ++i) MPTG_HIP(ctx, hipEventElapsedTime(&h_ms[i], ctx->phase_ev[i], ctx->phase_ev[i + 1]));
    return MPTG_OK;
}

extern "C" int mptg_diag_math(mptg_ctx* ctx, int op, const double* h_x, int64_t n, double* h_y) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (op < MPTG_DIAG_SIN || op > MPTG_DIAG_SQRT) return fail(ctx, MPTG_E_ARG, "unknown diagnostic op");
    if (n < 0 || (n > 0 && (!h_x || !h_y))) return fail(ctx, MPTG_E_ARG, "bad buffers");
    if (n == 0) return MPTG_OK;
    if ((rc = ensure(ctx, ctx->scratch[0], sizeof(double) * n))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[1], sizeof(double) * n))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scratch[0].p, h_x, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(diag_math_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, op,
                       (const double*)ctx->scratch[0].p, n, (double*)ctx->scratch[1].p);
    MPTG_HIP(ctx, hipGetLastError());
    MPTG_HIP(ctx, hipMemcpyAsync(h_y, ctx->scratch[1].p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}
