// Host scenes: the Scenario's own valid()/link() (src/mpt/scenario.hpp:93-130)
// called on the host for whole batches, while sampling, nearest neighbours,
// steering and the tree stay on the device.  The reference calls
// scenario_.valid(newState) and scenario_.link(near, newState) once per sample
// from each worker (src/mpt/impl/prrt/prrt.hpp:439,454-457); here every
// batch the planner needs decided is copied to the host, only its listed rows
// are passed to one callback, and the verdicts are copied back.
#include <string>
#include <vector>

#include "internal.hpp"

namespace mptg {

namespace {

int callback_failed(mptg_ctx* ctx, const char* which, int code) {
    return fail(ctx, MPTG_E_ARG, std::string("host scene ") + which + "() callback returned " + std::to_string(code));
}

}  // namespace

int host_valid_dev(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_ok) {
    const SceneDev& sc = ctx->scene;
    if (!sc.host_valid) return fail(ctx, MPTG_E_STATE, "host scene has no valid() callback");
    if (n <= 0) return MPTG_OK;
    const int D = ctx->dim;
    std::vector<double> q((size_t)n * D);
    std::vector<uint8_t> ok((size_t)n, 0);
    MPTG_HIP(ctx, hipMemcpyAsync(q.data(), d_q, sizeof(double) * q.size(), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int cb = sc.host_valid(sc.host_user, q.data(), n, ok.data());
    if (cb) return callback_failed(ctx, "valid", cb);
    for (auto& v : ok) v = v ? 1 : 0;
    MPTG_HIP(ctx, hipMemcpyAsync(d_ok, ok.data(), ok.size(), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int host_filter(mptg_ctx* ctx, const double* d_q, int64_t n, uint8_t* d_mask, uint8_t* d_status, uint8_t bad) {
    const SceneDev& sc = ctx->scene;
    if (!sc.host_valid) return fail(ctx, MPTG_E_STATE, "host scene has no valid() callback");
    if (n <= 0) return MPTG_OK;
    const int D = ctx->dim;
    std::vector<double> q((size_t)n * D);
    std::vector<uint8_t> mask((size_t)n), status((size_t)n);
    MPTG_HIP(ctx, hipMemcpyAsync(q.data(), d_q, sizeof(double) * q.size(), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(mask.data(), d_mask, mask.size(), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(status.data(), d_status, status.size(), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int64_t> rows;
    for (int64_t i = 0; i < n; ++i)
        if (mask[(size_t)i]) rows.push_back(i);
    if (rows.empty()) return MPTG_OK;
    std::vector<double> g(rows.size() * D);
    for (size_t r = 0; r < rows.size(); ++r)
        for (int j = 0; j < D; ++j) g[r * D + j] = q[(size_t)rows[r] * D + j];
    std::vector<uint8_t> ok(rows.size(), 0);
    const int cb = sc.host_valid(sc.host_user, g.data(), (int64_t)rows.size(), ok.data());
    if (cb) return callback_failed(ctx, "valid", cb);
    for (size_t r = 0; r < rows.size(); ++r)
        if (!ok[r]) {
            mask[(size_t)rows[r]] = 0;
            status[(size_t)rows[r]] = bad;
        }
    MPTG_HIP(ctx, hipMemcpyAsync(d_mask, mask.data(), mask.size(), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(d_status, status.data(), status.size(), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int host_link_dev(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, const uint8_t* d_mask,
                  uint8_t* d_verdict, uint32_t* d_checks) {
    const SceneDev& sc = ctx->scene;
    if (!sc.host_link) return fail(ctx, MPTG_E_STATE, "host scene has no link() callback");
    if (n <= 0) return MPTG_OK;
    const int D = ctx->dim;
    std::vector<double> a((size_t)n * D), b((size_t)n * D);
    std::vector<uint8_t> mask;
    MPTG_HIP(ctx, hipMemcpyAsync(a.data(), d_from, sizeof(double) * a.size(), hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(b.data(), d_to, sizeof(double) * b.size(), hipMemcpyDeviceToHost, ctx->stream));
    std::vector<uint8_t> verdict((size_t)n, 0);
    std::vector<uint32_t> checks((size_t)n, 0);
    if (d_mask) {
        // unlisted edges keep what the device holds
        mask.resize((size_t)n);
        MPTG_HIP(ctx, hipMemcpyAsync(mask.data(), d_mask, mask.size(), hipMemcpyDeviceToHost, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(verdict.data(), d_verdict, verdict.size(), hipMemcpyDeviceToHost, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(checks.data(), d_checks, sizeof(uint32_t) * checks.size(),
                                     hipMemcpyDeviceToHost, ctx->stream));
    }
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int64_t> rows;
    for (int64_t i = 0; i < n; ++i)
        if (!d_mask || mask[(size_t)i]) rows.push_back(i);
    if (!rows.empty()) {
        std::vector<double> ga(rows.size() * D), gb(rows.size() * D);
        for (size_t r = 0; r < rows.size(); ++r)
            for (int j = 0; j < D; ++j) {
                ga[r * D + j] = a[(size_t)rows[r] * D + j];
                gb[r * D + j] = b[(size_t)rows[r] * D + j];
            }
        std::vector<uint8_t> ok(rows.size(), 0);
        const int cb = sc.host_link(sc.host_user, ga.data(), gb.data(), (int64_t)rows.size(), ok.data());
        if (cb) return callback_failed(ctx, "link", cb);
        for (size_t r = 0; r < rows.size(); ++r) {
            verdict[(size_t)rows[r]] = ok[r] ? 1 : 0;
            checks[(size_t)rows[r]] = 0;
        }
    }
    MPTG_HIP(ctx, hipMemcpyAsync(d_verdict, verdict.data(), verdict.size(), hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(d_checks, checks.data(), sizeof(uint32_t) * checks.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

}  // namespace mptg

using namespace mptg;

extern "C" int mptg_scene_host(mptg_ctx* ctx, mptg_valid_fn valid_fn, mptg_link_fn link_fn, void* user) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (!valid_fn || !link_fn) return fail(ctx, MPTG_E_ARG, "host scenes need valid() and link() callbacks");
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called first");
    SceneDev& sc = ctx->scene;
    sc.kind = MPTG_SCENE_HOST;
    sc.host_valid = valid_fn;
    sc.host_link = link_fn;
    sc.host_user = user;
    ctx->valid_upto = 0;  // valid() changed: every node is checked again
    ctx->some_invalid = false;
    return MPTG_OK;
}
