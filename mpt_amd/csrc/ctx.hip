// Context lifecycle, space/tree/scene state and the host-pointer entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "space_dev.hpp"

namespace mptg {

thread_local std::string g_create_err;
void prrt_free(mptg_ctx* ctx);
void surrogate_free(mptg_ctx* ctx);

namespace {
__global__ __launch_bounds__(256) void zero_u32_kernel(uint32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ __launch_bounds__(64) void zero_pair_kernel(uint32_t* __restrict__ a, uint32_t* __restrict__ b) {
    if (threadIdx.x == 0) *a = 0u;
    if (threadIdx.x == 1) *b = 0u;
}
__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
}  // namespace

int launch_zero_u32(mptg_ctx* ctx, uint32_t* p, int64_t n) {
    if (n <= 0) return MPTG_OK;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

int launch_zero_pair(mptg_ctx* ctx, uint32_t* a, uint32_t* b) {
    hipLaunchKernelGGL(zero_pair_kernel, dim3(1), dim3(64), 0, ctx->stream, a, b);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

int launch_fill_u32(mptg_ctx* ctx, uint32_t* p, int64_t n, uint32_t v) {
    if (n <= 0) return MPTG_OK;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n, v);
    MPTG_HIP(ctx, hipGetLastError());
    return MPTG_OK;
}

int ensure(mptg_ctx* ctx, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return MPTG_OK;
    if (ctx->capturing) return fail(ctx, MPTG_E_STATE, "device buffer would grow during graph capture");
    if (b.p) {
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        MPTG_HIP(ctx, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    ++ctx->epoch;
    size_t want = std::max<size_t>(bytes, 256);
    MPTG_HIP(ctx, hipMalloc(&b.p, want));
    b.bytes = want;
    return MPTG_OK;
}

int ensure_pinned(mptg_ctx* ctx, size_t bytes) {
    if (ctx->pinned_bytes >= bytes) return MPTG_OK;
    if (ctx->pinned) {
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        MPTG_HIP(ctx, hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 1 << 16);
    MPTG_HIP(ctx, hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
    ctx->pinned_bytes = want;
    return MPTG_OK;
}

int check_device(mptg_ctx* ctx) {
    if (!ctx) return fail(nullptr, MPTG_E_ARG, "null context");
    MPTG_HIP(ctx, hipSetDevice(ctx->device));
    return MPTG_OK;
}

int check_device_ready(mptg_ctx* ctx) {
    if (ctx) ++ctx->epoch;  // the entry point may change what a captured graph holds
    return check_device(ctx);
}

static void free_tree(mptg_ctx* ctx) {
    (void)hipFree(ctx->x64);
    (void)hipFree(ctx->x32);
    (void)hipFree(ctx->parent);
    (void)hipFree(ctx->cost);
    (void)hipFree(ctx->child);
    ctx->x64 = nullptr;
    ctx->x32 = nullptr;
    ctx->parent = nullptr;
    ctx->cost = nullptr;
    ctx->child = ctx->next = ctx->prev = nullptr;
    ctx->cap = 0;
}

static int grow_tree(mptg_ctx* ctx, int64_t cap) {
    if (cap <= ctx->cap) return MPTG_OK;
    ++ctx->epoch;
    int d = ctx->dim, dp = pitch32(d);
    double* x64 = nullptr;
    float* x32 = nullptr;
    int32_t* par = nullptr;
    double* cost = nullptr;
    int32_t* links = nullptr;
    MPTG_HIP(ctx, hipMalloc(&x64, sizeof(double) * d * cap));
    MPTG_HIP(ctx, hipMalloc(&x32, sizeof(float) * dp * cap));
    MPTG_HIP(ctx, hipMalloc(&par, sizeof(int32_t) * cap));
    MPTG_HIP(ctx, hipMalloc(&cost, sizeof(double) * cap));
    MPTG_HIP(ctx, hipMalloc(&links, sizeof(int32_t) * 3 * cap));
    // zero the FP32 pad columns once so every row is fully defined
    MPTG_HIP(ctx, hipMemsetAsync(x32, 0, sizeof(float) * dp * cap, ctx->stream));
    if (ctx->size > 0) {
        const size_t n = (size_t)ctx->size;
        MPTG_HIP(ctx, hipMemcpyAsync(x64, ctx->x64, sizeof(double) * d * n, hipMemcpyDeviceToDevice, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(x32, ctx->x32, sizeof(float) * dp * n, hipMemcpyDeviceToDevice, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(par, ctx->parent, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, ctx->stream));
        MPTG_HIP(ctx, hipMemcpyAsync(cost, ctx->cost, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        for (int l = 0; l < 3; ++l)
            MPTG_HIP(ctx, hipMemcpyAsync(links + l * cap, ctx->child + l * ctx->cap, sizeof(int32_t) * n,
                                         hipMemcpyDeviceToDevice, ctx->stream));
    }
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_tree(ctx);
    ctx->x64 = x64;
    ctx->x32 = x32;
    ctx->parent = par;
    ctx->cost = cost;
    ctx->child = links;
    ctx->next = links + cap;
    ctx->prev = links + 2 * cap;
    ctx->cap = cap;
    return MPTG_OK;
}

int tree_reserve(mptg_ctx* ctx, int64_t cap) {
    if (ctx->dim <= 0) return fail(ctx, MPTG_E_STATE, "mptg_space must be called before the tree is used");
    if (cap < 0 || cap > (int64_t)INT32_MAX) return fail(ctx, MPTG_E_RANGE, "tree capacity must be < 2^31");
    if (cap <= ctx->cap) return MPTG_OK;
    int64_t c = std::max<int64_t>(cap, std::min<int64_t>((int64_t)INT32_MAX, ctx->cap * 2));
    return grow_tree(ctx, c);
}

}  // namespace mptg

using namespace mptg;

extern "C" {

int mptg_create(int device, mptg_ctx** out) {
    if (!out) return fail(nullptr, MPTG_E_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, MPTG_E_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(nullptr, MPTG_E_ARG, "device index out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, MPTG_E_HIP, hipGetErrorString(e));
    mptg_ctx* ctx = new mptg_ctx();
    ctx->device = device;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(nullptr, MPTG_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cus = prop.multiProcessorCount;
    if (ctx->cus <= 0) ctx->cus = 256;
    *out = ctx;
    return MPTG_OK;
}

void mptg_destroy(mptg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->phase_ev)
        if (e) (void)hipEventDestroy(e);
    free_tree(ctx);
    (void)hipFree(ctx->scene.boxes);
    (void)hipFree(ctx->scene.grid);
    (void)hipFree(ctx->scene.lengths);
    (void)hipFree(ctx->scene.disks);
    (void)hipFree(ctx->scene.circles);
    (void)hipFree(ctx->scene.arm_disks.p);
    (void)hipFree(ctx->scene.arm_lim.p);
    (void)hipFree(ctx->scene.arm_index.p);
    (void)hipFree(ctx->scene.mesh_nodes.p);
    (void)hipFree(ctx->scene.mesh_tris.p);
    (void)hipFree(ctx->scene.mesh_robot.p);
    for (auto& b : ctx->scratch) (void)hipFree(b.p);
    (void)hipFree(ctx->dmv_table16.p);
    (void)hipFree(ctx->rnn_buf.p);
    (void)hipFree(ctx->rnn_pairs.p);
    (void)hipFree(ctx->rnn_q.p);
    (void)hipFree(ctx->rnn_out.p);
    for (auto& b : ctx->knn_sel) (void)hipFree(b.p);
    for (auto& b : ctx->knn_grid) (void)hipFree(b.p);
    prrt_free(ctx);
    surrogate_free(ctx);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* mptg_last_error(const mptg_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int mptg_sync(mptg_ctx* ctx) {
    int rc = check_device(ctx);
    if (rc) return rc;
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

void* mptg_stream(mptg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mptg_dev_alloc(mptg_ctx* ctx, uint64_t bytes, void** d_out) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (!d_out) return fail(ctx, MPTG_E_ARG, "null output pointer");
    *d_out = nullptr;
    MPTG_HIP(ctx, hipMalloc(d_out, bytes ? bytes : 1));
    return MPTG_OK;
}

int mptg_dev_free(mptg_ctx* ctx, void* d_ptr) {
    int rc = check_device(ctx);
    if (rc) return rc;
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MPTG_HIP(ctx, hipFree(d_ptr));
    return MPTG_OK;
}

int mptg_dev_upload(mptg_ctx* ctx, void* d_dst, const void* h_src, uint64_t bytes) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (bytes == 0) return MPTG_OK;
    MPTG_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int mptg_dev_download(mptg_ctx* ctx, void* h_dst, const void* d_src, uint64_t bytes) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (bytes == 0) return MPTG_OK;
    MPTG_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int mptg_dev_memset(mptg_ctx* ctx, void* d_dst, int value, uint64_t bytes) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (bytes == 0) return MPTG_OK;
    MPTG_HIP(ctx, hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    return MPTG_OK;
}

int mptg_timer_start(mptg_ctx* ctx) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (!ctx->ev_start) MPTG_HIP(ctx, hipEventCreate(&ctx->ev_start));
    if (!ctx->ev_stop) MPTG_HIP(ctx, hipEventCreate(&ctx->ev_stop));
    MPTG_HIP(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
    return MPTG_OK;
}

int mptg_timer_stop(mptg_ctx* ctx, float* ms) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (!ctx->ev_start || !ms) return fail(ctx, MPTG_E_STATE, "mptg_timer_start was not called");
    MPTG_HIP(ctx, hipEventRecord(ctx->ev_stop, ctx->stream));
    MPTG_HIP(ctx, hipEventSynchronize(ctx->ev_stop));
    MPTG_HIP(ctx, hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
    return MPTG_OK;
}

int mptg_space(mptg_ctx* ctx, int metric, int dim, double w0, double w1) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (metric != MPTG_METRIC_L2 && metric != MPTG_METRIC_L1 && metric != MPTG_METRIC_SE3)
        return fail(ctx, MPTG_E_ARG, "unknown metric");
    if (metric == MPTG_METRIC_SE3 && dim != 7) return fail(ctx, MPTG_E_ARG, "SE3 states have 7 coordinates");
    if (dim < 1 || dim > kMaxDim) return fail(ctx, MPTG_E_ARG, "dimension must be in 1..32");
    if (metric == MPTG_METRIC_SE3 && !(w0 > 0 && w1 > 0)) return fail(ctx, MPTG_E_ARG, "SE3 weights must be > 0");
    if (ctx->size > 0 && (dim != ctx->dim || metric != ctx->metric))
        return fail(ctx, MPTG_E_STATE, "cannot change the space of a non-empty tree");
    if (dim != ctx->dim && ctx->x64) {
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        free_tree(ctx);
    }
    ctx->metric = metric;
    ctx->dim = dim;
    ctx->w_so3 = metric == MPTG_METRIC_SE3 ? w0 : 1.0;
    ctx->w_l2 = metric == MPTG_METRIC_SE3 ? w1 : 1.0;
    return MPTG_OK;
}

int mptg_tree_reserve(mptg_ctx* ctx, int64_t capacity) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    return tree_reserve(ctx, capacity);
}

int mptg_tree_clear(mptg_ctx* ctx) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->size = 0;
    ctx->star_ok = true;
    ctx->valid_upto = 0;
    ctx->some_invalid = false;
    // the prefilter's coordinate bound: the planner's sampling box and goal
    // (every node a planner adds lies within them), grown again by inserts
    ctx->coord_bound = planner_coord_bound(ctx);
    ctx->prrt.iteration = 0;
    ctx->prrt.goals = 0;
    ctx->prrt.goal_ids.clear();
    ctx->prrt.goal_cost0.clear();
    ctx->prrt.goal_parent0.clear();
    ctx->prrt.stats = mptg_prrt_stats{};
    ctx->prrt.counters_valid = false;
    return MPTG_OK;
}

int mptg_tree_insert(mptg_ctx* ctx, const double* h_states, const int32_t* h_parents, int64_t n) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !h_states)) return fail(ctx, MPTG_E_ARG, "bad states");
    if (n == 0) return MPTG_OK;
    rc = tree_reserve(ctx, ctx->size + n);
    if (rc) return rc;
    int d = ctx->dim, dp = pitch32(d);
    std::vector<float> x32((size_t)n * dp, 0.0f);
    double bound = ctx->coord_bound;
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < d; ++j) {
            double v = h_states[i * d + j];
            if (!std::isfinite(v)) return fail(ctx, MPTG_E_ARG, "non-finite state coordinate");
            x32[i * dp + j] = (float)v;
            bound = std::max(bound, std::fabs(v));
        }
    std::vector<int32_t> par((size_t)n, -1);
    if (h_parents)
        for (int64_t i = 0; i < n; ++i) {
            if (h_parents[i] < -1 || h_parents[i] >= ctx->size + i)
                return fail(ctx, MPTG_E_ARG, "parent id must be -1 or an earlier node");
            par[i] = h_parents[i];
        }
    int64_t s = ctx->size;
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->x64 + s * d, h_states, sizeof(double) * d * n, hipMemcpyHostToDevice,
                                 ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->x32 + s * dp, x32.data(), sizeof(float) * dp * n, hipMemcpyHostToDevice,
                                 ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->parent + s, par.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                                 ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->size += n;
    ctx->coord_bound = bound;
    ctx->star_ok = false;  // no path costs or child lists for these nodes
    return MPTG_OK;
}

int64_t mptg_tree_size(const mptg_ctx* ctx) { return ctx ? ctx->size : -1; }

int mptg_tree_read(mptg_ctx* ctx, int64_t first, int64_t n, double* h_states, int32_t* h_parents) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (first < 0 || n < 0 || first + n > ctx->size) return fail(ctx, MPTG_E_RANGE, "read beyond tree size");
    if (n == 0) return MPTG_OK;
    int d = ctx->dim;
    if (h_states)
        MPTG_HIP(ctx, hipMemcpyAsync(h_states, ctx->x64 + first * d, sizeof(double) * d * n,
                                     hipMemcpyDeviceToHost, ctx->stream));
    if (h_parents)
        MPTG_HIP(ctx, hipMemcpyAsync(h_parents, ctx->parent + first, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

// ---- scenes ---------------------------------------------------------------
static int upload(mptg_ctx* ctx, void** dst, const void* src, size_t bytes) {
    if (*dst) {
        MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        MPTG_HIP(ctx, hipFree(*dst));
        *dst = nullptr;
    }
    if (bytes == 0) return MPTG_OK;
    MPTG_HIP(ctx, hipMalloc(dst, bytes));
    MPTG_HIP(ctx, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int mptg_scene_boxes(mptg_ctx* ctx, const double* h_boxes, int n) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !h_boxes)) return fail(ctx, MPTG_E_ARG, "bad boxes");
    if (n > 4096) return fail(ctx, MPTG_E_RANGE, "at most 4096 boxes");
    rc = upload(ctx, (void**)&ctx->scene.boxes, h_boxes, sizeof(double) * 4 * n);
    if (rc) return rc;
    ctx->scene.kind = MPTG_SCENE_BOXES;
    ctx->valid_upto = 0;  // valid() changed: every node is checked again
    ctx->some_invalid = false;
    ctx->scene.nboxes = n;
    return MPTG_OK;
}

int mptg_scene_shapes(mptg_ctx* ctx, const double* h_circles, int ncircles, const double* h_rects, int nrects) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (ncircles < 0 || nrects < 0 || (ncircles > 0 && !h_circles) || (nrects > 0 && !h_rects))
        return fail(ctx, MPTG_E_ARG, "bad circles or rectangles");
    if (ncircles > 4096 || nrects > 4096) return fail(ctx, MPTG_E_RANGE, "at most 4096 circles and 4096 rectangles");
    std::vector<double> c((size_t)4 * ncircles);
    for (int i = 0; i < ncircles; ++i) {
        const double r = h_circles[3 * i + 2];
        if (!std::isfinite(h_circles[3 * i]) || !std::isfinite(h_circles[3 * i + 1]) || !std::isfinite(r))
            return fail(ctx, MPTG_E_ARG, "circles must be finite");
        c[4 * i] = h_circles[3 * i];
        c[4 * i + 1] = h_circles[3 * i + 1];
        c[4 * i + 2] = r * r;  // radius_ * radius_, as Circle compares
        c[4 * i + 3] = r;
    }
    if ((rc = upload(ctx, (void**)&ctx->scene.circles, c.data(), sizeof(double) * c.size()))) return rc;
    if ((rc = upload(ctx, (void**)&ctx->scene.boxes, h_rects, sizeof(double) * 4 * nrects))) return rc;
    ctx->scene.kind = MPTG_SCENE_SHAPES;
    ctx->scene.ncircles = ncircles;
    ctx->scene.nboxes = nrects;
    ctx->valid_upto = 0;  // valid() changed: every node is checked again
    ctx->some_invalid = false;
    return MPTG_OK;
}

int mptg_scene_grid(mptg_ctx* ctx, const uint8_t* h_cells, int width, int height) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (width <= 0 || height <= 0 || !h_cells) return fail(ctx, MPTG_E_ARG, "bad grid");
    if ((int64_t)width * height > ((int64_t)1 << 31)) return fail(ctx, MPTG_E_RANGE, "grid larger than 2^31 cells");
    // one bit per cell in 8 x 8 tiles (the layout grid_valid reads); cells of
    // the last tile row / column beyond the grid stay 0 and are never read
    // (grid_valid rejects them by the bounds check first)
    const int tx = (width + 7) / 8, ty = (height + 7) / 8;
    std::vector<uint64_t> bits((size_t)tx * ty, 0);
    for (int y = 0; y < height; ++y) {
        const uint8_t* row = h_cells + (size_t)width * y;
        uint64_t* trow = bits.data() + (size_t)(y >> 3) * tx;
        const int sh = (y & 7) << 3;
        for (int x = 0; x < width; ++x)
            if (row[x]) trow[x >> 3] |= 1ull << (sh | (x & 7));
    }
    rc = upload(ctx, (void**)&ctx->scene.grid, bits.data(), bits.size() * sizeof(uint64_t));
    if (rc) return rc;
    ctx->scene.kind = MPTG_SCENE_GRID;
    ctx->valid_upto = 0;  // valid() changed: every node is checked again
    ctx->some_invalid = false;
    ctx->scene.width = width;
    ctx->scene.height = height;
    ctx->scene.grid_tx = tx;
    return MPTG_OK;
}

int mptg_scene_arm(mptg_ctx* ctx, const double* h_lengths, int nlinks, double robot_radius,
                   const double* h_disks, int ndisks) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (nlinks < 1 || nlinks > kMaxDim || !h_lengths) return fail(ctx, MPTG_E_ARG, "arm needs 1..32 links");
    if (ndisks < 0 || ndisks > 65536 || (ndisks > 0 && !h_disks)) return fail(ctx, MPTG_E_ARG, "bad disks");
    // reference order: [cx, cy, (r+R)^2, r+R]
    std::vector<double> dk((size_t)ndisks * 4);
    for (int i = 0; i < ndisks; ++i) {
        dk[4 * i + 0] = h_disks[3 * i + 0];
        dk[4 * i + 1] = h_disks[3 * i + 1];
        double rr = h_disks[3 * i + 2] + robot_radius;  // (radius_ + robotRadius), squared as the reference
        dk[4 * i + 2] = rr * rr;
        dk[4 * i + 3] = rr;
    }
    rc = upload(ctx, (void**)&ctx->scene.lengths, h_lengths, sizeof(double) * nlinks);
    if (rc) return rc;
    rc = upload(ctx, (void**)&ctx->scene.disks, dk.data(), sizeof(double) * 4 * ndisks);
    if (rc) return rc;
    // kernel layout: [cx, cy, (r+R)^2, r+R] and a per-link segment-length
    // bound L (1 + 2^-30) + 1e-12 for the far-disk skip
    // then per group of kArmGroup links a bound on its half-length,
    // S_g/2 (1 + 2^-30) + 1e-12
    constexpr int kGroup = 5;
    const int ngroups = (nlinks + kGroup - 1) / kGroup;
    std::vector<double> ad((size_t)std::max(1, ndisks) * 4, 0.0), lim((size_t)(nlinks + ngroups));
    for (int i = 0; i < ndisks; ++i) {
        double rt = h_disks[3 * i + 2] + robot_radius;
        ad[4 * i + 0] = h_disks[3 * i + 0];
        ad[4 * i + 1] = h_disks[3 * i + 1];
        ad[4 * i + 2] = rt * rt;
        ad[4 * i + 3] = rt;
    }
    // disk grid: about one disk per cell on average (at most kArmGridMax per
    // axis) over the disk centres' bounding box; disks stored by cell,
    // row-major, ties by index.  The verdict is an AND over disks, so their
    // order changes nothing.
    int gnx = 1, gny = 1;
    double gx0 = 0.0, gy0 = 0.0, gix = 0.0, giy = 0.0, rtmax = 0.0;
    if (ndisks > 0) {
        double xmin = ad[0], xmax = ad[0], ymin = ad[1], ymax = ad[1];
        for (int i = 0; i < ndisks; ++i) {
            xmin = std::min(xmin, ad[4 * (size_t)i]);
            xmax = std::max(xmax, ad[4 * (size_t)i]);
            ymin = std::min(ymin, ad[4 * (size_t)i + 1]);
            ymax = std::max(ymax, ad[4 * (size_t)i + 1]);
            rtmax = std::max(rtmax, ad[4 * (size_t)i + 3]);
        }
        const int side = std::max(1, std::min(kArmGridMax, (int)std::lround(std::sqrt((double)ndisks))));
        gnx = xmax > xmin ? side : 1;
        gny = ymax > ymin ? side : 1;
        gx0 = xmin;
        gy0 = ymin;
        gix = xmax > xmin ? (double)gnx / (xmax - xmin) : 0.0;
        giy = ymax > ymin ? (double)gny / (ymax - ymin) : 0.0;
    }
    auto key = [&](size_t i) {
        return arm_cell(ad[4 * i + 1], gy0, giy, gny) * gnx + arm_cell(ad[4 * i], gx0, gix, gnx);
    };
    {
        std::vector<int> ord((size_t)ndisks);
        for (int i = 0; i < ndisks; ++i) ord[(size_t)i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return key((size_t)x) < key((size_t)y); });
        std::vector<double> sorted(ad.size(), 0.0);
        for (int i = 0; i < ndisks; ++i)
            for (int k = 0; k < 4; ++k) sorted[4 * (size_t)i + k] = ad[4 * (size_t)ord[(size_t)i] + k];
        ad.swap(sorted);
    }
    const int ncells = gnx * gny;
    std::vector<double> index(8 + (ncells + 1 + 1) / 2, 0.0);
    {
        index[0] = gx0;
        index[1] = gy0;
        index[2] = gix;
        index[3] = giy;
        index[4] = rtmax;
        index[5] = gnx;
        index[6] = gny;
        int32_t* starts = reinterpret_cast<int32_t*>(index.data() + 8);
        int j = 0;
        for (int c = 0; c <= ncells; ++c) {
            while (j < ndisks && key((size_t)j) < c) ++j;
            starts[c] = j;
        }
    }
    if ((rc = ensure(ctx, ctx->scene.arm_index, sizeof(double) * index.size()))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scene.arm_index.p, index.data(), sizeof(double) * index.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    for (int i = 0; i < nlinks; ++i) lim[i] = h_lengths[i] * (1.0 + 0x1.0p-30) + 1e-12;
    for (int g = 0; g < ngroups; ++g) {
        double S = 0.0;
        for (int i = g * kGroup; i < (g + 1) * kGroup && i < nlinks; ++i) S += h_lengths[i];
        lim[nlinks + g] = S / 2 * (1.0 + 0x1.0p-30) + 1e-12;
    }
    if ((rc = ensure(ctx, ctx->scene.arm_disks, sizeof(double) * ad.size()))) return rc;
    if ((rc = ensure(ctx, ctx->scene.arm_lim, sizeof(double) * lim.size()))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scene.arm_disks.p, ad.data(), sizeof(double) * ad.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scene.arm_lim.p, lim.data(), sizeof(double) * lim.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->scene.kind = MPTG_SCENE_ARM;
    ctx->valid_upto = 0;  // valid() changed: every node is checked again
    ctx->some_invalid = false;
    ctx->scene.total_length = 0;
    for (int i = 0; i < nlinks; ++i) ctx->scene.total_length += h_lengths[i];
    ctx->scene.nlinks = nlinks;
    ctx->scene.ndisks = ndisks;
    ctx->scene.robot_radius = robot_radius;
    return MPTG_OK;
}

int mptg_set_link(mptg_ctx* ctx, int mode, double param) {
    int rc = check_device_ready(ctx);
    if (rc) return rc;
    if (mode != MPTG_LINK_DMV && mode != MPTG_LINK_BISECT && mode != MPTG_LINK_BISECT_L2)
        return fail(ctx, MPTG_E_ARG, "unknown link mode");
    if (!(param > 0) || !std::isfinite(param)) return fail(ctx, MPTG_E_ARG, "link parameter must be > 0");
    ctx->link_mode = mode;
    ctx->link_param = param;
    return MPTG_OK;
}

// ---- host-pointer wrappers --------------------------------------------------
int mptg_knn(mptg_ctx* ctx, const double* h_q, int64_t nq, int k, double radius, int32_t* h_idx,
             double* h_dist) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (nq < 0 || (nq > 0 && (!h_q || !h_idx || !h_dist))) return fail(ctx, MPTG_E_ARG, "bad knn buffers");
    if (k < 1 || k > kMaxK) return fail(ctx, MPTG_E_ARG, "k must be in 1..64");
    if (nq == 0) return MPTG_OK;
    int d = ctx->dim;
    size_t qb = sizeof(double) * d * nq, ib = sizeof(int32_t) * k * nq, db = sizeof(double) * k * nq;
    if ((rc = ensure(ctx, ctx->scratch[0], qb))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[1], ib))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[2], db))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scratch[0].p, h_q, qb, hipMemcpyHostToDevice, ctx->stream));
    rc = launch_knn(ctx, (const double*)ctx->scratch[0].p, nq, k, radius, (int32_t*)ctx->scratch[1].p,
                    (double*)ctx->scratch[2].p, ctx->size);
    if (rc) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(h_idx, ctx->scratch[1].p, ib, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(h_dist, ctx->scratch[2].p, db, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int mptg_knn_dev(mptg_ctx* ctx, const double* d_q, int64_t nq, int k, double radius, int32_t* d_idx,
                 double* d_dist) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (nq < 0 || (nq > 0 && (!d_q || !d_idx || !d_dist))) return fail(ctx, MPTG_E_ARG, "bad knn buffers");
    if (k < 1 || k > kMaxK) return fail(ctx, MPTG_E_ARG, "k must be in 1..64");
    if (nq == 0) return MPTG_OK;
    return launch_knn(ctx, d_q, nq, k, radius, d_idx, d_dist, ctx->size);
}

int mptg_valid(mptg_ctx* ctx, const double* h_q, int64_t n, uint8_t* h_ok) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!h_q || !h_ok))) return fail(ctx, MPTG_E_ARG, "bad valid buffers");
    if (n == 0) return MPTG_OK;
    size_t qb = sizeof(double) * ctx->dim * n;
    if ((rc = ensure(ctx, ctx->scratch[0], qb))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[1], n))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scratch[0].p, h_q, qb, hipMemcpyHostToDevice, ctx->stream));
    rc = launch_valid(ctx, (const double*)ctx->scratch[0].p, n, (uint8_t*)ctx->scratch[1].p);
    if (rc) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(h_ok, ctx->scratch[1].p, n, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return MPTG_OK;
}

int mptg_link(mptg_ctx* ctx, const double* h_from, const double* h_to, int64_t n, int order,
              int endpoints_known_valid, uint8_t* h_verdict, uint32_t* h_checks) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!h_from || !h_to || !h_verdict || !h_checks)))
        return fail(ctx, MPTG_E_ARG, "bad link buffers");
    if (n == 0) return MPTG_OK;
    size_t qb = sizeof(double) * ctx->dim * n;
    if ((rc = ensure(ctx, ctx->scratch[0], qb))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[1], qb))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[2], n))) return rc;
    if ((rc = ensure(ctx, ctx->scratch[3], sizeof(uint32_t) * n))) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scratch[0].p, h_from, qb, hipMemcpyHostToDevice, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(ctx->scratch[1].p, h_to, qb, hipMemcpyHostToDevice, ctx->stream));
    rc = launch_link(ctx, (const double*)ctx->scratch[0].p, (const double*)ctx->scratch[1].p, n, order,
                     endpoints_known_valid, (uint8_t*)ctx->scratch[2].p, (uint32_t*)ctx->scratch[3].p);
    if (rc) return rc;
    MPTG_HIP(ctx, hipMemcpyAsync(h_verdict, ctx->scratch[2].p, n, hipMemcpyDeviceToHost, ctx->stream));
    MPTG_HIP(ctx, hipMemcpyAsync(h_checks, ctx->scratch[3].p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                                 ctx->stream));
    MPTG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < n; ++i)
        if (h_checks[i] == MPTG_CHECKS_OVER_LIMIT)
            return fail(ctx, MPTG_E_RANGE, "an edge exceeds the link() work limits (2^24 waypoints / 30 levels)");
    return MPTG_OK;
}

int mptg_link_dev(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                  int endpoints_known_valid, uint8_t* d_verdict, uint32_t* d_checks) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!d_from || !d_to || !d_verdict || !d_checks)))
        return fail(ctx, MPTG_E_ARG, "bad link buffers");
    if (n == 0) return MPTG_OK;
    return launch_link(ctx, d_from, d_to, n, order, endpoints_known_valid, d_verdict, d_checks);
}

int mptg_link_dev_noalloc(mptg_ctx* ctx, const double* d_from, const double* d_to, int64_t n, int order,
                          int endpoints_known_valid, uint8_t* d_verdict, uint32_t* d_checks) {
    int rc = check_device(ctx);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!d_from || !d_to || !d_verdict || !d_checks)))
        return fail(ctx, MPTG_E_ARG, "bad link buffers");
    return link_dev_noalloc(ctx, d_from, d_to, n, order, endpoints_known_valid, d_verdict, d_checks);
}

}  // extern "C"
