// nd_ctx.hip -- context, device memory, timers and specification checks.
#include "nd_internal.hpp"

namespace nd {

void set_error(nd_ctx *ctx, const std::string &m) {
    if (ctx) ctx->err = m;
}

void clear_status(nd_ctx *ctx) {
    ctx->status.reserve(64);
    launch_fill_i32(ctx->stream, ctx->status.as<int32_t>(), 16, 0);
}

int32_t read_status(nd_ctx *ctx) {
    int32_t st = 0;
    ND_HIP(hipMemcpyAsync(&st, ctx->status.p, sizeof(int32_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    ND_HIP(hipStreamSynchronize(ctx->stream));
    if (st) clear_status(ctx);
    return st;
}

}  // namespace nd

extern "C" {

int nd_version(void) { return 1; }

int nd_ctx_create(int32_t device, nd_ctx **out) {
    if (!out) return ND_ERR_ARG;
    *out = nullptr;
    nd_ctx *ctx = new nd_ctx();
    try {
        int count = 0;
        ND_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw nd::Error(ND_ERR_ARG, "nd_ctx_create: device index out of range");
        ctx->device = device;
        ND_HIP(hipSetDevice(device));
        ND_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        (void)nd::static_scratch(device);  // the device's scratch address, once per device
        ND_HIP(hipEventCreate(&ctx->ev0));
        ND_HIP(hipEventCreate(&ctx->ev1));
        nd::clear_status(ctx);
        ND_HIP(hipStreamSynchronize(ctx->stream));
    } catch (const nd::Error &e) {
        int code = e.code;
        delete ctx;
        return code;
    }
    *out = ctx;
    return ND_OK;
}

int nd_ctx_destroy(nd_ctx *ctx) {
    if (!ctx) return ND_OK;
    // teardown: errors here have no caller to report to
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    nd::DevBuf *bufs[] = {&ctx->status, &ctx->codes, &ctx->codes_sorted, &ctx->perm,
                          &ctx->xyz_sorted, &ctx->nodes, &ctx->store, &ctx->counts,
                          &ctx->offsets, &ctx->xyz_in, &ctx->pairs_out, &ctx->bbox,
                          &ctx->sort.keys_alt, &ctx->sort.vals_alt, &ctx->sort.hist,
                          &ctx->sort.flags, &ctx->cell_keys, &ctx->cell_vals, &ctx->cell_pos,
                          &ctx->cell_start, &ctx->trav_rec, &ctx->scan_tmp, &ctx->stage_store,
                          &ctx->rank_keys,
                          &ctx->force_acc, &ctx->force_part, &ctx->async_ws};
    for (auto *b : bufs) b->release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ND_OK;
}

const char *nd_last_error(const nd_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int nd_malloc(nd_ctx *ctx, size_t bytes, void **d_ptr) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && d_ptr, "nd_malloc: null argument");
    ND_HIP(hipSetDevice(ctx->device));
    *d_ptr = nullptr;
    ND_HIP(hipMalloc(d_ptr, bytes ? bytes : 256));
    ND_API_END(ctx)
}

int nd_free(nd_ctx *ctx, void *d_ptr) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_free: null context");
    if (d_ptr) {
        ND_HIP(hipStreamSynchronize(ctx->stream));
        ND_HIP(hipFree(d_ptr));
    }
    ND_API_END(ctx)
}

int nd_memcpy_h2d(nd_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_memcpy_h2d: null context");
    if (bytes) {
        ND_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
        ND_HIP(hipStreamSynchronize(ctx->stream));
    }
    ND_API_END(ctx)
}

int nd_memcpy_d2h(nd_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_memcpy_d2h: null context");
    if (bytes) {
        ND_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        ND_HIP(hipStreamSynchronize(ctx->stream));
    }
    ND_API_END(ctx)
}

int nd_memcpy_d2d(nd_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_memcpy_d2d: null context");
    if (bytes) ND_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    ND_API_END(ctx)
}

int nd_memset(nd_ctx *ctx, void *d_dst, int32_t byte_value, size_t bytes) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_memset: null context");
    if (bytes) ND_HIP(hipMemsetAsync(d_dst, byte_value, bytes, ctx->stream));
    ND_API_END(ctx)
}

int nd_ctx_stream(nd_ctx *ctx, void **stream) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && stream, "nd_ctx_stream: null argument");
    *stream = static_cast<void *>(ctx->stream);
    ND_API_END(ctx)
}

int nd_synchronize(nd_ctx *ctx) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_synchronize: null context");
    ND_HIP(hipStreamSynchronize(ctx->stream));
    ND_API_END(ctx)
}

int nd_timer_start(nd_ctx *ctx) {
    ND_API_BEGIN
    ND_REQUIRE(ctx, "nd_timer_start: null context");
    ND_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    ND_API_END(ctx)
}

int nd_timer_stop(nd_ctx *ctx, float *ms) {
    ND_API_BEGIN
    ND_REQUIRE(ctx && ms, "nd_timer_stop: null argument");
    ND_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ND_HIP(hipEventSynchronize(ctx->ev1));
    ND_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    ND_API_END(ctx)
}

int nd_bvh_spec_init(nd_bvh_spec *spec, float neighbor_distance, int64_t atom_count,
                     int64_t atomsperleaf) {
    if (!spec) return ND_ERR_ARG;
    // SpheresBVHSpecs (BVHTraverse.jl:71-94): atomsperleaf must divide
    // atom_count, and there must be at least two leaves.
    if (atomsperleaf < 1 || atom_count < 1 || atom_count > INT32_MAX) return ND_ERR_ARG;
    if (atom_count % atomsperleaf != 0) return ND_ERR_ARG;
    int64_t leaves = atom_count / atomsperleaf;
    if (leaves < 2) return ND_ERR_ARG;
    spec->neighbor_distance = neighbor_distance;
    spec->atom_count = (int32_t)atom_count;
    spec->leaves_count = (int32_t)leaves;
    spec->branches_count = (int32_t)(leaves - 1);
    spec->atomsperleaf = (int32_t)atomsperleaf;
    return ND_OK;
}

}  // extern "C"
