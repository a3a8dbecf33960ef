// Context, memory and error plumbing of the C ABI.
#include "common.h"

int set_error(mshqc_hip_ctx* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    return code;
}

void* workspace(mshqc_hip_ctx* ctx, int slot, size_t bytes) {
    auto& s = ctx->slabs[slot];
    if (bytes == 0) bytes = 256;
    if (s.bytes >= bytes) return s.ptr;
    if (s.ptr) {
        // Work queued on the stream may still read the old slab.
        hipStreamSynchronize(ctx->stream);
        hipFree(s.ptr);
        s.ptr = nullptr;
        s.bytes = 0;
    }
    size_t want = bytes + bytes / 8;  // headroom for slightly larger repeats
    if (hipMalloc(&s.ptr, want) != hipSuccess) {
        hipGetLastError();
        s.ptr = nullptr;
        if (hipMalloc(&s.ptr, bytes) != hipSuccess) {
            hipGetLastError();
            s.ptr = nullptr;
            set_error(ctx, MSHQC_HIP_ERR_ALLOC, "workspace slot %d: cannot allocate %.1f MiB",
                      slot, bytes / kMiB);
            return nullptr;
        }
        want = bytes;
    }
    s.bytes = want;
    return s.ptr;
}

int ctx_enter(mshqc_hip_ctx* ctx) {
    if (!ctx) return MSHQC_HIP_ERR_ARG;
    ctx->last_error.clear();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return MSHQC_HIP_OK;
}

static hipEvent_t pool_get(mshqc_hip_ctx* ctx) {
    if (!ctx->pool.empty()) {
        hipEvent_t e = ctx->pool.back();
        ctx->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        hipGetLastError();
        return nullptr;
    }
    return e;
}

ProfScope::ProfScope(mshqc_hip_ctx* c, int kind) : ctx(c) {
    if (!ctx->prof) return;
    hipEvent_t a = pool_get(ctx), b = pool_get(ctx);
    if (!a || !b) return;
    hipEventRecord(a, ctx->stream);
    ctx->pending.push_back({kind, a, b});
    idx = (int)ctx->pending.size() - 1;
}

ProfScope::~ProfScope() {
    if (idx >= 0) hipEventRecord(ctx->pending[idx].b, ctx->stream);
}

int ctx_finish(mshqc_hip_ctx* ctx) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipGetLastError());
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ctx->prof_ms[p.kind] += ms;
            ctx->prof_n[p.kind] += 1;
        }
        hipGetLastError();
        ctx->pool.push_back(p.a);
        ctx->pool.push_back(p.b);
    }
    ctx->pending.clear();
    return MSHQC_HIP_OK;
}

extern "C" {

int mshqc_hip_ctx_create(int device, mshqc_hip_ctx** out) {
    if (!out) return MSHQC_HIP_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        hipGetLastError();
        return MSHQC_HIP_ERR_DEVICE;
    }
    if (device < 0 || device >= ndev) return MSHQC_HIP_ERR_ARG;
    auto* ctx = new mshqc_hip_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        hipGetLastError();
        delete ctx;
        return MSHQC_HIP_ERR_DEVICE;
    }
    ctx->stream = ctx->own_stream;
    // the no-allocation entry point's scratch address on this GPU
    if (noalloc_init(device) != MSHQC_HIP_OK) {
        hipEventDestroy(ctx->ev0);
        hipEventDestroy(ctx->ev1);
        hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return MSHQC_HIP_ERR_DEVICE;
    }
    *out = ctx;
    return MSHQC_HIP_OK;
}

int mshqc_hip_ctx_destroy(mshqc_hip_ctx* ctx) {
    if (!ctx) return MSHQC_HIP_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto& s : ctx->slabs)
        if (s.ptr) hipFree(s.ptr);
    for (auto& p : ctx->pending) ctx->pool.push_back(p.a), ctx->pool.push_back(p.b);
    for (auto e : ctx->pool) hipEventDestroy(e);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return MSHQC_HIP_OK;
}

const char* mshqc_hip_last_error(const mshqc_hip_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int mshqc_hip_set_stream(mshqc_hip_ctx* ctx, void* stream) {
    if (!ctx) return MSHQC_HIP_ERR_ARG;
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return MSHQC_HIP_OK;
}

int mshqc_hip_last_ms(const mshqc_hip_ctx* ctx, float* ms) {
    if (!ctx || !ms) return MSHQC_HIP_ERR_ARG;
    *ms = ctx->last_ms;
    return MSHQC_HIP_OK;
}

int mshqc_hip_set_zblock(mshqc_hip_ctx* ctx, size_t elems) {
    if (!ctx) return MSHQC_HIP_ERR_ARG;
    ctx->zblock_elems = elems ? elems : ZBLOCK_DEFAULT;
    return MSHQC_HIP_OK;
}

int mshqc_hip_profile(mshqc_hip_ctx* ctx, int enable) {
    if (!ctx) return MSHQC_HIP_ERR_ARG;
    ctx->prof = enable != 0;
    for (int k = 0; k < 4; ++k) ctx->prof_ms[k] = 0.0, ctx->prof_n[k] = 0;
    return MSHQC_HIP_OK;
}

int mshqc_hip_profile_get(const mshqc_hip_ctx* ctx, int kind, long* launches, double* total_ms) {
    if (!ctx || kind < 0 || kind > 3 || !launches || !total_ms) return MSHQC_HIP_ERR_ARG;
    *launches = ctx->prof_n[kind];
    *total_ms = ctx->prof_ms[kind];
    return MSHQC_HIP_OK;
}

const char* mshqc_hip_version(void) { return "mshqc_hip gfx950 " __DATE__; }

int mshqc_hip_alloc(mshqc_hip_ctx* ctx, size_t bytes, void** dptr) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, dptr != nullptr, "alloc: null output pointer");
    *dptr = nullptr;
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
        hipGetLastError();
        *dptr = nullptr;
        return set_error(ctx, MSHQC_HIP_ERR_ALLOC, "alloc: cannot allocate %.1f MiB",
                         bytes / kMiB);
    }
    return MSHQC_HIP_OK;
}

int mshqc_hip_free(mshqc_hip_ctx* ctx, void* dptr) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    if (!dptr) return MSHQC_HIP_OK;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipFree(dptr));
    return MSHQC_HIP_OK;
}

int mshqc_hip_memcpy_h2d(mshqc_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, (dst && src) || bytes == 0, "memcpy_h2d: null pointer");
    if (bytes) HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return ctx_finish(ctx);
}

int mshqc_hip_memcpy_d2h(mshqc_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
    int rc = ctx_enter(ctx);
    if (rc) return rc;
    CHECK_ARG(ctx, (dst && src) || bytes == 0, "memcpy_d2h: null pointer");
    if (bytes) HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return ctx_finish(ctx);
}

}  // extern "C"
