// Context, workspace and error plumbing of the C ABI.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "internal.h"

std::mutex &solver_mutex() {
    static std::mutex mu;
    return mu;
}

int brpa_fail(brpa_ctx *ctx, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

bool is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int to_device(brpa_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return BRPA_OK;
    if (is_device_ptr(src)) {
        BRPA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return BRPA_OK;
    }
    // a pageable host source is copied to completion before returning
    return copy_sync(ctx, dst, src, bytes, hipMemcpyHostToDevice);
}

int from_device(brpa_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return BRPA_OK;
    hipMemcpyKind kind = is_device_ptr(dst) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    BRPA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    return BRPA_OK;
}

int copy_sync(brpa_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    if (bytes == 0) return BRPA_OK;
    // small transfers between pageable host memory and the device go through
    // the context's pinned bounce buffer: a DMA on this stream only, where a
    // pageable copy would take the runtime's shared staging path
    if (ctx->bounce && bytes <= kBounceBytes &&
        (kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToHost)) {
        if (kind == hipMemcpyHostToDevice) {
            std::memcpy(ctx->bounce, src, bytes);
            BRPA_HIP(ctx, hipMemcpyAsync(dst, ctx->bounce, bytes, kind, ctx->stream));
            BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        } else {
            BRPA_HIP(ctx, hipMemcpyAsync(ctx->bounce, src, bytes, kind, ctx->stream));
            BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
            std::memcpy(dst, ctx->bounce, bytes);
        }
        return BRPA_OK;
    }
    BRPA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BRPA_OK;
}

int ws_reserve(brpa_ctx *ctx, size_t bytes, Arena *arena) {
    if (bytes > ctx->ws_cap) {
        BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        graphs_release(ctx);   // captured graphs hold workspace addresses
        if (ctx->ws) {
            BRPA_HIP(ctx, hipFree(ctx->ws));
            ctx->ws = nullptr;
            ctx->ws_cap = 0;
        }
        size_t cap = bytes + bytes / 8 + (64u << 20);
        hipError_t e = hipMalloc(&ctx->ws, cap);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ctx->ws = nullptr;
            return brpa_fail(ctx, BRPA_ERR_OOM, "device workspace of %zu bytes: %s", cap,
                             hipGetErrorString(e));
        }
        ctx->ws_cap = cap;
    }
    arena->base = ctx->ws;
    arena->off = 0;
    return BRPA_OK;
}

int pin_reserve(brpa_ctx *ctx, size_t bytes) {
    if (bytes <= ctx->pin_cap) return BRPA_OK;
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // captured graphs hold the old staging address
    graphs_release(ctx);
    if (ctx->pin) BRPA_HIP(ctx, hipHostFree(ctx->pin));
    ctx->pin = nullptr;
    ctx->pin_cap = 0;
    size_t cap = bytes + bytes / 4 + (1u << 20);
    hipError_t e = hipHostMalloc((void **)&ctx->pin, cap, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->pin = nullptr;
        return brpa_fail(ctx, BRPA_ERR_OOM, "pinned host staging of %zu bytes: %s", cap,
                         hipGetErrorString(e));
    }
    ctx->pin_cap = cap;
    return BRPA_OK;
}

void graphs_release(brpa_ctx *ctx) {
    for (auto &g : ctx->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
}

void stage_reset(brpa_ctx *ctx) {
    for (int s = 0; s < BRPA_NSTAGES; ++s) {
        ctx->stage_ms[s] = 0.0;
        ctx->stage_launches[s] = 0;
    }
    for (int k = 0; k < BRPA_NKERNELS; ++k) {
        ctx->kst_ms[k] = ctx->kst_flops[k] = 0.0;
        ctx->kst_launches[k] = 0;
    }
}

int kstat_begin(brpa_ctx *ctx, size_t *slot) {
    while (ctx->gev.size() < ctx->gev_used + 2) {
        hipEvent_t e;
        BRPA_HIP(ctx, hipEventCreate(&e));
        ctx->gev.push_back(e);
    }
    *slot = ctx->gev_used;
    BRPA_HIP(ctx, hipEventRecord(ctx->gev[*slot], ctx->stream));
    ctx->gev_used += 2;
    ctx->gev_flops.push_back(0.0);
    ctx->gev_kind.push_back(-1);
    return BRPA_OK;
}

int kstat_end(brpa_ctx *ctx, size_t slot, int kernel, double flops) {
    BRPA_HIP(ctx, hipEventRecord(ctx->gev[slot + 1], ctx->stream));
    ctx->gev_flops[slot / 2] = flops;
    ctx->gev_kind[slot / 2] = kernel;
    return BRPA_OK;
}

int kstat_collect(brpa_ctx *ctx) {
    for (size_t i = 0; i + 1 < ctx->gev_used; i += 2) {
        const int k = ctx->gev_kind[i / 2];
        if (k < 0 || k >= BRPA_NKERNELS) continue;
        float ms = 0.f;
        BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->gev[i], ctx->gev[i + 1]));
        ctx->kst_ms[k] += ms;
        ctx->kst_flops[k] += ctx->gev_flops[i / 2];
        ctx->kst_launches[k] += 1;
    }
    kstat_drop(ctx);
    return BRPA_OK;
}

void kstat_drop(brpa_ctx *ctx) {
    ctx->gev_used = 0;
    ctx->gev_flops.clear();
    ctx->gev_kind.clear();
}

int stage_begin(brpa_ctx *ctx, int stage) {
    BRPA_HIP(ctx, hipEventRecord(ctx->sev[2 * stage], ctx->stream));
    return BRPA_OK;
}

int stage_end(brpa_ctx *ctx, int stage) {
    BRPA_HIP(ctx, hipEventRecord(ctx->sev[2 * stage + 1], ctx->stream));
    return BRPA_OK;
}

int stage_collect(brpa_ctx *ctx, int stage) {
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->sev[2 * stage], ctx->sev[2 * stage + 1]));
    ctx->stage_ms[stage] += ms;
    ctx->stage_launches[stage] += 1;
    return BRPA_OK;
}

int ctx_create_cu_share(int device, int part, int nparts, brpa_ctx **out) {
    if (!out) return BRPA_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return BRPA_ERR_HIP;
    }
    if (device < 0 || device >= n || (nparts > 1 && (part < 0 || part >= nparts)))
        return BRPA_ERR_ARG;
    brpa_ctx *ctx = new brpa_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        delete ctx;
        return BRPA_ERR_HIP;
    }
    hipError_t se = hipSuccess;
    int ncu = 0;
    if (nparts > 1 &&
        (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) !=
             hipSuccess || ncu < nparts)) {
        (void)hipGetLastError();
        nparts = 1;   // no share to give: the whole chip
    }
    if (nparts > 1) {
        // share `part`: a contiguous range of the CU mask bits
        std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
        const int c0 = (int)((int64_t)part * ncu / nparts);
        const int c1 = (int)((int64_t)(part + 1) * ncu / nparts);
        for (int c = c0; c < c1; ++c) mask[c / 32] |= 1u << (c % 32);
        ctx->ncu = c1 - c0;
        se = hipExtStreamCreateWithCUMask(&ctx->own_stream, (uint32_t)mask.size(), mask.data());
        if (se != hipSuccess) {
            // a runtime that refuses the mask: the share only serves performance,
            // so the context runs on the whole chip instead
            (void)hipGetLastError();
            ctx->own_stream = nullptr;
            ctx->ncu = 0;
            nparts = 1;
        }
    }
    if (nparts <= 1) se = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (se != hipSuccess || hipEventCreate(&ctx->ev0) != hipSuccess ||
        hipEventCreate(&ctx->ev1) != hipSuccess) {
        (void)hipGetLastError();
        delete ctx;
        return BRPA_ERR_HIP;
    }
    for (int i = 0; i < 2 * BRPA_NSTAGES; ++i)
        if (hipEventCreate(&ctx->sev[i]) != hipSuccess) {
        (void)hipGetLastError();
        delete ctx;
        return BRPA_ERR_HIP;
    }
    if (hipHostMalloc(&ctx->bounce, kBounceBytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        ctx->bounce = nullptr;
        brpa_ctx_destroy(ctx);
        return BRPA_ERR_OOM;
    }
    // device address of the allocation-free entry point's scratch array (once per
    // device, with the device current)
    if (noalloc_setup(device) != BRPA_OK) {
        brpa_ctx_destroy(ctx);
        return BRPA_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return BRPA_OK;
}

extern "C" {

int brpa_ctx_create(int device, brpa_ctx **out) { return ctx_create_cu_share(device, 0, 1, out); }

int brpa_ctx_destroy(brpa_ctx *ctx) {
    if (!ctx) return BRPA_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
    if (ctx->stream && ctx->stream != ctx->own_stream) (void)hipStreamSynchronize(ctx->stream);
    graphs_release(ctx);
    for (brpa_ctx *w : ctx->workers) brpa_ctx_destroy(w);
    ctx->workers.clear();
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->bounce) (void)hipHostFree(ctx->bounce);
    t2_release(ctx);
    mbpt3_release(ctx);
    for (hipEvent_t e : ctx->gev) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    for (int i = 0; i < 2 * BRPA_NSTAGES; ++i)
        if (ctx->sev[i]) (void)hipEventDestroy(ctx->sev[i]);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return BRPA_OK;
}

const char *brpa_last_error(const brpa_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int brpa_ctx_set_stream(brpa_ctx *ctx, void *s) {
    if (!ctx) return BRPA_ERR_ARG;
    ctx->stream = s ? static_cast<hipStream_t>(s) : ctx->own_stream;
    return BRPA_OK;
}

int brpa_ctx_set_graphs(brpa_ctx *ctx, int enable) {
    if (!ctx) return BRPA_ERR_ARG;
    ctx->use_graphs = enable ? 1 : 0;
    if (!enable) {
        BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        graphs_release(ctx);
    }
    return BRPA_OK;
}

int brpa_last_graph_stats(const brpa_ctx *ctx, int *captures, int *replays) {
    if (!ctx) return BRPA_ERR_ARG;
    if (captures) *captures = ctx->last_graph_captures;
    if (replays) *replays = ctx->last_graph_replays;
    return BRPA_OK;
}

int brpa_ctx_trim(brpa_ctx *ctx) {
    if (!ctx) return BRPA_ERR_ARG;
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    graphs_release(ctx);
    for (brpa_ctx *w : ctx->workers) brpa_ctx_destroy(w);
    ctx->workers.clear();
    if (ctx->ws) BRPA_HIP(ctx, hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_cap = 0;
    return BRPA_OK;
}

double brpa_last_kernel_ms(const brpa_ctx *ctx) { return ctx ? ctx->last_kernel_ms : 0.0; }

int brpa_last_gemm_i8_stats(const brpa_ctx *ctx, double *ms, double *flops, int64_t *launches) {
    if (!ctx) return BRPA_ERR_ARG;
    return brpa_last_kernel_stats(ctx, BRPA_KERNEL_GEMM_I8, ms, flops, launches);
}

int brpa_last_kernel_stats(const brpa_ctx *ctx, int kernel, double *ms, double *flops,
                           int64_t *launches) {
    if (!ctx || kernel < 0 || kernel >= BRPA_NKERNELS) return BRPA_ERR_ARG;
    if (ms) *ms = ctx->kst_ms[kernel];
    if (flops) *flops = ctx->kst_flops[kernel];
    if (launches) *launches = ctx->kst_launches[kernel];
    return BRPA_OK;
}

double brpa_last_stage_ms(const brpa_ctx *ctx, int stage) {
    return (ctx && stage >= 0 && stage < BRPA_NSTAGES) ? ctx->stage_ms[stage] : 0.0;
}

int brpa_last_stage_launches(const brpa_ctx *ctx, int stage) {
    return (ctx && stage >= 0 && stage < BRPA_NSTAGES) ? ctx->stage_launches[stage] : 0;
}

}  // extern "C"
