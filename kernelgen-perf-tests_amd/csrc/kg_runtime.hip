// kg_runtime.hip -- context, memory, transfer, timing and error handling of
// libkgstencil.so.  Replaces the CUDA runtime calls of the reference's `cuda`
// target and the launch wrapper of laplacian/cuda/cuda_profiling.cu:211-251
// (kernel time and register count per launch).
#include "kg_internal.h"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

static thread_local char g_last_error[512] = "";

int kg_set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* kg_last_error(void) { return g_last_error; }

extern "C" const char* kg_strerror(int status)
{
    switch (status) {
    case KG_OK: return "success";
    case KG_ERR_INVALID: return "invalid argument";
    case KG_ERR_HIP: return "HIP runtime error";
    case KG_ERR_NOMEM: return "out of device memory";
    case KG_ERR_NODEV: return "no such HIP device";
    case KG_ERR_STATE: return "invalid state";
    default: return "unknown error";
    }
}

// ---- kernel registry (for regcount) -------------------------------------
static std::mutex g_reg_mu;
static kg_kernel_entry g_reg[64];
static int g_nreg = 0;

void kg_register_kernel(const char* name, const void* func)
{
    std::lock_guard<std::mutex> lk(g_reg_mu);
    for (int i = 0; i < g_nreg; i++)
        if (strcmp(g_reg[i].name, name) == 0) return;
    if (g_nreg < 64) g_reg[g_nreg++] = {name, func};
}

extern "C" int kg_kernel_regcount(kg_ctx* ctx, const char* kernel, int* vgprs)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (!kernel || !vgprs) return kg_set_error(KG_ERR_INVALID, "kg_kernel_regcount: null argument");
    const void* f = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        for (int i = 0; i < g_nreg; i++)
            if (strcmp(g_reg[i].name, kernel) == 0) f = g_reg[i].func;
    }
    if (!f) return kg_set_error(KG_ERR_INVALID, "kg_kernel_regcount: unknown kernel '%s' (launch it first)", kernel);
    hipFuncAttributes attr;
    KG_HIP(hipSetDevice(ctx->device));
    KG_HIP(hipFuncGetAttributes(&attr, f));
    *vgprs = attr.numRegs;
    return KG_OK;
}

// ---- validation -----------------------------------------------------------
int kg_check_ctx(kg_ctx* ctx)
{
    if (!ctx) return kg_set_error(KG_ERR_INVALID, "null kg_ctx");
    return KG_OK;
}

int kg_check_dims3(int nx, int ny, int ns)
{
    if (nx < 0 || ny < 0 || ns < 0)
        return kg_set_error(KG_ERR_INVALID, "negative grid dimension (%d, %d, %d)", nx, ny, ns);
    // In-plane offsets are 32-bit in the kernels; plane offsets are 64-bit.
    if ((long long)nx * ny >= (1LL << 31))
        return kg_set_error(KG_ERR_INVALID, "xy-plane of %d x %d points exceeds 2^31", nx, ny);
    return KG_OK;
}

// The plane engine addresses each xy-plane through one buffer resource with
// 32-bit byte offsets: a plane may hold at most 2^31 bytes.
int kg_check_plane_bytes(long long points, int elem_bytes)
{
    if (points * elem_bytes > (1LL << 31))
        return kg_set_error(KG_ERR_INVALID, "xy-plane of %lld points x %d bytes exceeds 2^31 bytes",
                            points, elem_bytes);
    return KG_OK;
}

int kg_check_dims2(int nx, int ny)
{
    if (nx < 0 || ny < 0)
        return kg_set_error(KG_ERR_INVALID, "negative grid dimension (%d, %d)", nx, ny);
    return KG_OK;
}

// ---- context ----------------------------------------------------------------
extern "C" int kg_device_count(int* count)
{
    if (!count) return kg_set_error(KG_ERR_INVALID, "kg_device_count: null argument");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) { *count = 0; return KG_OK; }
    KG_HIP(e);
    *count = n;
    return KG_OK;
}

extern "C" int kg_init(kg_ctx** out, int device)
{
    if (!out) return kg_set_error(KG_ERR_INVALID, "kg_init: null argument");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return kg_set_error(KG_ERR_NODEV, "kg_init: no HIP device (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n)
        return kg_set_error(KG_ERR_NODEV, "kg_init: device %d out of range [0,%d)", device, n);
    KG_HIP(hipSetDevice(device));
    kg_ctx* c = (kg_ctx*)calloc(1, sizeof(kg_ctx));
    if (!c) return kg_set_error(KG_ERR_INVALID, "kg_init: host allocation failed");
    c->device = device;
    c->timing = 1;
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        hipDeviceGetAttribute(&c->lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                              device) != hipSuccess ||
        c->cus <= 0 || c->lds_cu <= 0) {
        free(c);
        return kg_set_error(KG_ERR_HIP, "kg_init: cannot read the compute unit count or LDS size");
    }
    c->ev_start = (hipEvent_t*)calloc(KG_TIMING_RING, sizeof(hipEvent_t));
    c->ev_stop = (hipEvent_t*)calloc(KG_TIMING_RING, sizeof(hipEvent_t));
    if (!c->ev_start || !c->ev_stop) {
        free(c->ev_start); free(c->ev_stop); free(c);
        return kg_set_error(KG_ERR_INVALID, "kg_init: host allocation failed");
    }
    // Default-priority, non-blocking stream: no implicit sync with the null stream.
    hipError_t se = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (se != hipSuccess) {
        free(c->ev_start); free(c->ev_stop); free(c);
        return kg_set_error(KG_ERR_HIP, "hipStreamCreateWithFlags: %s", hipGetErrorString(se));
    }
    c->last_kernel = "";
    // FP64 partials of kg_field_stats (a few KB, owned by the context)
    if (hipMalloc((void**)&c->stats, kg_stats_scratch_bytes()) != hipSuccess) {
        (void)hipGetLastError();
        c->stats = nullptr;
        kg_destroy(c);
        return kg_set_error(KG_ERR_NOMEM, "kg_init: statistics scratch allocation failed");
    }
    int (*preload[])() = {kg_preload_stencil3d, kg_preload_tricubic, kg_preload_stencil2d,
                          kg_preload_matmul, kg_preload_util, kg_preload_life,
                          kg_preload_typed, kg_preload_batch, kg_preload_poisson,
                          kg_preload_vecmath};
    for (auto f : preload) {
        if (int rc = f()) {
            kg_destroy(c);
            return rc;
        }
    }
    if (int rc = kg_poisson_static_setup(device)) {
        kg_destroy(c);
        return rc;
    }
    *out = c;
    return KG_OK;
}

int kg_preload_kernel(const void* func)
{
    hipFuncAttributes attr;
    KG_HIP(hipFuncGetAttributes(&attr, func));
    return KG_OK;
}

extern "C" int kg_destroy(kg_ctx* ctx)
{
    if (!ctx) return KG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < KG_TIMING_RING; i++) {
        if (ctx->ev_start[i]) (void)hipEventDestroy(ctx->ev_start[i]);
        if (ctx->ev_stop[i]) (void)hipEventDestroy(ctx->ev_stop[i]);
    }
    if (ctx->stats) (void)hipFree(ctx->stats);
    (void)hipStreamDestroy(ctx->stream);
    free(ctx->ev_start);
    free(ctx->ev_stop);
    free(ctx);
    return KG_OK;
}

// ---- memory -----------------------------------------------------------------
extern "C" int kg_alloc(kg_ctx* ctx, void** dptr, size_t bytes)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (!dptr) return kg_set_error(KG_ERR_INVALID, "kg_alloc: null argument");
    *dptr = nullptr;
    KG_HIP(hipSetDevice(ctx->device));
    if (bytes == 0) return KG_OK;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        *dptr = nullptr;
        return kg_set_error(KG_ERR_NOMEM, "kg_alloc: %zu bytes: %s", bytes, hipGetErrorString(e));
    }
    KG_HIP(e);
    return KG_OK;
}

extern "C" int kg_free(kg_ctx* ctx, void* dptr)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (!dptr) return KG_OK;
    KG_HIP(hipSetDevice(ctx->device));
    KG_HIP(hipStreamSynchronize(ctx->stream));
    KG_HIP(hipFree(dptr));
    return KG_OK;
}

static int kg_copy(kg_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (bytes == 0) return KG_OK;
    if (!dst || !src) return kg_set_error(KG_ERR_INVALID, "copy: null pointer");
    KG_HIP(hipSetDevice(ctx->device));
    KG_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
    return KG_OK;
}

extern "C" int kg_h2d(kg_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return kg_copy(ctx, dst, src, bytes, hipMemcpyHostToDevice);
}

extern "C" int kg_d2h(kg_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return kg_copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost);
}

extern "C" int kg_d2d(kg_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    return kg_copy(ctx, dst, src, bytes, hipMemcpyDeviceToDevice);
}

extern "C" int kg_sync(kg_ctx* ctx)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    KG_HIP(hipStreamSynchronize(ctx->stream));
    return KG_OK;
}

// ---- launch bookkeeping and timing -----------------------------------------
int kg_launch_begin(kg_ctx* ctx, const char* name)
{
    KG_HIP(hipSetDevice(ctx->device));
    ctx->last_kernel = name;
    if (!ctx->timing) return KG_OK;
    if (ctx->launches - ctx->drained >= KG_TIMING_RING)
        ctx->drained = ctx->launches - KG_TIMING_RING + 1;   // oldest pair is reused
    int slot = (int)(ctx->launches % KG_TIMING_RING);
    if (!ctx->ev_start[slot]) {
        KG_HIP(hipEventCreate(&ctx->ev_start[slot]));
        KG_HIP(hipEventCreate(&ctx->ev_stop[slot]));
    }
    KG_HIP(hipEventRecord(ctx->ev_start[slot], ctx->stream));
    return KG_OK;
}

int kg_launch_end(kg_ctx* ctx)
{
    KG_HIP(hipGetLastError());
    if (!ctx->timing) return KG_OK;
    int slot = (int)(ctx->launches % KG_TIMING_RING);
    KG_HIP(hipEventRecord(ctx->ev_stop[slot], ctx->stream));
    ctx->launches++;
    return KG_OK;
}

extern "C" int kg_timing_enable(kg_ctx* ctx, int on)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    ctx->timing = on ? 1 : 0;
    return KG_OK;
}

extern "C" int kg_last_kernel_time(kg_ctx* ctx, double* seconds)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (!seconds) return kg_set_error(KG_ERR_INVALID, "kg_last_kernel_time: null argument");
    if (!ctx->timing || ctx->launches == 0)
        return kg_set_error(KG_ERR_STATE, "kg_last_kernel_time: no timed launch");
    int slot = (int)((ctx->launches - 1) % KG_TIMING_RING);
    KG_HIP(hipEventSynchronize(ctx->ev_stop[slot]));
    float ms = 0.f;
    KG_HIP(hipEventElapsedTime(&ms, ctx->ev_start[slot], ctx->ev_stop[slot]));
    *seconds = ms * 1e-3;
    return KG_OK;
}

extern "C" int kg_timing_drain(kg_ctx* ctx, double* seconds, int cap, int* count)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (!count || (cap > 0 && !seconds)) return kg_set_error(KG_ERR_INVALID, "kg_timing_drain: null argument");
    *count = 0;
    KG_HIP(hipStreamSynchronize(ctx->stream));
    int n = 0;
    for (long long l = ctx->drained; l < ctx->launches && n < cap; l++, n++) {
        int slot = (int)(l % KG_TIMING_RING);
        float ms = 0.f;
        KG_HIP(hipEventElapsedTime(&ms, ctx->ev_start[slot], ctx->ev_stop[slot]));
        seconds[n] = ms * 1e-3;
    }
    ctx->drained = ctx->launches;
    *count = n;
    return KG_OK;
}

extern "C" const char* kg_last_kernel_name(kg_ctx* ctx)
{
    return ctx ? ctx->last_kernel : "";
}
