// kg_util.hip -- calibration kernel: a plain device-to-device copy whose HBM
// traffic is known exactly (n*8 bytes read, n*8 bytes written).  It is the
// yardstick for the achievable bandwidth of the stencils and for scaling the
// FETCH_SIZE / WRITE_SIZE counters of rocprofv3 to bytes.
#include "kg_internal.h"

namespace {

__global__ __launch_bounds__(256)
void stream_copy_k(double2* __restrict__ dst, const double2* __restrict__ src, long long n2)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
        dst[i] = src[i];
}

__global__ __launch_bounds__(64)
void stream_copy_tail_k(double* __restrict__ dst, const double* __restrict__ src, long long n)
{
    if (threadIdx.x == 0 && n > 0) dst[n - 1] = src[n - 1];
}

}  // namespace

extern "C" int kg_stream_copy(kg_ctx* ctx, double* dst, const double* src, size_t n)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (n && (!dst || !src)) return kg_set_error(KG_ERR_INVALID, "kg_stream_copy: null buffer");
    if (((uintptr_t)dst | (uintptr_t)src) & 15)
        return kg_set_error(KG_ERR_INVALID, "kg_stream_copy: buffers must be 16-byte aligned");
    kg_register_kernel("stream_copy", (const void*)stream_copy_k);
    if (int rc = kg_launch_begin(ctx, "stream_copy")) return rc;
    const long long n2 = (long long)(n / 2);
    if (n2 > 0) {
        long long blocks = (n2 + 255) / 256;
        if (blocks > 256 * 64) blocks = 256 * 64;   // 64 workgroups per CU, grid-stride
        hipLaunchKernelGGL(stream_copy_k, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                           (double2*)dst, (const double2*)src, n2);
    }
    if (n & 1)
        hipLaunchKernelGGL(stream_copy_tail_k, dim3(1), dim3(64), 0, ctx->stream, dst, src,
                           (long long)n);
    return kg_launch_end(ctx);
}

int kg_preload_util() { return kg_preload_kernel((const void*)stream_copy_k); }
