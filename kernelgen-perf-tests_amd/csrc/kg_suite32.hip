// kg_suite32.hip -- the suite's three single-precision tests (the reference
// builds them with real = float): vecadd, matvec and sincos.  They sit outside
// the stencil path and complete the benchmark table.
//
//   vecadd  w2 = w0 + w1 over every point (3-level rotation in the host loop)
//   matvec  y[j] = sum_i A[i + nx*j] * x[i], i ascending (matvec/matvec.c:38-70)
//   sincos  xy = sin(x) + cos(y) over every point (sincos/sincos.F90:14-74)
//
// matvec keeps the reference's sequential sum per row (bit-identical to an IEEE
// build): a workgroup owns 128 rows, stages 64-column chunks of A in LDS with
// coalesced loads, and each lane sums its own row from LDS in column order.
#include "kg_internal.h"

#include <algorithm>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256)
void vecadd_k(long long n4, long long n, const f4* __restrict__ w0, const f4* __restrict__ w1,
              f4* __restrict__ w2)
{
    const long long stride = (long long)gridDim.x * 256;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n4; q += stride)
        kg_st(&w2[q], w0[q] + w1[q]);
    // tail (n not a multiple of 4)
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(n - 4 * n4)) {
        const long long e = 4 * n4 + threadIdx.x;
        ((float*)w2)[e] = ((const float*)w0)[e] + ((const float*)w1)[e];
    }
}

constexpr int MV_ROWS = 128;   // rows (lanes) per workgroup; LDS tile 33 KB
constexpr int MV_COLS = 64;

__global__ __launch_bounds__(MV_ROWS)
void matvec_k(int nx, int ny, const float* __restrict__ A, const float* __restrict__ x,
              float* __restrict__ y)
{
    __shared__ float tile[MV_ROWS][MV_COLS + 1];
    __shared__ float xs[MV_COLS];
    const int j0 = blockIdx.x * MV_ROWS;
    const int t = threadIdx.x;
    const int j = j0 + t;
    float acc = 0.f;
    for (int c0 = 0; c0 < nx; c0 += MV_COLS) {
        const int nc = min(MV_COLS, nx - c0);
        // coalesced: 2 rows x 64 columns per pass of the 128 lanes
        for (int r = t / MV_COLS; r < MV_ROWS; r += MV_ROWS / MV_COLS) {
            const int c = t % MV_COLS;
            const int jj = j0 + r;
            tile[r][c] = (jj < ny && c < nc) ? A[(long long)jj * nx + c0 + c] : 0.f;
        }
        if (t < MV_COLS) xs[t] = t < nc ? x[c0 + t] : 0.f;
        __syncthreads();
        for (int c = 0; c < nc; c++) acc += tile[t][c] * xs[c];
        __syncthreads();
    }
    if (j < ny) y[j] = acc;
}

__global__ __launch_bounds__(256)
void sincos_k(long long n, const float* __restrict__ x, const float* __restrict__ y,
              float* __restrict__ xy)
{
    const long long stride = (long long)gridDim.x * 256;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += stride)
        kg_st(&xy[e], sinf(x[e]) + cosf(y[e]));
}

unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 256 * 32); }

#define KG_REQUIRE(cond, msg) \
    do { if (!(cond)) return kg_set_error(KG_ERR_INVALID, "%s", msg); } while (0)

}  // namespace

extern "C" int kg_vecadd(kg_ctx* ctx, int nx, int ny, int ns, const float* w0, const float* w1,
                         float* w2)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    KG_REQUIRE(w0 && w1 && w2, "kg_vecadd: null buffer");
    KG_REQUIRE(w2 != w0 && w2 != w1, "kg_vecadd: w2 aliases an input");
    KG_REQUIRE(!(((uintptr_t)w0 | (uintptr_t)w1 | (uintptr_t)w2) & 15),
               "kg_vecadd: buffers must be 16-byte aligned");
    const long long n = (long long)nx * ny * ns;
    kg_register_kernel("vecadd", (const void*)vecadd_k);
    if (int rc = kg_launch_begin(ctx, "vecadd")) return rc;
    if (n > 0)
        hipLaunchKernelGGL(vecadd_k, dim3(grid_for(std::max(n / 4, 1LL))), dim3(256), 0, ctx->stream,
                           n / 4, n, (const f4*)w0, (const f4*)w1, (f4*)w2);
    return kg_launch_end(ctx);
}

extern "C" int kg_matvec(kg_ctx* ctx, int nx, int ny, const float* A, const float* x, float* y)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims2(nx, ny)) return rc;
    KG_REQUIRE(A && x && y, "kg_matvec: null buffer");
    KG_REQUIRE(y != A && y != x, "kg_matvec: y aliases an input");
    kg_register_kernel("matvec", (const void*)matvec_k);
    if (int rc = kg_launch_begin(ctx, "matvec")) return rc;
    if (ny > 0)
        hipLaunchKernelGGL(matvec_k, dim3((ny + MV_ROWS - 1) / MV_ROWS), dim3(MV_ROWS), 0,
                           ctx->stream, nx, ny, A, x, y);
    return kg_launch_end(ctx);
}

extern "C" int kg_sincos(kg_ctx* ctx, int nx, int ny, int ns, const float* x, const float* y,
                         float* xy)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    KG_REQUIRE(x && y && xy, "kg_sincos: null buffer");
    KG_REQUIRE(xy != x && xy != y, "kg_sincos: xy aliases an input");
    const long long n = (long long)nx * ny * ns;
    kg_register_kernel("sincos", (const void*)sincos_k);
    if (int rc = kg_launch_begin(ctx, "sincos")) return rc;
    if (n > 0)
        hipLaunchKernelGGL(sincos_k, dim3(grid_for(n)), dim3(256), 0, ctx->stream, n, x, y, xy);
    return kg_launch_end(ctx);
}

int kg_preload_suite32() { return kg_preload_kernel((const void*)vecadd_k); }
