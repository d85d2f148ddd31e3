// kg_matmul.hip -- C(nx,ns) += A(nx,ny) * B(ny,ns), column-major FP64, on
// the FP64 matrix cores (v_mfma_f64_16x16x4f64).  The suite's only dense
// contraction (matmul/matmul.F90:56-68).
//
// One wave64 owns a 16x16 tile of C and walks k in steps of 4.  Per step the
// lane l supplies A(i0 + l%16, k + l/16) and B(k + l/16, j0 + l%16); the four
// accumulators of lane l are C(i0 + l/16 + 4*r, j0 + l%16), r = 0..3 (the
// accumulator layout of this instruction, checked on the device).  k
// tails that are not a multiple of 4 are padded with zero products.  Sums are
// formed by the matrix core in blocks of 4 fused products, so the result
// differs from a strictly sequential k loop by rounding.
#include "kg_internal.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int WAVES = 4;   // waves (16x16 tiles) per workgroup

__global__ __launch_bounds__(64 * WAVES)
void matmul_k(int nx, int ny, int ns, int ntile_i, int ntile_j,
              const double* __restrict__ A, const double* __restrict__ B,
              double* __restrict__ C)
{
    const int wave = (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    const int tile = (int)blockIdx.x * WAVES + wave;
    if (tile >= ntile_i * ntile_j) return;
    const int i0 = (tile % ntile_i) * 16;
    const int j0 = (tile / ntile_i) * 16;
    const int lr = lane & 15;   // row of A / column of B and C within the tile
    const int lk = lane >> 4;   // k offset 0..3 within a step

    v4d acc;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + lk + 4 * r, j = j0 + lr;
        acc[r] = (i < nx && j < ns) ? C[(long long)i + (long long)nx * j] : 0.0;
    }
    const int ia = i0 + lr;
    const int jb = j0 + lr;
    for (int k = 0; k < ny; k += 4) {
        const int kk = k + lk;
        const double av = (ia < nx && kk < ny) ? A[(long long)ia + (long long)nx * kk] : 0.0;
        const double bv = (jb < ns && kk < ny) ? B[(long long)kk + (long long)ny * jb] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + lk + 4 * r, j = j0 + lr;
        if (i < nx && j < ns) C[(long long)i + (long long)nx * j] = acc[r];
    }
}

}  // namespace

extern "C" int kg_matmul(kg_ctx* ctx, int nx, int ny, int ns, const double* A, const double* B,
                         double* C)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (nx < 0 || ny < 0 || ns < 0)
        return kg_set_error(KG_ERR_INVALID, "kg_matmul: negative dimension (%d, %d, %d)", nx, ny, ns);
    if (!A || !B || !C) return kg_set_error(KG_ERR_INVALID, "kg_matmul: null buffer");
    if (C == A || C == B) return kg_set_error(KG_ERR_INVALID, "kg_matmul: C aliases an input");
    const int ti = (nx + 15) / 16, tj = (ns + 15) / 16;
    const long long tiles = (long long)ti * tj;
    kg_register_kernel("matmul_", (const void*)matmul_k);
    if (int rc = kg_launch_begin(ctx, "matmul_")) return rc;
    if (tiles > 0) {
        const unsigned blocks = (unsigned)((tiles + WAVES - 1) / WAVES);
        hipLaunchKernelGGL(matmul_k, dim3(blocks), dim3(64 * WAVES), 0, ctx->stream,
                           nx, ny, ns, ti, tj, A, B, C);
    }
    return kg_launch_end(ctx);
}

int kg_preload_matmul() { return kg_preload_kernel((const void*)matmul_k); }
