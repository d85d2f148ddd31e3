// kg_matmul.hip -- C(nx,ns) += A(nx,ny) * B(ny,ns), column-major FP64, on
// the FP64 matrix cores (v_mfma_f64_16x16x4f64).  The suite's only dense
// contraction (matmul/matmul.F90:56-68).
//
// At the benchmark shape (512 x 256 x 256, 67 MFLOP) the step is latency
// work, not throughput work: at the FP64 matrix peak it is under a
// microsecond.  So the kernel is shaped for many short, independent waves
// whose loads are all in flight before the first product:
//
//   * a workgroup of KS = 4 waves owns one 16 x 16 tile of C (512 workgroups
//     at 512 x 256: two per compute unit);
//   * wave w takes k in [w*kq, (w+1)*kq), kq = ny/KS rounded up to a
//     multiple of 4, in chunks of KC = 64 (16 MFMA steps); every operand of a
//     chunk is loaded into registers before its first MFMA;
//   * the MFMA step s of a chunk uses the four k values k0 + 2*(l/16) + s%2 +
//     8*(s/2) of lane group l/16 (any four distinct k form a valid step), so
//     the B operands of steps 2t and 2t+1 are one 16-byte load of the
//     contiguous pair B(k, j), B(k+1, j) (B is k-contiguous);
//   * waves 1..3 leave their partial tiles in LDS and wave 0 adds them to
//     its own in that fixed order (the result does not depend on timing),
//     then writes C.
//
// Lane l supplies A(i0 + l%16, k) and B(k, j0 + l%16); its four accumulators
// are C(i0 + l/16 + 4*r, j0 + l%16), r = 0..3 (the accumulator layout of the
// instruction, checked on the device).  Products outside the matrices are
// zero.  Sums are formed by the matrix core in blocks of four fused products
// and the four k ranges are added at the end, so the result differs from a
// strictly sequential k loop by rounding (bounded norm-wise).
#include "kg_internal.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// Measured at 512 x 256 x 256 (40 back-to-back launches): 8.1 us (8.3
// TFLOP/s) against 61 us for one wave per tile walking all of k with each
// load waited for before its product; k split 2 ways 10.0 us, 8 ways 8.6,
// 16 ways 8.3, 64-k or 32-k chunks alike.
#ifndef KG_MM_KS
#define KG_MM_KS 4
#endif
#ifndef KG_MM_KC
#define KG_MM_KC 64
#endif
constexpr int KS = KG_MM_KS;   // waves per workgroup: the k range split KS ways
constexpr int KC = KG_MM_KC;   // k values per chunk (KC/4 MFMA steps)
constexpr int STEPS = KC / 4;

__global__ __launch_bounds__(64 * KS)
void matmul_k(int nx, int ny, int ns, int ntile_i, int kq, const double* __restrict__ A,
              const double* __restrict__ B, double* __restrict__ C)
{
    __shared__ v4d part[KS - 1][64];
    const int wave = (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    const int i0 = (int)(blockIdx.x % (unsigned)ntile_i) * 16;
    const int j0 = (int)(blockIdx.x / (unsigned)ntile_i) * 16;
    const int lr = lane & 15;        // row of A / column of B and C within the tile
    const int lg = lane >> 4;        // lane group 0..3
    const int ia = i0 + lr, jb = j0 + lr;
    const bool a_ok = ia < nx, b_ok = jb < ns;
    const double* Arow = A + ia;
    const double* Bcol = B + (long long)ny * jb;
    // B pairs are 16-byte loads when every column starts 16-byte aligned
    const bool b16 = !(ny & 1) && !((uintptr_t)B & 15);

    v4d acc = {0.0, 0.0, 0.0, 0.0};
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = i0 + lg + 4 * r;
            acc[r] = (i < nx && b_ok) ? C[(long long)i + (long long)nx * jb] : 0.0;
        }
    }
    const int kb = wave * kq, ke = min(kb + kq, ny);
    for (int k0 = kb; k0 < ke; k0 += KC) {
        double av[STEPS], bv[STEPS];
#pragma unroll
        for (int t = 0; t < STEPS / 2; t++) {
            const int k = k0 + 8 * t + 2 * lg;   // this lane group's pair k, k+1
            const bool k0ok = k < ke, k1ok = k + 1 < ke;
            av[2 * t] = (a_ok && k0ok) ? Arow[(long long)nx * k] : 0.0;
            av[2 * t + 1] = (a_ok && k1ok) ? Arow[(long long)nx * (k + 1)] : 0.0;
            if (b16 && b_ok && k1ok) {
                const v2d p = *(const v2d*)(Bcol + k);
                bv[2 * t] = p[0];
                bv[2 * t + 1] = p[1];
            } else {
                bv[2 * t] = (b_ok && k0ok) ? Bcol[k] : 0.0;
                bv[2 * t + 1] = (b_ok && k1ok) ? Bcol[k + 1] : 0.0;
            }
        }
#pragma unroll
        for (int s = 0; s < STEPS; s++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc, 0, 0, 0);
    }
    if (wave > 0) part[wave - 1][lane] = acc;
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int w = 0; w < KS - 1; w++) acc += part[w][lane];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = i0 + lg + 4 * r;
        if (i < nx && b_ok) C[(long long)i + (long long)nx * jb] = acc[r];
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
    if (tiles > 0xFFFFFFFFLL / 256)
        return kg_set_error(KG_ERR_INVALID, "kg_matmul: %lld tiles exceed one launch", tiles);
    // each wave's k range: a quarter of ny rounded up to whole MFMA steps
    const int kq = ((ny + KS - 1) / KS + 3) / 4 * 4;
    kg_register_kernel("matmul_", (const void*)matmul_k);
    if (int rc = kg_launch_begin(ctx, "matmul_")) return rc;
    if (tiles > 0)
        hipLaunchKernelGGL(matmul_k, dim3((unsigned)tiles), dim3(64 * KS), 0, ctx->stream,
                           nx, ny, ns, ti, kq, A, B, C);
    return kg_launch_end(ctx);
}

int kg_preload_matmul() { return kg_preload_kernel((const void*)matmul_k); }
