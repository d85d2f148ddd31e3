// kg_tricubic.hip -- FP64 tricubic interpolation (tricubic, tricubic2).
//
// Each point reads a 4x4x4 neighbourhood of u0 (offsets -1..+2) and its own
// a, b, c offsets.  A thread marches its (i, j) column through KC planes and
// keeps the 4x4 xy-neighbourhood of the 4 planes k-1..k+2 in registers, so
// each step loads one new 4x4 plane section (16 values, mostly L1 hits) in
// place of 64.  The k loop is unrolled by 4 and the window's plane slots are
// renamed at compile time (slot (r + kk) & 3 holds plane k-1+kk at phase r),
// so no register copies shift the window.  a, b, c are loaded two planes
// ahead of use, which hides their HBM latency at two waves per SIMD.  The 64
// products ((wa*wb)*wc)*u are summed in the reference's k-plane, j-row, i
// order without FMA contraction: bit-identical to the reference's IEEE build.
//
// The kernel is bound by FP64 issue rather than HBM: 180 multiplies and 72
// adds per point (the minimum for this summation order), i.e. 252 DP ops
// against 40 bytes.  Measured at 512x256x256 (cold caches): 376 us with this
// schedule, 496 us with a rolled loop shifting the window by copies, 925 us
// with the four planes staged in LDS (64 dependent LDS reads per point).
#include "kg_zmarch.h"

#include <type_traits>

namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int KC = 32;
constexpr int PFA = 2;  // a/b/c look-ahead in planes; plane k+2 < ns when k < ns-2

// Lagrange-cubic weights at nodes {-1,0,1,2} (tricubic/tricubic.c:1046-1059,
// FP64 branch; tricubic2/tricubic2.c:83-94), in source order.
__device__ __forceinline__ void weights(double a, double (&w)[4])
{
    w[0] = 1.0 / 6.0 * a * (a + 1.0) * (a + 2.0);
    w[1] = -0.5 * (a - 1.0) * (a + 1.0) * (a + 2.0);
    w[2] = 0.5 * (a - 1.0) * a * (a + 2.0);
    w[3] = -1.0 / 6.0 * (a - 1.0) * a * (a + 1.0);
}

// One point from the window at rotation phase R.  FROM_ZERO selects
// tricubic2's `u1 = 0; u1 += ...` accumulation (tricubic2/tricubic2.c:97-102);
// otherwise the first product starts the sum (tricubic/tricubic.c:1062-1145).
template <int R, bool FROM_ZERO>
__device__ __forceinline__ double point(const double (&win)[4][4][4], double av, double bv,
                                        double cv)
{
    double wa[4], wb[4], wc[4];
    weights(av, wa);
    weights(bv, wb);
    weights(cv, wc);
    double s;
    if (FROM_ZERO) {
        s = 0;
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
            for (int jj = 0; jj < 4; jj++)
#pragma unroll
                for (int ii = 0; ii < 4; ii++)
                    s += wa[ii] * wb[jj] * wc[kk] * win[(R + kk) & 3][jj][ii];
    } else {
        s = wa[0] * wb[0] * wc[0] * win[R & 3][0][0];
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
            for (int jj = 0; jj < 4; jj++)
#pragma unroll
                for (int ii = 0; ii < 4; ii++)
                    if (kk | jj | ii) s = s + wa[ii] * wb[jj] * wc[kk] * win[(R + kk) & 3][jj][ii];
    }
    return s;
}

template <int LO, int HI, bool FROM_ZERO>
__global__ __launch_bounds__(TX * TY)
void tricubic_k(int nx, int ny, int ns, int ntx, int nty,
                const double* __restrict__ u0, double* __restrict__ u1,
                const double* __restrict__ a, const double* __restrict__ b,
                const double* __restrict__ c)
{
    static_assert(HI >= PFA, "a/b/c look-ahead must stay inside the grid");
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, LO, HI, false>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    // At phase 0, win[p][jj][ii] holds u0 at (i+ii-1, j+jj-1, k+p-1).
    double win[4][4][4];
    const double* base = u0 + o - 1 - sy - sz;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
                win[p][jj][ii] = base[p * sz + jj * sy + ii];
    // qa[q] = a at plane k+q (likewise b, c).
    double qa[PFA + 1], qb[PFA + 1], qc[PFA + 1];
#pragma unroll
    for (int q = 0; q < PFA; q++) {
        qa[q] = a[o + q * sz];
        qb[q] = b[o + q * sz];
        qc[q] = c[o + q * sz];
    }
    int k = t.k0;
    auto step = [&](auto phase) {
        constexpr int R = decltype(phase)::value;
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
                win[(R + 3) & 3][jj][ii] = base[3 * sz + jj * sy + ii];
        qa[PFA] = a[o + PFA * sz];
        qb[PFA] = b[o + PFA * sz];
        qc[PFA] = c[o + PFA * sz];
        const double s = point<R, FROM_ZERO>(win, qa[0], qb[0], qc[0]);
#pragma unroll
        for (int q = 0; q < PFA; q++) {
            qa[q] = qa[q + 1];
            qb[q] = qb[q + 1];
            qc[q] = qc[q + 1];
        }
        kg_st(&u1[o], s);
        base += sz;
        o += sz;
        k++;
    };
    while (k + 4 <= t.k1) {
        step(std::integral_constant<int, 0>{});
        step(std::integral_constant<int, 1>{});
        step(std::integral_constant<int, 2>{});
        step(std::integral_constant<int, 3>{});
    }
    if (k < t.k1) step(std::integral_constant<int, 0>{});
    if (k < t.k1) step(std::integral_constant<int, 1>{});
    if (k < t.k1) step(std::integral_constant<int, 2>{});
}

}  // namespace

template <int LO, int HI, bool FROM_ZERO>
static int launch_tricubic(const char* name, kg_ctx* ctx, int nx, int ny, int ns,
                           const double* u0, double* u1, const double* a, const double* b,
                           const double* c)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    if (!u0 || !u1 || !a || !b || !c) return kg_set_error(KG_ERR_INVALID, "kg_%s: null buffer", name);
    if (u1 == u0 || u1 == a || u1 == b || u1 == c)
        return kg_set_error(KG_ERR_INVALID, "kg_%s: u1 aliases an input", name);
    const kg_zgrid g = kg_zgrid_make(nx, ny, ns, LO, HI, TX, TY, KC);
    auto kern = tricubic_k<LO, HI, FROM_ZERO>;
    kg_register_kernel(name, (const void*)kern);
    if (int rc = kg_launch_begin(ctx, name)) return rc;
    if (!g.empty)
        hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(TX, TY), 0, ctx->stream,
                           nx, ny, ns, g.ntx, g.nty, u0, u1, a, b, c);
    return kg_launch_end(ctx);
}

// tricubic/tricubic.c:138-155,509: k in [1,ns-2), j in [1,ny-2), i in [1,nx-2)
extern "C" int kg_tricubic(kg_ctx* ctx, int nx, int ny, int ns, const double* u0, double* u1,
                           const double* a, const double* b, const double* c)
{
    return launch_tricubic<1, 2, false>("tricubic", ctx, nx, ny, ns, u0, u1, a, b, c);
}

// tricubic2/tricubic2.c:69-80: [2,n-2)^3
extern "C" int kg_tricubic2(kg_ctx* ctx, int nx, int ny, int ns, const double* u0, double* u1,
                            const double* a, const double* b, const double* c)
{
    return launch_tricubic<2, 2, true>("tricubic2", ctx, nx, ny, ns, u0, u1, a, b, c);
}

int kg_preload_tricubic() { return kg_preload_kernel((const void*)tricubic_k<1, 2, false>); }
