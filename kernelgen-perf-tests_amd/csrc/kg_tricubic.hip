// kg_tricubic.hip -- FP64 tricubic interpolation (tricubic, tricubic2).
//
// Each point reads a 4x4x4 neighbourhood of u0 (offsets -1..+2) and its own
// a, b, c offsets.  A thread marches its (i, j) column through KC planes and
// keeps the 4x4 xy-neighbourhood of the 4 planes k-1..k+2 in registers, so
// each step loads one new 4x4 plane section (16 values, mostly L1 hits) in
// place of 64.  The 64 products ((wa*wb)*wc)*u are summed in the reference's
// k-plane, j-row, i order without FMA contraction: bit-identical to the
// reference's IEEE build.  (Staging the four planes in LDS instead -- the
// plane engine of kg_plane.h -- measured 925 us against 509 us for this
// register window at 512x256x256: 64 dependent LDS reads per point at two
// waves per SIMD.)
#include "kg_zmarch.h"

namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int KC = 16;

// Lagrange-cubic weights at nodes {-1,0,1,2} (tricubic/tricubic.c:1046-1059,
// FP64 branch; tricubic2/tricubic2.c:83-94), in source order.
__device__ __forceinline__ void weights(double a, double& w1, double& w2, double& w3, double& w4)
{
    w1 = 1.0 / 6.0 * a * (a + 1.0) * (a + 2.0);
    w2 = -0.5 * (a - 1.0) * (a + 1.0) * (a + 2.0);
    w3 = 0.5 * (a - 1.0) * a * (a + 2.0);
    w4 = -1.0 / 6.0 * (a - 1.0) * a * (a + 1.0);
}

// FROM_ZERO selects tricubic2's `u1 = 0; u1 += ...` accumulation
// (tricubic2/tricubic2.c:97-102); otherwise the first product starts the sum
// (tricubic/tricubic.c:1062-1145).
template <int LO, int HI, bool FROM_ZERO>
__global__ __launch_bounds__(TX * TY)
void tricubic_k(int nx, int ny, int ns, int ntx, int nty,
                const double* __restrict__ u0, double* __restrict__ u1,
                const double* __restrict__ a, const double* __restrict__ b,
                const double* __restrict__ c)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, LO, HI>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    // win[p][jj][ii] holds u0 at (i+ii-1, j+jj-1, k+p-1).
    double win[4][4][4];
    const double* base = u0 + o - 1 - sy - sz;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
                win[p][jj][ii] = base[p * sz + jj * sy + ii];
    for (int k = t.k0; k < t.k1; k++) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
#pragma unroll
            for (int ii = 0; ii < 4; ii++)
                win[3][jj][ii] = base[3 * sz + jj * sy + ii];
        double wa[4], wb[4], wc[4];
        weights(a[o], wa[0], wa[1], wa[2], wa[3]);
        weights(b[o], wb[0], wb[1], wb[2], wb[3]);
        weights(c[o], wc[0], wc[1], wc[2], wc[3]);
        double s;
        if (FROM_ZERO) {
            s = 0;
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
#pragma unroll
                    for (int ii = 0; ii < 4; ii++)
                        s += wa[ii] * wb[jj] * wc[kk] * win[kk][jj][ii];
        } else {
            s = wa[0] * wb[0] * wc[0] * win[0][0][0];
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
#pragma unroll
                    for (int ii = 0; ii < 4; ii++)
                        if (kk | jj | ii) s = s + wa[ii] * wb[jj] * wc[kk] * win[kk][jj][ii];
        }
        kg_st(&u1[o], s);
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int jj = 0; jj < 4; jj++)
#pragma unroll
                for (int ii = 0; ii < 4; ii++)
                    win[p][jj][ii] = win[p + 1][jj][ii];
        base += sz;
        o += sz;
    }
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
