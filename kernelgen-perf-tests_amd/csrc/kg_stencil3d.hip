// kg_stencil3d.hip -- FP64 3D stencils of the suite: laplacian, divergence,
// gradient, wave13pt, uxx1, lapgsrb.
//
// Each kernel marches a TX x TY column tile through KC z-planes (kg_zmarch.h),
// keeping the z-column of its main input in registers, so every input array
// streams through HBM about once per step; xy neighbours come from the
// L1/L2 lines the neighbouring lanes and rows of the same tile fetch.
// Arithmetic is written in the reference's source order and the library is
// built with -ffp-contract=off: results are bit-identical to an IEEE build of
// the reference kernels.
#include "kg_zmarch.h"

namespace {

constexpr int TX = 64;   // one wave64 per x-row segment
constexpr int TY = 4;    // 4 rows -> 256 threads per workgroup
constexpr int KC = 16;   // z-planes per workgroup

#define KG_REQUIRE(cond, msg) \
    do { if (!(cond)) return kg_set_error(KG_ERR_INVALID, "%s", msg); } while (0)

// ---------------------------------------------------------------- laplacian
// laplacian/laplacian.c:93-97
__global__ __launch_bounds__(TX * TY)
void laplacian_k(int nx, int ny, int ns, int ntx, int nty, double alpha, double beta,
                 const double* __restrict__ w0, double* __restrict__ w1)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 1, 1>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    const long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    const double* p = w0 + o;
    double* q = w1 + o;
    double zm = p[-sz];
    double c = p[0];
#pragma unroll 4
    for (int k = t.k0; k < t.k1; k++) {
        const double zp = p[sz];
        q[0] = alpha * c + beta * (
            p[1] + p[-1] +
            p[sy] + p[-sy] +
            zp + zm);
        zm = c;
        c = zp;
        p += sz;
        q += sz;
    }
}

// --------------------------------------------------------------- divergence
// divergence/divergence.c:93-96
__global__ __launch_bounds__(TX * TY)
void divergence_k(int nx, int ny, int ns, int ntx, int nty, double alpha, double beta, double gamma,
                  double* __restrict__ u, const double* __restrict__ ux,
                  const double* __restrict__ uy, const double* __restrict__ uz)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 1, 1>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    double zm = uz[o - sz];
    double zc = uz[o];
#pragma unroll 4
    for (int k = t.k0; k < t.k1; k++) {
        const double zp = uz[o + sz];
        u[o] = alpha * (ux[o + 1] - ux[o - 1]) +
               beta * (uy[o + sy] - uy[o - sy]) +
               gamma * (zp - zm);
        zm = zc;
        zc = zp;
        o += sz;
    }
}

// ----------------------------------------------------------------- gradient
// gradient/gradient.c:95-97
__global__ __launch_bounds__(TX * TY)
void gradient_k(int nx, int ny, int ns, int ntx, int nty, double alpha, double beta, double gamma,
                const double* __restrict__ u, double* __restrict__ ux,
                double* __restrict__ uy, double* __restrict__ uz)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 1, 1>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    double zm = u[o - sz];
    double zc = u[o];
#pragma unroll 4
    for (int k = t.k0; k < t.k1; k++) {
        const double zp = u[o + sz];
        ux[o] = alpha * (u[o + 1] - u[o - 1]);
        uy[o] = beta * (u[o + sy] - u[o - sy]);
        uz[o] = gamma * (zp - zm);
        zm = zc;
        zc = zp;
        o += sz;
    }
}

// ----------------------------------------------------------------- wave13pt
// wave13pt/wave13pt.c:640-651 (default path), interior [2,n-2)^3
__global__ __launch_bounds__(TX * TY)
void wave13pt_k(int nx, int ny, int ns, int ntx, int nty, double m0, double m1, double m2,
                const double* __restrict__ w0, const double* __restrict__ w1,
                double* __restrict__ w2)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 2, 2>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    const double* p = w1 + o;
    double zm2 = p[-2 * sz], zm1 = p[-sz], zc = p[0], zp1 = p[sz];
#pragma unroll 4
    for (int k = t.k0; k < t.k1; k++) {
        const double zp2 = p[2 * sz];
        w2[o] = m0 * zc - w0[o] +
            m1 * (
                p[1] + p[-1] +
                p[sy] + p[-sy] +
                zp1 + zm1) +
            m2 * (
                p[2] + p[-2] +
                p[2 * sy] + p[-2 * sy] +
                zp2 + zm2);
        zm2 = zm1;
        zm1 = zc;
        zc = zp1;
        zp1 = zp2;
        p += sz;
        o += sz;
    }
}

// --------------------------------------------------------------------- uxx1
// uxx1/uxx1.c:95-105, interior [2,n-1)^3, dth = 1/nx (:70)
__global__ __launch_bounds__(TX * TY)
void uxx1_k(int nx, int ny, int ns, int ntx, int nty, double c1, double c2, double dth,
            const double* __restrict__ u0, double* __restrict__ u1,
            const double* __restrict__ d1, const double* __restrict__ xx,
            const double* __restrict__ xy, const double* __restrict__ xz)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 2, 1>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    // z-columns kept in registers: d1 at (i,j) and (i,j-1), planes k-1..k;
    // xz at (i,j), planes k-2..k+1.
    double d_c_m = d1[o - sz], d_y_m = d1[o - sy - sz];
    double xz_m2 = xz[o - 2 * sz], xz_m1 = xz[o - sz], xz_c = xz[o];
#pragma unroll 2
    for (int k = t.k0; k < t.k1; k++) {
        const double d_c = d1[o], d_y = d1[o - sy];
        const double xz_p1 = xz[o + sz];
        const double d = 0.25 * (d_c + d_y + d_c_m + d_y_m);
        u1[o] = u0[o] + (dth / d) * (
            c1 * (xx[o] - xx[o - 1]) +
            c2 * (xx[o + 1] - xx[o - 2]) +
            c1 * (xy[o] - xy[o - sy]) +
            c2 * (xy[o + sy] - xy[o - 2 * sy]) +
            c1 * (xz_c - xz_m1) +
            c2 * (xz_p1 - xz_m2));
        d_c_m = d_c;
        d_y_m = d_y;
        xz_m2 = xz_m1;
        xz_m1 = xz_c;
        xz_c = xz_p1;
        o += sz;
    }
}

// ------------------------------------------------------------------ lapgsrb
// lapgsrb/lapgsrb.c:93-117, interior [2,n-2)^3
__global__ __launch_bounds__(TX * TY)
void lapgsrb_k(int nx, int ny, int ns, int ntx, int nty, double c0, double c1, double c2, double c3,
               const double* __restrict__ w0, double* __restrict__ w1)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 2, 2>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    const double* p = w0 + o;
    double zm2 = p[-2 * sz], zm1 = p[-sz], zc = p[0], zp1 = p[sz];
#pragma unroll 2
    for (int k = t.k0; k < t.k1; k++) {
        const double zp2 = p[2 * sz];
        w1[o] = c0 * zc +
            c1 * (
                p[1] + p[-1] +
                p[sy] + p[-sy] +
                zp1 + zm1) +
            c2 * (
                p[1 + sy] + p[1 + sz] + p[1 - sz] + p[1 - sy] +
                p[sy + sz] + p[sy - sz] + p[-sy + sz] + p[-sy - sz] +
                p[-1 + sy] + p[-1 + sz] + p[-1 - sz] + p[-1 - sy]) +
            c3 * (
                p[2] + p[-2] +
                p[2 * sy] + p[-2 * sy] +
                zp2 + zm2);
        zm2 = zm1;
        zm1 = zc;
        zc = zp1;
        zp1 = zp2;
        p += sz;
        o += sz;
    }
}

}  // namespace

// ============================================================ entry points

#define KG_LAUNCH3(NAME, KERNEL, LO, HI, ...)                                              \
    do {                                                                                   \
        if (int rc_ = kg_check_ctx(ctx)) return rc_;                                       \
        if (int rc_ = kg_check_dims3(nx, ny, ns)) return rc_;                              \
        const kg_zgrid g = kg_zgrid_make(nx, ny, ns, LO, HI, TX, TY, KC);                  \
        kg_register_kernel(NAME, (const void*)KERNEL);                                     \
        if (int rc_ = kg_launch_begin(ctx, NAME)) return rc_;                              \
        if (!g.empty)                                                                      \
            hipLaunchKernelGGL(KERNEL, dim3(g.blocks), dim3(TX, TY), 0, ctx->stream,       \
                               nx, ny, ns, g.ntx, g.nty, __VA_ARGS__);                     \
        return kg_launch_end(ctx);                                                         \
    } while (0)

extern "C" int kg_laplacian(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                            const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_laplacian: null buffer");
    KG_REQUIRE(w0 != w1, "kg_laplacian: w0 and w1 must be distinct (out-of-place update)");
    KG_LAUNCH3("laplacian", laplacian_k, 1, 1, alpha, beta, w0, w1);
}

extern "C" int kg_divergence(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                             double gamma, double* u, const double* ux, const double* uy,
                             const double* uz)
{
    KG_REQUIRE(u && ux && uy && uz, "kg_divergence: null buffer");
    KG_REQUIRE(u != ux && u != uy && u != uz, "kg_divergence: output aliases an input");
    KG_LAUNCH3("divergence", divergence_k, 1, 1, alpha, beta, gamma, u, ux, uy, uz);
}

extern "C" int kg_gradient(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                           double gamma, const double* u, double* ux, double* uy, double* uz)
{
    KG_REQUIRE(u && ux && uy && uz, "kg_gradient: null buffer");
    KG_REQUIRE(u != ux && u != uy && u != uz && ux != uy && ux != uz && uy != uz,
               "kg_gradient: buffers must be distinct");
    KG_LAUNCH3("gradient", gradient_k, 1, 1, alpha, beta, gamma, u, ux, uy, uz);
}

extern "C" int kg_wave13pt(kg_ctx* ctx, int nx, int ny, int ns, double m0, double m1, double m2,
                           const double* w0, const double* w1, double* w2)
{
    KG_REQUIRE(w0 && w1 && w2, "kg_wave13pt: null buffer");
    KG_REQUIRE(w2 != w0 && w2 != w1, "kg_wave13pt: w2 aliases an input");
    KG_LAUNCH3("wave13pt", wave13pt_k, 2, 2, m0, m1, m2, w0, w1, w2);
}

extern "C" int kg_uxx1(kg_ctx* ctx, int nx, int ny, int ns, double c1, double c2,
                       const double* u0, double* u1, const double* d1,
                       const double* xx, const double* xy, const double* xz)
{
    KG_REQUIRE(u0 && u1 && d1 && xx && xy && xz, "kg_uxx1: null buffer");
    KG_REQUIRE(u1 != u0 && u1 != d1 && u1 != xx && u1 != xy && u1 != xz,
               "kg_uxx1: u1 aliases an input");
    const double dth = 1. / nx;   // uxx1/uxx1.c:70
    KG_LAUNCH3("uxx1", uxx1_k, 2, 1, c1, c2, dth, u0, u1, d1, xx, xy, xz);
}

extern "C" int kg_lapgsrb(kg_ctx* ctx, int nx, int ny, int ns, double c0, double c1, double c2,
                          double c3, const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_lapgsrb: null buffer");
    KG_REQUIRE(w0 != w1, "kg_lapgsrb: w0 and w1 must be distinct");
    KG_LAUNCH3("lapgsrb", lapgsrb_k, 2, 2, c0, c1, c2, c3, w0, w1);
}
