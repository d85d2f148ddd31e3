/*
 * kg_cpu.c -- the `gcc` target of the suite: the C ABI of kg_stencil.h
 * implemented with plain host loops on host memory, built with the
 * reference's gcc flags (makefile.in:28-29: -O3 -ffast-math -march=native,
 * single-threaded unless -fopenmp is added).  It is the CPU baseline that
 * `./benchmark ... hip gcc` times beside the GPU; the `hip` target never
 * links it.  Loops follow the reference kernels (cited per function); whispering
 * lives in kg_cpu_whispering.c because its reference target is built -O0.
 */
#include <malloc.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/kg_stencil.h"

const int kg_target_has_device = 0;

struct kg_ctx {
    double last_time;
    const char* last_kernel;
};

static char g_err[256] = "";

#define IX3(i, j, k) ((size_t)(i) + (size_t)nx * (size_t)(j) + (size_t)nx * (size_t)ny * (size_t)(k))
#define IX2(i, j) ((size_t)(i) + (size_t)nx * (size_t)(j))

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_REALTIME, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void toc(kg_ctx* ctx, const char* name, double t0)
{
    ctx->last_time = now() - t0;
    ctx->last_kernel = name;
}

int kg_device_count(int* count) { if (!count) return KG_ERR_INVALID; *count = 0; return KG_OK; }
int kg_init(kg_ctx** ctx, int device)
{
    (void)device;
    if (!ctx) return KG_ERR_INVALID;
    *ctx = (kg_ctx*)calloc(1, sizeof(kg_ctx));
    (*ctx)->last_kernel = "";
    return *ctx ? KG_OK : KG_ERR_NOMEM;
}
int kg_destroy(kg_ctx* ctx) { free(ctx); return KG_OK; }
int kg_alloc(kg_ctx* ctx, void** p, size_t bytes)
{
    (void)ctx;
    *p = bytes ? memalign(4096, bytes) : NULL;
    return (*p || !bytes) ? KG_OK : KG_ERR_NOMEM;
}
int kg_free(kg_ctx* ctx, void* p) { (void)ctx; free(p); return KG_OK; }
int kg_h2d(kg_ctx* ctx, void* d, const void* s, size_t n) { (void)ctx; memcpy(d, s, n); return KG_OK; }
int kg_d2h(kg_ctx* ctx, void* d, const void* s, size_t n) { (void)ctx; memcpy(d, s, n); return KG_OK; }
int kg_d2d(kg_ctx* ctx, void* d, const void* s, size_t n) { (void)ctx; memmove(d, s, n); return KG_OK; }
int kg_sync(kg_ctx* ctx) { (void)ctx; return KG_OK; }
int kg_timing_enable(kg_ctx* ctx, int on) { (void)ctx; (void)on; return KG_OK; }
int kg_last_kernel_time(kg_ctx* ctx, double* s) { *s = ctx->last_time; return KG_OK; }
int kg_timing_drain(kg_ctx* ctx, double* s, int cap, int* count)
{
    (void)ctx; (void)s; (void)cap;
    *count = 0;
    return KG_OK;
}
const char* kg_last_kernel_name(kg_ctx* ctx) { return ctx ? ctx->last_kernel : ""; }
int kg_kernel_regcount(kg_ctx* ctx, const char* k, int* n) { (void)ctx; (void)k; *n = 0; return KG_ERR_STATE; }
const char* kg_strerror(int status)
{
    switch (status) {
    case KG_OK: return "success";
    case KG_ERR_INVALID: return "invalid argument";
    case KG_ERR_NOMEM: return "out of memory";
    default: return "error";
    }
}
const char* kg_last_error(void) { return g_err; }

int kg_stream_copy(kg_ctx* ctx, double* dst, const double* src, size_t n)
{
    {
        const double t0_ = now();
        memcpy(dst, src, n * sizeof(double));
        toc(ctx, "stream_copy", t0_);
    }
    return KG_OK;
}

/* laplacian/laplacian.c:81-100 */
int kg_laplacian(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                 const double* w0, double* w1)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int k = 1; k < ns - 1; k++)
            for (int j = 1; j < ny - 1; j++)
                for (int i = 1; i < nx - 1; i++)
                    w1[IX3(i, j, k)] = alpha * w0[IX3(i, j, k)] + beta * (
                        w0[IX3(i + 1, j, k)] + w0[IX3(i - 1, j, k)] +
                        w0[IX3(i, j + 1, k)] + w0[IX3(i, j - 1, k)] +
                        w0[IX3(i, j, k + 1)] + w0[IX3(i, j, k - 1)]);
        toc(ctx, "laplacian", t0_);
    }
    return KG_OK;
}

/* divergence/divergence.c:81-99 */
int kg_divergence(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta, double gamma,
                  double* u, const double* ux, const double* uy, const double* uz)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int k = 1; k < ns - 1; k++)
            for (int j = 1; j < ny - 1; j++)
                for (int i = 1; i < nx - 1; i++)
                    u[IX3(i, j, k)] =
                        alpha * (ux[IX3(i + 1, j, k)] - ux[IX3(i - 1, j, k)]) +
                        beta * (uy[IX3(i, j + 1, k)] - uy[IX3(i, j - 1, k)]) +
                        gamma * (uz[IX3(i, j, k + 1)] - uz[IX3(i, j, k - 1)]);
        toc(ctx, "divergence", t0_);
    }
    return KG_OK;
}

/* gradient/gradient.c:83-100 */
int kg_gradient(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta, double gamma,
                const double* u, double* ux, double* uy, double* uz)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int k = 1; k < ns - 1; k++)
            for (int j = 1; j < ny - 1; j++)
                for (int i = 1; i < nx - 1; i++) {
                    ux[IX3(i, j, k)] = alpha * (u[IX3(i + 1, j, k)] - u[IX3(i - 1, j, k)]);
                    uy[IX3(i, j, k)] = beta * (u[IX3(i, j + 1, k)] - u[IX3(i, j - 1, k)]);
                    uz[IX3(i, j, k)] = gamma * (u[IX3(i, j, k + 1)] - u[IX3(i, j, k - 1)]);
                }
        toc(ctx, "gradient", t0_);
    }
    return KG_OK;
}

/* wave13pt/wave13pt.c:143-160,640-651 */
int kg_wave13pt(kg_ctx* ctx, int nx, int ny, int ns, double m0, double m1, double m2,
                const double* restrict w0, const double* restrict w1, double* restrict w2)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int k = 2; k < ns - 2; k++)
            for (int j = 2; j < ny - 2; j++)
                for (int i = 2; i < nx - 2; i++)
                    w2[IX3(i, j, k)] = m0 * w1[IX3(i, j, k)] - w0[IX3(i, j, k)] +
                        m1 * (
                            w1[IX3(i + 1, j, k)] + w1[IX3(i - 1, j, k)] +
                            w1[IX3(i, j + 1, k)] + w1[IX3(i, j - 1, k)] +
                            w1[IX3(i, j, k + 1)] + w1[IX3(i, j, k - 1)]) +
                        m2 * (
                            w1[IX3(i + 2, j, k)] + w1[IX3(i - 2, j, k)] +
                            w1[IX3(i, j + 2, k)] + w1[IX3(i, j - 2, k)] +
                            w1[IX3(i, j, k + 2)] + w1[IX3(i, j, k - 2)]);
        toc(ctx, "wave13pt", t0_);
    }
    return KG_OK;
}

/* uxx1/uxx1.c:70,83-105 */
int kg_uxx1(kg_ctx* ctx, int nx, int ny, int ns, double c1, double c2,
            const double* u0, double* u1, const double* d1,
            const double* xx, const double* xy, const double* xz)
{
    {
        const double t0_ = now();
        const double dth = 1. / nx;
#pragma omp parallel for
        for (int k = 2; k < ns - 1; k++)
            for (int j = 2; j < ny - 1; j++)
                for (int i = 2; i < nx - 1; i++) {
                    double d = 0.25 * (d1[IX3(i, j, k)] + d1[IX3(i, j - 1, k)] +
                                       d1[IX3(i, j, k - 1)] + d1[IX3(i, j - 1, k - 1)]);
                    u1[IX3(i, j, k)] = u0[IX3(i, j, k)] + (dth / d) * (
                        c1 * (xx[IX3(i, j, k)] - xx[IX3(i - 1, j, k)]) +
                        c2 * (xx[IX3(i + 1, j, k)] - xx[IX3(i - 2, j, k)]) +
                        c1 * (xy[IX3(i, j, k)] - xy[IX3(i, j - 1, k)]) +
                        c2 * (xy[IX3(i, j + 1, k)] - xy[IX3(i, j - 2, k)]) +
                        c1 * (xz[IX3(i, j, k)] - xz[IX3(i, j, k - 1)]) +
                        c2 * (xz[IX3(i, j, k + 1)] - xz[IX3(i, j, k - 2)]));
                }
        toc(ctx, "uxx1", t0_);
    }
    return KG_OK;
}

#define TRI_WEIGHTS(a, w)                                        \
    w[0] = 1.0 / 6.0 * (a) * ((a) + 1.0) * ((a) + 2.0);           \
    w[1] = -0.5 * ((a) - 1.0) * ((a) + 1.0) * ((a) + 2.0);       \
    w[2] = 0.5 * ((a) - 1.0) * (a) * ((a) + 2.0);                \
    w[3] = -1.0 / 6.0 * ((a) - 1.0) * (a) * ((a) + 1.0)

/* tricubic/tricubic.c:138-155,509,1046-1145 */
int kg_tricubic(kg_ctx* ctx, int nx, int ny, int ns, const double* restrict u0,
                double* restrict u1, const double* restrict a, const double* restrict b,
                const double* restrict c)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int k = 1; k < ns - 2; k++)
            for (int j = 1; j < ny - 2; j++)
                for (int i = 1; i < nx - 2; i++) {
                    double wa[4], wb[4], wc[4];
                    TRI_WEIGHTS(a[IX3(i, j, k)], wa);
                    TRI_WEIGHTS(b[IX3(i, j, k)], wb);
                    TRI_WEIGHTS(c[IX3(i, j, k)], wc);
                    double s = wa[0] * wb[0] * wc[0] * u0[IX3(i - 1, j - 1, k - 1)];
                    for (int kk = 0; kk < 4; kk++)
                        for (int jj = 0; jj < 4; jj++)
                            for (int ii = 0; ii < 4; ii++)
                                if (kk | jj | ii)
                                    s = s + wa[ii] * wb[jj] * wc[kk] *
                                        u0[IX3(i + ii - 1, j + jj - 1, k + kk - 1)];
                    u1[IX3(i, j, k)] = s;
                }
        toc(ctx, "tricubic", t0_);
    }
    return KG_OK;
}

/* tricubic2/tricubic2.c:69-102 (accumulates into u1 through plain pointers) */
int kg_tricubic2(kg_ctx* ctx, int nx, int ny, int ns, const double* u0, double* u1,
                 const double* a, const double* b, const double* c)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int k = 2; k < ns - 2; k++)
            for (int j = 2; j < ny - 2; j++)
                for (int i = 2; i < nx - 2; i++) {
                    double wa[4], wb[4], wc[4];
                    TRI_WEIGHTS(a[IX3(i, j, k)], wa);
                    TRI_WEIGHTS(b[IX3(i, j, k)], wb);
                    TRI_WEIGHTS(c[IX3(i, j, k)], wc);
                    u1[IX3(i, j, k)] = 0;
                    for (int kk = 0; kk < 4; kk++)
                        for (int jj = 0; jj < 4; jj++)
                            for (int ii = 0; ii < 4; ii++)
                                u1[IX3(i, j, k)] += wa[ii] * wb[jj] * wc[kk] *
                                    u0[IX3(i + ii - 1, j + jj - 1, k + kk - 1)];
                }
        toc(ctx, "tricubic2", t0_);
    }
    return KG_OK;
}

/* lapgsrb/lapgsrb.c:81-117 */
int kg_lapgsrb(kg_ctx* ctx, int nx, int ny, int ns, double c0, double c1, double c2, double c3,
               const double* w0, double* w1)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int k = 2; k < ns - 2; k++)
            for (int j = 2; j < ny - 2; j++)
                for (int i = 2; i < nx - 2; i++) {
#define W(di, dj, dk) w0[IX3(i + (di), j + (dj), k + (dk))]
                    w1[IX3(i, j, k)] = c0 * W(0, 0, 0) +
                        c1 * (W(1, 0, 0) + W(-1, 0, 0) + W(0, 1, 0) + W(0, -1, 0) +
                              W(0, 0, 1) + W(0, 0, -1)) +
                        c2 * (W(1, 1, 0) + W(1, 0, 1) + W(1, 0, -1) + W(1, -1, 0) +
                              W(0, 1, 1) + W(0, 1, -1) + W(0, -1, 1) + W(0, -1, -1) +
                              W(-1, 1, 0) + W(-1, 0, 1) + W(-1, 0, -1) + W(-1, -1, 0)) +
                        c3 * (W(2, 0, 0) + W(-2, 0, 0) + W(0, 2, 0) + W(0, -2, 0) +
                              W(0, 0, 2) + W(0, 0, -2));
#undef W
                }
        toc(ctx, "lapgsrb", t0_);
    }
    return KG_OK;
}

/* jacobi/jacobi.F90:60-69 */
int kg_jacobi(kg_ctx* ctx, int nx, int ny, double c0, double c1, double c2,
              const double* w0, double* w1)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
#define W(di, dj) w0[IX2(i + (di), j + (dj))]
                w1[IX2(i, j)] = c0 * W(0, 0) +
                    c1 * (W(-1, 0) + W(0, -1) + W(1, 0) + W(0, 1)) +
                    c2 * (W(-1, -1) + W(-1, 1) + W(1, -1) + W(1, 1));
#undef W
            }
        toc(ctx, "jacobi_", t0_);
    }
    return KG_OK;
}

/* gaussblur/gaussblur.c:65,78-92 */
int kg_gaussblur(kg_ctx* ctx, int nx, int ny, double s0, double s1, double s2,
                 double s4, double s5, double s8, const double* w0, double* w1)
{
    {
        const double t0_ = now();
        const double f = 1. / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5);
#pragma omp parallel for
        for (int j = 2; j < ny - 2; j++)
            for (int i = 2; i < nx - 2; i++) {
#define W(di, dj) w0[IX2(i + (di), j + (dj))]
                w1[IX2(i, j)] = f * (
                    s0 * W(0, 0) +
                    s1 * (W(-1, 0) + W(1, 0) + W(0, -1) + W(0, 1)) +
                    s2 * (W(-1, -1) + W(1, -1) + W(-1, 1) + W(1, 1)) +
                    s4 * (W(-2, 0) + W(2, 0) + W(0, -2) + W(0, 2)) +
                    s5 * (W(-2, -1) + W(-1, -2) + W(1, -2) + W(2, -1) +
                          W(-2, 1) + W(-1, 2) + W(1, 2) + W(2, 1)) +
                    s8 * (W(-2, -2) + W(2, -2) + W(-2, 2) + W(2, 2)));
#undef W
            }
        toc(ctx, "gaussblur", t0_);
    }
    return KG_OK;
}

/* gameoflife/gameoflife.c:74-91 */
int kg_gameoflife(kg_ctx* ctx, int nx, int ny, const double* u0, double* u1)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
#define U(di, dj) u0[IX2(i + (di), j + (dj))]
                double C = 100000000000000000000.;
                double L = U(-1, -1) + U(0, -1) + U(1, -1) + U(-1, 0) + U(1, 0) +
                           U(-1, 1) + U(0, 1) + U(1, 1);
                u1[IX2(i, j)] = 1. / (1. + (U(0, 0) + L - 3.) * (L - 3.) * C);
#undef U
            }
        toc(ctx, "gameoflife", t0_);
    }
    return KG_OK;
}

/* matmul/matmul.F90:56-68 */
int kg_matmul(kg_ctx* ctx, int nx, int ny, int ns, const double* A, const double* B, double* C)
{
    {
        const double t0_ = now();
#pragma omp parallel for
        for (int j = 0; j < ns; j++)
            for (int i = 0; i < nx; i++)
                for (int k = 0; k < ny; k++)
                    C[(size_t)i + (size_t)nx * j] = C[(size_t)i + (size_t)nx * j] +
                        A[(size_t)i + (size_t)nx * k] * B[(size_t)k + (size_t)ny * j];
        toc(ctx, "matmul_", t0_);
    }
    return KG_OK;
}

/* ---- the suite's non-stencil tests (real = double: makefile.in:12) ---- */
/* vecadd/vecadd.c:69-84 */
int kg_vecadd(kg_ctx* ctx, int nx, int ny, int ns, const double* w0, const double* w1, double* w2)
{
    const double t0_ = now();
#pragma omp parallel for
    for (int k = 0; k < ns; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t c = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                w2[c] = w0[c] + w1[c];
            }
    toc(ctx, "vecadd", t0_);
    return KG_OK;
}

/* matvec/matvec.c:38-70 */
int kg_matvec(kg_ctx* ctx, int nx, int ny, const double* A, const double* x, double* y)
{
    const double t0_ = now();
#pragma omp parallel for
    for (int j = 0; j < ny; j++) {
        y[j] = 0;
        for (int i = 0; i < nx; i++) y[j] += A[(size_t)i + (size_t)nx * j] * x[i];
    }
    toc(ctx, "matvec", t0_);
    return KG_OK;
}

/* sincos/sincos.F90:62-70 */
int kg_sincos(kg_ctx* ctx, int nx, int ny, int ns, const double* x, const double* y, double* xy)
{
    const double t0_ = now();
#pragma omp parallel for
    for (int k = 0; k < ns; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const size_t c = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                xy[c] = sin(x[c]) + cos(y[c]);
            }
    toc(ctx, "sincos", t0_);
    return KG_OK;
}
