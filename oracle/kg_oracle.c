/*
 * kg_oracle.c -- CPU restatement of the kernelgen-perf-tests stencil kernels.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see kg_oracle.h).  Built with
 * -O2 -fno-fast-math -ffp-contract=off: every expression below keeps the
 * reference's operand order and association, so results are bit-identical to
 * an IEEE build of the reference sources.
 *
 * Reference citations are relative to the kernelgen-perf-tests tree.
 */
#include "kg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX3(i, j, k) ((size_t)(i) + (size_t)nx * (size_t)(j) + (size_t)nx * (size_t)ny * (size_t)(k))
#define IDX2(i, j) ((size_t)(i) + (size_t)nx * (size_t)(j))

/* laplacian/laplacian.c:81-100 -- interior [1,n-1)^3,
 * w1 = alpha*c + beta*(x+ + x- + y+ + y- + z+ + z-). */
void oracle_laplacian(int nx, int ny, int ns, double alpha, double beta,
                      const double* w0, double* w1)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    for (int k = 1; k < ns - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                size_t c = IDX3(i, j, k);
                w1[c] = alpha * w0[c] + beta * (
                    w0[c + 1] + w0[c - 1] +
                    w0[c + sy] + w0[c - sy] +
                    w0[c + sz] + w0[c - sz]);
            }
}

/* divergence/divergence.c:81-99 */
void oracle_divergence(int nx, int ny, int ns, double alpha, double beta, double gamma,
                       double* u, const double* ux, const double* uy, const double* uz)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    for (int k = 1; k < ns - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                size_t c = IDX3(i, j, k);
                u[c] = alpha * (ux[c + 1] - ux[c - 1]) +
                       beta * (uy[c + sy] - uy[c - sy]) +
                       gamma * (uz[c + sz] - uz[c - sz]);
            }
}

/* gradient/gradient.c:83-100 */
void oracle_gradient(int nx, int ny, int ns, double alpha, double beta, double gamma,
                     const double* u, double* ux, double* uy, double* uz)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    for (int k = 1; k < ns - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                size_t c = IDX3(i, j, k);
                ux[c] = alpha * (u[c + 1] - u[c - 1]);
                uy[c] = beta * (u[c + sy] - u[c - sy]);
                uz[c] = gamma * (u[c + sz] - u[c - sz]);
            }
}

/* wave13pt/wave13pt.c:143-160 (loop bounds: k in [2,ns-2), j and i skip the
 * two outer layers via `continue`, :153,:495) and :640-651 (update). */
void oracle_wave13pt(int nx, int ny, int ns, double m0, double m1, double m2,
                     const double* w0, const double* w1, double* w2)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    for (int k = 2; k < ns - 2; k++)
        for (int j = 2; j < ny - 2; j++)
            for (int i = 2; i < nx - 2; i++) {
                size_t c = IDX3(i, j, k);
                w2[c] = m0 * w1[c] - w0[c] +
                    m1 * (
                        w1[c + 1] + w1[c - 1] +
                        w1[c + sy] + w1[c - sy] +
                        w1[c + sz] + w1[c - sz]) +
                    m2 * (
                        w1[c + 2] + w1[c - 2] +
                        w1[c + 2 * sy] + w1[c - 2 * sy] +
                        w1[c + 2 * sz] + w1[c - 2 * sz]);
            }
}

/* uxx1/uxx1.c:70 (dth) and :83-105, interior [2,n-1)^3. */
void oracle_uxx1(int nx, int ny, int ns, double c1, double c2,
                 const double* u0, double* u1, const double* d1,
                 const double* xx, const double* xy, const double* xz)
{
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const double dth = 1. / nx;
    for (int k = 2; k < ns - 1; k++)
        for (int j = 2; j < ny - 1; j++)
            for (int i = 2; i < nx - 1; i++) {
                size_t c = IDX3(i, j, k);
                double d = 0.25 * (d1[c] + d1[c - sy] + d1[c - sz] + d1[c - sy - sz]);
                u1[c] = u0[c] + (dth / d) * (
                    c1 * (xx[c] - xx[c - 1]) +
                    c2 * (xx[c + 1] - xx[c - 2]) +
                    c1 * (xy[c] - xy[c - sy]) +
                    c2 * (xy[c + sy] - xy[c - 2 * sy]) +
                    c1 * (xz[c] - xz[c - sz]) +
                    c2 * (xz[c + sz] - xz[c - 2 * sz]));
            }
}

/* Lagrange-cubic weights at nodes {-1,0,1,2}, written exactly as the FP64
 * branch of tricubic/tricubic.c:1046-1059 and tricubic2/tricubic2.c:83-94. */
static inline void tricubic_weights(double a, double w[4])
{
    w[0] = 1.0 / 6.0 * a * (a + 1.0) * (a + 2.0);
    w[1] = -0.5 * (a - 1.0) * (a + 1.0) * (a + 2.0);
    w[2] = 0.5 * (a - 1.0) * a * (a + 2.0);
    w[3] = -1.0 / 6.0 * (a - 1.0) * a * (a + 1.0);
}

/* tricubic/tricubic.c:138-155,509 (k in [1,ns-2), j in [1,ny-2), i in
 * [1,nx-2)) and :1062-1145: 64 products (wa*wb)*wc*u summed in k-plane,
 * j-row, i order, the first product starting the sum. */
void oracle_tricubic(int nx, int ny, int ns, const double* u0, double* u1,
                     const double* a, const double* b, const double* c)
{
    for (int k = 1; k < ns - 2; k++)
        for (int j = 1; j < ny - 2; j++)
            for (int i = 1; i < nx - 2; i++) {
                size_t p = IDX3(i, j, k);
                double wa[4], wb[4], wc[4];
                tricubic_weights(a[p], wa);
                tricubic_weights(b[p], wb);
                tricubic_weights(c[p], wc);
                double s = 0.0;
                int first = 1;
                for (int kk = 0; kk < 4; kk++)
                    for (int jj = 0; jj < 4; jj++)
                        for (int ii = 0; ii < 4; ii++) {
                            double t = wa[ii] * wb[jj] * wc[kk] *
                                u0[IDX3(i + ii - 1, j + jj - 1, k + kk - 1)];
                            if (first) { s = t; first = 0; }
                            else s = s + t;
                        }
                u1[p] = s;
            }
}

/* tricubic2/tricubic2.c:69-102: same weights, interior [2,n-2)^3, the sum
 * starts from 0 and accumulates every product. */
void oracle_tricubic2(int nx, int ny, int ns, const double* u0, double* u1,
                      const double* a, const double* b, const double* c)
{
    for (int k = 2; k < ns - 2; k++)
        for (int j = 2; j < ny - 2; j++)
            for (int i = 2; i < nx - 2; i++) {
                size_t p = IDX3(i, j, k);
                double wa[4], wb[4], wc[4];
                tricubic_weights(a[p], wa);
                tricubic_weights(b[p], wb);
                tricubic_weights(c[p], wc);
                double s = 0;
                for (int kk = 0; kk < 4; kk++)
                    for (int jj = 0; jj < 4; jj++)
                        for (int ii = 0; ii < 4; ii++)
                            s += wa[ii] * wb[jj] * wc[kk] *
                                u0[IDX3(i + ii - 1, j + jj - 1, k + kk - 1)];
                u1[p] = s;
            }
}

/* lapgsrb/lapgsrb.c:81-117, interior [2,n-2)^3, 25 points. */
void oracle_lapgsrb(int nx, int ny, int ns, double c0, double c1, double c2, double c3,
                    const double* w0, double* w1)
{
    for (int k = 2; k < ns - 2; k++)
        for (int j = 2; j < ny - 2; j++)
            for (int i = 2; i < nx - 2; i++) {
#define W(di, dj, dk) w0[IDX3(i + (di), j + (dj), k + (dk))]
                w1[IDX3(i, j, k)] = c0 * W(0, 0, 0) +
                    c1 * (
                        W(1, 0, 0) + W(-1, 0, 0) +
                        W(0, 1, 0) + W(0, -1, 0) +
                        W(0, 0, 1) + W(0, 0, -1)) +
                    c2 * (
                        W(1, 1, 0) + W(1, 0, 1) + W(1, 0, -1) + W(1, -1, 0) +
                        W(0, 1, 1) + W(0, 1, -1) + W(0, -1, 1) + W(0, -1, -1) +
                        W(-1, 1, 0) + W(-1, 0, 1) + W(-1, 0, -1) + W(-1, -1, 0)) +
                    c3 * (
                        W(2, 0, 0) + W(-2, 0, 0) +
                        W(0, 2, 0) + W(0, -2, 0) +
                        W(0, 0, 2) + W(0, 0, -2));
#undef W
            }
}

/* jacobi/jacobi.F90:60-69: Fortran do j = 2, ny-1; do i = 2, nx-1 (1-based)
 * == 0-based [1,n-1)^2; operands evaluated left to right as written. */
void oracle_jacobi(int nx, int ny, double c0, double c1, double c2,
                   const double* w0, double* w1)
{
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
#define W(di, dj) w0[IDX2(i + (di), j + (dj))]
            w1[IDX2(i, j)] = c0 * W(0, 0) +
                c1 * (W(-1, 0) + W(0, -1) + W(1, 0) + W(0, 1)) +
                c2 * (W(-1, -1) + W(-1, 1) + W(1, -1) + W(1, 1));
#undef W
        }
}

/* gaussblur/gaussblur.c:65 (f) and :78-92, interior [2,n-2)^2. */
void oracle_gaussblur(int nx, int ny, double s0, double s1, double s2,
                      double s4, double s5, double s8, const double* w0, double* w1)
{
    const double f = 1. / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5);
    for (int j = 2; j < ny - 2; j++)
        for (int i = 2; i < nx - 2; i++) {
#define W(di, dj) w0[IDX2(i + (di), j + (dj))]
            w1[IDX2(i, j)] = f * (
                s0 * W(0, 0) +
                s1 * (W(-1, 0) + W(1, 0) + W(0, -1) + W(0, 1)) +
                s2 * (W(-1, -1) + W(1, -1) + W(-1, 1) + W(1, 1)) +
                s4 * (W(-2, 0) + W(2, 0) + W(0, -2) + W(0, 2)) +
                s5 * (W(-2, -1) + W(-1, -2) + W(1, -2) + W(2, -1) +
                      W(-2, 1) + W(-1, 2) + W(1, 2) + W(2, 1)) +
                s8 * (W(-2, -2) + W(2, -2) + W(-2, 2) + W(2, 2)));
#undef W
        }
}

/* whispering/whispering.c:87-106: the serial sweep with e updated in place,
 * j outer, i inner, exactly as the reference's gcc target executes it. */
void oracle_whispering(int nx, int ny, double mu, double epsilon, double* e,
                       const double* h0, double* h1, const double* u_em0, double* u_em1,
                       const double* ca, const double* cb, const double* da, const double* db)
{
    double* e_0 = e;
    double* e_1 = e + (size_t)nx * ny;
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            size_t c = IDX2(i, j), cy = IDX2(i, j - 1), cx = IDX2(i - 1, j);
            double e0 = ca[c] * e_0[c] + cb[c] * (h0[IDX2(i, j + 1)] - h0[c]);
            double e1 = ca[c] * e_1[c] - cb[c] * (h0[IDX2(i + 1, j)] - h0[c]);
            double ey = ca[cy] * e_0[cy] + cb[cy] * (h0[c] - h0[cy]);
            double ex = ca[cx] * e_1[cx] - cb[cx] * (h0[c] - h0[cx]);
            e_0[c] = e0;
            e_1[c] = e1;
            double h = da[c] * h0[c] + db[c] * (e0 - ey + ex - e1);
            h1[c] = h;
            u_em1[c] = u_em0[c] + 0.5 * (h * h / mu + epsilon * (e0 * e0 + e1 * e1));
        }
}

/* Closed form of the serial sweep above.  When (i,j) is visited, e(i,j-1,0)
 * has already been replaced by the (i,j-1) update if j-1 >= 1, and
 * e(i-1,j,1) by the (i-1,j) update if i-1 >= 1; both replacements are
 * recomputed here from e_in.  The (i,j-1) update's x-difference term is
 * h0(i,j)-h0(i,j-1), the same one ey uses; likewise for (i-1,j). */
void oracle_whispering_closed(int nx, int ny, double mu, double epsilon,
                              const double* e_in, double* e_out,
                              const double* h0, double* h1, const double* u_em0, double* u_em1,
                              const double* ca, const double* cb, const double* da, const double* db)
{
    const size_t n = (size_t)nx * ny;
    const double* ei0 = e_in;
    const double* ei1 = e_in + n;
    double* eo0 = e_out;
    double* eo1 = e_out + n;
    memcpy(e_out, e_in, 2 * n * sizeof(double));
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            size_t c = IDX2(i, j), cy = IDX2(i, j - 1), cx = IDX2(i - 1, j);
            double e0 = ca[c] * ei0[c] + cb[c] * (h0[IDX2(i, j + 1)] - h0[c]);
            double e1 = ca[c] * ei1[c] - cb[c] * (h0[IDX2(i + 1, j)] - h0[c]);
            double dy = h0[c] - h0[cy];
            double dx = h0[c] - h0[cx];
            double ey_src = (j - 1 >= 1) ? ca[cy] * ei0[cy] + cb[cy] * dy : ei0[cy];
            double ex_src = (i - 1 >= 1) ? ca[cx] * ei1[cx] - cb[cx] * dx : ei1[cx];
            double ey = ca[cy] * ey_src + cb[cy] * dy;
            double ex = ca[cx] * ex_src - cb[cx] * dx;
            eo0[c] = e0;
            eo1[c] = e1;
            double h = da[c] * h0[c] + db[c] * (e0 - ey + ex - e1);
            h1[c] = h;
            u_em1[c] = u_em0[c] + 0.5 * (h * h / mu + epsilon * (e0 * e0 + e1 * e1));
        }
}

/* gameoflife/gameoflife.c:74-91, interior [1,n-1)^2. */
void oracle_gameoflife(int nx, int ny, const double* u0, double* u1)
{
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
#define U(di, dj) u0[IDX2(i + (di), j + (dj))]
            double C = 100000000000000000000.;
            double L =
                U(-1, -1) + U(0, -1) + U(1, -1) +
                U(-1, 0) + U(1, 0) +
                U(-1, 1) + U(0, 1) + U(1, 1);
            u1[IDX2(i, j)] = 1. / (1. + (U(0, 0) + L - 3.) * (L - 3.) * C);
#undef U
        }
}

/* matmul/matmul.F90:56-68: do j = 1,ns; do i = 1,nx; do k = 1,ny:
 * C(i,j) = C(i,j) + A(i,k)*B(k,j), k strictly sequential. */
void oracle_matmul(int nx, int ny, int ns, const double* A, const double* B, double* C)
{
    for (int j = 0; j < ns; j++)
        for (int i = 0; i < nx; i++) {
            double acc = C[(size_t)i + (size_t)nx * j];
            for (int k = 0; k < ny; k++)
                acc = acc + A[(size_t)i + (size_t)nx * k] * B[(size_t)k + (size_t)ny * j];
            C[(size_t)i + (size_t)nx * j] = acc;
        }
}

/* Conway's rule on 0/1 cells, interior [1,n-1)^2, border cells fixed -- the
 * integer counterpart of the FP64 rule above (which reproduces it on exact
 * 0/1 inputs under a 0.5 threshold). */
void oracle_life_u8(int nx, int ny, const uint8_t* c0, uint8_t* c1)
{
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            int n = 0;
            for (int dj = -1; dj <= 1; dj++)
                for (int di = -1; di <= 1; di++)
                    if (di || dj) n += c0[IDX2(i + di, j + dj)] != 0;
            int alive = c0[IDX2(i, j)] != 0;
            c1[IDX2(i, j)] = (uint8_t)((n == 3) || (alive && n == 2));
        }
}

/* laplacian/laplacian.c:112 */
double oracle_real_rand(void)
{
    return (((double)(rand() / (double)RAND_MAX) - 0.5) * 2);
}

void oracle_srand(unsigned seed)
{
    srand(seed);
}

void oracle_fill_interleaved(size_t n, int narrays, double** arrays)
{
    for (size_t i = 0; i < n; i++)
        for (int a = 0; a < narrays; a++)
            arrays[a][i] = oracle_real_rand();
}

void oracle_fill_strided(size_t n, int narrays, double** bases, const long* strides)
{
    for (size_t i = 0; i < n; i++)
        for (int a = 0; a < narrays; a++)
            bases[a][i * (size_t)strides[a]] = oracle_real_rand();
}

double oracle_sum(size_t n, const double* a)
{
    double s = 0.0;
    for (size_t i = 0; i < n; i++)
        s += a[i];
    return s;
}

/* ---- precision variants: the same stencils in FP32 arithmetic ------------
 * FP32 storage (_f32) or bfloat16 storage (_bf16: uint16_t bit patterns,
 * widened exactly to float at load and rounded to nearest even at store).
 * Operand order as in the FP64 functions above (laplacian/laplacian.c:93-97,
 * wave13pt/wave13pt.c:640-651, gaussblur/gaussblur.c:65,85-92). */

float oracle_bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

uint16_t oracle_f32_to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u); /* quiet NaN */
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);                 /* nearest even */
}

#define LD_F32(a, c) ((a)[c])
#define ST_F32(a, c, v) ((a)[c] = (v))
#define LD_BF16(a, c) oracle_bf16_to_f32((a)[c])
#define ST_BF16(a, c, v) ((a)[c] = oracle_f32_to_bf16(v))

#define DEF_LAPLACIAN(SUFFIX, ST, LD, STORE)                                              \
    void oracle_laplacian_##SUFFIX(int nx, int ny, int ns, float alpha, float beta,       \
                                   const ST* w0, ST* w1)                                  \
    {                                                                                     \
        const size_t sy = (size_t)nx, sz = (size_t)nx * ny;                               \
        for (int k = 1; k < ns - 1; k++)                                                  \
            for (int j = 1; j < ny - 1; j++)                                              \
                for (int i = 1; i < nx - 1; i++) {                                        \
                    size_t c = IDX3(i, j, k);                                             \
                    float r = alpha * LD(w0, c) + beta * (                                \
                        LD(w0, c + 1) + LD(w0, c - 1) +                                   \
                        LD(w0, c + sy) + LD(w0, c - sy) +                                 \
                        LD(w0, c + sz) + LD(w0, c - sz));                                 \
                    STORE(w1, c, r);                                                      \
                }                                                                         \
    }
DEF_LAPLACIAN(f32, float, LD_F32, ST_F32)
DEF_LAPLACIAN(bf16, uint16_t, LD_BF16, ST_BF16)

#define DEF_WAVE(SUFFIX, ST, LD, STORE)                                                   \
    void oracle_wave13pt_##SUFFIX(int nx, int ny, int ns, float m0, float m1, float m2,   \
                                  const ST* w0, const ST* w1, ST* w2)                     \
    {                                                                                     \
        const size_t sy = (size_t)nx, sz = (size_t)nx * ny;                               \
        for (int k = 2; k < ns - 2; k++)                                                  \
            for (int j = 2; j < ny - 2; j++)                                              \
                for (int i = 2; i < nx - 2; i++) {                                        \
                    size_t c = IDX3(i, j, k);                                             \
                    float r = m0 * LD(w1, c) - LD(w0, c) +                                \
                        m1 * (                                                            \
                            LD(w1, c + 1) + LD(w1, c - 1) +                               \
                            LD(w1, c + sy) + LD(w1, c - sy) +                             \
                            LD(w1, c + sz) + LD(w1, c - sz)) +                            \
                        m2 * (                                                            \
                            LD(w1, c + 2) + LD(w1, c - 2) +                               \
                            LD(w1, c + 2 * sy) + LD(w1, c - 2 * sy) +                     \
                            LD(w1, c + 2 * sz) + LD(w1, c - 2 * sz));                     \
                    STORE(w2, c, r);                                                      \
                }                                                                         \
    }
DEF_WAVE(f32, float, LD_F32, ST_F32)
DEF_WAVE(bf16, uint16_t, LD_BF16, ST_BF16)

/* gaussblur/gaussblur.c:85-92 on every xy-plane k in [0,ns): xy-interior
 * [2,n-2). */
#define DEF_GAUSS3(SUFFIX, RT)                                                            \
    void oracle_gaussblur3_##SUFFIX(int nx, int ny, int ns, RT s0, RT s1, RT s2, RT s4,   \
                                    RT s5, RT s8, const RT* w0, RT* w1)                   \
    {                                                                                     \
        const RT f = (RT)1 / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5);                     \
        for (int k = 0; k < ns; k++)                                                      \
            for (int j = 2; j < ny - 2; j++)                                              \
                for (int i = 2; i < nx - 2; i++) {                                        \
                    const RT* p = w0 + IDX3(0, 0, k);                                     \
                    w1[IDX3(i, j, k)] = f * (                                             \
                        s0 * p[IDX2(i, j)] +                                              \
                        s1 * (p[IDX2(i - 1, j)] + p[IDX2(i + 1, j)] +                     \
                              p[IDX2(i, j - 1)] + p[IDX2(i, j + 1)]) +                    \
                        s2 * (p[IDX2(i - 1, j - 1)] + p[IDX2(i + 1, j - 1)] +             \
                              p[IDX2(i - 1, j + 1)] + p[IDX2(i + 1, j + 1)]) +             \
                        s4 * (p[IDX2(i - 2, j)] + p[IDX2(i + 2, j)] +                     \
                              p[IDX2(i, j - 2)] + p[IDX2(i, j + 2)]) +                    \
                        s5 * (p[IDX2(i - 2, j - 1)] + p[IDX2(i - 1, j - 2)] +             \
                              p[IDX2(i + 1, j - 2)] + p[IDX2(i + 2, j - 1)] +             \
                              p[IDX2(i - 2, j + 1)] + p[IDX2(i - 1, j + 2)] +             \
                              p[IDX2(i + 1, j + 2)] + p[IDX2(i + 2, j + 1)]) +             \
                        s8 * (p[IDX2(i - 2, j - 2)] + p[IDX2(i + 2, j - 2)] +             \
                              p[IDX2(i - 2, j + 2)] + p[IDX2(i + 2, j + 2)]));            \
                }                                                                         \
    }
DEF_GAUSS3(f64, double)
DEF_GAUSS3(f32, float)

/* ---- the suite's non-stencil tests (real = double: makefile.in:12) ------ */

/* vecadd/vecadd.c:69-84: w2 = w0 + w1 over every point. */
void oracle_vecadd(size_t n, const double* w0, const double* w1, double* w2)
{
    for (size_t i = 0; i < n; i++) w2[i] = w0[i] + w1[i];
}

/* matvec/matvec.c:59-69: y[j] = 0; y[j] += A(i,j) * x[i], i ascending. */
void oracle_matvec(int nx, int ny, const double* A, const double* x, double* y)
{
    for (int j = 0; j < ny; j++) {
        y[j] = 0;
        for (int i = 0; i < nx; i++) y[j] += A[IDX2(i, j)] * x[i];
    }
}

/* sincos/sincos.F90:62-70 with -fdefault-real-8: xy = sin(x) + cos(y). */
void oracle_sincos(size_t n, const double* x, const double* y, double* xy)
{
    for (size_t i = 0; i < n; i++) xy[i] = sin(x[i]) + cos(y[i]);
}

/* Single-precision forms of the same three (the library's _f32 entry
 * points; the reference does not build these). */
void oracle_vecadd_f32(size_t n, const float* w0, const float* w1, float* w2)
{
    for (size_t i = 0; i < n; i++) w2[i] = w0[i] + w1[i];
}

void oracle_matvec_f32(int nx, int ny, const float* A, const float* x, float* y)
{
    for (int j = 0; j < ny; j++) {
        y[j] = 0;
        for (int i = 0; i < nx; i++) y[j] += A[IDX2(i, j)] * x[i];
    }
}

void oracle_sincos_f32(size_t n, const float* x, const float* y, float* xy)
{
    for (size_t i = 0; i < n; i++) xy[i] = sinf(x[i]) + cosf(y[i]);
}
