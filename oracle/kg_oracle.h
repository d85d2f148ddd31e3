/*
 * kg_oracle.h -- CPU restatement of the kernelgen-perf-tests stencil hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product path (libkgstencil.so, the <test>/hip
 * targets) never links, calls or falls back to it.
 *
 * Every function restates one reference kernel in IEEE source order: the
 * operands are combined in exactly the order and association the reference C
 * (or Fortran) expression is written in, and the library is compiled with
 * -O2 -fno-fast-math -ffp-contract=off so that no FMA contraction and no
 * reassociation happens.  Under those rules the result is bit-identical to the
 * reference compiled the same way (checked against oracle/_ref and the golden
 * fixtures under tests/golden/).
 *
 * Layout: x fastest, a[i + nx*j + nx*ny*k] (reference macro _A,
 * laplacian/laplacian.c:30).  Out-of-place kernels write interior points only;
 * boundary cells of the output keep whatever they held.
 */
#ifndef KG_ORACLE_H
#define KG_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* laplacian/laplacian.c:44-102 */
void oracle_laplacian(int nx, int ny, int ns, double alpha, double beta,
                      const double* w0, double* w1);

/* divergence/divergence.c:44-101 */
void oracle_divergence(int nx, int ny, int ns, double alpha, double beta, double gamma,
                       double* u, const double* ux, const double* uy, const double* uz);

/* gradient/gradient.c:45-102 */
void oracle_gradient(int nx, int ny, int ns, double alpha, double beta, double gamma,
                     const double* u, double* ux, double* uy, double* uz);

/* wave13pt/wave13pt.c:80-677, default (no variant macro) path :640-651 */
void oracle_wave13pt(int nx, int ny, int ns, double m0, double m1, double m2,
                     const double* w0, const double* w1, double* w2);

/* uxx1/uxx1.c:45-110 */
void oracle_uxx1(int nx, int ny, int ns, double c1, double c2,
                 const double* u0, double* u1, const double* d1,
                 const double* xx, const double* xy, const double* xz);

/* tricubic/tricubic.c:71-1171, default FP64 path :1011-1145 */
void oracle_tricubic(int nx, int ny, int ns, const double* u0, double* u1,
                     const double* a, const double* b, const double* c);

/* tricubic2/tricubic2.c:38-106 */
void oracle_tricubic2(int nx, int ny, int ns, const double* u0, double* u1,
                      const double* a, const double* b, const double* c);

/* lapgsrb/lapgsrb.c:44-122 */
void oracle_lapgsrb(int nx, int ny, int ns, double c0, double c1, double c2, double c3,
                    const double* w0, double* w1);

/* jacobi/jacobi.F90:15-71 (2D, column-major (nx,ny) == x fastest) */
void oracle_jacobi(int nx, int ny, double c0, double c1, double c2,
                   const double* w0, double* w1);

/* gaussblur/gaussblur.c:44-96 (2D) */
void oracle_gaussblur(int nx, int ny, double s0, double s1, double s2,
                      double s4, double s5, double s8, const double* w0, double* w1);

/* whispering/whispering.c:48-110: serial sweep, e updated in place (planar
 * e[i + nx*j + nx*ny*c], c = 0,1). */
void oracle_whispering(int nx, int ny, double mu, double epsilon, double* e,
                       const double* h0, double* h1, const double* u_em0, double* u_em1,
                       const double* ca, const double* cb, const double* da, const double* db);

/* The same step written without in-place dependencies: reads e_in only and
 * writes e_out, reproducing the serial sweep's values in closed form.  Used to
 * pin the closed form the GPU kernel implements against the serial loop. */
void oracle_whispering_closed(int nx, int ny, double mu, double epsilon,
                              const double* e_in, double* e_out,
                              const double* h0, double* h1, const double* u_em0, double* u_em1,
                              const double* ca, const double* cb, const double* da, const double* db);

/* gameoflife/gameoflife.c:43-95 (2D, FP64 polynomial rule) */
void oracle_gameoflife(int nx, int ny, const double* u0, double* u1);

/* matmul/matmul.F90:14-70: C(nx,ns) += A(nx,ny) * B(ny,ns), column-major */
void oracle_matmul(int nx, int ny, int ns, const double* A, const double* B, double* C);

/* Integer Conway Life on a uint8 board (the uint8 8192x8192 extension of
 * gameoflife): cells 0/1, fixed dead border outside the board. */
void oracle_life_u8(int nx, int ny, const uint8_t* c0, uint8_t* c1);

/* Precision variants (FP32 arithmetic; _bf16 = bfloat16 storage as uint16_t
 * bit patterns, rounded to nearest even) and gaussblur per xy-plane of a 3D
 * grid (all planes, xy-interior [2,n-2)). */
float oracle_bf16_to_f32(uint16_t h);
uint16_t oracle_f32_to_bf16(float f);
void oracle_laplacian_f32(int nx, int ny, int ns, float alpha, float beta,
                          const float* w0, float* w1);
void oracle_laplacian_bf16(int nx, int ny, int ns, float alpha, float beta,
                           const uint16_t* w0, uint16_t* w1);
void oracle_wave13pt_f32(int nx, int ny, int ns, float m0, float m1, float m2,
                         const float* w0, const float* w1, float* w2);
void oracle_wave13pt_bf16(int nx, int ny, int ns, float m0, float m1, float m2,
                          const uint16_t* w0, const uint16_t* w1, uint16_t* w2);
void oracle_gaussblur3_f64(int nx, int ny, int ns, double s0, double s1, double s2,
                           double s4, double s5, double s8, const double* w0, double* w1);
void oracle_gaussblur3_f32(int nx, int ny, int ns, float s0, float s1, float s2,
                           float s4, float s5, float s8, const float* w0, float* w1);

/* The suite's non-stencil tests in FP64 (the reference's real = double):
 * vecadd (w2 = w0 + w1), matvec (y = A x, sequential row sums), sincos
 * (xy = sin(x) + cos(y)); and their single-precision forms. */
void oracle_vecadd(size_t n, const double* w0, const double* w1, double* w2);
void oracle_matvec(int nx, int ny, const double* A, const double* x, double* y);
void oracle_sincos(size_t n, const double* x, const double* y, double* xy);
void oracle_vecadd_f32(size_t n, const float* w0, const float* w1, float* w2);
void oracle_matvec_f32(int nx, int ny, const float* A, const float* x, float* y);
void oracle_sincos_f32(size_t n, const float* x, const float* y, float* xy);

/* glibc rand() based fill in the reference's order: values are drawn
 * interleaved per index, arrays[0][i], arrays[1][i], ..., (e.g.
 * laplacian/laplacian.c:158-164).  real_rand() = ((rand()/RAND_MAX)-0.5)*2
 * (laplacian/laplacian.c:112).  Does not reseed: the caller controls srand. */
double oracle_real_rand(void);
void oracle_srand(unsigned seed);
void oracle_fill_interleaved(size_t n, int narrays, double** arrays);
/* Same, writing bases[a][i*strides[a]] (whispering fills e[2i], e[2i+1],
 * whispering/whispering.c:176-177). */
void oracle_fill_strided(size_t n, int narrays, double** bases, const long* strides);

/* Plain left-to-right FP64 sum, as the reference's mean loops are written. */
double oracle_sum(size_t n, const double* a);

#ifdef __cplusplus
}
#endif

#endif
