/*
 * kg_stencil.h -- C ABI of libkgstencil.so, the MI355X (gfx950) stencil
 * library behind the `hip` target of kernelgen-perf-tests.
 *
 * Plain C: no HIP or torch types cross this boundary.  The caller owns every
 * buffer (host and device); kernels never allocate.  Every function returns
 * KG_OK (0) or a negative kg_status; kg_strerror() describes the code and
 * kg_last_error() gives the most recent detailed message.
 *
 * Launch functions are asynchronous on the context's stream (one HIP stream
 * per context, default priority) and perform no host synchronisation; call
 * kg_sync() to wait, as the reference does after every launch
 * (laplacian/laplacian.c:296-297).  One host thread per context.
 *
 * Grid layout for every stencil: x fastest, a[i + nx*j + nx*ny*k]
 * (laplacian/laplacian.c:30).  Out-of-place kernels write the interior points
 * the reference's loops visit and leave every other cell of the output as it
 * was.  FP64 kernels evaluate each expression in the reference's source order
 * without FMA contraction, so they are bit-identical to an IEEE build of the
 * reference.
 */
#ifndef KG_STENCIL_H
#define KG_STENCIL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kg_ctx kg_ctx;

typedef enum kg_status {
    KG_OK = 0,
    KG_ERR_INVALID = -1,   /* bad argument: null pointer, negative or oversized dims */
    KG_ERR_HIP = -2,       /* HIP runtime error; see kg_last_error() */
    KG_ERR_NOMEM = -3,     /* device allocation failed */
    KG_ERR_NODEV = -4,     /* no HIP device / bad device index */
    KG_ERR_STATE = -5      /* call not valid in the current state (e.g. timing off) */
} kg_status;

/* ---- context, memory, transfers, timing --------------------------------
 * These replace the CUDA runtime calls of the reference's `cuda` target:
 * cudaGetDeviceCount (laplacian/laplacian.c:192-196), cudaMalloc (:223-228),
 * cudaMemcpy H2D (:255-259), cudaMemcpy D2H (:334-337), cudaFree (:362-366),
 * cudaDeviceSynchronize (:297), and the per-launch timing and register count
 * printed by __wrap_cudaLaunchKernel (laplacian/cuda/cuda_profiling.cu:211-251). */
int kg_device_count(int* count);
int kg_init(kg_ctx** ctx, int device);
int kg_destroy(kg_ctx* ctx);
int kg_alloc(kg_ctx* ctx, void** dptr, size_t bytes);
int kg_free(kg_ctx* ctx, void* dptr);
int kg_h2d(kg_ctx* ctx, void* dst, const void* src, size_t bytes);
int kg_d2h(kg_ctx* ctx, void* dst, const void* src, size_t bytes);
int kg_d2d(kg_ctx* ctx, void* dst, const void* src, size_t bytes);
int kg_sync(kg_ctx* ctx);
/* Per-launch device timing with HIP events (on by default).  Each launch is
 * bracketed by an event pair from a ring of KG_TIMING_RING pairs. */
#define KG_TIMING_RING 8192
int kg_timing_enable(kg_ctx* ctx, int on);
/* Duration in seconds of the most recent launch (waits for it). */
int kg_last_kernel_time(kg_ctx* ctx, double* seconds);
/* Waits, then copies the durations (seconds) of the launches made since the
 * previous drain, oldest first, up to cap; *count receives how many. */
int kg_timing_drain(kg_ctx* ctx, double* seconds, int cap, int* count);
/* Name of the most recent launch's kernel and the registers it uses. */
const char* kg_last_kernel_name(kg_ctx* ctx);
/* Architected VGPRs per lane of the named entry point's main kernel (the
 * reference prints numRegs the same way: "<kernel> regcount = N"). */
int kg_kernel_regcount(kg_ctx* ctx, const char* kernel, int* vgprs);
const char* kg_strerror(int status);
const char* kg_last_error(void);

/* Calibration kernel, no counterpart in the reference: dst[0,n) = src[0,n)
 * with 16-byte vector accesses (buffers 16-byte aligned).  Its HBM traffic is
 * exactly 16*n bytes, the yardstick for achievable bandwidth and for scaling
 * rocprofv3's FETCH_SIZE / WRITE_SIZE counters. */
int kg_stream_copy(kg_ctx* ctx, double* dst, const double* src, size_t n);

/* ---- FP64 drop-in stencils (one entry per reference kernel) ------------- */

/* laplacian/laplacian.c:44-102.  w1 = alpha*w0 + beta*(6 face neighbours) on
 * [1,n-1)^3. */
int kg_laplacian(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                 const double* w0, double* w1);

/* divergence/divergence.c:44-101.  u = alpha*dx(ux) + beta*dy(uy) +
 * gamma*dz(uz), central differences, [1,n-1)^3. */
int kg_divergence(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta, double gamma,
                  double* u, const double* ux, const double* uy, const double* uz);

/* gradient/gradient.c:45-102.  ux = alpha*dx(u), uy = beta*dy(u),
 * uz = gamma*dz(u), [1,n-1)^3. */
int kg_gradient(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta, double gamma,
                const double* u, double* ux, double* uy, double* uz);

/* wave13pt/wave13pt.c:80-677 (default path :640-651).  w2 = m0*w1 - w0 +
 * m1*(6 nbrs, r=1) + m2*(6 nbrs, r=2) on [2,n-2)^3. */
int kg_wave13pt(kg_ctx* ctx, int nx, int ny, int ns, double m0, double m1, double m2,
                const double* w0, const double* w1, double* w2);

/* uxx1/uxx1.c:45-110.  Staggered velocity update on [2,n-1)^3, dth = 1/nx. */
int kg_uxx1(kg_ctx* ctx, int nx, int ny, int ns, double c1, double c2,
            const double* u0, double* u1, const double* d1,
            const double* xx, const double* xy, const double* xz);

/* tricubic/tricubic.c:71-1171 (FP64 path :1011-1145).  64-point Lagrange
 * cubic interpolation; k in [1,ns-2), j in [1,ny-2), i in [1,nx-2). */
int kg_tricubic(kg_ctx* ctx, int nx, int ny, int ns, const double* u0, double* u1,
                const double* a, const double* b, const double* c);

/* tricubic2/tricubic2.c:38-106.  Same interpolation, loop form summed from 0,
 * on [2,n-2)^3. */
int kg_tricubic2(kg_ctx* ctx, int nx, int ny, int ns, const double* u0, double* u1,
                 const double* a, const double* b, const double* c);

/* lapgsrb/lapgsrb.c:44-122.  25-point update on [2,n-2)^3. */
int kg_lapgsrb(kg_ctx* ctx, int nx, int ny, int ns, double c0, double c1, double c2, double c3,
               const double* w0, double* w1);

/* jacobi/jacobi.F90:15-71, called from C as jacobi_(&nx,&ny,&c0,&c1,&c2,w0,w1)
 * (jacobi/main.c:245).  2D 9-point sweep on [1,n-1)^2. */
int kg_jacobi(kg_ctx* ctx, int nx, int ny, double c0, double c1, double c2,
              const double* w0, double* w1);

/* gaussblur/gaussblur.c:44-96.  2D 25-point weighted blur on [2,n-2)^2,
 * normalised by f = 1/(s0 + 4(s1+s2+s4+s8) + 8 s5). */
int kg_gaussblur(kg_ctx* ctx, int nx, int ny, double s0, double s1, double s2,
                 double s4, double s5, double s8, const double* w0, double* w1);

/* whispering/whispering.c:48-110.  The reference updates e in place during a
 * serial sweep; this entry point reproduces that sweep's values exactly while
 * reading e_in and writing e_out (both planar 2*nx*ny arrays), so the caller
 * ping-pongs e as it does h and u_em.  The border cells of e are never
 * updated: keep them equal in e_in and e_out (copy e into both at load). */
int kg_whispering(kg_ctx* ctx, int nx, int ny, double mu, double epsilon,
                  const double* e_in, double* e_out,
                  const double* h0, double* h1, const double* u_em0, double* u_em1,
                  const double* ca, const double* cb, const double* da, const double* db);

/* gameoflife/gameoflife.c:43-95.  FP64 rule u1 = 1/(1+(u+L-3)(L-3)*1e20) on
 * [1,n-1)^2. */
int kg_gameoflife(kg_ctx* ctx, int nx, int ny, const double* u0, double* u1);

/* matmul/matmul.F90:14-70, called as matmul_(&nx,&ny,&ns,A,B,C)
 * (matmul/main.c:235).  C(nx,ns) += A(nx,ny)*B(ny,ns), column-major, on the
 * FP64 matrix cores.  Products are accumulated with fused multiply-adds in
 * blocks of k, so results differ from the sequential-k reference by rounding
 * (bounded norm-wise, not bit-exact). */
int kg_matmul(kg_ctx* ctx, int nx, int ny, int ns, const double* A, const double* B, double* C);

#ifdef __cplusplus
}
#endif

#endif
