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

/* gameoflife/gameoflife.c:82-91 on uint8 cells (X mode): c[i + nx*j], nonzero
 * = alive; c1 = Conway's rule (alive next iff 3 neighbours, or 2 and alive)
 * on [1,n-1)^2, written as 0/1; the border of c1 is left as it was.  For 0/1
 * cells this is the FP64 kernel's rule exactly.  Buffers 16-byte aligned. */
int kg_life_u8(kg_ctx* ctx, int nx, int ny, const uint8_t* c0, uint8_t* c1);

/* ---- the suite's non-stencil tests (real = double in the reference:
 * makefile.in:12 overrides the `real ?= float` of each gcc makefile) ---- */
/* vecadd/vecadd.c:38-84: w2 = w0 + w1 over all nx*ny*ns points (the host
 * rotates w0 <- w1 <- w2 <- w0 between steps). */
int kg_vecadd(kg_ctx* ctx, int nx, int ny, int ns, const double* w0, const double* w1, double* w2);
/* matvec/matvec.c:38-70: y[j] = sum_i A[i + nx*j] * x[i], i ascending, for
 * j < ny (the 2D test runs at nx x ny*ns); sequential sums, bit-identical to
 * an IEEE build of the reference. */
int kg_matvec(kg_ctx* ctx, int nx, int ny, const double* A, const double* x, double* y);
/* sincos/sincos.F90:14-74 (-fdefault-real-8): xy = sin(x) + cos(y) over all
 * points, with the device's FP64 sin and cos. */
int kg_sincos(kg_ctx* ctx, int nx, int ny, int ns, const double* x, const double* y, double* xy);
/* Single-precision forms of the three, beside them (no reference
 * counterpart: the reference's `real` is double for these tests). */
int kg_vecadd_f32(kg_ctx* ctx, int nx, int ny, int ns, const float* w0, const float* w1, float* w2);
int kg_matvec_f32(kg_ctx* ctx, int nx, int ny, const float* A, const float* x, float* y);
int kg_sincos_f32(kg_ctx* ctx, int nx, int ny, int ns, const float* x, const float* y, float* xy);

/* ---- precision variants (beside the FP64 drop-in kernels) -------------------
 * No reference counterpart: the reference computes in FP64 only.  Same
 * stencils, interiors and source order as the FP64 entry points above, with
 * FP32 storage and FP32 arithmetic (_f32) or bfloat16 storage -- uint16_t bit
 * patterns -- and FP32 arithmetic (_bf16, results rounded to nearest even).
 * Means, norms and errors of such fields are taken in FP64 with
 * kg_field_stats. */
typedef enum { KG_F64 = 0, KG_F32 = 1, KG_BF16 = 2 } kg_dtype;

/* laplacian/laplacian.c:93-97 */
int kg_laplacian_f32(kg_ctx* ctx, int nx, int ny, int ns, float alpha, float beta,
                     const float* w0, float* w1);
int kg_laplacian_bf16(kg_ctx* ctx, int nx, int ny, int ns, float alpha, float beta,
                      const uint16_t* w0, uint16_t* w1);
/* wave13pt/wave13pt.c:640-651 */
int kg_wave13pt_f32(kg_ctx* ctx, int nx, int ny, int ns, float m0, float m1, float m2,
                    const float* w0, const float* w1, float* w2);
int kg_wave13pt_bf16(kg_ctx* ctx, int nx, int ny, int ns, float m0, float m1, float m2,
                     const uint16_t* w0, const uint16_t* w1, uint16_t* w2);
/* gaussblur/gaussblur.c:65,85-92 applied to every xy-plane of a 3D grid:
 * xy-interior [2,n-2), all z-planes. */
int kg_gaussblur3_f64(kg_ctx* ctx, int nx, int ny, int ns, double s0, double s1, double s2,
                      double s4, double s5, double s8, const double* w0, double* w1);
int kg_gaussblur3_f32(kg_ctx* ctx, int nx, int ny, int ns, float s0, float s1, float s2,
                      float s4, float s5, float s8, const float* w0, float* w1);

/* n values from src (src_dtype) to dst (dst_dtype), rounded to nearest even;
 * asynchronous on the context stream. */
int kg_convert(kg_ctx* ctx, int dst_dtype, void* dst, int src_dtype, const void* src, size_t n);

/* Fills n values (KG_F64 or KG_F32) at p with uniform [-1, 1) values from a
 * counter-based hash of (seed, index); asynchronous on the context stream. */
int kg_fill_uniform(kg_ctx* ctx, int dtype, size_t n, void* p, unsigned long long seed);

/* ---- 2D Poisson solve (mixed precision) ------------------------------------------
 * A u = f on the unit square, u = 0 on the boundary, (n+1) x (n+1) vertices
 * (n cells per side, n a power of two up to 16384, h = 1/n), x fastest.  A is
 * the 9-point operator whose damped Jacobi sweep has the form of the
 * reference's jacobi kernel (jacobi/jacobi.F90:65-67) plus a right-hand side:
 * (A u) = (20 u - 4 sum_axis - sum_diag) / (6 h^2).  Defect correction: the
 * FP64 residual r = f - A u, then `vcycles` FP32 multigrid V-cycles on A e = r,
 * u += e, until ||r||_2 / ||f||_2 < tol or max_outer corrections.  u (device,
 * FP64) holds the initial guess and receives the solution; f is not changed.
 * Returns KG_ERR_STATE if tol was not reached. */
#define KG_POISSON_HISTORY 32
typedef struct kg_poisson_report {
    int outer;                 /* corrections applied */
    int vcycles;               /* FP32 V-cycles in total */
    int levels;                /* multigrid levels */
    double rel_residual;       /* final ||f - A u||_2 / ||f||_2 (FP64) */
    double seconds;            /* device time of the solve */
    double history[KG_POISSON_HISTORY];   /* relative residual before each correction */
} kg_poisson_report;

int kg_poisson_solve(kg_ctx* ctx, int n, const double* f, double* u, double tol, int max_outer,
                     int vcycles, kg_poisson_report* report);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * solve as kg_poisson_solve -- the same kernels in the same order, so u and
 * the report's residuals are bit-identical -- with its multigrid levels and
 * partial sums in a workspace reserved once per device inside the library
 * (a static device array of about 17 MB, held until the process ends,
 * located by kg_init).  It calls no hipMalloc, hipFree or other allocating
 * routine.  n is limited to KG_POISSON_NOALLOC_MAX_N (larger n: KG_ERR_INVALID).
 * report->seconds is host wall time from the first residual pass to the end.
 * Every call on a device shares that device's workspace: calls that overlap
 * on one device (two host threads, or two contexts on the same device) race
 * and must be serialised by the caller. */
#define KG_POISSON_NOALLOC_MAX_N 1024
int kg_poisson_solve_noalloc(kg_ctx* ctx, int n, const double* f, double* u, double tol,
                             int max_outer, int vcycles, kg_poisson_report* report);

/* FP64 statistics of a device field x of n values of type dtype, optionally
 * against a reference field ref of type ref_dtype (NULL: none).  out (host,
 * 5 doubles) = { sum x, sum x^2, max |x|, sum (x-ref)^2, max |x-ref| }.
 * Deterministic (fixed reduction order).  Synchronises the context stream. */
int kg_field_stats(kg_ctx* ctx, int dtype, size_t n, const void* x, int ref_dtype,
                   const void* ref, double* out);

/* matmul/matmul.F90:14-70, called as matmul_(&nx,&ny,&ns,A,B,C)
 * (matmul/main.c:235).  C(nx,ns) += A(nx,ny)*B(ny,ns), column-major, on the
 * FP64 matrix cores.  Products are accumulated with fused multiply-adds in
 * blocks of k, so results differ from the sequential-k reference by rounding
 * (bounded norm-wise, not bit-exact). */
int kg_matmul(kg_ctx* ctx, int nx, int ny, int ns, const double* A, const double* B, double* C);

/* ---- batched grids -----------------------------------------------------------
 * No reference counterpart (the reference runs one grid per process).  A batch
 * keeps many independent 3D grids resident in HBM and advances all of them by
 * one step per persistent launch; kg_batch_step(b, n, 1) captures n launches
 * in a HIP graph and replays it.
 *
 * Levels: laplacian and gaussblur3 grids have 2 (ping-pong), wave13pt 3
 * (w0, w1, w2).  Before the first step, level 0 is the input (wave13pt: w0,
 * and level 1 is w1).  A grid's levels share one boundary: the cells outside
 * the kind's interior box hold the same values in every level (a Dirichlet
 * boundary).  kg_batch_fill writes them so; uploads must do the same.  A
 * step writes whole rows of the output level -- the interior points and,
 * copied from the input level, the x-boundary points of those rows -- so that
 * no cache line of a row is written in part (partly written row-edge lines
 * made narrow grids up to 2.3x slower); every other boundary cell is never
 * written.  With a shared boundary, each grid's result is bit-identical to
 * stepping it alone with the single-grid entry point of its kind and type
 * (kg_laplacian / _f32, kg_wave13pt / _f32, kg_gaussblur3_f64 / _f32),
 * whose boundary cells are never written.  kg_batch_newest_level tells which
 * level holds the latest values.  Host buffers are dense (x fastest); device
 * rows are padded to 16 bytes. */
/* Kinds: the 3D stencils, and the 2D ones applied to every xy-plane (a 2D
 * nx x ny problem is a grid with ns = 1): jacobi (jacobi/jacobi.F90:65-67,
 * interior [1,n-1)) and gameoflife's FP rule (gameoflife/gameoflife.c:82-91). */
typedef enum {
    KG_KIND_LAPLACIAN = 0, KG_KIND_WAVE13PT = 1, KG_KIND_GAUSSBLUR3 = 2,
    KG_KIND_JACOBI2 = 3, KG_KIND_LIFE2 = 4
} kg_kind;

typedef struct kg_grid_spec {
    int kind;        /* kg_kind */
    int dtype;       /* KG_F64 or KG_F32 */
    int nx, ny, ns;
    double coef[6];  /* laplacian: alpha, beta; wave13pt: m0, m1, m2;
                        gaussblur3: s0, s1, s2, s4, s5, s8; jacobi2: c0, c1, c2;
                        life2: none */
} kg_grid_spec;

typedef struct kg_batch kg_batch;

int kg_batch_create(kg_ctx* ctx, kg_batch** out);
int kg_batch_destroy(kg_batch* b);
/* Adds a grid (its levels are allocated, contents undefined until uploaded);
 * it joins at the batch's current step. */
int kg_batch_add(kg_batch* b, const kg_grid_spec* spec, int* id);
/* Releases a grid's levels for reuse by later grids (waits for queued work). */
int kg_batch_remove(kg_batch* b, int id);
int kg_batch_upload(kg_batch* b, int id, int level, const void* host);
int kg_batch_download(kg_batch* b, int id, int level, void* host);
int kg_batch_newest_level(kg_batch* b, int id, int* level);
int kg_batch_steps_done(kg_batch* b, int id, long long* steps);
/* n steps of every grid; use_graph != 0 replays a captured graph of n launches. */
int kg_batch_step(kg_batch* b, int nsteps, int use_graph);
/* Grid points and algorithmic HBM bytes per step, work items, workgroups. */
int kg_batch_info(kg_batch* b, double* points_per_step, double* bytes_per_step,
                  int* work_items, int* workgroups);
/* Autotuner: for every size class present (kind, type, nx and ny bucket),
 * times each tile shape -- the plane engine with 16/32/64 lanes per row, with
 * whole rows packed back to back across the workgroup (lanes_per_row -1, for
 * rows narrower than a wave), the direct row body of the planar kinds
 * (gaussblur3, jacobi2, life2; lanes_per_row -2: rows loaded through the
 * caches, one plane per tile), the column march of the planar kinds
 * (lanes_per_row -3: each wave marches one row piece of one plane along y in
 * registers, planes_per_tile then counting the rows per wave), or
 * the flat body (lanes_per_row 0: points swept from cache, for small grids)
 * -- with 8/16/32/64 planes per tile, over reps launches on the batch's own
 * grids, and keeps the fastest.  Untuned classes take the direct row body
 * (jacobi2, life2), packed rows (rows up to 64 lanes) or 64 lanes per row.  Grid
 * contents are preserved. */
int kg_batch_tune(kg_batch* b, int reps, int* classes_tuned);
int kg_batch_grid_shape(kg_batch* b, int id, int* lanes_per_row, int* planes_per_tile);
/* Forces the tile shape of the size class of grid id. */
int kg_batch_set_shape(kg_batch* b, int id, int lanes_per_row, int planes_per_tile);
/* Fills every level of every grid on the device with uniform [-1, 1) values
 * from a counter-based hash: interior cells of (seed, grid, level, element),
 * boundary cells of (seed, grid, element), shared by the levels. */
int kg_batch_fill(kg_batch* b, unsigned long long seed);
/* The comparison path: n steps as one plain launch per grid and step (same
 * tile bodies and results as kg_batch_step). */
int kg_batch_step_sequential(kg_batch* b, int nsteps);
/* HBM reserved by the batch's arena and bytes held by live grids. */
int kg_batch_memory(kg_batch* b, size_t* reserved, size_t* in_use);
/* Fills the levels of one grid like kg_batch_fill (asynchronous); both levels
 * of a jacobi2 grid get the same values throughout. */
int kg_batch_fill_grid(kg_batch* b, int id, unsigned long long seed);

/* ---- serving -------------------------------------------------------------------
 * No reference counterpart.  A request loop over a batch: jobs arrive at
 * given times, are admitted in arrival order while at most max_active run and
 * their levels fit in hbm_budget bytes, stay resident until done, and advance
 * one step per tick together (continuous batching: one persistent launch per
 * tick over every active job).  KG_SERVE_SEQUENTIAL runs one job at a time
 * with one plain launch per step, for comparison.  Initial data come from the
 * job's seed (kg_batch_fill_grid).  A jacobi2 job with tol > 0 stops early
 * once max |w_new - w_old| < tol (checked every 10 steps). */
enum { KG_SERVE_CONTINUOUS = 0, KG_SERVE_SEQUENTIAL = 1 };

typedef struct kg_job {
    int kind, dtype, nx, ny, ns;
    int steps;
    double tol;
    double coef[6];
    double arrival;            /* seconds after the start of the run */
    unsigned long long seed;
    void* out;                 /* optional host buffer (dense nx*ny*ns) for the final field */
} kg_job;

typedef struct kg_job_result {
    double admitted, finished; /* seconds after the start of the run */
    long long steps_run;
    double checksum[2];        /* FP64 sum and sum of squares of the final field's level */
    int status;                /* KG_OK, or KG_ERR_NOMEM if it exceeds the budget */
} kg_job_result;

typedef struct kg_serve_stats {
    double wall;               /* seconds */
    double points;             /* grid points x steps computed */
    double p50, p99, mean_latency;   /* finished - arrival, seconds */
    int peak_active;
    size_t peak_bytes;
} kg_serve_stats;

int kg_serve_run(kg_ctx* ctx, const kg_job* jobs, int njobs, size_t hbm_budget, int max_active,
                 int mode, kg_job_result* results, kg_serve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif
