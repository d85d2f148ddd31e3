/*
 * kg_host.h -- host-side helpers shared by the per-test programs of the
 * suite (<test>/<test>.c), for both the `hip` target (linked with
 * libkgstencil.so) and the `gcc` target (linked with the CPU backend).
 *
 * They reproduce the host conventions of the reference's test programs:
 * argument parsing and messages (laplacian/laplacian.c:104-139), glibc rand()
 * inputs (:112, :141-164), CLOCK_REALTIME phase timing and the stdout lines
 * the benchmark driver parses (:165-377), and CUDA_SAFE_CALL-style error exits
 * (laplacian/cuda/cuda_profiling.h:11-15).
 */
#ifndef KG_HOST_H
#define KG_HOST_H

#include <malloc.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "../../include/kg_stencil.h"

#define KG_MEMALIGN 4096

/* 1 in the `hip` target, 0 in the `gcc` target (defined by the target's
 * backend object).  With 0 the device phases (init, alloc, load, save, free)
 * are neither run nor printed, as in the reference's gcc target. */
extern const int kg_target_has_device;

/* Print the library error and exit(-1), as CUDA_SAFE_CALL does. */
#define KG_CALL(call)                                                                   \
    do {                                                                                \
        int kg_rc_ = (call);                                                            \
        if (kg_rc_ != KG_OK) {                                                          \
            fprintf(stderr, "%s:%d: %s failed: %s: %s\n", __FILE__, __LINE__, #call,   \
                    kg_strerror(kg_rc_), kg_last_error());                              \
            exit(-1);                                                                   \
        }                                                                               \
    } while (0)

/* atoi + the reference's negative-value check. */
int kg_parse_arg(const char* name, const char* arg);
#define KG_PARSE_ARG(name, arg) int name = kg_parse_arg(#name, arg)

/* ((rand() / RAND_MAX) - 0.5) * 2, glibc rand() with its default seed. */
double kg_real_rand(void);
/* The same with real = float: (((real)(rand() / (double)RAND_MAX) - 0.5) * 2)
 * -- the quotient is rounded to float before the shift (matvec/matvec.c:81). */
float kg_real_rand_f(void);

/* Wall-clock timer (CLOCK_REALTIME, seconds). */
void kg_get_time(struct timespec* t);
double kg_time_diff(const struct timespec* a, const struct timespec* b);

/* Allocates, or prints the reference's message and exits. */
double* kg_host_array(size_t n);

/* Phase state shared by a test program. */
typedef struct kg_run {
    kg_ctx* ctx;
    int no_timing;          /* NO_TIMING set: suppress timing lines */
    const char* profile;    /* PROFILING_FNAME: print kernel time / regcount */
    struct timespec t0;
} kg_run;

/* Reads NO_TIMING and PROFILING_FNAME. */
void kg_run_begin(kg_run* r);
/* init phase: device count + context ("init time = %f sec"). */
void kg_run_init(kg_run* r);
/* Start / stop a phase timer. */
void kg_phase_start(kg_run* r);
double kg_phase_stop(kg_run* r);
/* Device allocation of n doubles (alloc phase). */
double* kg_dev_array(kg_run* r, size_t n);
void kg_dev_free(kg_run* r, double* p);
void kg_to_dev(kg_run* r, double* d, const double* h, size_t n);
void kg_to_host(kg_run* r, double* h, const double* d, size_t n);
/* After each launch: wait for it (the reference synchronises after every
 * launch, laplacian/laplacian.c:296-297) and, when PROFILING_FNAME is set,
 * print "<kernel> regcount = N" and "<kernel> kernel time = T" as the
 * reference's launch wrapper does (laplacian/cuda/cuda_profiling.cu:239-249);
 * T is the kernel's device time in seconds. */
void kg_after_launch(kg_run* r);
void kg_run_end(kg_run* r);

/* The reference's device phases over n arrays (host[i] has len[i] doubles).
 * In the `gcc` target dev[i] simply aliases host[i] and nothing is printed.
 *   alloc: "device buffer alloc time = %f sec"
 *   load:  "data load time = %f sec (%f GB/sec)"  -- GB/s of shown_bytes,
 *          the byte count each reference program prints (some count more or
 *          fewer arrays than they copy, e.g. laplacian/laplacian.c:262)
 *   save:  "data save time = %f sec (%f GB/sec)"  -- copies dev[i] to host[i]
 *          for the listed arrays
 *   free:  "device buffer free time = %f sec" */
void kg_alloc_phase(kg_run* r, int n, double** dev, double* const* host, const size_t* len);
void kg_load_phase(kg_run* r, int n, double* const* dev, double* const* host, const size_t* len,
                   double shown_bytes);
void kg_save_phase(kg_run* r, int n, double* const* host, double* const* dev, const size_t* len,
                   double shown_bytes);
void kg_free_phase(kg_run* r, int n, double* const* dev);
/* Plain left-to-right sum, as the reference's mean loops are written. */
double kg_sum(const double* a, size_t n);

/* The same phases for the single-precision tests (real = float in the
 * reference: vecadd, matvec, sincos); len[i] counts floats. */
float* kg_host_array_f(size_t n);
void kg_alloc_phase_f(kg_run* r, int n, float** dev, float* const* host, const size_t* len);
void kg_load_phase_f(kg_run* r, int n, float* const* dev, float* const* host, const size_t* len,
                     double shown_bytes);
void kg_save_phase_f(kg_run* r, int n, float* const* host, float* const* dev, const size_t* len,
                     double shown_bytes);
void kg_free_phase_f(kg_run* r, int n, float* const* dev);

#endif
