/*
 * kg_cpu_whispering.c -- whispering for the `gcc` target, built -g -O0 as the
 * reference's gcc target is (whispering/gcc/makefile:24).  With e_in == e_out
 * it runs the reference's serial sweep, e updated in place
 * (whispering/whispering.c:87-106); with distinct buffers it runs the same
 * sweep on a copy (e_out = e_in, then in place on e_out).
 */
#include <string.h>
#include <time.h>

#include "../../include/kg_stencil.h"

struct kg_ctx {
    double last_time;
    const char* last_kernel;
};

#define IX2(i, j) ((size_t)(i) + (size_t)nx * (size_t)(j))

int kg_whispering(kg_ctx* ctx, int nx, int ny, double mu, double epsilon,
                  const double* e_in, double* e, const double* h0, double* h1,
                  const double* u_em0, double* u_em1, const double* ca, const double* cb,
                  const double* da, const double* db)
{
    struct timespec t0, t1;
    clock_gettime(CLOCK_REALTIME, &t0);
    const size_t n = (size_t)nx * ny;
    if (e != e_in) memcpy(e, e_in, 2 * n * sizeof(double));
    double* e0p = e;
    double* e1p = e + n;
#pragma omp parallel for
    for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
            double e0 = ca[IX2(i, j)] * e0p[IX2(i, j)] + cb[IX2(i, j)] * (h0[IX2(i, j + 1)] - h0[IX2(i, j)]);
            double e1 = ca[IX2(i, j)] * e1p[IX2(i, j)] - cb[IX2(i, j)] * (h0[IX2(i + 1, j)] - h0[IX2(i, j)]);
            double ey = ca[IX2(i, j - 1)] * e0p[IX2(i, j - 1)] + cb[IX2(i, j - 1)] * (h0[IX2(i, j)] - h0[IX2(i, j - 1)]);
            double ex = ca[IX2(i - 1, j)] * e1p[IX2(i - 1, j)] - cb[IX2(i - 1, j)] * (h0[IX2(i, j)] - h0[IX2(i - 1, j)]);
            e0p[IX2(i, j)] = e0;
            e1p[IX2(i, j)] = e1;
            double h = da[IX2(i, j)] * h0[IX2(i, j)] + db[IX2(i, j)] * (e0 - ey + ex - e1);
            h1[IX2(i, j)] = h;
            u_em1[IX2(i, j)] = u_em0[IX2(i, j)] + 0.5 * (h * h / mu + epsilon * (e0 * e0 + e1 * e1));
        }
    clock_gettime(CLOCK_REALTIME, &t1);
    ctx->last_time = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    ctx->last_kernel = "whispering";
    return KG_OK;
}
