/*
 * matmul -- host program of the test (reference: matmul/main.c:56-324, whose
 * kernel is the Fortran matmul_ called by reference, matmul/main.c:235).
 * C(nx,ns) += A(nx,ny) * B(ny,ns) every step, column-major.  The reference
 * never initialises C (matmul/main.c:90); a fresh allocation reads as zeros,
 * which this program makes explicit.
 */
#include <string.h>

#include "../kernelgen-perf-tests_amd/host/kg_host.h"

int main(int argc, char* argv[])
{
    if (argc != 5) {
        printf("Usage: %s <nx> <ny> <ns> <nt>\n", argv[0]);
        exit(1);
    }
    kg_run r;
    kg_run_begin(&r);
    KG_PARSE_ARG(nx, argv[1]);
    KG_PARSE_ARG(ny, argv[2]);
    KG_PARSE_ARG(ns, argv[3]);
    KG_PARSE_ARG(nt, argv[4]);

    size_t na = (size_t)nx * ny, nb = (size_t)ny * ns, nc = (size_t)nx * ns;
    double* A = kg_host_array(na);
    double* B = kg_host_array(nb);
    double* C = kg_host_array(nc);
    memset(C, 0, nc * sizeof(double));
    double mean_a = 0.0, mean_b = 0.0;
    for (size_t i = 0; i < na; i++) {
        A[i] = kg_real_rand();
        mean_a += A[i];
    }
    for (size_t i = 0; i < nb; i++) {
        B[i] = kg_real_rand();
        mean_b += B[i];
    }
    printf("initial mean = %f\n", (mean_a / na + mean_b / nb));

    kg_run_init(&r);
    double* host[3] = {A, B, C};
    double* dev[3];
    size_t len[3] = {na, nb, nc};
    kg_alloc_phase(&r, 3, dev, host, len);
    kg_load_phase(&r, 3, dev, host, len, (double)(na + nb) * sizeof(double));

    kg_phase_start(&r);
    for (int it = 0; it < nt; it++) {
        KG_CALL(kg_matmul(r.ctx, nx, ny, ns, dev[0], dev[1], dev[2]));
        kg_after_launch(&r);
    }
    double t = kg_phase_stop(&r);
    if (!r.no_timing) printf("compute time = %f sec\n", t);

    kg_save_phase(&r, 1, &host[2], &dev[2], &nc, (double)nc * sizeof(double));
    kg_free_phase(&r, 3, dev);

    printf("final mean = %f\n", kg_sum(C, nc) / nc);
    free(A);
    free(B);
    free(C);
    kg_run_end(&r);
    return 0;
}
