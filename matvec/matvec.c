/*
 * matvec -- host program of the test (reference: matvec/matvec.c:72-350).
 * real = double, as the reference builds it (makefile.in:12).  2D: run as
 * ./matvec <nx> <ny*ns> <nt>.  y = A x, nt times, on the same inputs.
 */
#include "../kernelgen-perf-tests_amd/host/kg_host.h"

int main(int argc, char* argv[])
{
    if (argc != 4) {
        printf("Usage: %s <nx> <ny> <nt>\n", argv[0]);
        exit(1);
    }
    kg_run r;
    kg_run_begin(&r);
    KG_PARSE_ARG(nx, argv[1]);
    KG_PARSE_ARG(ny, argv[2]);
    KG_PARSE_ARG(nt, argv[3]);

    double* A = kg_host_array((size_t)nx * ny);
    double* x = kg_host_array(nx);
    double* y = kg_host_array(ny);
    /* three means, each accumulated separately (matvec/matvec.c:118-133) */
    double amean = 0.0, xmean = 0.0, ymean = 0.0;
    for (int i = 0; i < nx * ny; i++) {
        A[i] = kg_real_rand();
        amean += A[i];
    }
    for (int i = 0; i < nx; i++) {
        x[i] = kg_real_rand();
        xmean += x[i];
    }
    for (int i = 0; i < ny; i++) {
        y[i] = kg_real_rand();
        ymean += y[i];
    }
    if (!r.no_timing) printf("initial mean = %f\n", amean / (nx * ny) + xmean / nx + ymean / ny);

    kg_run_init(&r);
    double* host[3] = {A, x, y};
    double* dev[3];
    size_t len[3] = {(size_t)nx * ny, (size_t)nx, (size_t)ny};
    kg_alloc_phase(&r, 3, dev, host, len);
    kg_load_phase(&r, 3, dev, host, len, ((double)nx * ny + nx + ny) * sizeof(double));

    kg_phase_start(&r);
    for (int it = 0; it < nt; it++) {
        KG_CALL(kg_matvec(r.ctx, nx, ny, dev[0], dev[1], dev[2]));
        kg_after_launch(&r);
    }
    double t = kg_phase_stop(&r);
    if (!r.no_timing) printf("compute time = %f sec\n", t);

    double* yh = y;
    double* yd = dev[2];
    size_t ny_ = (size_t)ny;
    kg_save_phase(&r, 1, &yh, &yd, &ny_, (double)ny * sizeof(double));
    kg_free_phase(&r, 3, dev);

    printf("final mean = %f\n", kg_sum(y, (size_t)ny) / ny);
    free(A);
    free(x);
    free(y);
    kg_run_end(&r);
    return 0;
}
