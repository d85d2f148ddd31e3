/*
 * sincos -- host program of the test (reference: sincos/main.c:67-326 with the
 * Fortran kernel sincos/sincos.F90:14-74).  real = double, as the reference
 * builds it (makefile.in:12; the kernel with -fdefault-real-8).
 * xy = sin(x) + cos(y), nt times, on the same inputs.
 */
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

    size_t n = (size_t)nx * ny * ns;
    double* x = kg_host_array(n);
    double* y = kg_host_array(n);
    double* xy = kg_host_array(n);
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) {
        x[i] = kg_real_rand();
        y[i] = kg_real_rand();
        xy[i] = kg_real_rand();
        mean += x[i] + y[i] + xy[i];
    }
    /* printed whether or not NO_TIMING is set (sincos/main.c:107) */
    printf("initial mean = %f\n", mean / n / 3);

    kg_run_init(&r);
    double* host[3] = {x, y, xy};
    double* dev[3];
    size_t len[3] = {n, n, n};
    kg_alloc_phase(&r, 3, dev, host, len);
    /* three arrays copied, two counted in the printed rate (sincos/main.c:209) */
    kg_load_phase(&r, 3, dev, host, len, 2.0 * n * sizeof(double));

    kg_phase_start(&r);
    for (int it = 0; it < nt; it++) {
        KG_CALL(kg_sincos(r.ctx, nx, ny, ns, dev[0], dev[1], dev[2]));
        kg_after_launch(&r);
    }
    double t = kg_phase_stop(&r);
    if (!r.no_timing) printf("compute time = %f sec\n", t);

    double* oh = xy;
    double* od = dev[2];
    kg_save_phase(&r, 1, &oh, &od, &n, (double)n * sizeof(double));
    kg_free_phase(&r, 3, dev);

    printf("final mean = %f\n", kg_sum(xy, n) / n);
    free(x);
    free(y);
    free(xy);
    kg_run_end(&r);
    return 0;
}
