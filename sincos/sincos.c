/*
 * sincos -- host program of the test (reference: sincos/main.c:67-326 with the
 * Fortran kernel sincos/sincos.F90:14-74).  Single precision (real = float).
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
    float* x = kg_host_array_f(n);
    float* y = kg_host_array_f(n);
    float* xy = kg_host_array_f(n);
    float mean = 0.0f;   /* accumulated in real (sincos/main.c:120-128) */
    for (size_t i = 0; i < n; i++) {
        x[i] = kg_real_rand_f();
        y[i] = kg_real_rand_f();
        xy[i] = kg_real_rand_f();
        mean += x[i] + y[i] + xy[i];
    }
    printf("initial mean = %f\n", mean / n / 3);

    kg_run_init(&r);
    float* host[3] = {x, y, xy};
    float* dev[3];
    size_t len[3] = {n, n, n};
    kg_alloc_phase_f(&r, 3, dev, host, len);
    kg_load_phase_f(&r, 3, dev, host, len, 3.0 * n * sizeof(float));

    kg_phase_start(&r);
    for (int it = 0; it < nt; it++) {
        KG_CALL(kg_sincos(r.ctx, nx, ny, ns, dev[0], dev[1], dev[2]));
        kg_after_launch(&r);
    }
    double t = kg_phase_stop(&r);
    if (!r.no_timing) printf("compute time = %f sec\n", t);

    float* oh = xy;
    float* od = dev[2];
    kg_save_phase_f(&r, 1, &oh, &od, &n, (double)n * sizeof(float));
    kg_free_phase_f(&r, 3, dev);

    mean = 0.0f;
    for (size_t i = 0; i < n; i++) mean += xy[i];
    printf("final mean = %f\n", mean / n);
    free(x);
    free(y);
    free(xy);
    kg_run_end(&r);
    return 0;
}
