/*
 * gaussblur -- host program of the test (reference: gaussblur/gaussblur.c:107-383).
 * 2D: argv nx ny nt; w0/w1 ping-pong; six random weights.
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

    double s0 = kg_real_rand();
    double s1 = kg_real_rand();
    double s2 = kg_real_rand();
    double s4 = kg_real_rand();
    double s5 = kg_real_rand();
    double s8 = kg_real_rand();
    printf("s0 = %f, s1 = %f, s2 = %f\n", s0, s1, s2);
    printf("s4 = %f, s5 = %f, s8 = %f\n", s4, s5, s8);

    size_t n = (size_t)nx * ny;
    double* w0 = kg_host_array(n);
    double* w1 = kg_host_array(n);
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) {
        w0[i] = kg_real_rand();
        w1[i] = kg_real_rand();
        mean += w0[i] + w1[i];
    }
    printf("initial mean = %f\n", mean / n / 2);

    kg_run_init(&r);
    double* host[2] = {w0, w1};
    double* dev[2];
    size_t len[2] = {n, n};
    kg_alloc_phase(&r, 2, dev, host, len);
    kg_load_phase(&r, 2, dev, host, len, 2.0 * n * sizeof(double));

    int idxs[2] = {0, 1};
    kg_phase_start(&r);
    double *w0p = dev[0], *w1p = dev[1];
    for (int it = 0; it < nt; it++) {
        KG_CALL(kg_gaussblur(r.ctx, nx, ny, s0, s1, s2, s4, s5, s8, w0p, w1p));
        kg_after_launch(&r);
        double* w = w0p; w0p = w1p; w1p = w;
        int idx = idxs[0]; idxs[0] = idxs[1]; idxs[1] = idx;
    }
    double t = kg_phase_stop(&r);
    if (!r.no_timing) printf("compute time = %f sec\n", t);

    double* out_h = host[idxs[1]];
    double* out_d = dev[idxs[1]];
    kg_save_phase(&r, 1, &out_h, &out_d, &n, (double)n * sizeof(double));
    kg_free_phase(&r, 2, dev);

    printf("final mean = %f\n", kg_sum(out_h, n) / n);
    free(w0);
    free(w1);
    kg_run_end(&r);
    return 0;
}
