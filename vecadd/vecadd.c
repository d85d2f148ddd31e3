/*
 * vecadd -- host program of the test (reference: vecadd/vecadd.c:86-373).
 * real = double, as the reference builds it (makefile.in:12).  Inputs drawn
 * interleaved with real_rand(), w2 = w0 + w1 with the three levels rotated
 * each step, the final mean over the newest level.
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
    double* w0 = kg_host_array(n);
    double* w1 = kg_host_array(n);
    double* w2 = kg_host_array(n);
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) {
        w0[i] = kg_real_rand();
        w1[i] = kg_real_rand();
        w2[i] = kg_real_rand();
        mean += w0[i] + w1[i] + w2[i];
    }
    if (!r.no_timing) printf("initial mean = %f\n", mean / n / 3);

    kg_run_init(&r);
    double* host[3] = {w0, w1, w2};
    double* dev[3];
    size_t len[3] = {n, n, n};
    kg_alloc_phase(&r, 3, dev, host, len);
    kg_load_phase(&r, 3, dev, host, len, 3.0 * n * sizeof(double));

    int idxs[3] = {0, 1, 2};
    kg_phase_start(&r);
    double *w0p = dev[0], *w1p = dev[1], *w2p = dev[2];
    for (int it = 0; it < nt; it++) {
        KG_CALL(kg_vecadd(r.ctx, nx, ny, ns, w0p, w1p, w2p));
        kg_after_launch(&r);
        double* w = w0p; w0p = w1p; w1p = w2p; w2p = w;
        int idx = idxs[0]; idxs[0] = idxs[1]; idxs[1] = idxs[2]; idxs[2] = idx;
    }
    double t = kg_phase_stop(&r);
    if (!r.no_timing) printf("compute time = %f sec\n", t);

    /* newest level: the one idxs[1] names after the rotation */
    double* out_h = host[idxs[1]];
    double* out_d = dev[idxs[1]];
    kg_save_phase(&r, 1, &out_h, &out_d, &n, (double)n * sizeof(double));
    kg_free_phase(&r, 3, dev);

    printf("final mean = %f\n", kg_sum(out_h, n) / n);
    free(w0);
    free(w1);
    free(w2);
    kg_run_end(&r);
    return 0;
}
