/*
 * uxx1 -- host program of the test (reference: uxx1/uxx1.c:121-442).
 * u0/u1 ping-pong; d1, xx, xy, xz are read-only.
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

    double c1 = kg_real_rand();
    double c2 = kg_real_rand();
    printf("c1 = %f, c2 = %f\n", c1, c2);

    size_t n = (size_t)nx * ny * ns;
    double* host[6];
    for (int a = 0; a < 6; a++) host[a] = kg_host_array(n);   /* u0 u1 d1 xx xy xz */
    double mean = 0.0;
    for (size_t i = 0; i < n; i++) {
        for (int a = 0; a < 6; a++) host[a][i] = kg_real_rand();
        mean += host[0][i] + host[1][i] + host[2][i] + host[3][i] + host[4][i] + host[5][i];
    }
    if (!r.no_timing) printf("initial mean = %f\n", mean / n / 6);

    kg_run_init(&r);
    double* dev[6];
    size_t len[6] = {n, n, n, n, n, n};
    kg_alloc_phase(&r, 6, dev, host, len);
    /* the reference prints five arrays' worth for six copies (uxx1.c:300) */
    kg_load_phase(&r, 6, dev, host, len, 5.0 * n * sizeof(double));

    int idxs[2] = {0, 1};
    kg_phase_start(&r);
    double *u0p = dev[0], *u1p = dev[1];
    for (int it = 0; it < nt; it++) {
        KG_CALL(kg_uxx1(r.ctx, nx, ny, ns, c1, c2, u0p, u1p, dev[2], dev[3], dev[4], dev[5]));
        kg_after_launch(&r);
        double* u = u0p; u0p = u1p; u1p = u;
        int idx = idxs[0]; idxs[0] = idxs[1]; idxs[1] = idx;
    }
    double t = kg_phase_stop(&r);
    if (!r.no_timing) printf("compute time = %f sec\n", t);

    double* out_h = host[idxs[1]];
    double* out_d = dev[idxs[1]];
    kg_save_phase(&r, 1, &out_h, &out_d, &n, (double)n * sizeof(double));
    kg_free_phase(&r, 6, dev);

    printf("final mean = %f\n", kg_sum(out_h, n) / n);
    for (int a = 0; a < 6; a++) free(host[a]);
    kg_run_end(&r);
    return 0;
}
