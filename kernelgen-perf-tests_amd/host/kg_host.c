/* kg_host.c -- see kg_host.h. */
#include "kg_host.h"

#include <string.h>

int kg_parse_arg(const char* name, const char* arg)
{
    int v = atoi(arg);
    if (v < 0) {
        printf("Value for %s is invalid: %d\n", name, v);
        exit(1);
    }
    return v;
}

double kg_real_rand(void)
{
    return (((double)(rand() / (double)RAND_MAX) - 0.5) * 2);
}

float kg_real_rand_f(void)
{
    return (float)((((float)(rand() / (double)RAND_MAX)) - 0.5) * 2);
}

void kg_get_time(struct timespec* t)
{
    clock_gettime(CLOCK_REALTIME, t);
}

double kg_time_diff(const struct timespec* a, const struct timespec* b)
{
    long long sec = (long long)b->tv_sec - (long long)a->tv_sec;
    long long nsec = (long long)b->tv_nsec - (long long)a->tv_nsec;
    if (nsec < 0) {
        sec--;
        nsec += 1000000000LL;
    }
    return (double)sec + 1e-9 * (double)nsec;
}

double* kg_host_array(size_t n)
{
    double* p = (double*)memalign(KG_MEMALIGN, n * sizeof(double));
    if (!p) {
        printf("Error allocating memory for arrays: %p\n", (void*)p);
        exit(1);
    }
    return p;
}

void kg_run_begin(kg_run* r)
{
    memset(r, 0, sizeof(*r));
    r->no_timing = getenv("NO_TIMING") != NULL;
    r->profile = getenv("PROFILING_FNAME");
}

void kg_phase_start(kg_run* r)
{
    kg_get_time(&r->t0);
}

double kg_phase_stop(kg_run* r)
{
    struct timespec t1;
    kg_get_time(&t1);
    return kg_time_diff(&r->t0, &t1);
}

void kg_run_init(kg_run* r)
{
    int count = 0;
    if (kg_target_has_device) {
        kg_phase_start(r);
        KG_CALL(kg_device_count(&count));
        KG_CALL(kg_init(&r->ctx, 0));
        double t = kg_phase_stop(r);
        if (!r->no_timing) printf("init time = %f sec\n", t);
    } else {
        KG_CALL(kg_init(&r->ctx, 0));
    }
}

double* kg_dev_array(kg_run* r, size_t n)
{
    void* p = NULL;
    KG_CALL(kg_alloc(r->ctx, &p, n * sizeof(double)));
    return (double*)p;
}

void kg_dev_free(kg_run* r, double* p)
{
    KG_CALL(kg_free(r->ctx, p));
}

void kg_to_dev(kg_run* r, double* d, const double* h, size_t n)
{
    KG_CALL(kg_h2d(r->ctx, d, h, n * sizeof(double)));
}

void kg_to_host(kg_run* r, double* h, const double* d, size_t n)
{
    KG_CALL(kg_d2h(r->ctx, h, d, n * sizeof(double)));
}

void kg_after_launch(kg_run* r)
{
    KG_CALL(kg_sync(r->ctx));
    if (!r->profile) return;
    const char* name = kg_last_kernel_name(r->ctx);
    int regs = 0;
    if (kg_kernel_regcount(r->ctx, name, &regs) == KG_OK) printf("%s regcount = %d\n", name, regs);
    double t = 0.0;
    KG_CALL(kg_last_kernel_time(r->ctx, &t));
    printf("%s kernel time = %.9f\n", name, t);
}

static const double GIB = 1024.0 * 1024.0 * 1024.0;

void kg_alloc_phase(kg_run* r, int n, double** dev, double* const* host, const size_t* len)
{
    if (!kg_target_has_device) {
        for (int i = 0; i < n; i++) dev[i] = host[i];
        return;
    }
    kg_phase_start(r);
    for (int i = 0; i < n; i++) dev[i] = kg_dev_array(r, len[i]);
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("device buffer alloc time = %f sec\n", t);
}

void kg_load_phase(kg_run* r, int n, double* const* dev, double* const* host, const size_t* len,
                   double shown_bytes)
{
    if (!kg_target_has_device) return;
    kg_phase_start(r);
    for (int i = 0; i < n; i++) kg_to_dev(r, dev[i], host[i], len[i]);
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("data load time = %f sec (%f GB/sec)\n", t, shown_bytes / (t * GIB));
}

void kg_save_phase(kg_run* r, int n, double* const* host, double* const* dev, const size_t* len,
                   double shown_bytes)
{
    if (!kg_target_has_device) return;
    kg_phase_start(r);
    for (int i = 0; i < n; i++) kg_to_host(r, host[i], dev[i], len[i]);
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("data save time = %f sec (%f GB/sec)\n", t, shown_bytes / (t * GIB));
}

void kg_free_phase(kg_run* r, int n, double* const* dev)
{
    if (!kg_target_has_device) return;
    kg_phase_start(r);
    for (int i = 0; i < n; i++) kg_dev_free(r, dev[i]);
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("device buffer free time = %f sec\n", t);
}

double kg_sum(const double* a, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += a[i];
    return s;
}

float* kg_host_array_f(size_t n)
{
    float* p = (float*)memalign(KG_MEMALIGN, n * sizeof(float));
    if (!p) {
        printf("Error allocating memory for arrays: %p\n", (void*)p);
        exit(1);
    }
    return p;
}

void kg_alloc_phase_f(kg_run* r, int n, float** dev, float* const* host, const size_t* len)
{
    if (!kg_target_has_device) {
        for (int i = 0; i < n; i++) dev[i] = host[i];
        return;
    }
    kg_phase_start(r);
    for (int i = 0; i < n; i++) {
        void* p = NULL;
        KG_CALL(kg_alloc(r->ctx, &p, len[i] * sizeof(float)));
        dev[i] = (float*)p;
    }
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("device buffer alloc time = %f sec\n", t);
}

void kg_load_phase_f(kg_run* r, int n, float* const* dev, float* const* host, const size_t* len,
                     double shown_bytes)
{
    if (!kg_target_has_device) return;
    kg_phase_start(r);
    for (int i = 0; i < n; i++) KG_CALL(kg_h2d(r->ctx, dev[i], host[i], len[i] * sizeof(float)));
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("data load time = %f sec (%f GB/sec)\n", t, shown_bytes / (t * GIB));
}

void kg_save_phase_f(kg_run* r, int n, float* const* host, float* const* dev, const size_t* len,
                     double shown_bytes)
{
    if (!kg_target_has_device) return;
    kg_phase_start(r);
    for (int i = 0; i < n; i++) KG_CALL(kg_d2h(r->ctx, host[i], dev[i], len[i] * sizeof(float)));
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("data save time = %f sec (%f GB/sec)\n", t, shown_bytes / (t * GIB));
}

void kg_free_phase_f(kg_run* r, int n, float* const* dev)
{
    if (!kg_target_has_device) return;
    kg_phase_start(r);
    for (int i = 0; i < n; i++) KG_CALL(kg_free(r->ctx, dev[i]));
    double t = kg_phase_stop(r);
    if (!r->no_timing) printf("device buffer free time = %f sec\n", t);
}

void kg_run_end(kg_run* r)
{
    kg_destroy(r->ctx);
    r->ctx = NULL;
    fflush(stdout);
}
