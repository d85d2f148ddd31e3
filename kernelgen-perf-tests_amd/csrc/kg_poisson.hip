// kg_poisson.hip -- 2D Poisson solve with FP32 multigrid sweeps inside an FP64
// defect-correction loop.
//
// Problem: A u = f on the unit square with u = 0 on the boundary, on an
// (n+1) x (n+1) vertex grid (n cells per side, n a power of two, h = 1/n),
// x fastest.  A is the 9-point operator whose Jacobi sweep is the reference's
// jacobi kernel (jacobi/jacobi.F90:65-67: c0*w + c1*(4 axis neighbours) +
// c2*(4 diagonal neighbours)) plus a right-hand-side term:
//
//   (A u)(i,j) = (20 u - 4 (u_E + u_W + u_N + u_S) - (u_NE + u_NW + u_SE + u_SW)) / (6 h^2)
//
// damped Jacobi: u' = (1-w) u + (w/5) sum_axis + (w/20) sum_diag + (3 w h^2 / 10) f.
//
// Outer loop (FP64): r = f - A u; stop when ||r||_2 / ||f||_2 < tol; else
// solve A e = r approximately in FP32 (V-cycles: full-weighting restriction,
// bilinear prolongation, damped-Jacobi smoothing, rediscretised coarse
// operators down to 3 x 3 vertices) and set u += e.  The FP32 inner solve
// needs only a few digits per correction; FP64 residuals recover the rest.
#include "kg_internal.h"

#include <cmath>
#include <vector>

namespace {

constexpr int PT = 16;   // 16 x 16 threads per workgroup

// damped Jacobi sweep (interior only; boundary rows/columns of uo untouched)
template <class T>
__global__ __launch_bounds__(PT * PT)
void pj_smooth_k(int N, T c0, T c1, T c2, T cf, const T* __restrict__ u,
                 const T* __restrict__ f, T* __restrict__ uo)
{
    const int i = blockIdx.x * PT + threadIdx.x, j = blockIdx.y * PT + threadIdx.y;
    if (i < 1 || j < 1 || i >= N - 1 || j >= N - 1) return;
    const long long o = (long long)j * N + i;
    uo[o] = c0 * u[o] + c1 * (u[o - 1] + u[o - N] + u[o + 1] + u[o + N]) +
            c2 * (u[o - 1 - N] + u[o - 1 + N] + u[o + 1 - N] + u[o + 1 + N]) + cf * f[o];
}

// r = f - A u on the interior, 0 on the boundary; TA: arithmetic type
template <class T, class TR>
__global__ __launch_bounds__(PT * PT)
void pj_residual_k(int N, T inv6h2, const T* __restrict__ u, const T* __restrict__ f,
                   TR* __restrict__ r)
{
    const int i = blockIdx.x * PT + threadIdx.x, j = blockIdx.y * PT + threadIdx.y;
    if (i >= N || j >= N) return;
    const long long o = (long long)j * N + i;
    if (i == 0 || j == 0 || i == N - 1 || j == N - 1) {
        r[o] = (TR)0;
        return;
    }
    const T au = (T(20) * u[o] - T(4) * (u[o - 1] + u[o - N] + u[o + 1] + u[o + N]) -
                  (u[o - 1 - N] + u[o - 1 + N] + u[o + 1 - N] + u[o + 1 + N])) * inv6h2;
    r[o] = (TR)(f[o] - au);
}

// full weighting: fc(I,J) from rf around (2I, 2J); coarse boundary = 0
__global__ __launch_bounds__(PT * PT)
void pj_restrict_k(int Nc, const float* __restrict__ rf, float* __restrict__ fc)
{
    const int I = blockIdx.x * PT + threadIdx.x, J = blockIdx.y * PT + threadIdx.y;
    if (I >= Nc || J >= Nc) return;
    const long long oc = (long long)J * Nc + I;
    if (I == 0 || J == 0 || I == Nc - 1 || J == Nc - 1) {
        fc[oc] = 0.f;
        return;
    }
    const int Nf = 2 * Nc - 1;
    const long long o = (long long)(2 * J) * Nf + 2 * I;
    fc[oc] = (4.f * rf[o] + 2.f * (rf[o - 1] + rf[o + 1] + rf[o - Nf] + rf[o + Nf]) +
              (rf[o - 1 - Nf] + rf[o + 1 - Nf] + rf[o - 1 + Nf] + rf[o + 1 + Nf])) * 0.0625f;
}

// u_f += bilinear interpolation of e_c (interior of the fine grid)
__global__ __launch_bounds__(PT * PT)
void pj_prolong_k(int Nf, const float* __restrict__ ec, float* __restrict__ uf)
{
    const int i = blockIdx.x * PT + threadIdx.x, j = blockIdx.y * PT + threadIdx.y;
    if (i < 1 || j < 1 || i >= Nf - 1 || j >= Nf - 1) return;
    const int Nc = (Nf + 1) / 2;
    const int I = i >> 1, J = j >> 1;
    const long long oc = (long long)J * Nc + I;
    float e;
    if (!(i & 1) && !(j & 1)) e = ec[oc];
    else if (i & 1 && !(j & 1)) e = 0.5f * (ec[oc] + ec[oc + 1]);
    else if (!(i & 1)) e = 0.5f * (ec[oc] + ec[oc + Nc]);
    else e = 0.25f * (ec[oc] + ec[oc + 1] + ec[oc + Nc] + ec[oc + Nc + 1]);
    uf[(long long)j * Nf + i] += e;
}

__global__ __launch_bounds__(256)
void pj_zero_k(long long n, float* __restrict__ p)
{
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        p[e] = 0.f;
}

__global__ __launch_bounds__(256)
void pj_to_f32_k(long long n, const double* __restrict__ s, float* __restrict__ d)
{
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        d[e] = (float)s[e];
}

__global__ __launch_bounds__(256)
void pj_add_k(long long n, const float* __restrict__ e, double* __restrict__ u)
{
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256)
        u[q] += (double)e[q];
}

struct Level {
    int N;
    double h;
    float *u, *t, *f, *r;
};

dim3 grid2(int N) { return dim3((N + PT - 1) / PT, (N + PT - 1) / PT); }
unsigned blocks1(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 256 * 64); }

// n damped-Jacobi sweeps on level L (result in L.u)
void smooth(hipStream_t st, Level& L, int n, double omega)
{
    const float c0 = (float)(1 - omega), c1 = (float)(omega / 5), c2 = (float)(omega / 20);
    const float cf = (float)(3 * omega * L.h * L.h / 10);
    for (int s = 0; s < n; s++) {
        hipLaunchKernelGGL(pj_smooth_k<float>, grid2(L.N), dim3(PT, PT), 0, st, L.N, c0, c1, c2, cf,
                           L.u, L.f, L.t);
        // the boundary of t must be 0 like u's: copy-free by swapping after
        // both were zeroed at the start of the cycle
        std::swap(L.u, L.t);
    }
}

void vcycle(hipStream_t st, std::vector<Level>& lv, size_t l, int pre, int post, double omega)
{
    Level& L = lv[l];
    if (l + 1 == lv.size()) {
        smooth(st, L, 32, omega);
        return;
    }
    smooth(st, L, pre, omega);
    hipLaunchKernelGGL((pj_residual_k<float, float>), grid2(L.N), dim3(PT, PT), 0, st, L.N,
                       (float)(1.0 / (6.0 * L.h * L.h)), L.u, L.f, L.r);
    Level& C = lv[l + 1];
    hipLaunchKernelGGL(pj_restrict_k, grid2(C.N), dim3(PT, PT), 0, st, C.N, L.r, C.f);
    const long long nc = (long long)C.N * C.N;
    hipLaunchKernelGGL(pj_zero_k, dim3(blocks1(nc)), dim3(256), 0, st, nc, C.u);
    hipLaunchKernelGGL(pj_zero_k, dim3(blocks1(nc)), dim3(256), 0, st, nc, C.t);
    vcycle(st, lv, l + 1, pre, post, omega);
    hipLaunchKernelGGL(pj_prolong_k, grid2(L.N), dim3(PT, PT), 0, st, L.N, C.u, L.u);
    smooth(st, L, post, omega);
}

}  // namespace

extern "C" int kg_poisson_solve(kg_ctx* ctx, int n, const double* f, double* u, double tol,
                                int max_outer, int vcycles, kg_poisson_report* rep)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (n < 2 || (n & (n - 1)) || n > (1 << 14))
        return kg_set_error(KG_ERR_INVALID, "kg_poisson_solve: n = %d must be a power of two in [2, 16384]", n);
    if (!f || !u) return kg_set_error(KG_ERR_INVALID, "kg_poisson_solve: null buffer");
    if ((const void*)f == (const void*)u)
        return kg_set_error(KG_ERR_INVALID, "kg_poisson_solve: f and u must be distinct");
    if (max_outer < 1) max_outer = 1;
    if (vcycles < 1) vcycles = 1;
    hipStream_t st = ctx->stream;
    const int N = n + 1;
    const long long np = (long long)N * N;
    const double h = 1.0 / n;
    const double omega = 0.8;
    const int pre = 2, post = 2;
    std::vector<Level> lv;
    for (int m = n; m >= 2; m /= 2) lv.push_back(Level{m + 1, 1.0 / m, nullptr, nullptr, nullptr, nullptr});
    std::vector<void*> allocs;
    auto fail = [&](int rc) {
        (void)hipStreamSynchronize(st);
        for (void* p : allocs) (void)hipFree(p);
        return rc;
    };
    auto alloc = [&](size_t bytes, void** p) -> int {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return kg_set_error(KG_ERR_NOMEM, "kg_poisson_solve: %zu bytes: %s", bytes, hipGetErrorString(e));
        }
        allocs.push_back(*p);
        return KG_OK;
    };
    for (auto& L : lv) {
        const size_t b = (size_t)L.N * L.N * sizeof(float);
        for (float** p : {&L.u, &L.t, &L.f, &L.r})
            if (int rc = alloc(b, (void**)p)) return fail(rc);
    }
    double* r64 = nullptr;
    if (int rc = alloc((size_t)np * sizeof(double), (void**)&r64)) return fail(rc);

    kg_register_kernel("poisson_smooth", (const void*)pj_smooth_k<float>);
    double st_f[5], st_r[5];
    if (int rc = kg_field_stats(ctx, KG_F64, (size_t)np, f, KG_F64, nullptr, st_f)) return fail(rc);
    const double fnorm = std::sqrt(st_f[1]);
    int outer = 0, cycles = 0;
    double rel = 0;
    hipEvent_t e0, e1;
    KG_HIP(hipEventCreate(&e0));
    KG_HIP(hipEventCreate(&e1));
    (void)hipEventRecord(e0, st);
    for (;;) {
        hipLaunchKernelGGL((pj_residual_k<double, double>), grid2(N), dim3(PT, PT), 0, st, N,
                           1.0 / (6.0 * h * h), (const double*)u, f, r64);
        if (int rc = kg_field_stats(ctx, KG_F64, (size_t)np, r64, KG_F64, nullptr, st_r)) return fail(rc);
        rel = fnorm > 0 ? std::sqrt(st_r[1]) / fnorm : 0.0;
        if (rep && outer < KG_POISSON_HISTORY) rep->history[outer] = rel;
        if (rel < tol || outer >= max_outer) break;
        // inner FP32 solve of A e = r, e0 = 0
        Level& F = lv[0];
        hipLaunchKernelGGL(pj_to_f32_k, dim3(blocks1(np)), dim3(256), 0, st, np, (const double*)r64, F.f);
        hipLaunchKernelGGL(pj_zero_k, dim3(blocks1(np)), dim3(256), 0, st, np, F.u);
        hipLaunchKernelGGL(pj_zero_k, dim3(blocks1(np)), dim3(256), 0, st, np, F.t);
        for (int c = 0; c < vcycles; c++) vcycle(st, lv, 0, pre, post, omega);
        cycles += vcycles;
        hipLaunchKernelGGL(pj_add_k, dim3(blocks1(np)), dim3(256), 0, st, np, (const float*)F.u, u);
        outer++;
    }
    (void)hipEventRecord(e1, st);
    float ms = 0;
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    KG_HIP(hipGetLastError());
    if (rep) {
        rep->outer = outer;
        rep->vcycles = cycles;
        rep->rel_residual = rel;
        rep->seconds = ms * 1e-3;
        rep->levels = (int)lv.size();
    }
    return fail(rel < tol ? KG_OK : kg_set_error(KG_ERR_STATE, "kg_poisson_solve: relative residual "
                                                 "%.3e after %d corrections (tol %.1e)", rel, outer, tol));
}

int kg_preload_poisson() { return kg_preload_kernel((const void*)pj_smooth_k<float>); }
