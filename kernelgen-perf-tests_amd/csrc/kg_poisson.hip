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
// Outer loop (FP64): r = f - A u (one pass that also writes r in FP32 as the
// inner right-hand side and sums r^2); stop when ||r||_2 / ||f||_2 < tol; else
// solve A e = r approximately in FP32 (V-cycles: full-weighting restriction,
// bilinear prolongation, damped-Jacobi smoothing, rediscretised coarse
// operators down to 3 x 3 vertices) and set u += e.  The FP32 inner solve
// needs only a few digits per correction; FP64 residuals recover the rest.
//
// The FP32 levels are stored with rows padded to P = N rounded up to 4
// floats, so that the sweeps and the prolongation work on 16-byte pieces:
// a lane owns 4 points of PY rows, loads the rows it needs (pieces at x-4, x,
// x+4) through the caches, and writes whole rows -- interior values, and 0 at
// the boundary and padding points, which those arrays hold anyway -- so that
// no cache line is written in part.
#include "kg_internal.h"

#include <chrono>
#include <cmath>
#include <mutex>

namespace {

constexpr int PT = 16;   // 16 x 16 threads per workgroup (scalar kernels)
// Tile order of the sweeps, the FP64 residual and the prolongation
// (kg_tile_order): chunks of 2 tile rows per XCD, 8192^2 solve 17.7 ms
// against 17.9-18.1 in hardware order (chunks of 4 or 8 rows 17.6-17.9),
// bit-identical (the residual's partial sums stay indexed by tile).
#ifndef KG_ORD_PJ
#define KG_ORD_PJ 2
#endif
constexpr int VW = 4;    // waves per workgroup of the 4-point kernels, stacked in y
constexpr int PY = 4;    // rows per lane of the 4-point kernels

typedef float f4 __attribute__((ext_vector_type(4)));

typedef __attribute__((address_space(1))) const f4 gf4;

__device__ __forceinline__ f4 ldf4(const float* p) { return *(gf4*)p; }

// The rows j0-1 .. j0+PY of a lane's 4-point column x (pieces at x-4, x, x+4
// clamped into the padded row; clamped pieces feed only boundary points).
struct Win4 {
    float w[PY + 2][12];
};
__device__ __forceinline__ void load_win4(const float* __restrict__ u, int N, int P, int x, int j0,
                                          Win4& W)
{
    const int xl = x >= 4 ? x - 4 : x, xr = x + 4 < P ? x + 4 : x;
#pragma unroll
    for (int q = 0; q < PY + 2; q++) {
        const float* row = u + (long long)min(j0 - 1 + q, N - 1) * P;
        const f4 a = ldf4(row + xl), b = ldf4(row + x), c = ldf4(row + xr);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            W.w[q][v] = a[v];
            W.w[q][4 + v] = b[v];
            W.w[q][8 + v] = c[v];
        }
    }
}

// damped Jacobi sweep on 4-point pieces, u' = c0 u + c1 (W + S + E + N) +
// c2 (SW + NW + SE + NE) + cf f; writes rows 1 .. N-2 whole (0 at the
// boundary and padding)
__global__ __launch_bounds__(64 * VW)
void pj_smooth4_k(int N, int P, int ntx, float c0, float c1, float c2, float cf,
                  const float* __restrict__ u, const float* __restrict__ f, float* __restrict__ uo)
{
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned tb = kg_tile_order<KG_ORD_PJ>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int x = (int)(tb % (unsigned)ntx) * 256 + lane * 4;
    const int j0 = 1 + ((int)(tb / (unsigned)ntx) * VW + w) * PY;
    if (x >= P || j0 >= N - 1) return;
    Win4 W;
    load_win4(u, N, P, x, j0, W);
#pragma unroll
    for (int r = 0; r < PY; r++) {
        const int j = j0 + r;
        if (j >= N - 1) break;
        const f4 fv = ldf4(f + (long long)j * P + x);
        f4 o;
#pragma unroll
        for (int v = 0; v < 4; v++) {
#define U(dx, dy) W.w[r + 1 + (dy)][4 + v + (dx)]
            const float val = c0 * U(0, 0) + c1 * (U(-1, 0) + U(0, -1) + U(1, 0) + U(0, 1)) +
                              c2 * (U(-1, -1) + U(-1, 1) + U(1, -1) + U(1, 1)) + cf * fv[v];
#undef U
            o[v] = (x + v >= 1 && x + v < N - 1) ? val : 0.f;
        }
        __builtin_nontemporal_store(o, (__attribute__((address_space(1))) f4*)(uo + (long long)j * P + x));
    }
}

// Two damped Jacobi sweeps in one pass, bit-identical to two pj_smooth4_k
// launches (same expression and order, no contraction): the lane forms the
// first sweep's values on the block one point wider on every side than its
// 4 points x PY rows (columns x-1 .. x+4, rows j0-1 .. j0+PY; 0 outside
// rows/columns 1 .. N-2,
// as the first launch would have written them), then the second sweep on
// its own points.  u rows j0-2 .. j0+PY+1 and f rows j0-1 .. j0+PY come
// through the caches; HBM sees each array once per two sweeps.
#ifndef KG_SM2_PY
// rows per lane: 8192^2 solve 19.0 ms (3 rows 19.1, 4 rows 19.4, 6 rows 19.0);
// with the chunked tile order 17.0-17.2 ms (3 rows 17.2-17.3, 4 rows 17.5,
// 6 rows 17.2-17.3)
#define KG_SM2_PY 2
#endif
template <int PY>
__global__ __launch_bounds__(64 * VW)
void pj_smooth4x2_k(int N, int P, int ntx, float c0, float c1, float c2, float cf,
                    const float* __restrict__ u, const float* __restrict__ f,
                    float* __restrict__ uo)
{
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned tb = kg_tile_order<KG_ORD_PJ>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int x = (int)(tb % (unsigned)ntx) * 256 + lane * 4;
    const int j0 = 1 + ((int)(tb / (unsigned)ntx) * VW + w) * PY;
    if (x >= P || j0 >= N - 1) return;
    const int xl = x >= 4 ? x - 4 : x, xr = x + 4 < P ? x + 4 : x;
    // uu[q][c]: u at row j0-2+q, column x-2+c (c = 0 .. 7)
    float uu[PY + 4][8];
#pragma unroll
    for (int q = 0; q < PY + 4; q++) {
        const float* row = u + (long long)min(max(j0 - 2 + q, 0), N - 1) * P;
        const f4 a = ldf4(row + xl), b = ldf4(row + x), c = ldf4(row + xr);
        uu[q][0] = a[2];
        uu[q][1] = a[3];
#pragma unroll
        for (int v = 0; v < 4; v++) uu[q][2 + v] = b[v];
        uu[q][6] = c[0];
        uu[q][7] = c[1];
    }
    // ff[q][c]: f at row j0-1+q, column x-1+c (c = 0 .. 5)
    float ff[PY + 2][6];
#pragma unroll
    for (int q = 0; q < PY + 2; q++) {
        const float* row = f + (long long)min(max(j0 - 1 + q, 0), N - 1) * P;
        const f4 a = ldf4(row + xl), b = ldf4(row + x), c = ldf4(row + xr);
        ff[q][0] = a[3];
#pragma unroll
        for (int v = 0; v < 4; v++) ff[q][1 + v] = b[v];
        ff[q][5] = c[0];
    }
    // first sweep: s1[q][c] at row j0-1+q, column x-1+c
    float s1[PY + 2][6];
#pragma unroll
    for (int q = 0; q < PY + 2; q++) {
        const int j = j0 - 1 + q;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const int xc = x - 1 + c;
#define U(dx, dy) uu[q + 1 + (dy)][c + 1 + (dx)]
            const float val = c0 * U(0, 0) + c1 * (U(-1, 0) + U(0, -1) + U(1, 0) + U(0, 1)) +
                              c2 * (U(-1, -1) + U(-1, 1) + U(1, -1) + U(1, 1)) + cf * ff[q][c];
#undef U
            s1[q][c] = (xc >= 1 && xc < N - 1 && j >= 1 && j < N - 1) ? val : 0.f;
        }
    }
    // second sweep on the lane's points
#pragma unroll
    for (int r = 0; r < PY; r++) {
        const int j = j0 + r;
        if (j >= N - 1) break;
        f4 o;
#pragma unroll
        for (int v = 0; v < 4; v++) {
#define U(dx, dy) s1[r + 1 + (dy)][v + 1 + (dx)]
            const float val = c0 * U(0, 0) + c1 * (U(-1, 0) + U(0, -1) + U(1, 0) + U(0, 1)) +
                              c2 * (U(-1, -1) + U(-1, 1) + U(1, -1) + U(1, 1)) + cf * ff[r + 1][v + 1];
#undef U
            o[v] = (x + v >= 1 && x + v < N - 1) ? val : 0.f;
        }
        __builtin_nontemporal_store(o, (__attribute__((address_space(1))) f4*)(uo + (long long)j * P + x));
    }
}

// The outer correction's FP64 residual, fused with its FP32 copy and its
// squared norm: r = f - (20 u - 4 (u_W + u_S + u_E + u_N) - (u_SW + u_NW +
// u_SE + u_NE)) / (6 h^2), summed in that order in FP64, written as
// FP32 rows of pitch P, 0 on the boundary and the padding, and e = 0 (the
// inner solve's initial guess) at the same points; the sum of r^2 over the
// workgroup's points goes to part[blockIdx.x] (rows of a lane in order, then
// lanes and waves by a fixed tree).  A workgroup of 4 waves takes 256
// consecutive columns of RES_RY rows; a lane loads the RES_RY + 2 rows of
// its column and both neighbours (side values from L1) and the RES_RY
// values of f before the first use.
constexpr int RES_RY = 8;

__global__ __launch_bounds__(256)
void pj_residual64_k(int N, int P, int ntx, double inv6h2, const double* __restrict__ u,
                     const double* __restrict__ f, float* __restrict__ r, float* __restrict__ e,
                     double* __restrict__ part)
{
    __shared__ double wsum[4];
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned tb = kg_tile_order<KG_ORD_PJ>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int x = (int)(tb % (unsigned)ntx) * 256 + w * 64 + lane;
    const int j0 = (int)(tb / (unsigned)ntx) * RES_RY;
    // clamped columns and rows: they only feed points that are not computed
    const int xc = min(x, N - 1), xl = max(xc - 1, 0), xr = min(xc + 1, N - 1);
    double v[RES_RY + 2][3], fv[RES_RY];
#pragma unroll
    for (int q = 0; q < RES_RY + 2; q++) {
        const long long o = (long long)min(max(j0 - 1 + q, 0), N - 1) * N;
        v[q][0] = u[o + xl];
        v[q][1] = u[o + xc];
        v[q][2] = u[o + xr];
    }
#pragma unroll
    for (int q = 0; q < RES_RY; q++) fv[q] = f[(long long)min(j0 + q, N - 1) * N + xc];
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < RES_RY; q++) {
        const int j = j0 + q;
        if (j >= N) break;
        float out = 0.f;
        if (x >= 1 && x < N - 1 && j >= 1 && j < N - 1) {
            const double(&a)[3] = v[q];
            const double(&b)[3] = v[q + 1];
            const double(&c)[3] = v[q + 2];
            const double au = (20.0 * b[1] - 4.0 * (b[0] + a[1] + b[2] + c[1]) -
                               (a[0] + c[0] + a[2] + c[2])) * inv6h2;
            const double rv = fv[q] - au;
            acc += rv * rv;
            out = (float)rv;
        }
        if (x < P) {
            __builtin_nontemporal_store(out, r + (long long)j * P + x);
            __builtin_nontemporal_store(0.f, e + (long long)j * P + x);
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m);
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[tb] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// part[n] = sum of part[0 .. n) in a fixed order (one workgroup of 1024
// threads: strided per-thread sums, then a tree).
__global__ __launch_bounds__(1024)
void pj_sum_k(int n, double* __restrict__ part)
{
    __shared__ double t[1024];
    double s = 0.0;
    for (int i = (int)threadIdx.x; i < n; i += 1024) s += part[i];
    t[threadIdx.x] = s;
    __syncthreads();
    for (int m = 512; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) t[threadIdx.x] += t[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[n] = t[0];
}

// Residual and full-weighting restriction in one pass (FP32): the coarse
// right-hand side
//   fc(I,J) = (4 r_C + 2 (r_W + r_E + r_S + r_N) + (r_SW + r_SE + r_NW + r_NE)) / 16
// of the fine residual r = f - (20 u - 4 (u_W + u_S + u_E + u_N) - (u_SW +
// u_NW + u_SE + u_NE)) / (6 h^2) at the 3 x 3 fine points around (2I, 2J),
// summed in those orders; the fine residual is never stored.  Coarse
// boundary 0; uc = 0 at the same points (the coarse solve's initial guess).
template <int ORD>
__global__ __launch_bounds__(PT * PT)
void pj_resrestrict_k(int Nc, int Pc, int Pf, int ntx, float inv6h2, const float* __restrict__ u,
                      const float* __restrict__ f, float* __restrict__ fc, float* __restrict__ uc)
{
    const unsigned tb = kg_tile_order<ORD>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int I = (int)(tb % (unsigned)ntx) * PT + (int)threadIdx.x;
    const int J = (int)(tb / (unsigned)ntx) * PT + (int)threadIdx.y;
    if (I >= Nc || J >= Nc) return;
    const long long oc = (long long)J * Pc + I;
    uc[oc] = 0.f;
    if (I == 0 || J == 0 || I == Nc - 1 || J == Nc - 1) {
        fc[oc] = 0.f;
        return;
    }
    // fine rows 2J-2 .. 2J+2 and columns 2I-2 .. 2I+2 lie in 0 .. Nf-1
    const long long o = (long long)(2 * J) * Pf + 2 * I;
    float uu[5][5];
#pragma unroll
    for (int dy = -2; dy <= 2; dy++)
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) uu[dy + 2][dx + 2] = u[o + (long long)dy * Pf + dx];
    float rr[3][3];
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
#define U(a, b) uu[dy + 2 + (b)][dx + 2 + (a)]
            const float au = (20.f * U(0, 0) - 4.f * (U(-1, 0) + U(0, -1) + U(1, 0) + U(0, 1)) -
                              (U(-1, -1) + U(-1, 1) + U(1, -1) + U(1, 1))) * inv6h2;
#undef U
            rr[dy + 1][dx + 1] = f[o + (long long)dy * Pf + dx] - au;
        }
#define R(a, b) rr[(b) + 1][(a) + 1]
    fc[oc] = (4.f * R(0, 0) + 2.f * (R(-1, 0) + R(1, 0) + R(0, -1) + R(0, 1)) +
              (R(-1, -1) + R(1, -1) + R(-1, 1) + R(1, 1))) * 0.0625f;
#undef R
}

// u_f += bilinear interpolation of e_c on 4-point pieces of fine rows
// (x multiple of 4, so fine points x, x+2 sit on coarse columns x/2, x/2+1);
// e = e_c at even/even points, the mean of the 2 or 4 coarse neighbours
// elsewhere; writes rows 1 .. Nf-2 whole (0 at the
// boundary and padding).  One row per lane, 4 rows per workgroup.
__global__ __launch_bounds__(256)
void pj_prolong4_k(int Nf, int Pf, int Pc, int ntx, const float* __restrict__ ec,
                   float* __restrict__ uf)
{
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned tb = kg_tile_order<KG_ORD_PJ>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int x = (int)(tb % (unsigned)ntx) * 256 + lane * 4;
    const int j = 1 + (int)(tb / (unsigned)ntx) * 4 + w;
    if (x >= Pf || j >= Nf - 1) return;
    const int Nc = (Nf + 1) / 2;
    const int I0 = x >> 1, J = j >> 1;
    // coarse columns I0 .. I0+2 (clamped; clamped ones feed boundary or
    // padding points only) of rows J and J+1 (odd j)
    float a[3], b[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int I = min(I0 + q, Nc - 1);
        a[q] = ec[(long long)J * Pc + I];
        b[q] = (j & 1) ? ec[(long long)(J + 1) * Pc + I] : 0.f;
    }
    float* row = uf + (long long)j * Pf + x;
    const f4 u = ldf4(row);
    f4 o;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int q = v >> 1;   // coarse column I0 + q of fine point x + v
        float e;
        if (!(v & 1) && !(j & 1)) e = a[q];
        else if ((v & 1) && !(j & 1)) e = 0.5f * (a[q] + a[q + 1]);
        else if (!(v & 1)) e = 0.5f * (a[q] + b[q]);
        else e = 0.25f * (a[q] + a[q + 1] + b[q] + b[q + 1]);
        o[v] = (x + v >= 1 && x + v < Nf - 1) ? u[v] + e : 0.f;
    }
    __builtin_nontemporal_store(o, (__attribute__((address_space(1))) f4*)row);
}

// exact solve on the coarsest level (3 x 3 vertices: one unknown)
__global__ void pj_exact3_k(int P, float s, const float* __restrict__ f, float* __restrict__ u)
{
    if (threadIdx.x == 0) u[P + 1] = s * f[P + 1];
}

__global__ __launch_bounds__(256)
void pj_zero_k(long long n, float* __restrict__ p)
{
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256)
        p[e] = 0.f;
}

// dense N x N FP64 u += FP32 rows of pitch P
__global__ __launch_bounds__(256)
void pj_add_k(int N, int P, const float* __restrict__ e, double* __restrict__ u)
{
    const long long n = (long long)N * N;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        const long long j = q / N, i = q - j * N;
        u[q] += (double)e[j * P + i];
    }
}

struct Level {
    int N, P;   // vertices per side, row pitch (N rounded up to 4)
    double h;
    float *u, *t, *f;
};

unsigned blocks1(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 256 * 64); }

// workgroups of the 4-point kernels on level L
struct Grid4 {
    int ntx;
    unsigned blocks;
};
Grid4 grid4(const Level& L)
{
    Grid4 g;
    g.ntx = (L.P + 255) / 256;
    g.blocks = (unsigned)g.ntx * (unsigned)((L.N - 2 + VW * PY - 1) / (VW * PY));
    return g;
}

// n damped-Jacobi sweeps on level L (result in L.u)
void smooth(hipStream_t st, Level& L, int n, double omega)
{
    const float c0 = (float)(1 - omega), c1 = (float)(omega / 5), c2 = (float)(omega / 20);
    const float cf = (float)(3 * omega * L.h * L.h / 10);
    const Grid4 g = grid4(L);
    for (int s = 0; s < n; s++) {
        if (s + 1 < n) {   // two sweeps in one pass
            constexpr int PY2 = KG_SM2_PY;
            const unsigned b2 = (unsigned)g.ntx * (unsigned)((L.N - 2 + VW * PY2 - 1) / (VW * PY2));
            hipLaunchKernelGGL(pj_smooth4x2_k<PY2>, dim3(b2), dim3(64 * VW), 0, st, L.N, L.P, g.ntx,
                               c0, c1, c2, cf, L.u, L.f, L.t);
            std::swap(L.u, L.t);
            s++;
            continue;
        }
        hipLaunchKernelGGL(pj_smooth4_k, dim3(g.blocks), dim3(64 * VW), 0, st, L.N, L.P, g.ntx, c0,
                           c1, c2, cf, L.u, L.f, L.t);
        // the boundary rows of t are 0 like u's (zeroed at allocation, never
        // written), the boundary columns written as 0
        std::swap(L.u, L.t);
    }
}

void vcycle(hipStream_t st, Level* lv, int nl, int l, int pre, int post, double omega)
{
    Level& L = lv[l];
    if (l + 1 == nl) {
        // 3 x 3 vertices: (A u)(1,1) = 20 u / (6 h^2) = f
        hipLaunchKernelGGL(pj_exact3_k, dim3(1), dim3(64), 0, st, L.P,
                           (float)(6.0 * L.h * L.h / 20.0), L.f, L.u);
        return;
    }
    smooth(st, L, pre, omega);
    Level& C = lv[l + 1];
    // C.f = restricted residual, C.u = 0; C.t needs no clearing: every sweep
    // rewrites its rows 1 .. N-2 whole and its boundary rows stay 0 from the
    // workspace's set-up
    // tiles in chunks of 2 tile rows per XCD on the large levels (8192^2
    // solve 17.0 ms against 17.6 in hardware order, the same u bit for bit),
    // hardware order on the others (1024^2: 2.2 ms against 2.3-2.4 chunked)
    const int ntb = (C.N + PT - 1) / PT;   // tiles per side of the coarse level
    auto rr = C.N > 1024 ? pj_resrestrict_k<2> : pj_resrestrict_k<0>;
    hipLaunchKernelGGL(rr, dim3((unsigned)ntb * (unsigned)ntb), dim3(PT, PT), 0, st, C.N, C.P, L.P,
                       ntb, (float)(1.0 / (6.0 * L.h * L.h)), L.u, L.f, C.f, C.u);
    vcycle(st, lv, nl, l + 1, pre, post, omega);
    {
        const int ntx = (L.P + 255) / 256;
        const unsigned blocks = (unsigned)ntx * (unsigned)((L.N - 2 + 3) / 4);
        hipLaunchKernelGGL(pj_prolong4_k, dim3(blocks), dim3(256), 0, st, L.N, L.P, C.P, ntx, C.u, L.u);
    }
    smooth(st, L, post, omega);
}

// Levels n, n/2, ..., 2 cells per side: at most 14 for n <= 16384.
constexpr int MAX_LEVELS = 14;
constexpr int MAX_N = 1 << 14;

// Device memory of one solve, carved in a fixed order from one or more
// regions: per level (finest first) u, t and f (N x P floats each), then the
// residual pass's partial sums (res_blocks + 1 doubles), then the partials of
// the norm of f (kg_stats_scratch_bytes()).  Every block starts at a
// multiple of 256 bytes.
struct Workspace {
    int nl;
    Level lv[MAX_LEVELS];
    int res_ntx, res_blocks;
    double* part;
    double* stats;
};

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr int pitch4(int N) { return (N + 3) / 4 * 4; }
constexpr int res_blocks_of(int n)
{
    return (pitch4(n + 1) + 255) / 256 * ((n + 1 + RES_RY - 1) / RES_RY);
}

// Bytes of the workspace of an n-cell solve (n a power of two >= 2).
constexpr size_t workspace_bytes(int n, size_t stats_bytes)
{
    size_t b = 0;
    for (int m = n; m >= 2; m /= 2) {
        const size_t ne = (size_t)(m + 1) * (size_t)pitch4(m + 1);
        b += 3 * align256(ne * sizeof(float));
    }
    b += align256((size_t)(res_blocks_of(n) + 1) * sizeof(double));
    b += align256(stats_bytes);
    return b;
}

// Level sizes of an n-cell solve; pointers left null.
void workspace_shape(Workspace& ws, int n)
{
    ws.nl = 0;
    for (int m = n; m >= 2; m /= 2)
        ws.lv[ws.nl++] = Level{m + 1, pitch4(m + 1), 1.0 / m, nullptr, nullptr, nullptr};
    ws.res_ntx = (ws.lv[0].P + 255) / 256;
    ws.res_blocks = res_blocks_of(n);
    ws.part = ws.stats = nullptr;
}

// Points the workspace's blocks into the region at base (workspace_bytes(n)
// bytes, 256-byte aligned).
void workspace_carve(Workspace& ws, char* base)
{
    size_t off = 0;
    for (int l = 0; l < ws.nl; l++) {
        Level& L = ws.lv[l];
        const size_t b = align256((size_t)L.N * L.P * sizeof(float));
        for (float** p : {&L.u, &L.t, &L.f}) {
            *p = (float*)(base + off);
            off += b;
        }
    }
    ws.part = (double*)(base + off);
    off += align256((size_t)(ws.res_blocks + 1) * sizeof(double));
    ws.stats = (double*)(base + off);
}

// The levels' arrays start at 0: sweeps rely on boundary rows and padding
// that they never write.
void workspace_clear(hipStream_t st, Workspace& ws)
{
    for (int l = 0; l < ws.nl; l++) {
        Level& L = ws.lv[l];
        const long long ne = (long long)L.N * L.P;
        for (float* p : {L.u, L.t, L.f})
            hipLaunchKernelGGL(pj_zero_k, dim3(blocks1(ne)), dim3(256), 0, st, ne, p);
    }
}

int check_args(kg_ctx* ctx, const char* who, int n, int nmax, const double* f, const double* u)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (n < 2 || (n & (n - 1)) || n > nmax)
        return kg_set_error(KG_ERR_INVALID, "%s: n = %d must be a power of two in [2, %d]", who, n, nmax);
    if (!f || !u) return kg_set_error(KG_ERR_INVALID, "%s: null buffer", who);
    if ((const void*)f == (const void*)u)
        return kg_set_error(KG_ERR_INVALID, "%s: f and u must be distinct", who);
    return KG_OK;
}

// The defect-correction loop on a carved and cleared workspace; the stream
// is synchronised on return.  Timing (events or the host clock) is the
// caller's: t_start is called once the norm of f is known, right before the
// first residual pass.
template <class Start>
int solve(kg_ctx* ctx, int n, const double* f, double* u, double tol, int max_outer, int vcycles,
          Workspace& ws, kg_poisson_report* rep, int* outer_out, int* cycles_out, double* rel_out,
          Start&& t_start)
{
    hipStream_t st = ctx->stream;
    const int N = n + 1;
    const long long np = (long long)N * N;
    const double h = 1.0 / n;
    const double omega = 0.8;
    const int pre = 2, post = 2;
    if (max_outer < 1) max_outer = 1;
    if (vcycles < 1) vcycles = 1;
    double st_f[5];
    if (int rc = kg_field_stats_launch(st, KG_F64, (size_t)np, f, KG_F64, nullptr, ws.stats)) return rc;
    KG_HIP(hipMemcpyAsync(st_f, ws.stats + kg_stats_scratch_bytes() / sizeof(double) - 5,
                          sizeof st_f, hipMemcpyDeviceToHost, st));
    KG_HIP(hipStreamSynchronize(st));
    const double fnorm = std::sqrt(st_f[1]);
    int outer = 0, cycles = 0;
    double rel = 0;
    Level& F = ws.lv[0];
    t_start();
    for (;;) {
        // r = f - A u in FP64: its FP32 copy is the right-hand side of the
        // inner solve (F.f), its squared norm decides the stop
        hipLaunchKernelGGL(pj_residual64_k, dim3(ws.res_blocks), dim3(256), 0, st, N, F.P, ws.res_ntx,
                           1.0 / (6.0 * h * h), (const double*)u, f, F.f, F.u, ws.part);
        hipLaunchKernelGGL(pj_sum_k, dim3(1), dim3(1024), 0, st, ws.res_blocks, ws.part);
        double sumsq = 0.0;
        KG_HIP(hipMemcpyAsync(&sumsq, ws.part + ws.res_blocks, sizeof(double), hipMemcpyDeviceToHost, st));
        KG_HIP(hipStreamSynchronize(st));
        rel = fnorm > 0 ? std::sqrt(sumsq) / fnorm : 0.0;
        if (rep && outer < KG_POISSON_HISTORY) rep->history[outer] = rel;
        if (rel < tol || outer >= max_outer) break;
        // inner FP32 solve of A e = r from e = 0 (F.u, set by the residual pass)
        for (int c = 0; c < vcycles; c++) vcycle(st, ws.lv, ws.nl, 0, pre, post, omega);
        cycles += vcycles;
        hipLaunchKernelGGL(pj_add_k, dim3(blocks1(np)), dim3(256), 0, st, F.N, F.P, (const float*)F.u, u);
        outer++;
    }
    KG_HIP(hipGetLastError());
    *outer_out = outer;
    *cycles_out = cycles;
    *rel_out = rel;
    return KG_OK;
}

int finish(const char* who, kg_poisson_report* rep, const Workspace& ws, int outer, int cycles,
           double rel, double seconds, double tol)
{
    if (rep) {
        rep->outer = outer;
        rep->vcycles = cycles;
        rep->rel_residual = rel;
        rep->seconds = seconds;
        rep->levels = ws.nl;
    }
    return rel < tol ? KG_OK : kg_set_error(KG_ERR_STATE, "%s: relative residual %.3e after %d "
                                            "corrections (tol %.1e)", who, rel, outer, tol);
}

// ---- the static workspace of kg_poisson_solve_noalloc ----------------------
// Sized for the largest grid that entry point accepts; one copy per device
// (a __device__ array exists once on every GPU the code object is loaded
// on), held until the process ends.
constexpr int NOALLOC_MAX_N = KG_POISSON_NOALLOC_MAX_N;
constexpr size_t STATS_BYTES = (1024 * 5 + 5) * sizeof(double);   // kg_stats_scratch_bytes()
constexpr size_t NOALLOC_BYTES = workspace_bytes(NOALLOC_MAX_N, STATS_BYTES);

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[MAX_DEVICES];   // device address of `scratch` on each device, once set up

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// Called by kg_init with `device` current: looks up this device's copy of
// the static workspace once.  Host code uses only the address found here;
// `scratch` itself names a host-side stand-in.
int kg_poisson_static_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return kg_set_error(KG_ERR_NODEV, "kg_poisson_static_setup: device %d beyond %d", device, MAX_DEVICES);
    if (kg_stats_scratch_bytes() != STATS_BYTES)
        return kg_set_error(KG_ERR_STATE, "kg_poisson_static_setup: statistics scratch size mismatch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return KG_OK;
    void* p = nullptr;
    KG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p || ((uintptr_t)p & 255))
        return kg_set_error(KG_ERR_HIP, "kg_poisson_static_setup: bad workspace address %p", p);
    g_scratch_ptr[device] = (char*)p;
    return KG_OK;
}

extern "C" int kg_poisson_solve(kg_ctx* ctx, int n, const double* f, double* u, double tol,
                                int max_outer, int vcycles, kg_poisson_report* rep)
{
    if (int rc = check_args(ctx, "kg_poisson_solve", n, MAX_N, f, u)) return rc;
    KG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    Workspace ws;
    workspace_shape(ws, n);
    const size_t bytes = workspace_bytes(n, kg_stats_scratch_bytes());
    char* base = nullptr;
    hipError_t e = hipMalloc((void**)&base, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return kg_set_error(KG_ERR_NOMEM, "kg_poisson_solve: %zu bytes: %s", bytes, hipGetErrorString(e));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // every exit from here on waits for the stream and releases what was taken
    auto done = [&](int rc) {
        (void)hipStreamSynchronize(st);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(base);
        return rc;
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipGetLastError();
        return done(kg_set_error(KG_ERR_HIP, "kg_poisson_solve: hipEventCreate failed"));
    }
    workspace_carve(ws, base);
    workspace_clear(st, ws);
    kg_register_kernel("poisson_smooth", (const void*)pj_smooth4_k);
    int outer = 0, cycles = 0;
    double rel = 0;
    if (int rc = solve(ctx, n, f, u, tol, max_outer, vcycles, ws, rep, &outer, &cycles, &rel,
                       [&] { (void)hipEventRecord(e0, st); }))
        return done(rc);
    float ms = 0;
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        (void)hipGetLastError();
        return done(kg_set_error(KG_ERR_HIP, "kg_poisson_solve: event timing failed"));
    }
    return done(finish("kg_poisson_solve", rep, ws, outer, cycles, rel, ms * 1e-3, tol));
}

extern "C" int kg_poisson_solve_noalloc(kg_ctx* ctx, int n, const double* f, double* u, double tol,
                                        int max_outer, int vcycles, kg_poisson_report* rep)
{
    if (int rc = check_args(ctx, "kg_poisson_solve_noalloc", n, NOALLOC_MAX_N, f, u)) return rc;
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES || !g_scratch_ptr[ctx->device])
        return kg_set_error(KG_ERR_STATE, "kg_poisson_solve_noalloc: workspace of device %d not set up",
                            ctx->device);
    KG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    Workspace ws;
    workspace_shape(ws, n);
    if (workspace_bytes(n, STATS_BYTES) > NOALLOC_BYTES)
        return kg_set_error(KG_ERR_INVALID, "kg_poisson_solve_noalloc: n = %d needs %zu bytes of "
                            "workspace, more than the %zu reserved", n, workspace_bytes(n, STATS_BYTES),
                            NOALLOC_BYTES);
    workspace_carve(ws, g_scratch_ptr[ctx->device]);
    // the previous call's contents are still there
    workspace_clear(st, ws);
    kg_register_kernel("poisson_smooth", (const void*)pj_smooth4_k);
    int outer = 0, cycles = 0;
    double rel = 0;
    std::chrono::steady_clock::time_point t0;
    int rc = solve(ctx, n, f, u, tol, max_outer, vcycles, ws, rep, &outer, &cycles, &rel,
                   [&] { t0 = std::chrono::steady_clock::now(); });
    (void)hipStreamSynchronize(st);
    if (rc) return rc;
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return finish("kg_poisson_solve_noalloc", rep, ws, outer, cycles, rel, sec, tol);
}

int kg_preload_poisson() { return kg_preload_kernel((const void*)pj_smooth4_k); }
