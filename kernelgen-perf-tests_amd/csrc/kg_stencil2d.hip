// kg_stencil2d.hip -- FP64 2D stencils of the suite: jacobi, gaussblur,
// gameoflife, whispering (the reference runs them at nx x (ny*ns), e.g.
// 512 x 65536).
//
// A workgroup of TX threads (four wave64s side by side along x) marches its
// x-segment through a chunk of YC interior rows.  Each thread keeps the rows
// j-r..j+r of its x-neighbourhood in registers and loads one new row per step,
// so each array streams through HBM about once per step.  Arithmetic is in the
// reference's source order and the library is built with -ffp-contract=off.
#include "kg_internal.h"

namespace {

constexpr int TX = 256;
constexpr int YC = 32;

struct ytile {
    int i, j0, j1;
};

// Interior [lo, n-hi) on both axes.
template <int LO, int HI>
__device__ __forceinline__ bool ytile_get(int nx, int ny, int ntx, ytile& t)
{
    const unsigned b = kg_xcd_swizzle(blockIdx.x, gridDim.x);
    const int bx = (int)(b % (unsigned)ntx);
    const int by = (int)(b / (unsigned)ntx);
    t.i = bx * TX + (int)threadIdx.x;
    t.j0 = LO + by * YC;
    t.j1 = min(t.j0 + YC, ny - HI);
    return t.i >= LO && t.i < nx - HI && t.j0 < t.j1;
}

struct ygrid {
    int ntx, nyc;
    unsigned blocks;
    bool empty;
};

static inline ygrid ygrid_make(int nx, int ny, int lo, int hi)
{
    ygrid g;
    const int nyi = ny - hi - lo;
    g.empty = (nx - hi - lo <= 0) || nyi <= 0;
    g.ntx = (nx + TX - 1) / TX;
    g.nyc = g.empty ? 0 : (nyi + YC - 1) / YC;
    g.blocks = (unsigned)g.ntx * (unsigned)g.nyc;
    return g;
}

// ------------------------------------------------------------------- jacobi
// jacobi/jacobi.F90:65-67, interior [1,n-1)^2 (Fortran 2..n-1)
__global__ __launch_bounds__(TX)
void jacobi_k(int nx, int ny, int ntx, double c0, double c1, double c2,
              const double* __restrict__ w0, double* __restrict__ w1)
{
    ytile t;
    if (!ytile_get<1, 1>(nx, ny, ntx, t)) return;
    const long long sy = nx;
    long long o = (long long)t.j0 * sy + t.i;
    // r[m][d]: row j+m-1, column i+d-1
    double r[3][3];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int d = 0; d < 3; d++) r[m][d] = w0[o + (m - 1) * sy + d - 1];
#pragma unroll 4
    for (int j = t.j0; j < t.j1; j++) {
#pragma unroll
        for (int d = 0; d < 3; d++) r[2][d] = w0[o + sy + d - 1];
        kg_st(&w1[o], c0 * r[1][1] +
                      c1 * (r[1][0] + r[0][1] + r[1][2] + r[2][1]) +
                      c2 * (r[0][0] + r[2][0] + r[0][2] + r[2][2]));
#pragma unroll
        for (int d = 0; d < 3; d++) { r[0][d] = r[1][d]; r[1][d] = r[2][d]; }
        o += sy;
    }
}

// ---------------------------------------------------------------- gaussblur
// gaussblur/gaussblur.c:85-92, interior [2,n-2)^2; f from :65 (host-computed)
__global__ __launch_bounds__(TX)
void gaussblur_k(int nx, int ny, int ntx, double f, double s0, double s1, double s2,
                 double s4, double s5, double s8,
                 const double* __restrict__ w0, double* __restrict__ w1)
{
    ytile t;
    if (!ytile_get<2, 2>(nx, ny, ntx, t)) return;
    const long long sy = nx;
    long long o = (long long)t.j0 * sy + t.i;
    // r[m][d]: row j+m-2, column i+d-2
    double r[5][5];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int d = 0; d < 5; d++) r[m][d] = w0[o + (m - 2) * sy + d - 2];
#pragma unroll 2
    for (int j = t.j0; j < t.j1; j++) {
#pragma unroll
        for (int d = 0; d < 5; d++) r[4][d] = w0[o + 2 * sy + d - 2];
#define W(di, dj) r[(dj) + 2][(di) + 2]
        kg_st(&w1[o], f * (
            s0 * W(0, 0) +
            s1 * (W(-1, 0) + W(1, 0) + W(0, -1) + W(0, 1)) +
            s2 * (W(-1, -1) + W(1, -1) + W(-1, 1) + W(1, 1)) +
            s4 * (W(-2, 0) + W(2, 0) + W(0, -2) + W(0, 2)) +
            s5 * (W(-2, -1) + W(-1, -2) + W(1, -2) + W(2, -1) +
                  W(-2, 1) + W(-1, 2) + W(1, 2) + W(2, 1)) +
            s8 * (W(-2, -2) + W(2, -2) + W(-2, 2) + W(2, 2))));
#undef W
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int d = 0; d < 5; d++) r[m][d] = r[m + 1][d];
        o += sy;
    }
}

// --------------------------------------------------------------- gameoflife
// gameoflife/gameoflife.c:82-91, interior [1,n-1)^2
__global__ __launch_bounds__(TX)
void gameoflife_k(int nx, int ny, int ntx, const double* __restrict__ u0, double* __restrict__ u1)
{
    ytile t;
    if (!ytile_get<1, 1>(nx, ny, ntx, t)) return;
    const long long sy = nx;
    long long o = (long long)t.j0 * sy + t.i;
    double r[3][3];
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int d = 0; d < 3; d++) r[m][d] = u0[o + (m - 1) * sy + d - 1];
    const double C = 100000000000000000000.;
#pragma unroll 4
    for (int j = t.j0; j < t.j1; j++) {
#pragma unroll
        for (int d = 0; d < 3; d++) r[2][d] = u0[o + sy + d - 1];
        const double L =
            r[0][0] + r[0][1] + r[0][2] +
            r[1][0] + r[1][2] +
            r[2][0] + r[2][1] + r[2][2];
        kg_st(&u1[o], 1. / (1. + (r[1][1] + L - 3.) * (L - 3.) * C));
#pragma unroll
        for (int d = 0; d < 3; d++) { r[0][d] = r[1][d]; r[1][d] = r[2][d]; }
        o += sy;
    }
}

// --------------------------------------------------------------- whispering
// whispering/whispering.c:94-106 with the serial sweep's in-place e updates
// made explicit.  Visiting (i,j), the reference reads e(i,j-1,0) after the
// (i,j-1) update replaced it (when j-1 >= 1) and e(i-1,j,1) after the (i-1,j)
// update replaced it (when i-1 >= 1).  Those updates are recomputed here from
// e_in; their difference terms h0(i,j)-h0(i,j-1) and h0(i,j)-h0(i-1,j) are the
// ones ey and ex use.  e is planar: component c at e[i + nx*j + nx*ny*c].
__global__ __launch_bounds__(TX)
void whispering_k(int nx, int ny, int ntx, double mu, double epsilon,
                  const double* __restrict__ e_in, double* __restrict__ e_out,
                  const double* __restrict__ h0, double* __restrict__ h1,
                  const double* __restrict__ u_em0, double* __restrict__ u_em1,
                  const double* __restrict__ ca, const double* __restrict__ cb,
                  const double* __restrict__ da, const double* __restrict__ db)
{
    ytile t;
    if (!ytile_get<1, 1>(nx, ny, ntx, t)) return;
    const long long sy = nx;
    const long long n = (long long)nx * ny;
    const double* ei0 = e_in;
    const double* ei1 = e_in + n;
    double* eo0 = e_out;
    double* eo1 = e_out + n;
    const int i = t.i;
    long long o = (long long)t.j0 * sy + i;
    // Row j-1 values of this column carried from the previous step.
    double h_m = h0[o - sy], ca_m = ca[o - sy], cb_m = cb[o - sy], e0_m = ei0[o - sy];
    double h_c = h0[o];
#pragma unroll 2
    for (int j = t.j0; j < t.j1; j++) {
        const double ca_c = ca[o], cb_c = cb[o], e0_c = ei0[o], e1_c = ei1[o];
        const double h_p = h0[o + sy];
        const double h_xp = h0[o + 1], h_xm = h0[o - 1];
        const double ca_x = ca[o - 1], cb_x = cb[o - 1], e1_x = ei1[o - 1];
        const double e0 = ca_c * e0_c + cb_c * (h_p - h_c);
        const double e1 = ca_c * e1_c - cb_c * (h_xp - h_c);
        const double dy = h_c - h_m;
        const double dx = h_c - h_xm;
        const double ey_src = (j - 1 >= 1) ? ca_m * e0_m + cb_m * dy : e0_m;
        const double ex_src = (i - 1 >= 1) ? ca_x * e1_x - cb_x * dx : e1_x;
        const double ey = ca_m * ey_src + cb_m * dy;
        const double ex = ca_x * ex_src - cb_x * dx;
        kg_st(&eo0[o], e0);
        kg_st(&eo1[o], e1);
        const double h = da[o] * h_c + db[o] * (e0 - ey + ex - e1);
        kg_st(&h1[o], h);
        kg_st(&u_em1[o], u_em0[o] + 0.5 * (h * h / mu + epsilon * (e0 * e0 + e1 * e1)));
        h_m = h_c;
        h_c = h_p;
        ca_m = ca_c;
        cb_m = cb_c;
        e0_m = e0_c;
        o += sy;
    }
}

}  // namespace

#define KG_LAUNCH2(NAME, KERNEL, LO, HI, ...)                                              \
    do {                                                                                   \
        if (int rc_ = kg_check_ctx(ctx)) return rc_;                                       \
        if (int rc_ = kg_check_dims2(nx, ny)) return rc_;                                  \
        const ygrid g = ygrid_make(nx, ny, LO, HI);                                        \
        kg_register_kernel(NAME, (const void*)KERNEL);                                     \
        if (int rc_ = kg_launch_begin(ctx, NAME)) return rc_;                              \
        if (!g.empty)                                                                      \
            hipLaunchKernelGGL(KERNEL, dim3(g.blocks), dim3(TX), 0, ctx->stream,           \
                               nx, ny, g.ntx, __VA_ARGS__);                                \
        return kg_launch_end(ctx);                                                         \
    } while (0)

#define KG_REQUIRE(cond, msg) \
    do { if (!(cond)) return kg_set_error(KG_ERR_INVALID, "%s", msg); } while (0)

extern "C" int kg_jacobi(kg_ctx* ctx, int nx, int ny, double c0, double c1, double c2,
                         const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_jacobi: null buffer");
    KG_REQUIRE(w0 != w1, "kg_jacobi: w0 and w1 must be distinct");
    KG_LAUNCH2("jacobi_", jacobi_k, 1, 1, c0, c1, c2, w0, w1);
}

extern "C" int kg_gaussblur(kg_ctx* ctx, int nx, int ny, double s0, double s1, double s2,
                            double s4, double s5, double s8, const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_gaussblur: null buffer");
    KG_REQUIRE(w0 != w1, "kg_gaussblur: w0 and w1 must be distinct");
    const double f = 1. / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5);   // gaussblur/gaussblur.c:65
    KG_LAUNCH2("gaussblur", gaussblur_k, 2, 2, f, s0, s1, s2, s4, s5, s8, w0, w1);
}

extern "C" int kg_gameoflife(kg_ctx* ctx, int nx, int ny, const double* u0, double* u1)
{
    KG_REQUIRE(u0 && u1, "kg_gameoflife: null buffer");
    KG_REQUIRE(u0 != u1, "kg_gameoflife: u0 and u1 must be distinct");
    KG_LAUNCH2("gameoflife", gameoflife_k, 1, 1, u0, u1);
}

extern "C" int kg_whispering(kg_ctx* ctx, int nx, int ny, double mu, double epsilon,
                             const double* e_in, double* e_out,
                             const double* h0, double* h1, const double* u_em0, double* u_em1,
                             const double* ca, const double* cb, const double* da,
                             const double* db)
{
    KG_REQUIRE(e_in && e_out && h0 && h1 && u_em0 && u_em1 && ca && cb && da && db,
               "kg_whispering: null buffer");
    KG_REQUIRE(e_in != e_out && h0 != h1 && u_em0 != u_em1,
               "kg_whispering: e, h and u_em need distinct in/out buffers");
    KG_LAUNCH2("whispering", whispering_k, 1, 1, mu, epsilon, e_in, e_out, h0, h1, u_em0, u_em1,
               ca, cb, da, db);
}
