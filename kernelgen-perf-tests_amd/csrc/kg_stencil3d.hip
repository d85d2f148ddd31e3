// kg_stencil3d.hip -- FP64 3D stencils of the suite: laplacian, divergence,
// gradient, wave13pt, uxx1, lapgsrb.
//
// laplacian, gradient, wave13pt and lapgsrb run on the plane-staged engine
// (kg_plane.h): xy-neighbours from LDS, z-neighbours from a register column,
// every global load prefetched a few planes ahead.  divergence and uxx1 read
// each of their inputs at only a few offsets along one axis: with an even nx
// and 16-byte aligned fields they run the direct form (two points and a few
// planes per lane, every operand a 16-byte load through the caches, all
// issued up front), otherwise the plain z-march (kg_zmarch.h): a TX x TY
// column tile marched through KC planes, one point per thread.
// Arithmetic is written in the reference's source order and the library is
// built with -ffp-contract=off: results are bit-identical to an IEEE build of
// the reference kernels.
#include "kg_ops3d.h"
#include "kg_zmarch.h"

#include <algorithm>
#include <type_traits>

// Tile order per kernel (kg_tile_order: 0 hardware, 1 XCD-contiguous, R >= 2
// chunks of R tile rows per XCD; plane_kernel's -R: z-chunk fastest, runs of
// R whole z-columns per XCD), measured at 512x256x256.  wave13pt in chunks of
// 2 tile rows 141.5-142 us on the first ten launches against 144.7-145.3
// XCD-contiguous (z-column runs of 1, 2, 4: 148-153); the z-column runs were
// slower for laplacian too (111-121 us against 98-99) and alike for lapgsrb.
#ifndef KG_ORD_LAP
#define KG_ORD_LAP 2
#endif
#ifndef KG_ORD_GRAD
#define KG_ORD_GRAD 2
#endif
#ifndef KG_ORD_WAVE
#define KG_ORD_WAVE 2
#endif
#ifndef KG_ORD_GSRB
#define KG_ORD_GSRB 0
#endif
#ifndef KG_ORD_DIV
#define KG_ORD_DIV 0
#endif
#ifndef KG_ORD_UXX
#define KG_ORD_UXX 0
#endif

namespace {

using kgops::LaplacianOp;
using kgops::Wave13ptOp;

constexpr int TX = 64;   // one wave64 per x-row segment
constexpr int TY = 4;    // 4 rows -> 256 threads per workgroup
constexpr int KC = 16;   // z-planes per workgroup

#define KG_REQUIRE(cond, msg) \
    do { if (!(cond)) return kg_set_error(KG_ERR_INVALID, "%s", msg); } while (0)

// --------------------------------------------------------------- divergence
// divergence/divergence.c:93-96
__global__ __launch_bounds__(TX * TY)
void divergence_k(int nx, int ny, int ns, int ntx, int nty, double alpha, double beta, double gamma,
                  double* __restrict__ u, const double* __restrict__ ux,
                  const double* __restrict__ uy, const double* __restrict__ uz)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 1, 1, false>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    double zm = uz[o - sz];
    double zc = uz[o];
#pragma unroll 4
    for (int k = t.k0; k < t.k1; k++) {
        const double zp = uz[o + sz];
        kg_st(&u[o], alpha * (ux[o + 1] - ux[o - 1]) +
                     beta * (uy[o + sy] - uy[o - sy]) +
                     gamma * (zp - zm));
        zm = zc;
        zc = zp;
        o += sz;
    }
}

// Direct form (even nx, 16-byte aligned fields): a lane owns points x, x+1
// (x even) of row j and PZ consecutive planes; every operand is a 16-byte
// load through the caches (ux as the aligned pairs at x-2, x, x+2; uy at rows
// j-1, j+1; uz at planes k0-1 .. k0+PZ), all issued before the first use.
// Workgroups stay in hardware order, at most DIV_OCC (16 waves) per compute
// unit: 183 us at 512x256x256 (unlimited 187, 3 per CU 196; uxx1 gains
// nothing from such a cap).  Other tile orders (kg_tile_order) measured no
// better: chunks of 2, 4 or 8 tile rows per XCD 183-186 us (uxx1 286-295
// against 284-286 in hardware order), XCD-contiguous eighths 188 us (uxx1
// 310).  Same expression as divergence_k.
#ifndef KG_DIV_PZ
#define KG_DIV_PZ 4
#endif
#ifndef KG_DIV_ROWS
#define KG_DIV_ROWS 4
#endif
constexpr int DIV_PZ = KG_DIV_PZ, DIV_ROWS = KG_DIV_ROWS, DIV_OCC = 4;

typedef double d2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ d2v ld2(const double* p)
{
    return *(const kgp::gmem<d2v>*)kgp::as_global(p);
}

__global__ __launch_bounds__(64 * DIV_ROWS)
void divergence2_k(int nx, int ny, int ns, int ntx, int nty, double alpha, double beta,
                   double gamma, double* __restrict__ u, const double* __restrict__ ux,
                   const double* __restrict__ uy, const double* __restrict__ uz)
{
    unsigned b = kg_tile_order<KG_ORD_DIV>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int bx = (int)(b % (unsigned)ntx);
    b /= (unsigned)ntx;
    const int by = (int)(b % (unsigned)nty);
    const int bz = (int)(b / (unsigned)nty);
    const int lane = (int)(threadIdx.x & 63), r = (int)(threadIdx.x >> 6);
    const int x = bx * 128 + 2 * lane;
    const int j = 1 + by * DIV_ROWS + r;
    const int k0 = 1 + bz * DIV_PZ;
    if (x >= nx || j >= ny - 1 || k0 >= ns - 1) return;
    // pairs at x-2 / x+2 clamped into the row at the x-edges (they then feed
    // boundary points only); planes past ns-2 clamped (never stored)
    const int xl = x >= 2 ? x - 2 : x;
    const int xr = x + 2 < nx ? x + 2 : x;
    const long long sz = (long long)nx * ny;
    const long long oj = (long long)j * nx;
    d2v xm[DIV_PZ], xc[DIV_PZ], xp[DIV_PZ], ym[DIV_PZ], yp[DIV_PZ], z[DIV_PZ + 2];
#pragma unroll
    for (int p = 0; p < DIV_PZ; p++) {
        const long long o = (long long)min(k0 + p, ns - 2) * sz + oj;
        xm[p] = ld2(ux + o + xl);
        xc[p] = ld2(ux + o + x);
        xp[p] = ld2(ux + o + xr);
        ym[p] = ld2(uy + o - nx + x);
        yp[p] = ld2(uy + o + nx + x);
    }
#pragma unroll
    for (int p = 0; p < DIV_PZ + 2; p++) z[p] = ld2(uz + (long long)min(k0 - 1 + p, ns - 1) * sz + oj + x);
    const bool m0 = x >= 1 && x < nx - 1, m1 = x + 1 < nx - 1;
#pragma unroll
    for (int p = 0; p < DIV_PZ; p++) {
        if (k0 + p >= ns - 1) break;
        double res[2];
        res[0] = alpha * (xc[p].y - xm[p].y) + beta * (yp[p].x - ym[p].x) + gamma * (z[p + 2].x - z[p].x);
        res[1] = alpha * (xp[p].x - xc[p].x) + beta * (yp[p].y - ym[p].y) + gamma * (z[p + 2].y - z[p].y);
        double* o = u + (long long)(k0 + p) * sz + oj + x;
        if (m0 && m1) {
            d2v t = {res[0], res[1]};
            kg_st((d2v*)o, t);
        } else {
            if (m0) kg_st(o, res[0]);
            if (m1) kg_st(o + 1, res[1]);
        }
    }
}

// --------------------------------------------------------------------- uxx1
// uxx1/uxx1.c:95-105, interior [2,n-1)^3, dth = 1/nx (:70)
__global__ __launch_bounds__(TX * TY)
void uxx1_k(int nx, int ny, int ns, int ntx, int nty, double c1, double c2, double dth,
            const double* __restrict__ u0, double* __restrict__ u1,
            const double* __restrict__ d1, const double* __restrict__ xx,
            const double* __restrict__ xy, const double* __restrict__ xz)
{
    kg_ztile t;
    if (!kg_ztile_get<TX, TY, KC, 2, 1>(nx, ny, ns, ntx, nty, t)) return;
    const long long sz = (long long)nx * ny;
    const int sy = nx;
    long long o = kg_off3(t.i, t.j, t.k0, nx, sz);
    // z-columns kept in registers: d1 at (i,j) and (i,j-1), planes k-1..k;
    // xz at (i,j), planes k-2..k+1.
    double d_c_m = d1[o - sz], d_y_m = d1[o - sy - sz];
    double xz_m2 = xz[o - 2 * sz], xz_m1 = xz[o - sz], xz_c = xz[o];
#pragma unroll 2
    for (int k = t.k0; k < t.k1; k++) {
        const double d_c = d1[o], d_y = d1[o - sy];
        const double xz_p1 = xz[o + sz];
        const double d = 0.25 * (d_c + d_y + d_c_m + d_y_m);
        kg_st(&u1[o], u0[o] + (dth / d) * (
            c1 * (xx[o] - xx[o - 1]) +
            c2 * (xx[o + 1] - xx[o - 2]) +
            c1 * (xy[o] - xy[o - sy]) +
            c2 * (xy[o + sy] - xy[o - 2 * sy]) +
            c1 * (xz_c - xz_m1) +
            c2 * (xz_p1 - xz_m2)));
        d_c_m = d_c;
        d_y_m = d_y;
        xz_m2 = xz_m1;
        xz_m1 = xz_c;
        xz_c = xz_p1;
        o += sz;
    }
}

// Direct uxx1 (even nx, 16-byte aligned fields): a lane owns points x, x+1
// of row j and PZ planes; xx as the aligned pairs at x-2, x, x+2; xy at rows
// j-2 .. j+1; d1 at rows j-1, j of planes k0-1 .. k0+PZ-1; xz at planes
// k0-2 .. k0+PZ; u0 at the point.  Same expression as uxx1_k.
#ifndef KG_UXX_PZ
#define KG_UXX_PZ 4
#endif
#ifndef KG_UXX_ROWS
#define KG_UXX_ROWS 4
#endif
constexpr int UXX_PZ = KG_UXX_PZ, UXX_ROWS = KG_UXX_ROWS;

__global__ __launch_bounds__(64 * UXX_ROWS)
void uxx1_2_k(int nx, int ny, int ns, int ntx, int nty, double c1, double c2, double dth,
              const double* __restrict__ u0, double* __restrict__ u1,
              const double* __restrict__ d1, const double* __restrict__ xx,
              const double* __restrict__ xy, const double* __restrict__ xz)
{
    unsigned b = kg_tile_order<KG_ORD_UXX>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int bx = (int)(b % (unsigned)ntx);
    b /= (unsigned)ntx;
    const int by = (int)(b % (unsigned)nty);
    const int bz = (int)(b / (unsigned)nty);
    const int lane = (int)(threadIdx.x & 63), r = (int)(threadIdx.x >> 6);
    const int x = bx * 128 + 2 * lane;
    const int j = 2 + by * UXX_ROWS + r;
    const int k0 = 2 + bz * UXX_PZ;
    if (x >= nx || j >= ny - 1 || k0 >= ns - 1) return;
    const int xl = x >= 2 ? x - 2 : x;
    const int xr = x + 2 < nx ? x + 2 : x;
    const long long sz = (long long)nx * ny;
    const long long oj = (long long)j * nx;
    d2v uc[UXX_PZ], xa[UXX_PZ], xc[UXX_PZ], xb[UXX_PZ], y2[UXX_PZ], y1[UXX_PZ], y0[UXX_PZ],
        y3[UXX_PZ], dc[UXX_PZ + 1], dy[UXX_PZ + 1], z[UXX_PZ + 3];
#pragma unroll
    for (int p = 0; p < UXX_PZ; p++) {
        const long long o = (long long)min(k0 + p, ns - 2) * sz + oj;
        uc[p] = ld2(u0 + o + x);
        xa[p] = ld2(xx + o + xl);
        xc[p] = ld2(xx + o + x);
        xb[p] = ld2(xx + o + xr);
        y2[p] = ld2(xy + o - 2 * nx + x);
        y1[p] = ld2(xy + o - nx + x);
        y0[p] = ld2(xy + o + x);
        y3[p] = ld2(xy + o + nx + x);
    }
#pragma unroll
    for (int p = 0; p < UXX_PZ + 1; p++) {
        const long long o = (long long)min(k0 - 1 + p, ns - 2) * sz + oj + x;
        dc[p] = ld2(d1 + o);
        dy[p] = ld2(d1 + o - nx);
    }
#pragma unroll
    for (int p = 0; p < UXX_PZ + 3; p++) z[p] = ld2(xz + (long long)min(k0 - 2 + p, ns - 1) * sz + oj + x);
    const bool m0 = x >= 2 && x < nx - 1, m1 = x + 1 >= 2 && x + 1 < nx - 1;
#pragma unroll
    for (int p = 0; p < UXX_PZ; p++) {
        if (k0 + p >= ns - 1) break;
        double res[2];
#pragma unroll
        for (int v = 0; v < 2; v++) {
            // xx at x+v, x+v-1, x+v+1, x+v-2 from the pairs at x-2, x, x+2
            const double xx0 = v ? xc[p].y : xc[p].x;
            const double xxm1 = v ? xc[p].x : xa[p].y;
            const double xxp1 = v ? xb[p].x : xc[p].y;
            const double xxm2 = v ? xa[p].y : xa[p].x;
            const double d = 0.25 * ((v ? dc[p + 1].y : dc[p + 1].x) + (v ? dy[p + 1].y : dy[p + 1].x) +
                                     (v ? dc[p].y : dc[p].x) + (v ? dy[p].y : dy[p].x));
            res[v] = (v ? uc[p].y : uc[p].x) + (dth / d) * (
                c1 * (xx0 - xxm1) +
                c2 * (xxp1 - xxm2) +
                c1 * ((v ? y0[p].y : y0[p].x) - (v ? y1[p].y : y1[p].x)) +
                c2 * ((v ? y3[p].y : y3[p].x) - (v ? y2[p].y : y2[p].x)) +
                c1 * ((v ? z[p + 2].y : z[p + 2].x) - (v ? z[p + 1].y : z[p + 1].x)) +
                c2 * ((v ? z[p + 3].y : z[p + 3].x) - (v ? z[p].y : z[p].x)));
        }
        double* o = u1 + (long long)(k0 + p) * sz + oj + x;
        if (m0 && m1) {
            d2v t = {res[0], res[1]};
            kg_st((d2v*)o, t);
        } else {
            if (m0) kg_st(o, res[0]);
            if (m1) kg_st(o + 1, res[1]);
        }
    }
}

// ------------------------------------------------- plane-engine stencil ops
// Each op computes the VEC points of one lane at plane k (see kgp::View).

// gradient/gradient.c:95-97
template <class C>
struct GradientOp {
    const double* src;
    const double* side[1];
    double *ux, *uy, *uz;
    double alpha, beta, gamma;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long ob, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        double rx[C::VEC], ry[C::VEC], rz[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
            rx[v] = alpha * (kgp::xrow<C>(p, v, 1) - kgp::xrow<C>(p, v, -1));
            ry[v] = beta * (p.at(v, 0, 1, 0) - p.at(v, 0, -1, 0));
            rz[v] = gamma * (p.col(v, 1) - p.col(v, -1));
        }
        kgp::store<C::VEC>(ux, ob, x, rx, xm);
        kgp::store<C::VEC>(uy, ob, x, ry, xm);
        kgp::store<C::VEC>(uz, ob, x, rz, xm);
    }
};

// lapgsrb/lapgsrb.c:93-117
template <class C>
struct LapgsrbOp {
    const double* src;
    const double* side[1];
    double* dst;
    double c0, c1, c2, c3;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long ob, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        double r[C::VEC];
        // same-row x-neighbours (planes k-1 .. k+1): from registers in the
        // x-shift mode, else from LDS
        auto X = [&](int v, int dx, int dz) {
            if constexpr (C::XS) return p.xat(v, dx, dz);
            else return p.at(v, dx, 0, dz);
        };
#pragma unroll
        for (int v = 0; v < C::VEC; v++)
            r[v] = c0 * p.col(v, 0) +
                c1 * (
                    X(v, 1, 0) + X(v, -1, 0) +
                    p.at(v, 0, 1, 0) + p.at(v, 0, -1, 0) +
                    p.col(v, 1) + p.col(v, -1)) +
                c2 * (
                    p.at(v, 1, 1, 0) + X(v, 1, 1) + X(v, 1, -1) + p.at(v, 1, -1, 0) +
                    p.at(v, 0, 1, 1) + p.at(v, 0, 1, -1) + p.at(v, 0, -1, 1) + p.at(v, 0, -1, -1) +
                    p.at(v, -1, 1, 0) + X(v, -1, 1) + X(v, -1, -1) + p.at(v, -1, -1, 0)) +
                c3 * (
                    X(v, 2, 0) + X(v, -2, 0) +
                    p.at(v, 0, 2, 0) + p.at(v, 0, -2, 0) +
                    p.col(v, 2) + p.col(v, -2));
        kgp::store<C::VEC>(dst, ob, x, r, xm);
    }
};

// Tile shapes from sweeps at 512x256x256 (cold caches, 10-launch means), with
// the prefetch ring kept in flight (rotating slots, buffer loads): laplacian in
// chunks of 2 tile rows per XCD, W=12/D=2/KC=8 100 us (W=8 102, W=16 102,
// D=3 102, KC=16 105, KC=4 107; XCD-contiguous order W=8 106; two or four
// rows per wave, 64- and 32-point tiles with fewer y-halo rows, 101-107;
// z-chunks lengthened to 2-8 workgroups per CU, 14-64 planes, 102-108;
// 32-byte lanes, 256-point tiles with half the x-halo lines, 113-128 over
// W=8-16, D=2-3, KC=8-16);
// wave13pt
// W=12/D=3 141 us (W=16 148, W=14 169; z-chunks lengthened to 2 workgroups
// per CU 172); lapgsrb with 8-byte lanes, W=12/D=2, hardware order, z-chunks
// lengthened to about 2 workgroups per CU (126 planes here) 127-130 us (fixed
// 32 planes 136, 64 130, 256 154; 1 workgroup per CU 154, 4 128; W=16 157,
// W=14 183, D=3 179; chunked orders 136-142; 16-byte lanes, 130 VGPRs:
// 142-145).
//                 VEC W   H  LL LR ZL ZR D  NAUX KC
#ifndef KG_LAP_W
#define KG_LAP_W 12
#endif
#ifndef KG_LAP_D
#define KG_LAP_D 2
#endif
#ifndef KG_LAP_KC
#define KG_LAP_KC 8
#endif
#ifndef KG_LAP_XS
#define KG_LAP_XS 0
#endif
#ifndef KG_GRAD_XS
#define KG_GRAD_XS 0
#endif
#ifndef KG_WAVE_XS
#define KG_WAVE_XS 0
#endif
template <int V> using LapCfg = kgp::Cfg<V, KG_LAP_W, 1, 0, 0, 1, 1, KG_LAP_D, 0, KG_LAP_KC,
                                         kgp::F64, 64, false, V == 2 && KG_LAP_XS>;
// gradient: W=12/D=2/KC=8 in chunks of 2 tile rows, one workgroup (12 waves)
// per CU: 196 us (W=8 with 2 workgroups 201, 3 workgroups 208; W=12 with 2
// 213; W=16 with 1 214; KC=16 224, D=3 209, XCD-contiguous 219).
#ifndef KG_GRAD_W
#define KG_GRAD_W 12
#endif
#ifndef KG_GRAD_OCC
#define KG_GRAD_OCC 1
#endif
template <int V> using GradCfg = kgp::Cfg<V, KG_GRAD_W, 1, 0, 0, 1, 1, 2, 0, 8, kgp::F64, 64, false,
                                          V == 2 && KG_GRAD_XS>;
template <int V> using WaveCfg = kgp::Cfg<V, 12, 2, 0, 0, 2, 2, 3, 1, 32, kgp::F64, 64, false,
                                          V == 2 && KG_WAVE_XS>;
// lapgsrb: 16-byte lanes (even nx, 16-byte aligned fields) take every
// same-row x-neighbour, in planes k-1 .. k+1, from the neighbouring lanes'
// registers by DPP shifts (x-shift mode), which leaves LDS only the 12
// y-offset reads of the 25-point body (20 before: LDS busy 56 -> 37 % of CU
// cycles, LDS waits 19 -> 7 %).  Sweep at 512x256x256 (30-40 back-to-back
// launches, means): W=16/D=2 with z-chunks lengthened to about 2 workgroups
// per CU 120 us (W=12 134, W=8 137; D=3 134-170; fixed 32-plane chunks
// 129-137, 4 workgroups per CU 131-139); the same run timed the previous
// shape (8-byte lanes, every neighbour from LDS, W=12/D=2) at 136 us and
// 8-byte lanes in the x-shift mode at 149-178 us (twice the DPP moves per
// point), so odd or unaligned grids keep the previous shape.  Back-to-back
// launches of this kernel slow down after about a dozen (101-107 us, then
// 125-165 us, with the clock held down under the sustained load; laplacian
// stays at 98-100 us), so shapes are compared on their first ten launches
// with cold caches: chunks for one workgroup per CU (a single round of 252
// workgroups, each with 16 waves) 101-102 us, two per CU 106-107 us; tile
// orders in chunks of 2 or 4 tile rows alike, per-XCD regions 145-177 us.
// 32-byte lanes (4 points, two 16-byte accesses, half the DPP moves per
// point; 12 waves, 155 VGPRs: 16 waves spill) 122-128 us, bit-identical.
// The arithmetic costs little: timing copies (wrong values, same loads, LDS
// staging, barriers and stores) ran at 102 us with every x-neighbour taken
// from the lane's own registers instead of by DPP, 102-103 us with the
// y-neighbours from registers instead of LDS, and 98-99 us storing the
// centre value with no arithmetic at all, against 103-104 us.  The tile's
// data movement sets the time: it reads 1.67x its input (16 rows loaded per
// 12 computed, 2 x-halo lines per 8 of a 128-point row).
#ifndef KG_GSRB_W
#define KG_GSRB_W 16
#endif
#ifndef KG_GSRB_D
#define KG_GSRB_D 2
#endif
#ifndef KG_GSRB_KC
#define KG_GSRB_KC 32
#endif
#ifndef KG_GSRB_WGPC
#define KG_GSRB_WGPC 1
#endif
template <int V> using GsrbCfg = kgp::Cfg<V, (V == 2 ? KG_GSRB_W : 12), 2, 1, 1, 2, 2, KG_GSRB_D, 0,
                                          KG_GSRB_KC, kgp::F64, 64, false, V == 2>;
#ifndef KG_WAVE_WGPC
#define KG_WAVE_WGPC 0
#endif

// Launch Op on the plane engine; 16-byte lanes (VEC=2) need an even nx and
// 16-byte aligned fields (al).
// ORD: tile order (kg_tile_order).  WGPC > 0: z-chunks longer than C::KC
// planes when the launch would otherwise exceed about WGPC workgroups per
// compute unit (each workgroup then marches further and fills its pipeline
// fewer times); WGPC = 0 keeps C::KC.  OCC > 0: at most OCC resident
// workgroups per compute unit, enforced by reserving unused dynamic LDS.
template <template <int> class CfgT, template <class> class OpT, int LO, int HI, int ORD = 1,
          int WGPC = 0, int OCC = 0, class Fill>
static int launch_plane(kg_ctx* ctx, const char* name, int nx, int ny, int ns, bool al, Fill fill)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    if (int rc = kg_check_plane_bytes((long long)nx * ny, 8)) return rc;
    auto go = [&](auto vtag) -> int {
        constexpr int V = decltype(vtag)::value;
        using C = CfgT<V>;
        using Op = OpT<C>;
        Op op{};
        fill(op);
        int kc = C::KC;
        if (WGPC > 0) {
            const kgp::Grid g0 = kgp::make_grid<C>(nx, ny, ns, LO, HI);
            const long long xy = (long long)g0.ntx * g0.nty, want = (long long)WGPC * ctx->cus;
            if (!g0.empty && xy * g0.nkc > want) {
                const long long chunks = want > xy ? want / xy : 1;
                const long long nzi = ns - LO - HI;
                kc = (int)std::max<long long>(C::KC, (nzi + chunks - 1) / chunks);
            }
        }
        const kgp::Grid g = kgp::make_grid<C>(nx, ny, ns, LO, HI, LO, HI, kc);
        const size_t smem = kg_occupancy_lds(ctx, OCC, C::LDS_BYTES);
        auto kern = kgp::plane_kernel<C, LO, HI, Op, LO, HI, ORD>;
        kg_register_kernel(name, (const void*)kern);
        if (int rc = kg_launch_begin(ctx, name)) return rc;
        const unsigned blocks = ORD == kgp::KG_ORD_REGION
            ? kgp::region_blocks((unsigned)g.ntx * (unsigned)g.nty, (unsigned)g.nkc) : g.blocks;
        if (!g.empty)
            hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * C::W), smem, ctx->stream,
                               nx, ny, ns, g.ntx, g.nty, kc, op);
        return kg_launch_end(ctx);
    };
    if (nx % 2 == 0 && al) return go(std::integral_constant<int, 2>{});
    return go(std::integral_constant<int, 1>{});
}

}  // namespace

// ============================================================ entry points

#define KG_LAUNCH3(NAME, KERNEL, LO, HI, ...)                                              \
    do {                                                                                   \
        if (int rc_ = kg_check_ctx(ctx)) return rc_;                                       \
        if (int rc_ = kg_check_dims3(nx, ny, ns)) return rc_;                              \
        const kg_zgrid g = kg_zgrid_make(nx, ny, ns, LO, HI, TX, TY, KC);                  \
        kg_register_kernel(NAME, (const void*)KERNEL);                                     \
        if (int rc_ = kg_launch_begin(ctx, NAME)) return rc_;                              \
        if (!g.empty)                                                                      \
            hipLaunchKernelGGL(KERNEL, dim3(g.blocks), dim3(TX, TY), 0, ctx->stream,       \
                               nx, ny, ns, g.ntx, g.nty, __VA_ARGS__);                     \
        return kg_launch_end(ctx);                                                         \
    } while (0)

extern "C" int kg_laplacian(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                            const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_laplacian: null buffer");
    KG_REQUIRE(w0 != w1, "kg_laplacian: w0 and w1 must be distinct (out-of-place update)");
    return launch_plane<LapCfg, LaplacianOp, 1, 1, KG_ORD_LAP>(ctx, "laplacian", nx, ny, ns,
                                                   kg_aligned16(w0, w1), [&](auto& op) {
        op.src = w0;
        op.dst = w1;
        op.alpha = alpha;
        op.beta = beta;
    });
}

extern "C" int kg_divergence(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                             double gamma, double* u, const double* ux, const double* uy,
                             const double* uz)
{
    KG_REQUIRE(u && ux && uy && uz, "kg_divergence: null buffer");
    KG_REQUIRE(u != ux && u != uy && u != uz, "kg_divergence: output aliases an input");
    if (nx % 2 == 0 && !(((uintptr_t)u | (uintptr_t)ux | (uintptr_t)uy | (uintptr_t)uz) & 15)) {
        if (int rc = kg_check_ctx(ctx)) return rc;
        if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
        const kg_zgrid g = kg_zgrid_make(nx, ny, ns, 1, 1, 128, DIV_ROWS, DIV_PZ);
        kg_register_kernel("divergence", (const void*)divergence2_k);
        if (int rc = kg_launch_begin(ctx, "divergence")) return rc;
        if (!g.empty)
            hipLaunchKernelGGL(divergence2_k, dim3(g.blocks), dim3(64 * DIV_ROWS), kg_occupancy_lds(ctx, DIV_OCC, 0), ctx->stream,
                               nx, ny, ns, g.ntx, g.nty, alpha, beta, gamma, u, ux, uy, uz);
        return kg_launch_end(ctx);
    }
    KG_LAUNCH3("divergence", divergence_k, 1, 1, alpha, beta, gamma, u, ux, uy, uz);
}

extern "C" int kg_gradient(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                           double gamma, const double* u, double* ux, double* uy, double* uz)
{
    KG_REQUIRE(u && ux && uy && uz, "kg_gradient: null buffer");
    KG_REQUIRE(u != ux && u != uy && u != uz && ux != uy && ux != uz && uy != uz,
               "kg_gradient: buffers must be distinct");
    return launch_plane<GradCfg, GradientOp, 1, 1, KG_ORD_GRAD, 0, KG_GRAD_OCC>(ctx, "gradient", nx, ny, ns,
                                                         kg_aligned16(u, ux, uy, uz), [&](auto& op) {
        op.src = u;
        op.ux = ux;
        op.uy = uy;
        op.uz = uz;
        op.alpha = alpha;
        op.beta = beta;
        op.gamma = gamma;
    });
}

extern "C" int kg_wave13pt(kg_ctx* ctx, int nx, int ny, int ns, double m0, double m1, double m2,
                           const double* w0, const double* w1, double* w2)
{
    KG_REQUIRE(w0 && w1 && w2, "kg_wave13pt: null buffer");
    KG_REQUIRE(w2 != w0 && w2 != w1, "kg_wave13pt: w2 aliases an input");
    return launch_plane<WaveCfg, Wave13ptOp, 2, 2, KG_ORD_WAVE, KG_WAVE_WGPC>(ctx, "wave13pt", nx, ny, ns,
                                                   kg_aligned16(w0, w1, w2), [&](auto& op) {
        op.src = w1;
        op.side[0] = w0;
        op.dst = w2;
        op.m0 = m0;
        op.m1 = m1;
        op.m2 = m2;
    });
}

extern "C" int kg_uxx1(kg_ctx* ctx, int nx, int ny, int ns, double c1, double c2,
                       const double* u0, double* u1, const double* d1,
                       const double* xx, const double* xy, const double* xz)
{
    KG_REQUIRE(u0 && u1 && d1 && xx && xy && xz, "kg_uxx1: null buffer");
    KG_REQUIRE(u1 != u0 && u1 != d1 && u1 != xx && u1 != xy && u1 != xz,
               "kg_uxx1: u1 aliases an input");
    const double dth = 1. / nx;   // uxx1/uxx1.c:70
    if (nx % 2 == 0 && !(((uintptr_t)u0 | (uintptr_t)u1 | (uintptr_t)d1 | (uintptr_t)xx |
                          (uintptr_t)xy | (uintptr_t)xz) & 15)) {
        if (int rc = kg_check_ctx(ctx)) return rc;
        if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
        // rows/planes from 2: the grid helper counts tiles from 0, one spare
        const kg_zgrid g = kg_zgrid_make(nx, ny, ns, 2, 1, 128, UXX_ROWS, UXX_PZ);
        kg_register_kernel("uxx1", (const void*)uxx1_2_k);
        if (int rc = kg_launch_begin(ctx, "uxx1")) return rc;
        if (!g.empty)
            hipLaunchKernelGGL(uxx1_2_k, dim3(g.blocks), dim3(64 * UXX_ROWS), 0, ctx->stream,
                               nx, ny, ns, g.ntx, g.nty, c1, c2, dth, u0, u1, d1, xx, xy, xz);
        return kg_launch_end(ctx);
    }
    KG_LAUNCH3("uxx1", uxx1_k, 2, 1, c1, c2, dth, u0, u1, d1, xx, xy, xz);
}

extern "C" int kg_lapgsrb(kg_ctx* ctx, int nx, int ny, int ns, double c0, double c1, double c2,
                          double c3, const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_lapgsrb: null buffer");
    KG_REQUIRE(w0 != w1, "kg_lapgsrb: w0 and w1 must be distinct");
    return launch_plane<GsrbCfg, LapgsrbOp, 2, 2, KG_ORD_GSRB, KG_GSRB_WGPC>(ctx, "lapgsrb", nx, ny, ns,
                                                         kg_aligned16(w0, w1), [&](auto& op) {
        op.src = w0;
        op.dst = w1;
        op.c0 = c0;
        op.c1 = c1;
        op.c2 = c2;
        op.c3 = c3;
    });
}

int kg_preload_stencil3d() { return kg_preload_kernel((const void*)divergence_k); }
