// kg_stencil3d.hip -- FP64 3D stencils of the suite: laplacian, divergence,
// gradient, wave13pt, uxx1, lapgsrb.
//
// laplacian, gradient, wave13pt and lapgsrb run on the plane-staged engine
// (kg_plane.h): xy-neighbours from LDS, z-neighbours from a register column,
// every global load prefetched a few planes ahead.  divergence and uxx1 read
// each of their inputs at only a few offsets along one axis and use the plain
// z-march (kg_zmarch.h): a TX x TY column tile marched through KC planes.
// Arithmetic is written in the reference's source order and the library is
// built with -ffp-contract=off: results are bit-identical to an IEEE build of
// the reference kernels.
#include "kg_ops3d.h"
#include "kg_zmarch.h"

#include <type_traits>

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
            rx[v] = alpha * (p.at(v, 1, 0, 0) - p.at(v, -1, 0, 0));
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
#pragma unroll
        for (int v = 0; v < C::VEC; v++)
            r[v] = c0 * p.col(v, 0) +
                c1 * (
                    p.at(v, 1, 0, 0) + p.at(v, -1, 0, 0) +
                    p.at(v, 0, 1, 0) + p.at(v, 0, -1, 0) +
                    p.col(v, 1) + p.col(v, -1)) +
                c2 * (
                    p.at(v, 1, 1, 0) + p.at(v, 1, 0, 1) + p.at(v, 1, 0, -1) + p.at(v, 1, -1, 0) +
                    p.at(v, 0, 1, 1) + p.at(v, 0, 1, -1) + p.at(v, 0, -1, 1) + p.at(v, 0, -1, -1) +
                    p.at(v, -1, 1, 0) + p.at(v, -1, 0, 1) + p.at(v, -1, 0, -1) + p.at(v, -1, -1, 0)) +
                c3 * (
                    p.at(v, 2, 0, 0) + p.at(v, -2, 0, 0) +
                    p.at(v, 0, 2, 0) + p.at(v, 0, -2, 0) +
                    p.col(v, 2) + p.col(v, -2));
        kgp::store<C::VEC>(dst, ob, x, r, xm);
    }
};

// Tile shapes from sweeps at 512x256x256 (cold caches, 40-launch means), with
// the prefetch ring kept in flight (rotating slots, buffer loads): laplacian
// W=8/D=2/KC=8 101 us (KC=16 104, W=16 113); wave13pt W=12/D=3 141 us (W=16
// 148, W=14 169); lapgsrb with 8-byte lanes, W=12/D=2 137 us (W=16 157; with
// 16-byte lanes the 25-point body needs too many registers: 307 us).
//                 VEC W   H  LL LR ZL ZR D  NAUX KC
template <int V> using LapCfg = kgp::Cfg<V, 8, 1, 0, 0, 1, 1, 2, 0, 8>;
template <int V> using WaveCfg = kgp::Cfg<V, 12, 2, 0, 0, 2, 2, 3, 1, 32>;
template <int V> using GsrbCfg = kgp::Cfg<1, 12, 2, 1, 1, 2, 2, 2, 0, 32>;

// Launch Op on the plane engine; 16-byte lanes (VEC=2) need an even nx.
// SWZ: XCD-contiguous tile order (measured per kernel at 512x256x256).
template <template <int> class CfgT, template <class> class OpT, int LO, int HI, bool SWZ = true,
          class Fill>
static int launch_plane(kg_ctx* ctx, const char* name, int nx, int ny, int ns, Fill fill)
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
        const kgp::Grid g = kgp::make_grid<C>(nx, ny, ns, LO, HI);
        auto kern = kgp::plane_kernel<C, LO, HI, Op, LO, HI, SWZ>;
        kg_register_kernel(name, (const void*)kern);
        if (int rc = kg_launch_begin(ctx, name)) return rc;
        if (!g.empty)
            hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(64 * C::W), 0, ctx->stream,
                               nx, ny, ns, g.ntx, g.nty, op);
        return kg_launch_end(ctx);
    };
    if (nx % 2 == 0) return go(std::integral_constant<int, 2>{});
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
    return launch_plane<LapCfg, LaplacianOp, 1, 1>(ctx, "laplacian", nx, ny, ns, [&](auto& op) {
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
    KG_LAUNCH3("divergence", divergence_k, 1, 1, alpha, beta, gamma, u, ux, uy, uz);
}

extern "C" int kg_gradient(kg_ctx* ctx, int nx, int ny, int ns, double alpha, double beta,
                           double gamma, const double* u, double* ux, double* uy, double* uz)
{
    KG_REQUIRE(u && ux && uy && uz, "kg_gradient: null buffer");
    KG_REQUIRE(u != ux && u != uy && u != uz && ux != uy && ux != uz && uy != uz,
               "kg_gradient: buffers must be distinct");
    return launch_plane<LapCfg, GradientOp, 1, 1, false>(ctx, "gradient", nx, ny, ns, [&](auto& op) {
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
    return launch_plane<WaveCfg, Wave13ptOp, 2, 2>(ctx, "wave13pt", nx, ny, ns, [&](auto& op) {
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
    KG_LAUNCH3("uxx1", uxx1_k, 2, 1, c1, c2, dth, u0, u1, d1, xx, xy, xz);
}

extern "C" int kg_lapgsrb(kg_ctx* ctx, int nx, int ny, int ns, double c0, double c1, double c2,
                          double c3, const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_lapgsrb: null buffer");
    KG_REQUIRE(w0 != w1, "kg_lapgsrb: w0 and w1 must be distinct");
    return launch_plane<GsrbCfg, LapgsrbOp, 2, 2, false>(ctx, "lapgsrb", nx, ny, ns, [&](auto& op) {
        op.src = w0;
        op.dst = w1;
        op.c0 = c0;
        op.c1 = c1;
        op.c2 = c2;
        op.c3 = c3;
    });
}

int kg_preload_stencil3d() { return kg_preload_kernel((const void*)divergence_k); }
