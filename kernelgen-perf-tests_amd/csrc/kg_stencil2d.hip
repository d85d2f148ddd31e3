// kg_stencil2d.hip -- FP64 2D stencils of the suite: jacobi, gaussblur,
// gameoflife, whispering (the reference runs them at nx x (ny*ns), e.g.
// 512 x 65536).  The single-field stencils (jacobi, gaussblur, gameoflife)
// run on the direct row march of kg_row.h (every neighbour loaded through
// the caches, a few output rows per lane); whispering, with five fields and
// three side streams, on the row-staged engine: x-neighbours and diagonal
// neighbours from LDS rows, the lane's own column from registers, every load
// prefetched a few rows ahead.  Arithmetic is in the reference's source order
// and the library is built with -ffp-contract=off.
#include "kg_row.h"

#include <type_traits>

namespace {

// jacobi/jacobi.F90:65-67, interior [1,n-1)^2 (Fortran 2..n-1)
template <class C>
struct JacobiOp {
    const double* src[1];
    const double* side[1];
    double* dst;
    double c0, c1, c2;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long o, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        double r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++)
            r[v] = c0 * p.at(0, v, 0, 0) +
                   c1 * (p.at(0, v, -1, 0) + p.at(0, v, 0, -1) + p.at(0, v, 1, 0) + p.at(0, v, 0, 1)) +
                   c2 * (p.at(0, v, -1, -1) + p.at(0, v, -1, 1) + p.at(0, v, 1, -1) + p.at(0, v, 1, 1));
        kgr::store<C::VEC>(dst, o, x, r, xm);
    }
};

// gaussblur/gaussblur.c:85-92, interior [2,n-2)^2; f (:65) computed on the host
template <class C>
struct GaussblurOp {
    const double* src[1];
    const double* side[1];
    double* dst;
    double f, s0, s1, s2, s4, s5, s8;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long o, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        double r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
#define W(di, dj) p.at(0, v, di, dj)
            r[v] = f * (
                s0 * W(0, 0) +
                s1 * (W(-1, 0) + W(1, 0) + W(0, -1) + W(0, 1)) +
                s2 * (W(-1, -1) + W(1, -1) + W(-1, 1) + W(1, 1)) +
                s4 * (W(-2, 0) + W(2, 0) + W(0, -2) + W(0, 2)) +
                s5 * (W(-2, -1) + W(-1, -2) + W(1, -2) + W(2, -1) +
                      W(-2, 1) + W(-1, 2) + W(1, 2) + W(2, 1)) +
                s8 * (W(-2, -2) + W(2, -2) + W(-2, 2) + W(2, 2)));
#undef W
        }
        kgr::store<C::VEC>(dst, o, x, r, xm);
    }
};

// gameoflife/gameoflife.c:82-91, interior [1,n-1)^2
template <class C>
struct GameoflifeOp {
    const double* src[1];
    const double* side[1];
    double* dst;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long o, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        double r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
#define U(di, dj) p.at(0, v, di, dj)
            const double Cbig = 100000000000000000000.;
            const double L =
                U(-1, -1) + U(0, -1) + U(1, -1) +
                U(-1, 0) + U(1, 0) +
                U(-1, 1) + U(0, 1) + U(1, 1);
            r[v] = 1. / (1. + (U(0, 0) + L - 3.) * (L - 3.) * Cbig);
#undef U
        }
        kgr::store<C::VEC>(dst, o, x, r, xm);
    }
};

// whispering/whispering.c:94-106 with the serial sweep's in-place e updates
// made explicit.  Visiting (i,j), the reference reads e(i,j-1,0) after the
// (i,j-1) update replaced it (when j-1 >= 1) and e(i-1,j,1) after the (i-1,j)
// update replaced it (when i-1 >= 1).  Those updates are recomputed here from
// e_in; their difference terms h0(i,j)-h0(i,j-1) and h0(i,j)-h0(i-1,j) are the
// ones ey and ex use.  e is planar: component c at e[i + nx*j + nx*ny*c].
// Fields: 0 h0, 1 ca, 2 cb, 3 e_in component 0, 4 e_in component 1;
// side streams: 0 da, 1 db, 2 u_em0.
template <class C>
struct WhisperingOp {
    const double* src[5];
    const double* side[3];
    double *eo0, *eo1, *h1, *u_em1;
    double mu, epsilon;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long o, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        double r0[C::VEC], r1[C::VEC], rh[C::VEC], ru[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
            const int i = x + v;
            const double h_c = p.col_(0, v, 0), h_p = p.col_(0, v, 1), h_m = p.col_(0, v, -1);
            const double h_xp = p.at(0, v, 1, 0), h_xm = p.at(0, v, -1, 0);
            const double ca_c = p.col_(1, v, 0), cb_c = p.col_(2, v, 0);
            const double e0_c = p.col_(3, v, 0), e1_c = p.col_(4, v, 0);
            const double ca_m = p.col_(1, v, -1), cb_m = p.col_(2, v, -1), e0_m = p.col_(3, v, -1);
            const double ca_x = p.at(1, v, -1, 0), cb_x = p.at(2, v, -1, 0), e1_x = p.at(4, v, -1, 0);
            const double e0 = ca_c * e0_c + cb_c * (h_p - h_c);
            const double e1 = ca_c * e1_c - cb_c * (h_xp - h_c);
            const double dy = h_c - h_m;
            const double dx = h_c - h_xm;
            const double ey_src = (p.j - 1 >= 1) ? ca_m * e0_m + cb_m * dy : e0_m;
            const double ex_src = (i - 1 >= 1) ? ca_x * e1_x - cb_x * dx : e1_x;
            const double ey = ca_m * ey_src + cb_m * dy;
            const double ex = ca_x * ex_src - cb_x * dx;
            const double h = p.side(0, v) * h_c + p.side(1, v) * (e0 - ey + ex - e1);
            r0[v] = e0;
            r1[v] = e1;
            rh[v] = h;
            ru[v] = p.side(2, v) + 0.5 * (h * h / mu + epsilon * (e0 * e0 + e1 * e1));
        }
        kgr::store<C::VEC>(eo0, o, x, r0, xm);
        kgr::store<C::VEC>(eo1, o, x, r1, xm);
        kgr::store<C::VEC>(h1, o, x, rh, xm);
        kgr::store<C::VEC>(u_em1, o, x, ru, xm);
    }
};

// Row shapes from sweeps at 512x65536 (cold caches, 40-launch means) with
// the prefetch rings kept in flight: jacobi/gameoflife D=2/YC=64 113 us (D=4
// 141, W=8 118); gaussblur D=2/YC=32 116 us (D=3/YC=64 140); whispering with
// 16-byte lanes, W=4/D=2/YC=8 589 us (8-byte lanes 610-643; later, first ten
// launches of each: YC=4 584 us, 6 616, 12 644; D=1 633, D=3 584; hardware
// and chunked tile orders 606 and 586, against 585-588 for this shape).
// Whispering takes 258 registers, one wave per SIMD; held to 256 (two waves
// per SIMD, 3 values spilled) it ran slower: 614-618 us against 572-573
// (YC=6 615-628, 12 631-634, 16 660; D=1 605-622).
//                     VEC W      H  NF LL LR ZL ZR D  NAUX YC
template <int V> using J9Cfg = kgr::Cfg<V, 4, 1, 1, 1, 1, 0, 0, 2, 0, 64>;
template <int V> using G25Cfg = kgr::Cfg<V, 4, 2, 1, 2, 2, 0, 0, 2, 0, 32>;
template <int V> using WhCfg = kgr::Cfg<V, 8 / V, 1, 5, 0, 0, 1, 1, 2, 3, 8>;

// Launch Op on the row engine; 16-byte lanes (VEC=2) need an even nx and
// 16-byte aligned fields (al).
template <template <int> class CfgT, template <class> class OpT, int LO, int HI, int VEC_EVEN,
          int ORD = 1, class Fill>
static int launch_row(kg_ctx* ctx, const char* name, int nx, int ny, bool al, Fill fill)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims2(nx, ny)) return rc;
    auto go = [&](auto vtag) -> int {
        constexpr int V = decltype(vtag)::value;
        using C = CfgT<V>;
        using Op = OpT<C>;
        Op op{};
        fill(op);
        const kgr::Grid g = kgr::make_grid<C>(nx, ny, LO, HI);
        auto kern = kgr::row_kernel<C, LO, HI, Op, ORD>;
        kg_register_kernel(name, (const void*)kern);
        if (int rc = kg_launch_begin(ctx, name)) return rc;
        if (!g.empty)
            hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(64 * C::W), 0, ctx->stream, nx, ny,
                               g.ntx, op);
        return kg_launch_end(ctx);
    };
    if (VEC_EVEN == 2 && nx % 2 == 0 && al) return go(std::integral_constant<int, VEC_EVEN>{});
    return go(std::integral_constant<int, 1>{});
}

// Direct march shapes (output rows per lane, waves per workgroup) from
// sweeps at 512x65536 (cold caches, against the row engine's 113-117 us),
// in chunked tile order: jacobi PY=8/4 waves 102 us (PY=4: 103, 8 waves
// 104), gameoflife PY=4/4 waves 101 us (PY=8: 107); gaussblur PY=8/4 waves
// 95 us (PY=4/8 waves 111, PY=8/8 waves 106).  Rechecked on the final
// kernels: jacobi PY=12 112 us, PY=16 105, 2 or 8 waves 101-103 against
// 100; gameoflife PY=8 104, 2 or 8 waves 102-104 against 99; gaussblur
// PY=12 99-100 against 94.
#ifndef KG_J9_PY
#define KG_J9_PY 4
#endif
#ifndef KG_JAC_PY
#define KG_JAC_PY 8
#endif
#ifndef KG_J9_ROWS
#define KG_J9_ROWS 4
#endif
#ifndef KG_G25_PY
#define KG_G25_PY 8
#endif
#ifndef KG_G25_ROWS
#define KG_G25_ROWS 4
#endif
// Tile order of the direct kernels (kg_tile_order), measured at 512x65536.
#ifndef KG_ORD_JAC
#define KG_ORD_JAC 4
#endif
#ifndef KG_ORD_GAUSS
#define KG_ORD_GAUSS 4
#endif
#ifndef KG_ORD_LIFE
#define KG_ORD_LIFE 4
#endif
#ifndef KG_ORD_WH
#define KG_ORD_WH 1
#endif
using J9D = kgr::DCfg<KG_J9_PY, KG_J9_ROWS, 1>;
using JacD = kgr::DCfg<KG_JAC_PY, KG_J9_ROWS, 1>;
using G25D = kgr::DCfg<KG_G25_PY, KG_G25_ROWS, 2>;

// Single-field stencils: the direct march when nx is even and both fields
// are 16-byte aligned, else the row engine.
template <class DC, template <int> class CfgT, template <class> class OpT, int LO, int HI,
          int ORD, class Fill>
static int launch_direct(kg_ctx* ctx, const char* name, int nx, int ny, const void* in,
                         const void* out, Fill fill)
{
    const bool al = kg_aligned16(in, out);
    if (nx % 2 != 0 || !al) return launch_row<CfgT, OpT, LO, HI, 2>(ctx, name, nx, ny, al, fill);
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims2(nx, ny)) return rc;
    using Op = OpT<DC>;
    Op op{};
    fill(op);
    const kgr::DGrid<DC> g = kgr::make_dgrid<DC>(nx, ny, LO, HI);
    auto kern = kgr::direct_kernel<DC, LO, HI, Op, ORD>;
    kg_register_kernel(name, (const void*)kern);
    if (int rc = kg_launch_begin(ctx, name)) return rc;
    if (!g.empty)
        hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(64 * DC::ROWS), 0, ctx->stream, nx, ny,
                           g.ntx, op);
    return kg_launch_end(ctx);
}

#define KG_REQUIRE(cond, msg) \
    do { if (!(cond)) return kg_set_error(KG_ERR_INVALID, "%s", msg); } while (0)

}  // namespace

extern "C" int kg_jacobi(kg_ctx* ctx, int nx, int ny, double c0, double c1, double c2,
                         const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_jacobi: null buffer");
    KG_REQUIRE(w0 != w1, "kg_jacobi: w0 and w1 must be distinct");
    return launch_direct<JacD, J9Cfg, JacobiOp, 1, 1, KG_ORD_JAC>(ctx, "jacobi_", nx, ny, w0, w1, [&](auto& op) {
        op.src[0] = w0;
        op.dst = w1;
        op.c0 = c0;
        op.c1 = c1;
        op.c2 = c2;
    });
}

extern "C" int kg_gaussblur(kg_ctx* ctx, int nx, int ny, double s0, double s1, double s2,
                            double s4, double s5, double s8, const double* w0, double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_gaussblur: null buffer");
    KG_REQUIRE(w0 != w1, "kg_gaussblur: w0 and w1 must be distinct");
    const double f = 1. / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5);   // gaussblur/gaussblur.c:65
    return launch_direct<G25D, G25Cfg, GaussblurOp, 2, 2, KG_ORD_GAUSS>(ctx, "gaussblur", nx, ny, w0, w1, [&](auto& op) {
        op.src[0] = w0;
        op.dst = w1;
        op.f = f;
        op.s0 = s0;
        op.s1 = s1;
        op.s2 = s2;
        op.s4 = s4;
        op.s5 = s5;
        op.s8 = s8;
    });
}

extern "C" int kg_gameoflife(kg_ctx* ctx, int nx, int ny, const double* u0, double* u1)
{
    KG_REQUIRE(u0 && u1, "kg_gameoflife: null buffer");
    KG_REQUIRE(u0 != u1, "kg_gameoflife: u0 and u1 must be distinct");
    return launch_direct<J9D, J9Cfg, GameoflifeOp, 1, 1, KG_ORD_LIFE>(ctx, "gameoflife", nx, ny, u0, u1, [&](auto& op) {
        op.src[0] = u0;
        op.dst = u1;
    });
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
    const long long n = (long long)nx * ny;
    const bool al = kg_aligned16(e_in, e_out, h0, h1, u_em0, u_em1, ca, cb, da, db) &&
                    kg_aligned16(e_in + n, e_out + n);
    return launch_row<WhCfg, WhisperingOp, 1, 1, 2, KG_ORD_WH>(ctx, "whispering", nx, ny, al, [&](auto& op) {
        op.src[0] = h0;
        op.src[1] = ca;
        op.src[2] = cb;
        op.src[3] = e_in;
        op.src[4] = e_in + n;
        op.side[0] = da;
        op.side[1] = db;
        op.side[2] = u_em0;
        op.eo0 = e_out;
        op.eo1 = e_out + n;
        op.h1 = h1;
        op.u_em1 = u_em1;
        op.mu = mu;
        op.epsilon = epsilon;
    });
}

int kg_preload_stencil2d()
{
    return kg_preload_kernel((const void*)kgr::row_kernel<J9Cfg<2>, 1, 1, JacobiOp<J9Cfg<2>>>);
}
