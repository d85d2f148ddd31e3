// kg_ops3d.h -- stencil bodies that exist in several storage types: the FP64
// drop-in kernels, the FP32/BF16 precision variants and the batched/served
// grids all run these same bodies.  Each op computes the VEC points of one
// lane at plane k in the reference's source order, in the arithmetic type
// C::F, and stores them in the storage type C::S.  Operands come through a
// view: kgp::View (the plane engine: LDS and registers) or kgops::FlatView
// (global memory, for small grids in the batch); both give the same values,
// so the results are the same bit for bit.
#pragma once

#include "kg_plane.h"

namespace kgops {

// laplacian/laplacian.c:93-97 (interior [1,n-1)^3)
template <class C>
struct LaplacianOp {
    using S = typename C::S;
    using F = typename C::F;
    const S* src;
    const S* side[1];
    S* dst;
    F alpha, beta;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long ob, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        F r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++)
            r[v] = alpha * p.col(v, 0) + beta * (
                kgp::xrow<C>(p, v, 1) + kgp::xrow<C>(p, v, -1) +
                p.at(v, 0, 1, 0) + p.at(v, 0, -1, 0) +
                p.col(v, 1) + p.col(v, -1));
        p.put(dst, ob, x, r, xm);
    }
};

// wave13pt/wave13pt.c:640-651 (interior [2,n-2)^3); main field w1, side
// stream w0, output w2
template <class C>
struct Wave13ptOp {
    using S = typename C::S;
    using F = typename C::F;
    const S* src;
    const S* side[1];
    S* dst;
    F m0, m1, m2;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long ob, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        F r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++)
            r[v] = m0 * p.col(v, 0) - p.side(0, v) +
                m1 * (
                    kgp::xrow<C>(p, v, 1) + kgp::xrow<C>(p, v, -1) +
                    p.at(v, 0, 1, 0) + p.at(v, 0, -1, 0) +
                    p.col(v, 1) + p.col(v, -1)) +
                m2 * (
                    kgp::xrow<C>(p, v, 2) + kgp::xrow<C>(p, v, -2) +
                    p.at(v, 0, 2, 0) + p.at(v, 0, -2, 0) +
                    p.col(v, 2) + p.col(v, -2));
        p.put(dst, ob, x, r, xm);
    }
};

// gaussblur/gaussblur.c:85-92 applied to every xy-plane of a 3D grid:
// xy-interior [2,n-2), all planes.  f = 1/(s0 + 4(s1+s2+s4+s8) + 8 s5)
// (gaussblur/gaussblur.c:65) is computed by the caller.
template <class C>
struct Gaussblur3Op {
    using S = typename C::S;
    using F = typename C::F;
    const S* src;
    const S* side[1];
    S* dst;
    F f, s0, s1, s2, s4, s5, s8;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long ob, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        F r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
#define W(di, dj) p.at(v, di, dj, 0)
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
        p.put(dst, ob, x, r, xm);
    }
};

// jacobi/jacobi.F90:65-67 applied to every xy-plane (a 2D nx x ny problem
// is a grid with ns = 1): xy-interior [1,n-1), all planes.  Term order as the
// FP64 2D entry point kg_jacobi.
template <class C>
struct Jacobi2Op {
    using S = typename C::S;
    using F = typename C::F;
    const S* src;
    const S* side[1];
    S* dst;
    F c0, c1, c2;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long ob, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        F r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++)
            r[v] = c0 * p.at(v, 0, 0, 0) +
                   c1 * (p.at(v, -1, 0, 0) + p.at(v, 0, -1, 0) + p.at(v, 1, 0, 0) + p.at(v, 0, 1, 0)) +
                   c2 * (p.at(v, -1, -1, 0) + p.at(v, -1, 1, 0) + p.at(v, 1, -1, 0) + p.at(v, 1, 1, 0));
        p.put(dst, ob, x, r, xm);
    }
};

// gameoflife/gameoflife.c:82-91 applied to every xy-plane: xy-interior
// [1,n-1), all planes.  Same expression as the FP64 2D entry point.
template <class C>
struct Life2Op {
    using S = typename C::S;
    using F = typename C::F;
    const S* src;
    const S* side[1];
    S* dst;
    template <class V>
    __device__ __forceinline__ void run(const V& p, long long ob, int x,
                                        const bool (&xm)[C::VEC]) const
    {
        F r[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
#define U(di, dj) p.at(v, di, dj, 0)
            const F Cbig = (F)100000000000000000000.;
            const F L = U(-1, -1) + U(0, -1) + U(1, -1) +
                        U(-1, 0) + U(1, 0) +
                        U(-1, 1) + U(0, 1) + U(1, 1);
            r[v] = (F)1. / ((F)1. + (U(0, 0) + L - (F)3.) * (L - (F)3.) * Cbig);
#undef U
        }
        p.put(dst, ob, x, r, xm);
    }
};

// Interior bounds of each op: xy [LO, n-HI), z [ZLO, ns-ZHI).
template <template <class> class Op> struct Bounds;
template <> struct Bounds<LaplacianOp> { static constexpr int LO = 1, HI = 1, ZLO = 1, ZHI = 1; };
template <> struct Bounds<Wave13ptOp> { static constexpr int LO = 2, HI = 2, ZLO = 2, ZHI = 2; };
template <> struct Bounds<Gaussblur3Op> { static constexpr int LO = 2, HI = 2, ZLO = 0, ZHI = 0; };
template <> struct Bounds<Jacobi2Op> { static constexpr int LO = 1, HI = 1, ZLO = 0, ZHI = 0; };
template <> struct Bounds<Life2Op> { static constexpr int LO = 1, HI = 1, ZLO = 0, ZHI = 0; };

}  // namespace kgops

namespace kgops {

// Operands straight from global memory (VEC = 1): the point at offset o,
// neighbours at o + dx + dy*sy + dz*sz.  For grids small enough that their
// planes stay in the caches while a workgroup sweeps them.
template <class C>
struct FlatView {
    using S = typename C::S;
    using F = typename C::F;
    using T = typename C::T;
    const kgp::gmem<const S>* src;
    const kgp::gmem<const S>* aux[C::NA];
    long long o, sy, sz;
    __device__ __forceinline__ F at(int, int dx, int dy, int dz) const
    {
        return kgp::ld_cvt<T>(src[o + dx + dy * sy + dz * sz]);
    }
    __device__ __forceinline__ F col(int, int dz) const { return kgp::ld_cvt<T>(src[o + dz * sz]); }
    __device__ __forceinline__ F side(int a, int) const { return kgp::ld_cvt<T>(aux[a][o]); }
    template <class D>
    __device__ __forceinline__ void put(D* dst, long long ob, int x, F (&r)[1], const bool (&xm)[1]) const
    {
        kgp::store<T, 1>(dst, ob, x, r, xm);
    }
};

// Engine shape each op needs: halo width, LDS planes, register column and
// side streams (kgp::Cfg parameters H, LL, LR, ZL, ZR, NAUX).
template <template <class> class Op> struct Shape;
template <> struct Shape<LaplacianOp> { static constexpr int H = 1, LL = 0, LR = 0, ZL = 1, ZR = 1, NAUX = 0; };
template <> struct Shape<Wave13ptOp> { static constexpr int H = 2, LL = 0, LR = 0, ZL = 2, ZR = 2, NAUX = 1; };
template <> struct Shape<Gaussblur3Op> { static constexpr int H = 2, LL = 0, LR = 0, ZL = 0, ZR = 0, NAUX = 0; };
template <> struct Shape<Jacobi2Op> { static constexpr int H = 1, LL = 0, LR = 0, ZL = 0, ZR = 0, NAUX = 0; };
template <> struct Shape<Life2Op> { static constexpr int H = 1, LL = 0, LR = 0, ZL = 0, ZR = 0, NAUX = 0; };

// Engine configuration of op Op with storage type T, VEC points per lane, W
// waves, prefetch depth D, KC planes per single-grid tile, LPR lanes per row,
// XS the x-shift mode (LPR = 64 only).
template <template <class> class Op, class T, int VEC, int W, int D, int KC, int LPR = 64,
          bool XS = false>
using OpCfg = kgp::Cfg<VEC, W, Shape<Op>::H, Shape<Op>::LL, Shape<Op>::LR, Shape<Op>::ZL,
                       Shape<Op>::ZR, D, Shape<Op>::NAUX, KC, T, LPR, false, XS>;

}  // namespace kgops
