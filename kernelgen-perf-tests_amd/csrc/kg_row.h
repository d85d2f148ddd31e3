// kg_row.h -- row-staged y-marching engine for the 2D stencils.
//
// The 2D counterpart of kg_plane.h.  A workgroup of W wave64s covers a tile
// TXW = W*64*VEC points wide (VEC = 2: two consecutive x points per lane, one
// 16-byte load) and marches it along y through a chunk of YC rows:
//
//   * NF fields are read with neighbours.  For each, the rows j-LL .. j+LR sit
//     in LDS (LL+LR+2 row slots, one barrier per row) with an H-point x-halo,
//     and the lane's own columns of rows j-ZL .. j+ZR sit in registers;
//   * every global load is issued D rows ahead of use (a register ring per
//     field, plus a ring for NAUX side streams read only at the point), so the
//     loop's vector-memory queue holds only this prefetch stream and the
//     output stores.
#pragma once

#include "kg_internal.h"

namespace kgr {

typedef double d2 __attribute__((ext_vector_type(2)));

template <int VEC_, int W_, int H_, int NF_, int LL_, int LR_, int ZL_, int ZR_, int D_, int NAUX_,
          int YC_>
struct Cfg {
    static constexpr int VEC = VEC_;   // points per lane along x (1 or 2)
    static constexpr int W = W_;       // waves side by side along x
    static constexpr int H = H_;       // x-halo width
    static constexpr int NF = NF_;     // fields staged with neighbours
    static constexpr int LL = LL_;     // LDS rows below j
    static constexpr int LR = LR_;     // LDS rows above j
    static constexpr int ZL = ZL_;     // register rows below j
    static constexpr int ZR = ZR_;     // register rows above j
    static constexpr int D = D_;       // prefetch distance (rows)
    static constexpr int NAUX = NAUX_; // side streams read at the point only
    static constexpr int YC = YC_;     // rows per workgroup
    static constexpr int TXW = W * 64 * VEC;
    static constexpr int LW = TXW + 2 * H;
    static constexpr int NLS = LL + LR + 2;
    static constexpr int ZHI = (ZR > LR ? ZR : LR);
    static constexpr int RS = ZL + ZHI + D;          // ring rows j-ZL .. j+ZHI+D-1
    static constexpr int NA = NAUX > 0 ? NAUX : 1;
};

template <int VEC>
struct Slot {
    double v[VEC];
    double h;
};

template <int VEC, int NA>
struct ASlot {
    double a[NA][VEC];
};

struct Grid {
    int ntx, nyc;
    unsigned blocks;
    bool empty;
};

template <class C>
static inline Grid make_grid(int nx, int ny, int lo, int hi)
{
    Grid g;
    const int nyi = ny - hi - lo, nxi = nx - hi - lo;
    g.empty = nxi <= 0 || nyi <= 0;
    g.ntx = (nx + C::TXW - 1) / C::TXW;
    g.nyc = g.empty ? 0 : (nyi + C::YC - 1) / C::YC;
    g.blocks = (unsigned)g.ntx * (unsigned)g.nyc;
    return g;
}

template <class C>
struct View {
    const double* lds;                        // [NF][NLS][LW]
    const Slot<C::VEC> (*ring)[C::RS];        // [NF][RS]
    const ASlot<C::VEC, C::NA>* aux;          // aux[0] = side streams at row j
    int j, col;                               // col: LDS column of the lane's first point
    // field f at (x + v + dx, j + dy), |dx| <= H, -LL <= dy <= LR
    __device__ __forceinline__ double at(int f, int v, int dx, int dy) const
    {
        return lds[(f * C::NLS + (j + dy) % C::NLS) * C::LW + col + v + dx];
    }
    // field f at (x + v, j + dy), -ZL <= dy <= ZHI
    __device__ __forceinline__ double col_(int f, int v, int dy) const { return ring[f][C::ZL + dy].v[v]; }
    __device__ __forceinline__ double side(int a, int v) const { return aux[0].a[a][v]; }
};

// Op supplies src[NF], side[NAUX] and run(view, o, x, xmask) which computes
// and stores the lane's VEC points of row j (o = j*nx).
template <class C, int LO, int HI, class Op>
__global__ __launch_bounds__(64 * C::W)
void row_kernel(int nx, int ny, int ntx, Op op)
{
    __shared__ double lds[C::NF * C::NLS * C::LW];
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned b = kg_xcd_swizzle(blockIdx.x, gridDim.x);
    const int bx = (int)(b % (unsigned)ntx);
    const int by = (int)(b / (unsigned)ntx);
    const int x0 = bx * C::TXW;
    const int col = C::H + w * 64 * C::VEC + lane * C::VEC;
    const int x = x0 + col - C::H;
    const int j0 = LO + by * C::YC;
    const int j1 = min(j0 + C::YC, ny - HI);
    int hx = -1;
    if (w == 0 && lane < C::H) hx = x0 - C::H + lane;
    else if (w == C::W - 1 && lane >= 64 - C::H) hx = x0 + C::TXW + (lane - (64 - C::H));
    const int hcol = hx < 0 ? -1 : hx - x0 + C::H;
    const bool h_ok = hx >= 0 && hx < nx;
    bool p_ok[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) p_ok[v] = x + v < nx;
    const bool vec_ok = C::VEC == 2 && p_ok[0] && p_ok[C::VEC - 1];

    auto load_row = [&](const double* base, int row, double (&v)[C::VEC]) {
        const double* pp = base + (long long)row * nx;
        if (C::VEC == 2 && vec_ok) {
            const d2 t = *(const d2*)(pp + x);
            v[0] = t.x;
            v[C::VEC - 1] = t.y;
        } else {
#pragma unroll
            for (int q = 0; q < C::VEC; q++) v[q] = p_ok[q] ? pp[x + q] : 0.0;
        }
    };
    auto load = [&](int f, int row, Slot<C::VEC>& s) {
        if (row < 0 || row >= ny) return;
        load_row(op.src[f], row, s.v);
        s.h = h_ok ? op.src[f][(long long)row * nx + hx] : 0.0;
    };
    auto load_aux = [&](int row, ASlot<C::VEC, C::NA>& s) {
        if (C::NAUX == 0 || row >= ny) return;
#pragma unroll
        for (int a = 0; a < C::NAUX; a++) load_row(op.side[a], row, s.a[a]);
    };
    auto stage = [&](int f, int row, const Slot<C::VEC>& s) {
        double* L = lds + (f * C::NLS + row % C::NLS) * C::LW;
#pragma unroll
        for (int v = 0; v < C::VEC; v++) L[col + v] = s.v[v];
        if (hcol >= 0) L[hcol] = s.h;
    };

    Slot<C::VEC> ring[C::NF][C::RS];
    ASlot<C::VEC, C::NA> aux[C::D];
#pragma unroll
    for (int f = 0; f < C::NF; f++)
#pragma unroll
        for (int r = 0; r < C::RS; r++) {
#pragma unroll
            for (int v = 0; v < C::VEC; v++) ring[f][r].v[v] = 0.0;
            ring[f][r].h = 0.0;
        }
#pragma unroll
    for (int f = 0; f < C::NF; f++)
#pragma unroll
        for (int r = 0; r < C::RS; r++) load(f, j0 - C::ZL + r, ring[f][r]);
#pragma unroll
    for (int d = 0; d < C::D; d++) load_aux(j0 + d, aux[d]);
#pragma unroll
    for (int dy = -C::LL; dy < C::LR; dy++)
#pragma unroll
        for (int f = 0; f < C::NF; f++) {
            if (dy >= -C::ZL) {
                stage(f, j0 + dy, ring[f][C::ZL + dy]);
            } else {
                Slot<C::VEC> t;
                t.h = 0.0;
#pragma unroll
                for (int v = 0; v < C::VEC; v++) t.v[v] = 0.0;
                load(f, j0 + dy, t);
                stage(f, j0 + dy, t);
            }
        }
    bool xmask[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) xmask[v] = x + v >= LO && x + v < nx - HI;

    for (int j = j0; j < j1; j++) {
#pragma unroll
        for (int f = 0; f < C::NF; f++) stage(f, j + C::LR, ring[f][C::ZL + C::LR]);
        Slot<C::VEC> nq[C::NF];
        const int jp = j + C::ZHI + C::D;
#pragma unroll
        for (int f = 0; f < C::NF; f++) {
            nq[f].h = 0.0;
#pragma unroll
            for (int v = 0; v < C::VEC; v++) nq[f].v[v] = 0.0;
            if (jp <= j1 - 1 + C::ZHI) load(f, jp, nq[f]);
        }
        ASlot<C::VEC, C::NA> na;
        if (j + C::D < j1) load_aux(j + C::D, na);
        __syncthreads();
        View<C> vw{lds, (const Slot<C::VEC>(*)[C::RS])ring, aux, j, col};
        op.run(vw, (long long)j * nx, x, xmask);
#pragma unroll
        for (int f = 0; f < C::NF; f++) {
#pragma unroll
            for (int r = 0; r < C::RS - 1; r++) ring[f][r] = ring[f][r + 1];
            ring[f][C::RS - 1] = nq[f];
        }
#pragma unroll
        for (int d = 0; d < C::D - 1; d++) aux[d] = aux[d + 1];
        aux[C::D - 1] = na;
    }
}

template <int VEC>
__device__ __forceinline__ void store(double* dst, long long o, int x, const double (&r)[VEC],
                                      const bool (&xmask)[VEC])
{
    if (VEC == 2 && xmask[0] && xmask[VEC - 1]) {
        d2 t;
        t.x = r[0];
        t.y = r[VEC - 1];
        kg_st((d2*)(dst + o + x), t);
    } else {
#pragma unroll
        for (int v = 0; v < VEC; v++)
            if (xmask[v]) kg_st(dst + o + x + v, r[v]);
    }
}

}  // namespace kgr
