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
//     output stores.  As in kg_plane.h the rings rotate by renaming (the row
//     loop is unrolled RS+1 times) and loads go through buffer resources with
//     out-of-range offsets for idle lanes, so no load sits in conditional code
//     and the compiler waits only for the row it is about to use.
#pragma once

#include "kg_plane.h"

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
    static constexpr int NP = RS + 1;                // physical slots (one being filled)
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

// View at ring phase P: logical ring row r (row j - ZL + r) of field f is
// physical slot (P + r) % NP; logical side-stream row d (row j + d) is
// physical slot (P + d) % NP.
template <class C, int P = 0>
struct View {
    const double* lds;                        // [NF][NLS][LW]
    const Slot<C::VEC> (*ring)[C::NP];        // [NF][NP]
    const ASlot<C::VEC, C::NA>* aux;          // [NP]
    int j, col;                               // col: LDS column of the lane's first point
    // field f at (x + v + dx, j + dy), |dx| <= H, -LL <= dy <= LR
    __device__ __forceinline__ double at(int f, int v, int dx, int dy) const
    {
        return lds[(f * C::NLS + (j + dy) % C::NLS) * C::LW + col + v + dx];
    }
    // field f at (x + v, j + dy), -ZL <= dy <= ZHI
    __device__ __forceinline__ double col_(int f, int v, int dy) const
    {
        return ring[f][(P + C::ZL + dy) % C::NP].v[v];
    }
    __device__ __forceinline__ double side(int a, int v) const { return aux[P % C::NP].a[a][v]; }
};

// Op supplies src[NF], side[NAUX] and run(view, o, x, xmask) which computes
// and stores the lane's VEC points of row j (o = j*nx).
// ORD: tile order (kg_tile_order).
template <class C, int LO, int HI, class Op, int ORD = 1>
__global__ __launch_bounds__(64 * C::W)
void row_kernel(int nx, int ny, int ntx, Op op)
{
    __shared__ double lds[C::NF * C::NLS * C::LW];
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const unsigned b = kg_tile_order<ORD>(blockIdx.x, gridDim.x, (unsigned)ntx);
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
    // Per-lane byte offsets within a row (kgp::KG_OOB: nothing to load).  With
    // VEC = 2 callers guarantee an even nx, so a lane with x < nx may read two.
    const unsigned off_v = x < nx ? (unsigned)x * 8u : kgp::KG_OOB;
    const unsigned off_h = h_ok ? (unsigned)hx * 8u : kgp::KG_OOB;
    const unsigned row_bytes = (unsigned)nx * 8u;
    // Resource for row `row` of array base; empty unless want and 0 <= row < ny.
    auto row_rsrc = [&](const double* base, int row, bool want) {
        const bool ok = want && row >= 0 && row < ny;
        return kgp::make_rsrc(base + (long long)(ok ? row : 0) * nx, ok ? row_bytes : 0u);
    };
    auto load = [&](int f, int row, Slot<C::VEC>& s, bool want = true) {
        const __amdgpu_buffer_rsrc_t r = row_rsrc(op.src[f], row, want);
        kgp::bload<kgp::F64, C::VEC>(r, off_v, s.v);
        double h[1];
        kgp::bload<kgp::F64, 1>(r, off_h, h);
        s.h = h[0];
    };
    auto load_aux = [&](int row, ASlot<C::VEC, C::NA>& s, bool want = true) {
#pragma unroll
        for (int a = 0; a < C::NAUX; a++)
            kgp::bload<kgp::F64, C::VEC>(row_rsrc(op.side[a], row, want), off_v, s.a[a]);
    };
    auto stage = [&](int f, int row, const Slot<C::VEC>& s) {
        double* L = lds + (f * C::NLS + row % C::NLS) * C::LW;
#pragma unroll
        for (int v = 0; v < C::VEC; v++) L[col + v] = s.v[v];
        if (hcol >= 0) L[hcol] = s.h;
    };

    Slot<C::VEC> ring[C::NF][C::NP];
    ASlot<C::VEC, C::NA> aux[C::NP];
#pragma unroll
    for (int f = 0; f < C::NF; f++)
#pragma unroll
        for (int r = 0; r < C::NP; r++) {
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
                load(f, j0 + dy, t);
                stage(f, j0 + dy, t);
            }
        }
    bool xmask[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) xmask[v] = x + v >= LO && x + v < nx - HI;

    // Row j at phase P (rows the ring and LDS need are all < ny).
    auto row_step = [&](auto phase, int j) {
        constexpr int P = decltype(phase)::value;
#pragma unroll
        for (int f = 0; f < C::NF; f++) stage(f, j + C::LR, ring[f][(P + C::ZL + C::LR) % C::NP]);
        const int jp = j + C::ZHI + C::D;
#pragma unroll
        for (int f = 0; f < C::NF; f++) load(f, jp, ring[f][(P + C::RS) % C::NP], jp <= j1 - 1 + C::ZHI);
        load_aux(j + C::D, aux[(P + C::D) % C::NP], j + C::D < j1);
        __syncthreads();
        View<C, P> vw{lds, (const Slot<C::VEC>(*)[C::NP])ring, aux, j, col};
        op.run(vw, (long long)j * nx, x, xmask);
    };
    int j = j0;
    for (; j + C::NP <= j1; j += C::NP)
        kg_static_for<C::NP>([&](auto p) { row_step(p, j + decltype(p)::value); });
    kg_static_for<C::NP - 1>([&](auto p) {
        if (j + decltype(p)::value < j1) row_step(p, j + decltype(p)::value);
    });
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

// ------------------------------------------------------------ direct march
// For single-field 2D stencils (no side streams): no LDS and no ring.  A lane
// owns points x, x+1 (x even) of PY consecutive rows j0 .. j0+PY-1 and loads
// the rows j0-H .. j0+PY-1+H it needs straight from global memory, each as
// three aligned 16-byte loads covering points x-2 .. x+3, all of them before
// the first use; the PY output rows are then computed from registers.  The
// x- and y-neighbour loads of adjacent lanes and workgroups hit in the
// caches, so HBM still sees each input once.  With the workgroups in
// hardware order (no XCD swizzle) the whole chip streams one contiguous
// window of the grid, which measured faster than the LDS-staged engine for
// jacobi, gameoflife and gaussblur at 512 x 65536.  Loading only the lane's
// own piece and taking the pieces at x-2 and x+2 from the neighbouring lanes
// by DPP shifts (one load per row instead of three) was bit-identical but
// slower: jacobi 100 -> 110 us, gameoflife 100 -> 103, gaussblur 94 -> 122.
template <int PY_, int ROWS_, int H_>
struct DCfg {
    static constexpr int VEC = 2;
    static constexpr int PY = PY_;       // output rows per lane
    static constexpr int ROWS = ROWS_;   // waves per workgroup, stacked in y
    static constexpr int H = H_;         // stencil radius (<= 2)
    static constexpr int NR = PY + 2 * H;
    static constexpr int TXW = 128;
    static_assert(H >= 1 && H <= 2, "points x-2 .. x+3 cover a radius of 2");
};

// Output row R of the lane's window (row j = j0 + R).
template <class C, int R>
struct DView {
    const double (&win)[C::NR][6];   // rows j0-H .., points x-2 .. x+3
    int j;
    __device__ __forceinline__ double at(int, int v, int dx, int dy) const
    {
        return win[R + C::H + dy][2 + v + dx];
    }
};

template <class C>
struct DGrid {
    int ntx;
    unsigned blocks;
    bool empty;
};

template <class C>
static inline DGrid<C> make_dgrid(int nx, int ny, int lo, int hi)
{
    DGrid<C> g;
    const int nyi = ny - hi - lo, nxi = nx - hi - lo;
    g.empty = nxi <= 0 || nyi <= 0;
    g.ntx = (nx + C::TXW - 1) / C::TXW;
    const int per = C::ROWS * C::PY;
    g.blocks = g.empty ? 0u : (unsigned)g.ntx * (unsigned)((nyi + per - 1) / per);
    return g;
}

// Requires an even nx and 16-byte aligned fields (checked by the host).
// ORD: tile order (kg_tile_order).
template <class C, int LO, int HI, class Op, int ORD = 0>
__global__ __launch_bounds__(64 * C::ROWS)
void direct_kernel(int nx, int ny, int ntx, Op op)
{
    const int lane = (int)(threadIdx.x & 63), r = (int)(threadIdx.x >> 6);
    const unsigned b = kg_tile_order<ORD>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int bx = (int)(b % (unsigned)ntx);
    const int by = (int)(b / (unsigned)ntx);
    const int x = bx * C::TXW + 2 * lane;
    const int j0 = LO + (by * C::ROWS + r) * C::PY;
    if (x >= nx || j0 >= ny - HI) return;   // no barriers in this kernel
    // The pairs at x-2 and x+2 are clamped into the row at the grid's
    // x-edges; there they only feed points outside the interior.  Rows past
    // ny-1 are clamped likewise (they feed rows >= ny-HI, never stored).
    const int xl = x >= 2 ? x - 2 : x;
    const int xr = x + 2 < nx ? x + 2 : x;
    const kgp::gmem<double>* src = kgp::as_global(op.src[0]);
    double win[C::NR][6];
#pragma unroll
    for (int q = 0; q < C::NR; q++) {
        const kgp::gmem<double>* row = src + (long long)min(j0 - C::H + q, ny - 1) * nx;
        const d2 a = *(const kgp::gmem<d2>*)(row + xl);
        const d2 c = *(const kgp::gmem<d2>*)(row + x);
        const d2 e = *(const kgp::gmem<d2>*)(row + xr);
        win[q][0] = a.x;
        win[q][1] = a.y;
        win[q][2] = c.x;
        win[q][3] = c.y;
        win[q][4] = e.x;
        win[q][5] = e.y;
    }
    const bool xm[2] = {x >= LO && x < nx - HI, x + 1 >= LO && x + 1 < nx - HI};
    kg_static_for<C::PY>([&](auto rr) {
        constexpr int R = decltype(rr)::value;
        const int j = j0 + R;
        if (j < ny - HI) op.run(DView<C, R>{win, j}, (long long)j * nx, x, xm);
    });
}

}  // namespace kgr
