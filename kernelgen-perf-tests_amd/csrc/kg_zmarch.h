// kg_zmarch.h -- tiling shared by the 3D stencil kernels.
//
// A workgroup of TX x TY threads owns a TX x TY column tile of the xy-plane
// and marches it through a chunk of KC interior z-planes.  x is the fastest
// (coalesced) axis: one wave64 covers 64 consecutive x points of one row.
// Tiles are aligned to x = 0 / y = 0 and cover the whole plane; threads whose
// (i, j) lies outside the interior return at once (no barriers are used).
// With SWZ the hardware workgroup id goes through kg_xcd_swizzle so that each
// XCD's L2 holds one contiguous band of tiles (consecutive y-tiles of one
// z-chunk); without it consecutive tiles go to different XCDs and the whole
// chip streams one contiguous window of the grid, which measured faster for
// the kernels whose neighbours are few or already in registers.
#pragma once

#include "kg_internal.h"

struct kg_ztile {
    int i, j;        // column owned by this thread
    int k0, k1;      // interior planes [k0, k1) this thread computes
};

// lo / hi: interior is [lo, n - hi) along every axis.
template <int TX, int TY, int KC, int LO, int HI, bool SWZ = true>
__device__ __forceinline__ bool kg_ztile_get(int nx, int ny, int ns, int ntx, int nty, kg_ztile& t)
{
    unsigned b = SWZ ? kg_xcd_swizzle(blockIdx.x, gridDim.x) : blockIdx.x;
    const int bx = (int)(b % (unsigned)ntx);
    b /= (unsigned)ntx;
    const int by = (int)(b % (unsigned)nty);
    const int bz = (int)(b / (unsigned)nty);
    t.i = bx * TX + (int)threadIdx.x;
    t.j = by * TY + (int)threadIdx.y;
    t.k0 = LO + bz * KC;
    t.k1 = min(t.k0 + KC, ns - HI);
    return t.i >= LO && t.i < nx - HI && t.j >= LO && t.j < ny - HI && t.k0 < t.k1;
}

struct kg_zgrid {
    int ntx, nty, nkc;
    unsigned blocks;
    bool empty;
};

// Host-side grid for a kernel with interior [lo, n-hi) on each axis.
static inline kg_zgrid kg_zgrid_make(int nx, int ny, int ns, int lo, int hi, int TX, int TY, int KC)
{
    kg_zgrid g;
    const int nzi = ns - hi - lo;
    g.empty = (nx - hi - lo <= 0) || (ny - hi - lo <= 0) || nzi <= 0;
    g.ntx = (nx + TX - 1) / TX;
    g.nty = (ny + TY - 1) / TY;
    g.nkc = g.empty ? 0 : (nzi + KC - 1) / KC;
    g.blocks = (unsigned)g.ntx * (unsigned)g.nty * (unsigned)g.nkc;
    return g;
}

// 64-bit plane offset, 32-bit in-plane offset.
__device__ __forceinline__ long long kg_off3(int i, int j, int k, int nx, long long sz)
{
    return (long long)k * sz + (long long)(j * nx + i);
}
