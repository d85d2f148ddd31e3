// kg_life.hip -- Conway's Game of Life on uint8 cells (one byte per cell,
// nonzero = alive), the integer form of gameoflife/gameoflife.c:82-91: that
// FP64 kernel's 1/(1+(u+L-3)(L-3)*1e20) is 1 exactly when L == 3 or
// (u == 1 and L == 2) for 0/1 cells and 0 to within 1e-20 otherwise, i.e. the
// Conway rule.  Interior [1,n-1)^2 is written; boundary cells of c1 keep their
// contents, as in the FP64 kernel.
//
// Layout: c[i + nx*j], x fastest.  Each lane owns 16 consecutive cells of a
// row (one 16-byte load) and marches down a chunk of LIFE_YC rows with the
// rows j-1, j, j+1 in registers; the cells at x-1 and x+16 come from two
// 1-byte loads that hit the lines the neighbouring lanes just fetched.  The
// neighbour count is done four cells at a time in 32-bit words (SWAR): counts
// are at most 8, so byte lanes never carry into each other.  Per row the
// horizontal triple sum l+m+r is formed once and reused as the upper, middle
// (minus m) and lower contribution of three output rows.
//
// HBM bytes per cell and step: 1 read + 1 written (the x-1/x+16 bytes and the
// rows shared by neighbouring chunks are cache hits).
#include "kg_internal.h"

namespace {

constexpr int LIFE_LANES = 256;   // 4 wave64s per workgroup
constexpr int LIFE_CW = 16;       // cells per lane
constexpr int LIFE_TX = LIFE_LANES * LIFE_CW;
constexpr int LIFE_YC = 16;       // rows per workgroup

typedef unsigned u4 __attribute__((ext_vector_type(4)));

// 0x01 in every byte lane whose byte is nonzero, 0x00 elsewhere.
__device__ __forceinline__ unsigned nz_bytes(unsigned x)
{
    return ((((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u) >> 7;
}

struct LifeRow {
    unsigned m[4];   // own 16 cells as 0/1 bytes
    unsigned h[4];   // per-cell sum of the cells at x-1, x, x+1 of this row
};

__device__ __forceinline__ LifeRow life_row(const uint8_t* p)
{
    const u4 raw = *(const u4*)p;
    const unsigned l = nz_bytes((unsigned)p[-1]);
    const unsigned r = nz_bytes((unsigned)p[LIFE_CW]);
    LifeRow row;
    row.m[0] = nz_bytes(raw.x);
    row.m[1] = nz_bytes(raw.y);
    row.m[2] = nz_bytes(raw.z);
    row.m[3] = nz_bytes(raw.w);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        // byte b of `left` is the cell at x-1 of byte b, of `right` at x+1
        const unsigned prev = q == 0 ? (l << 24) : row.m[q - 1];
        const unsigned next = q == 3 ? r : row.m[q + 1];
        const unsigned left = (row.m[q] << 8) | (prev >> 24);
        const unsigned right = (row.m[q] >> 8) | (next << 24);
        row.h[q] = left + row.m[q] + right;
    }
    return row;
}

// Next state from the neighbour count s (0..8 per byte) and the own cell m
// (0/1): alive iff (s | m) == 3.
__device__ __forceinline__ unsigned life_rule(unsigned s, unsigned m)
{
    const unsigned x = (s | m) ^ 0x03030303u;                 // bytes 0..11
    const unsigned t = (x + 0x7f7f7f7fu) & 0x80808080u;       // 0x80 where x != 0
    return (t ^ 0x80808080u) >> 7;
}

__device__ __forceinline__ uint8_t life_cell(const uint8_t* c0, long long o, int nx)
{
    int n = 0;
#pragma unroll
    for (int dj = -1; dj <= 1; dj++)
#pragma unroll
        for (int di = -1; di <= 1; di++)
            if (di || dj) n += c0[o + (long long)dj * nx + di] != 0;
    const int alive = c0[o] != 0;
    return (uint8_t)((n == 3) || (alive && n == 2));
}

__global__ __launch_bounds__(LIFE_LANES)
void life_u8_k(int nx, int ny, int ntx, const uint8_t* __restrict__ c0, uint8_t* __restrict__ c1)
{
    const unsigned b = kg_xcd_swizzle(blockIdx.x, gridDim.x);
    const int bx = (int)(b % (unsigned)ntx);
    const int by = (int)(b / (unsigned)ntx);
    const int x = bx * LIFE_TX + (int)threadIdx.x * LIFE_CW;
    const int j0 = 1 + by * LIFE_YC;
    const int j1 = min(j0 + LIFE_YC, ny - 1);
    if (x >= nx || j0 >= j1) return;   // no barriers below
    // Fast lanes: all 16 cells interior (x >= 1, x+15 <= nx-2) and 16-byte
    // aligned; others (grid edges, ragged or unaligned rows) go cell by cell.
    const bool fast = x >= 1 && x + LIFE_CW <= nx - 1 && (nx & (LIFE_CW - 1)) == 0;
    if (!fast) {
        const int xa = max(x, 1), xb = min(x + LIFE_CW, nx - 1);
        for (int j = j0; j < j1; j++)
            for (int i = xa; i < xb; i++) {
                const long long o = (long long)j * nx + i;
                c1[o] = life_cell(c0, o, nx);
            }
        return;
    }
    const uint8_t* src = c0 + x;
    LifeRow up = life_row(src + (long long)(j0 - 1) * nx);
    LifeRow mid = life_row(src + (long long)j0 * nx);
    LifeRow dn = life_row(src + (long long)(j0 + 1) * nx);   // row j0+1 <= j1 <= ny-1
    for (int j = j0; j < j1; j++) {
        LifeRow nxt;
        if (j + 2 <= j1) nxt = life_row(src + (long long)(j + 2) * nx);
        u4 out;
        unsigned r[4];
#pragma unroll
        for (int q = 0; q < 4; q++) r[q] = life_rule(up.h[q] + (mid.h[q] - mid.m[q]) + dn.h[q], mid.m[q]);
        out.x = r[0];
        out.y = r[1];
        out.z = r[2];
        out.w = r[3];
        kg_st((u4*)(c1 + (long long)j * nx + x), out);
        up = mid;
        mid = dn;
        dn = nxt;
    }
}

}  // namespace

// gameoflife/gameoflife.c:82-91 on uint8 cells (see the file comment)
extern "C" int kg_life_u8(kg_ctx* ctx, int nx, int ny, const uint8_t* c0, uint8_t* c1)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims2(nx, ny)) return rc;
    if (!c0 || !c1) return kg_set_error(KG_ERR_INVALID, "kg_life_u8: null buffer");
    if (c0 == c1) return kg_set_error(KG_ERR_INVALID, "kg_life_u8: c0 and c1 must be distinct");
    if (((uintptr_t)c0 | (uintptr_t)c1) & (LIFE_CW - 1))
        return kg_set_error(KG_ERR_INVALID, "kg_life_u8: buffers must be 16-byte aligned");
    kg_register_kernel("life_u8", (const void*)life_u8_k);
    if (int rc = kg_launch_begin(ctx, "life_u8")) return rc;
    if (nx > 2 && ny > 2) {
        const int ntx = (nx + LIFE_TX - 1) / LIFE_TX;
        const int nty = (ny - 2 + LIFE_YC - 1) / LIFE_YC;
        hipLaunchKernelGGL(life_u8_k, dim3((unsigned)ntx * (unsigned)nty), dim3(LIFE_LANES), 0,
                           ctx->stream, nx, ny, ntx, c0, c1);
    }
    return kg_launch_end(ctx);
}

int kg_preload_life() { return kg_preload_kernel((const void*)life_u8_k); }
