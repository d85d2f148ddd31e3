// kg_life.hip -- Conway's Game of Life on uint8 cells (one byte per cell,
// nonzero = alive), the integer form of gameoflife/gameoflife.c:82-91: that
// FP64 kernel's 1/(1+(u+L-3)(L-3)*1e20) is 1 exactly when L == 3 or
// (u == 1 and L == 2) for 0/1 cells and 0 to within 1e-20 otherwise, i.e. the
// Conway rule.  Interior [1,n-1)^2 is written; boundary cells of c1 keep their
// contents, as in the FP64 kernel.
//
// Layout: c[i + nx*j], x fastest.  Each lane owns 16 consecutive cells of a
// row (one 16-byte load), a wave 1024 cells, a workgroup LIFE_WAVES such segments
// side by side and a band of LIFE_YC output rows, which it marches downwards.
// Rows are loaded LIFE_D rows ahead into a ring of LIFE_D+1 slots whose
// phase is fixed at compile time (the row loop is unrolled LIFE_D+1 times),
// through buffer loads: a row or lane outside the band gets an out-of-range
// offset and reads 0, so no load sits in conditional code and the compiler
// can wait for one row while the later ones stay in flight.  The cell at x-1
// comes from the lane below (ds_bpermute), the cell at x+16 from the lane
// above; the wave's end lanes load the 4 bytes beside their segment instead.
// The neighbour count is done four cells at a time in 32-bit words (SWAR):
// counts are at most 8, so byte lanes never carry into each other.  Per row
// the horizontal triple sum l+m+r is formed once and reused as the upper,
// middle (minus m) and lower contribution of three output rows.
//
// HBM bytes per cell and step: 1 read + 1 written (the 4-byte side loads and
// the two rows shared by neighbouring bands are cache hits).
#include "kg_plane.h"

namespace {

#ifndef KG_ORD_LIFE8
#define KG_ORD_LIFE8 8                        // tile order (kg_tile_order)
#endif
#ifndef LIFE_WAVES
#define LIFE_WAVES 4                          // segments per workgroup
#endif
constexpr int LIFE_CW = 16;                   // cells per lane
constexpr int LIFE_SEG = 64 * LIFE_CW;        // cells per wave segment
constexpr int LIFE_TX = LIFE_WAVES * LIFE_SEG;
#ifndef LIFE_D
#define LIFE_D 3                              // rows in flight per lane
#endif
#ifndef LIFE_U
#define LIFE_U 1                              // ring turns per band
#endif
constexpr int LIFE_NP = LIFE_D + 1;           // ring slots
constexpr int LIFE_YC = LIFE_NP * LIFE_U;     // output rows per band
constexpr int LIFE_SLOW_YC = 16;              // rows per workgroup, cell-wise kernel

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

// One row of a lane: raw = its 16 cells, side = the 4 bytes beside the wave's
// segment (lane 0: cells x-4..x-1, lane 63: cells x+16..x+19, else unused).
__device__ __forceinline__ LifeRow life_row(u4 raw, unsigned side, int lane)
{
    LifeRow row;
    row.m[0] = nz_bytes(raw.x);
    row.m[1] = nz_bytes(raw.y);
    row.m[2] = nz_bytes(raw.z);
    row.m[3] = nz_bytes(raw.w);
    const unsigned s = nz_bytes(side);
    // cell x-1: top byte of the lane below; cell x+16: low byte of the lane above
    const unsigned below = (unsigned)__builtin_amdgcn_ds_bpermute((lane - 1) * 4, (int)row.m[3]);
    const unsigned above = (unsigned)__builtin_amdgcn_ds_bpermute((lane + 1) * 4, (int)row.m[0]);
    const unsigned l = (lane == 0 ? s : below) >> 24;
    const unsigned r = (lane == 63 ? s : above) & 0xffu;
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

// Rows with nx % 16 != 0 (the vector kernel below needs aligned rows): a
// workgroup of 256 threads takes LIFE_SLOW_YC rows of 256 cells, cell by cell.
__global__ __launch_bounds__(256)
void life_u8_cells_k(int nx, int ny, int ntx, const uint8_t* __restrict__ c0,
                     uint8_t* __restrict__ c1)
{
    const unsigned b = kg_xcd_swizzle(blockIdx.x, gridDim.x);
    const int i = 1 + (int)(b % (unsigned)ntx) * 256 + (int)threadIdx.x;
    const int j0 = 1 + (int)(b / (unsigned)ntx) * LIFE_SLOW_YC;
    const int j1 = min(j0 + LIFE_SLOW_YC, ny - 1);
    if (i >= nx - 1) return;
    for (int j = j0; j < j1; j++) {
        const long long o = (long long)j * nx + i;
        c1[o] = life_cell(c0, o, nx);
    }
}

// nx % 16 == 0 and 1 <= nx <= 2^24 (offsets within a band fit 31 bits).
// Bands of 4 rows measured fastest at 4096^2 .. 16384^2 (16-row bands 15 %
// slower).  Tile order KG_ORD_LIFE8: chunks of 8 bands per XCD, so a band's
// neighbour rows and segments share its L2 (ping-pong steps, 50-launch means:
// 8192^2 30.7 us, 16384^2 100 us; hardware order 31.4 / 107, XCD-contiguous
// 37.9 / 108, chunks of 2, 4 or 16 bands within 2 %).  At 8192^2, 2 or 8
// segments per workgroup ran 33.5 and 32.2 us against 31.0 for 4; 2 rows in
// flight the same, 5 rows 32.4 us; two ring turns per band (D = 1) 31.1 us.
__global__ __launch_bounds__(64 * LIFE_WAVES)
void life_u8_k(int nx, int ny, int nsx, const uint8_t* __restrict__ c0, uint8_t* __restrict__ c1)
{
    const unsigned b = kg_tile_order<KG_ORD_LIFE8>(blockIdx.x, gridDim.x, (unsigned)nsx);
    const int sx = (int)(b % (unsigned)nsx);
    const int band = (int)(b / (unsigned)nsx);
    const int lane = (int)(threadIdx.x & 63);
    const int x = (sx * LIFE_WAVES + (int)(threadIdx.x >> 6)) * LIFE_SEG + lane * LIFE_CW;
    const int j0 = 1 + band * LIFE_YC;            // first output row (< ny-1)
    const int j1 = min(j0 + LIFE_YC, ny - 1);     // output rows j0 .. j1-1
    // Resources cover exactly the band: input rows j0-1 .. j1, output rows
    // j0 .. j1-1; rows past them read 0 / are dropped by the hardware.
    const __amdgpu_buffer_rsrc_t rin =
        kgp::make_rsrc(c0 + (long long)(j0 - 1) * nx, (unsigned)(j1 - j0 + 2) * (unsigned)nx);
    const __amdgpu_buffer_rsrc_t rout =
        kgp::make_rsrc(c1 + (long long)j0 * nx, (unsigned)(j1 - j0) * (unsigned)nx);
    const bool xin = x < nx;
    const unsigned vx = xin ? (unsigned)x : kgp::KG_OOB;
    const unsigned sx_off = lane == 0 && xin && x >= 4 ? (unsigned)(x - 4)
                          : lane == 63 && x + LIFE_CW < nx ? (unsigned)(x + LIFE_CW) : kgp::KG_OOB;
    // lanes whose 16 cells are all interior store a vector; the two lanes
    // holding cell 0 or cell nx-1 store their other 15 cells one by one
    const bool full = xin && x >= 1 && x + LIFE_CW <= nx - 1;
    const bool edge = xin && !full;
    const unsigned vst = full ? (unsigned)x : kgp::KG_OOB;

    u4 ring[LIFE_NP];
    unsigned side[LIFE_NP];
    // input row r of the band (r = row - (j0-1)); OOB offsets stay >= 2^31
    auto load = [&](int r, u4& raw, unsigned& sd) {
        const unsigned ro = (unsigned)r * (unsigned)nx;
        raw = __builtin_bit_cast(u4, kgp::buf_load<16>(rin, vx == kgp::KG_OOB ? vx : ro + vx));
        sd = kgp::buf_load<4>(rin, sx_off == kgp::KG_OOB ? sx_off : ro + sx_off);
    };
    u4 raw0, raw1;
    unsigned s0, s1;
    load(0, raw0, s0);
    load(1, raw1, s1);
#pragma unroll
    for (int d = 0; d < LIFE_D; d++) load(2 + d, ring[d], side[d]);
    LifeRow up = life_row(raw0, s0, lane);
    LifeRow mid = life_row(raw1, s1, lane);
    // Enter the loop with no load in flight (vmcnt(0)).  The ring's registers
    // before the loop need not be the ones the loop uses; a load still pending
    // there would make the compiler's wait at the loop head assume it was the
    // latest load, i.e. wait for nearly the whole queue on every pass.
    __builtin_amdgcn_s_waitcnt(0x0f70);

    // output row j = j0 + t at phase P = t % NP: slot P holds input row t+2
    auto step = [&](auto phase, int t) {
        constexpr int P = decltype(phase)::value;
        load(t + 2 + LIFE_D, ring[(P + LIFE_D) % LIFE_NP], side[(P + LIFE_D) % LIFE_NP]);
        const LifeRow dn = life_row(ring[P], side[P], lane);
        unsigned res[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
            res[q] = life_rule(up.h[q] + (mid.h[q] - mid.m[q]) + dn.h[q], mid.m[q]);
        const unsigned ro = (unsigned)t * (unsigned)nx;
        const u4 out = {res[0], res[1], res[2], res[3]};
        __builtin_amdgcn_raw_buffer_store_b128(out, rout, vst == kgp::KG_OOB ? vst : ro + vst, 0, 2);
        if (edge && j0 + t < j1) {
            uint8_t bytes[LIFE_CW];
            __builtin_memcpy(bytes, res, sizeof bytes);
            uint8_t* row = c1 + (long long)(j0 + t) * nx;
#pragma unroll
            for (int q = 0; q < LIFE_CW; q++)
                if (x + q >= 1 && x + q < nx - 1) row[x + q] = bytes[q];
        }
        up = mid;
        mid = dn;
    };
    for (int t = 0; t < j1 - j0; t += LIFE_NP)
        kg_static_for<LIFE_NP>([&](auto p) { step(p, t + decltype(p)::value); });
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
    const bool vec = nx % LIFE_CW == 0 && nx <= (1 << 24);
    kg_register_kernel("life_u8", vec ? (const void*)life_u8_k : (const void*)life_u8_cells_k);
    if (int rc = kg_launch_begin(ctx, "life_u8")) return rc;
    if (nx > 2 && ny > 2) {
        if (vec) {
            const unsigned nsx = (unsigned)((nx + LIFE_TX - 1) / LIFE_TX);
            const unsigned nb = (unsigned)((ny - 2 + LIFE_YC - 1) / LIFE_YC);
            hipLaunchKernelGGL(life_u8_k, dim3(nsx * nb), dim3(64 * LIFE_WAVES), 0, ctx->stream, nx,
                               ny, (int)nsx, c0, c1);
        } else {
            const unsigned ntx = (unsigned)((nx - 2 + 255) / 256);
            const unsigned nb = (unsigned)((ny - 2 + LIFE_SLOW_YC - 1) / LIFE_SLOW_YC);
            hipLaunchKernelGGL(life_u8_cells_k, dim3(ntx * nb), dim3(256), 0, ctx->stream, nx, ny,
                               (int)ntx, c0, c1);
        }
    }
    return kg_launch_end(ctx);
}

int kg_preload_life()
{
    if (int rc = kg_preload_kernel((const void*)life_u8_k)) return rc;
    return kg_preload_kernel((const void*)life_u8_cells_k);
}
