// kg_plane.h -- plane-staged z-marching engine for the 3D stencils.
//
// Workgroup = W wave64s.  A wave covers RPW = 64/LPR rows of LPR lanes; each
// lane holds VEC consecutive x points of its row, loaded with one access of
// VEC * sizeof(storage) bytes (16 bytes for FP64 VEC=2, FP32 VEC=4).  A tile
// is LPR*VEC points wide and W*RPW rows tall; the H rows at either end are
// y-halo rows that are loaded and staged but not computed.  LPR = 64 (one row
// per wave) is the shape for wide grids; LPR = 32 / 16 fold two / four short
// rows into a wave so that narrow grids keep their lanes busy.  The workgroup
// marches its tile through a chunk of z-planes:
//
//   * every global load of the main field is issued D planes ahead of use
//     into a register ring (own row only, plus the H x-halo columns on the
//     first and last H lanes of each row), so the only vector-memory traffic
//     in the loop is this prefetch stream plus the side streams and the output
//     stores -- neighbour values never wait behind an HBM miss;
//   * planes k-LL .. k+LR of the main field are staged in LDS (LL+LR+2 plane
//     slots, one barrier per plane) and give every xy-neighbour;
//   * the register ring keeps the own-row column k-ZL .. k+ZR for z-neighbours;
//   * up to NAUX side streams (arrays read only at the point itself) ride in
//     a second prefetch ring of depth D.
//
// Storage and arithmetic types come from a type pair T (F64: double/double,
// F32: float/float, BF16: bfloat16 storage with float arithmetic); values are
// converted at load and rounded to nearest-even at store.  Rows are `sy`
// elements apart and planes `sz` apart (sy = nx, sz = nx*ny for the dense
// layout of the C ABI; batched grids pad sy to a multiple of VEC).
//
// Measured on MI355X at 512x256x256 FP64 (cold caches): laplacian 99-102 us
// with this engine against 143 us for per-thread z-marching with
// cache-served neighbours.  HBM traffic (rocprofv3 FETCH_SIZE + WRITE_SIZE
// against the algorithmic bytes) is 1.03-1.09x for wave13pt (by tile
// order) but 1.17x for laplacian and 1.29x for lapgsrb: tiles 512 bytes wide
// re-fetch the x-halo lines their neighbours on other XCDs also read, and
// the y-halo rows of a tile are read again by the tiles above and below it.
//
// The header also holds the engines for stencils that read only their own
// xy-plane: packed rows (narrow grids), the direct row body (rows loaded
// through the caches) and the column march (a wave marches a row piece
// along y in registers, x-neighbours by DPP lane shifts, no LDS).
#pragma once

#include "kg_internal.h"

namespace kgp {

// ---------------------------------------------------------------- types
struct F64 {
    using S = double;
    using F = double;
};
struct F32 {
    using S = float;
    using F = float;
};
struct BF16 {
    using S = uint16_t;   // bfloat16 bit pattern
    using F = float;
};

__device__ __forceinline__ float bf16_in(uint16_t s) { return __uint_as_float((unsigned)s << 16); }

// float -> bfloat16, round to nearest even; a NaN keeps its sign and top
// payload bits with the quiet bit set ((u >> 16) | 0x40).  gfx950 converts
// in hardware (v_cvt_pk_bf16_f32, two values per instruction); over all 2^32
// float patterns it returns exactly (u + 0x7fff + ((u >> 16) & 1)) >> 16 for
// non-NaNs and the NaN rule above (checked exhaustively on an MI355X).
__device__ __forceinline__ uint16_t bf16_out(float f)
{
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}

template <class T>
__device__ __forceinline__ typename T::F ld_cvt(typename T::S s)
{
    if constexpr (sizeof(typename T::S) == 2) return bf16_in(s);
    else return s;
}
template <class T>
__device__ __forceinline__ typename T::S st_cvt(typename T::F f)
{
    if constexpr (sizeof(typename T::S) == 2) return bf16_out(f);
    else return f;
}

// Global-memory (address space 1) view of a pointer.  Pointers that a kernel
// loads from memory (the batched kernel's grid descriptors) are generic to the
// compiler, which would emit flat loads and stores through them; accessing
// through an address-space-1 pointer gives global instructions.
template <class T> using gmem = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ gmem<T>* as_global(T* p)
{
    return (gmem<T>*)p;
}

// An unsigned vector of BYTES bytes (one global access).
template <int BYTES> struct Raw;
template <> struct Raw<16> { typedef unsigned t __attribute__((ext_vector_type(4))); };
template <> struct Raw<8> { typedef unsigned t __attribute__((ext_vector_type(2))); };
template <> struct Raw<4> { typedef unsigned t; };
template <> struct Raw<2> { typedef unsigned short t; };

// VEC consecutive storage values at p (aligned to the access size).
template <class T, int VEC>
__device__ __forceinline__ void vload(const gmem<typename T::S>* p, typename T::F (&out)[VEC])
{
    using S = typename T::S;
    using R = typename Raw<VEC * (int)sizeof(S)>::t;
    const R r = *(const gmem<R>*)p;
    S s[VEC];
    __builtin_memcpy(s, &r, sizeof r);
#pragma unroll
    for (int v = 0; v < VEC; v++) out[v] = ld_cvt<T>(s[v]);
}

// Buffer-resource access.  A buffer load whose offset lies at or past the
// resource's num_records returns 0 without touching memory, so lanes outside
// the grid, and planes outside the grid or past the tile, are handled by the
// offset or by an empty resource instead of by a branch.  Keeping every load
// out of conditional code matters: the compiler waits for an earlier load by
// counting the memory operations issued after it on every path, and a load
// that may be skipped does not count -- with all loads conditional it would
// wait for the whole prefetch queue (vmcnt(0)) every plane.
constexpr unsigned KG_OOB = 0x80000000u;          // offset past any resource used here
constexpr int KG_RSRC_WORD3 = 0x00020000;         // gfx9 raw buffer: 32-bit data format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, KG_RSRC_WORD3);
}

template <int BYTES>
__device__ __forceinline__ typename Raw<BYTES>::t buf_load(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    if constexpr (BYTES == 16) return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    else if constexpr (BYTES == 8) return __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    else if constexpr (BYTES == 4) return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
    else return __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
}

// VEC consecutive storage values at byte offset off of resource r.
template <class T, int VEC>
__device__ __forceinline__ void bload(__amdgpu_buffer_rsrc_t r, unsigned off,
                                      typename T::F (&out)[VEC])
{
    using S = typename T::S;
    const auto raw = buf_load<VEC * (int)sizeof(S)>(r, off);
    S s[VEC];
    __builtin_memcpy(s, &raw, sizeof raw);
#pragma unroll
    for (int v = 0; v < VEC; v++) out[v] = ld_cvt<T>(s[v]);
}

// ---------------------------------------------------------------- config
template <class C, class V>
__device__ __forceinline__ typename C::F xrow(const V& p, int v, int dx);
template <int VEC_, int W_, int H_, int LL_, int LR_, int ZL_, int ZR_, int D_, int NAUX_, int KC_,
          class T_ = F64, int LPR_ = 64, bool FULL_ = false, bool XS_ = false>
struct Cfg {
    using T = T_;
    using S = typename T::S;           // storage type
    using F = typename T::F;           // arithmetic type
    static constexpr int VEC = VEC_;   // points per lane along x
    static constexpr int W = W_;       // waves per workgroup
    static constexpr int H = H_;       // x and y halo width
    static constexpr int LL = LL_;     // LDS planes below k
    static constexpr int LR = LR_;     // LDS planes above k
    static constexpr int ZL = ZL_;     // register column below k
    static constexpr int ZR = ZR_;     // register column above k
    static constexpr int D = D_;       // prefetch distance (planes)
    static constexpr int NAUX = NAUX_; // side streams read at the point only
    static constexpr int KC = KC_;     // planes per workgroup (single-grid launches)
    static constexpr int LPR = LPR_;   // lanes per row (64, 32 or 16)
    // FULL: stores write every point of a computed row (x-boundary points get
    // the centre value of the input level), so no cache line of the row is
    // written in part; for fields whose levels share their boundary values.
    static constexpr bool FULL = FULL_;
    // XS: x-neighbours along the lane's own row (any plane of the register
    // ring) come from the neighbouring lanes' registers through wave-wide DPP
    // shifts (View::xat) instead of LDS.  The first lane of the row then holds
    // the H halo values left of the tile and the last lane the H values right
    // of it.  One row per wave (LPR = 64).
    static constexpr bool XS = XS_;
    static constexpr int NH = XS ? H : 1;             // halo values per ring slot
    static constexpr int RPW = 64 / LPR;             // rows per wave
    static constexpr int ROWS = W * RPW;             // tile rows, halo rows included
    static constexpr int TXW = LPR * VEC;            // tile width in points
    static constexpr int HL = (H + VEC - 1) / VEC;   // x-halo lanes per side
    static constexpr int LQ = LPR + 2 * HL;          // LDS lane slots per row and component
    static constexpr int NLS = LL + LR + 2;          // LDS plane slots
    static constexpr int ZLO = ZL;                   // ring starts at k-ZL
    static constexpr int ZHI = (ZR > LR ? ZR : LR);  // ring reaches k+ZHI before prefetch
    static constexpr int RS = ZLO + ZHI + D;         // ring: planes k-ZL .. k+ZHI+D-1
    static constexpr int NP = RS + 1;                // physical slots (one being filled)
    static constexpr int TY = ROWS - 2 * H;          // computed rows per tile
    static constexpr int NA = NAUX > 0 ? NAUX : 1;
    static constexpr size_t LDS_BYTES = sizeof(F) * NLS * ROWS * VEC * LQ;
    static_assert(LPR == 64 || LPR == 32 || LPR == 16, "lanes per row");
    static_assert(LPR > 2 * H, "halo lanes of a row must not overlap");
    static_assert(!XS || (LPR == 64 && H <= 2 * VEC), "x-shift mode: one row per wave, H <= 2 VEC");
};

template <class F, int VEC, int NH = 1>
struct Slot {
    F v[VEC];
    // x-halo: one value carried by each of the first and last H lanes of a
    // row, or (x-shift mode) H values on the first and on the last lane
    F h[NH];
};

// Wave-wide DPP lane shifts (GFX9 family): wave_shr:1 gives lane i the value
// of lane i-1, wave_shl:1 that of lane i+1; the lane without a source (0 or
// 63) keeps `old`.
constexpr int KG_DPP_WAVE_SHL1 = 0x130;
constexpr int KG_DPP_WAVE_SHR1 = 0x138;

template <int CTRL>
__device__ __forceinline__ double dpp_shift(double old, double src)
{
    const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
    const unsigned long long s = __builtin_bit_cast(unsigned long long, src);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)s, CTRL,
                                                              0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32),
                                                              (int)(unsigned)(s >> 32), CTRL, 0xF, 0xF,
                                                              false);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float old, float src)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, src), CTRL,
                                                                 0xF, 0xF, false));
}

template <class F, int VEC, int NA>
struct ASlot {
    F a[NA][VEC];
};

struct Grid {
    int ntx, nty, nkc;
    unsigned blocks;
    bool empty;
};

// Tiles of a grid whose xy-interior is [lo, n-hi) and z-interior [zlo, ns-zhi).
template <class C>
static inline Grid make_grid(int nx, int ny, int ns, int lo, int hi, int zlo, int zhi, int kc)
{
    Grid g;
    const int nzi = ns - zhi - zlo, nyi = ny - hi - lo, nxi = nx - hi - lo;
    g.empty = nxi <= 0 || nyi <= 0 || nzi <= 0;
    g.ntx = (nx + C::TXW - 1) / C::TXW;
    g.nty = g.empty ? 0 : (nyi + C::TY - 1) / C::TY;
    g.nkc = g.empty ? 0 : (nzi + kc - 1) / kc;
    g.blocks = (unsigned)g.ntx * (unsigned)g.nty * (unsigned)g.nkc;
    return g;
}
template <class C>
static inline Grid make_grid(int nx, int ny, int ns, int lo, int hi)
{
    return make_grid<C>(nx, ny, ns, lo, hi, lo, hi, C::KC);
}

// LDS layout of a staged plane: component-major, lane-contiguous.  Point
// x0 + q*VEC + v of tile row r (lane q of the row holds its VEC points) sits at
// [slot][r][v][HL + q], the x-halo points at lanes -HL .. -1 and LPR ..
// LPR+HL-1.  A neighbour (v + dx) of a lane is then component (v+dx) mod VEC
// of lane q + floor((v+dx)/VEC): consecutive lanes read consecutive words,
// free of bank conflicts (a point-contiguous row, read at a stride of VEC
// values per lane, put several lanes of a wave on one bank).
__host__ __device__ constexpr int floor_div(int e, int d) { return (e + 8 * d) / d - 8; }

// Per-lane view handed to a stencil's compute function at ring phase P: the
// register ring's logical slot r (plane k - ZL + r) is physical slot
// (P + r) % NP, and the side-stream ring's logical slot d (plane k + d) is
// physical slot (P + d) % NP.
// Point x0 + e of a row held in a ring slot in the x-shift mode (x0 the
// piece's first column, e = v + dx, |dx| <= H): component e mod VEC of lane
// floor(e / VEC); a value m lanes away takes m wave-wide shifts, the lane at
// the wave's end receiving the halo point the shift brings in (left halo
// h[i] = x0 - H + i on lane 0, right halo h[i] = x0 + TXW + i on lane 63).
// All 64 lanes must be active.
template <class C, class F, int VEC, int NH>
__device__ __forceinline__ F slot_xat(const Slot<F, VEC, NH>& s, int v, int dx)
{
    const int e = v + dx;
    const int q = floor_div(e, VEC), r = e - q * VEC;
    F t = s.v[r];
#pragma unroll
    for (int m = 1; m <= -q; m++) t = dpp_shift<KG_DPP_WAVE_SHR1>(s.h[C::H - m * VEC + r], t);
#pragma unroll
    for (int m = 1; m <= q; m++) t = dpp_shift<KG_DPP_WAVE_SHL1>(s.h[(m - 1) * VEC + r], t);
    return t;
}

template <class C, int P = 0>
struct View {
    using F = typename C::F;
    const F* lds;                              // [NLS][ROWS][VEC][LQ]
    const Slot<F, C::VEC, C::NH>* ring;        // [NP]
    const ASlot<F, C::VEC, C::NA>* aux;        // [NP]
    int k, row, cl;   // plane, tile row and lane-in-row of this lane
    bool wr;          // the lane's points lie in the (padded) row
    // main field at (x + v + dx, j + dy, k + dz) from LDS, |dx|,|dy| <= H,
    // -LL <= dz <= LR
    __device__ __forceinline__ F at(int v, int dx, int dy, int dz) const
    {
        const int s = (k + dz) % C::NLS;
        const int lq = floor_div(v + dx, C::VEC), vq = v + dx - lq * C::VEC;
        return lds[((s * C::ROWS + row + dy) * C::VEC + vq) * C::LQ + C::HL + cl + lq];
    }
    // store the lane's results (see Cfg::FULL)
    template <class S>
    __device__ __forceinline__ void put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                        const bool (&xm)[C::VEC]) const;
    // main field at (x + v, j, k + dz) from registers, -ZL <= dz <= ZHI
    __device__ __forceinline__ F col(int v, int dz) const
    {
        return ring[(P + C::ZLO + dz) % C::NP].v[v];
    }
    // x-shift mode: main field at (x + v + dx, j, k + dz) from registers,
    // |dx| <= H, -ZL <= dz <= ZHI.  Point x0 + e of the row (x0 the tile's
    // first column) is component e mod VEC of lane floor(e / VEC); a value m
    // lanes away takes m shifts, the row's end lane receiving the halo point
    // the shift brings in (left halo h[i] = x0 - H + i on lane 0, right halo
    // h[i] = x0 + TXW + i on lane 63).  All 64 lanes must be active.
    __device__ __forceinline__ F xat(int v, int dx, int dz) const
    {
        static_assert(C::XS, "xat needs the x-shift mode");
        return slot_xat<C>(ring[(P + C::ZLO + dz) % C::NP], v, dx);
    }
    __device__ __forceinline__ F side(int a, int v) const { return aux[P % C::NP].a[a][v]; }
};

// Main field at (x + v + dx, j, k) -- a neighbour along the lane's own row:
// from registers (in-lane components, DPP shifts across lanes) in the x-shift
// mode, from LDS otherwise.  Any view type of the engines and the batch.
template <class C, class V>
__device__ __forceinline__ typename C::F xrow(const V& p, int v, int dx)
{
    if constexpr (C::XS) return p.xat(v, dx, 0);
    else return p.at(v, dx, 0, 0);
}

// The plane loop shared by the tile shapes, planes k0 .. k1-1.  load(k, slot,
// want) fills a ring slot with plane k (want false: an empty resource, the
// slot reads 0), load_aux(k, aslot, want) the side streams, stage(k, slot)
// writes a slot to the LDS plane of k, compute(phase, k, ring, aux) runs the
// stencil of plane k after the barrier.
//
// Register rings of NP = RS + 1 physical slots.  The plane loop is unrolled
// NP times and each copy uses compile-time slot indices (phase P: logical
// slot r is physical (P + r) % NP), so a prefetch stays in flight in its own
// registers until it is used D planes later: no register copies, hence no
// wait for loads issued in the same iteration.  The planes the ring and LDS
// need are all < ns (the stencil reach ZHI <= the z-boundary width), so
// slots whose load is skipped are never read.
template <class C, bool SYNC = true, class Load, class LoadAux, class Stage, class Compute>
__device__ __forceinline__ void march(int k0, int k1, Load&& load, LoadAux&& load_aux,
                                      Stage&& stage, Compute&& compute)
{
    using F = typename C::F;
    Slot<F, C::VEC, C::NH> ring[C::NP];
    ASlot<F, C::VEC, C::NA> aux[C::NP];
#pragma unroll
    for (int r = 0; r < C::NP; r++) {
#pragma unroll
        for (int v = 0; v < C::VEC; v++) ring[r].v[v] = F(0);
#pragma unroll
        for (int i = 0; i < C::NH; i++) ring[r].h[i] = F(0);
    }
#pragma unroll
    for (int r = 0; r < C::RS; r++) load(k0 - C::ZLO + r, ring[r], true);
#pragma unroll
    for (int d = 0; d < C::D; d++) load_aux(k0 + d, aux[d], true);
    // LDS planes k0-LL .. k0+LR-1 (outside the ring when LL > ZL)
#pragma unroll
    for (int dz = -C::LL; dz < C::LR; dz++) {
        if (dz >= -C::ZLO) {
            stage(k0 + dz, ring[C::ZLO + dz]);
        } else {
            Slot<F, C::VEC, C::NH> t;
#pragma unroll
            for (int i = 0; i < C::NH; i++) t.h[i] = F(0);
#pragma unroll
            for (int v = 0; v < C::VEC; v++) t.v[v] = F(0);
            load(k0 + dz, t, true);
            stage(k0 + dz, t);
        }
    }
    auto plane = [&](auto phase, int k) {
        constexpr int P = decltype(phase)::value;
        stage(k + C::LR, ring[(P + C::ZLO + C::LR) % C::NP]);
        const int kp = k + C::ZHI + C::D;   // plane entering the ring
        load(kp, ring[(P + C::RS) % C::NP], kp <= k1 - 1 + C::ZHI);
        load_aux(k + C::D, aux[(P + C::D) % C::NP], k + C::D < k1);
        if constexpr (SYNC) __syncthreads();
        compute(phase, k, ring, aux);
    };
    int k = k0;
    for (; k + C::NP <= k1; k += C::NP)
        kg_static_for<C::NP>([&](auto p) { plane(p, k + decltype(p)::value); });
    kg_static_for<C::NP - 1>([&](auto p) {
        if (k + decltype(p)::value < k1) plane(p, k + decltype(p)::value);
    });
}

// One tile (bx, by, bz) of a grid: xy-interior [LO, n-HI), z-interior
// [ZB0, ns-ZB1), planes ZB0 + bz*kc .. +kc.  Every thread of the workgroup
// calls it (it contains barriers).  Op supplies
//   const S* src;            main field
//   const S* side[NAUX];     side streams (may be unused)
//   void run(const View<C>&, long long ob, int x, const bool (&xmask)[VEC]) const;
// run computes the VEC points of this lane at plane k, row j (ob = k*sz +
// j*sy, x the first point's column) and stores them; xmask[v] says whether
// point x+v is inside the interior.
template <class C, int LO, int HI, int ZB0, int ZB1, class Op>
__device__ __forceinline__ void plane_tile(typename C::F* lds, int nx, int ny,
                                           int ns, long long sy, long long sz, int bx, int by,
                                           int bz, int kc, const Op& op)
{
    static_assert(C::TY > 0, "tile has no computed rows");
    using S = typename C::S;
    using F = typename C::F;
    using T = typename C::T;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const int c = lane % C::LPR;
    const int row = w * C::RPW + lane / C::LPR;
    const int x0 = bx * C::TXW;
    const int y0 = LO + by * C::TY;
    const int k0 = ZB0 + bz * kc;
    const int k1 = min(k0 + kc, ns - ZB1);
    const int j = y0 - C::H + row;
    const int x = x0 + c * C::VEC;
    const bool row_ok = j >= 0 && j < ny;
    const bool comp_row = row >= C::H && row < C::ROWS - C::H && j >= LO && j < ny - HI;
    // x-halo columns of this lane (if any): hx + i for i < NH
    int hx = -1;
    if constexpr (C::XS) {
        if (c == 0) hx = x0 - C::H;
        else if (c == C::LPR - 1) hx = x0 + C::TXW;
    } else {
        if (c < C::H) hx = x0 - C::H + c;
        else if (c >= C::LPR - C::H) hx = x0 + C::TXW + (c - (C::LPR - C::H));
    }
    bool p_ok[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) p_ok[v] = row_ok && x + v < nx;
    // Per-lane byte offsets within a plane, fixed for the whole tile (KG_OOB:
    // nothing to load).  Planes are addressed through one resource each:
    // base + k*sz, sz*sizeof(S) bytes (at most 2^31; the launchers check).
    const unsigned plane_bytes = (unsigned)(sz * (long long)sizeof(S));
    const unsigned off_v = p_ok[0] ? (unsigned)(((long long)j * sy + x) * (long long)sizeof(S)) : KG_OOB;
    unsigned off_h[C::NH];
#pragma unroll
    for (int i = 0; i < C::NH; i++) {
        const bool h_ok = row_ok && hx >= 0 && hx + i >= 0 && hx + i < nx;
        off_h[i] = h_ok ? (unsigned)(((long long)j * sy + hx + i) * (long long)sizeof(S)) : KG_OOB;
    }
    const unsigned off_a = comp_row ? off_v : KG_OOB;
    // Resource for plane k of array base; empty unless want and 0 <= k < ns.
    auto plane_rsrc = [&](const S* base, int k, bool want) {
        const bool ok = want && k >= 0 && k < ns;
        return make_rsrc(base + (long long)(ok ? k : 0) * sz, ok ? plane_bytes : 0u);
    };

    // One access per lane and slot (VEC values, then the halo value); values
    // a lane does not load (zeros) are never read for an interior point.  With
    // VEC > 1 every lane with x < nx reads VEC values: callers guarantee that
    // rows hold a multiple of VEC elements (nx % VEC == 0 in the dense
    // layout, padded rows otherwise).
    auto load = [&](int k, Slot<F, C::VEC, C::NH>& s, bool want) {
        const __amdgpu_buffer_rsrc_t r = plane_rsrc(op.src, k, want);
        bload<T, C::VEC>(r, off_v, s.v);
#pragma unroll
        for (int i = 0; i < C::NH; i++) {
            F h[1];
            bload<T, 1>(r, off_h[i], h);
            s.h[i] = h[0];
        }
    };
    auto load_aux = [&](int k, ASlot<F, C::VEC, C::NA>& s, bool want) {
#pragma unroll
        for (int a = 0; a < C::NAUX; a++) bload<T, C::VEC>(plane_rsrc(op.side[a], k, want), off_a, s.a[a]);
    };
    auto stage = [&](int k, const Slot<F, C::VEC, C::NH>& s) {
        F* L = lds + ((k % C::NLS) * C::ROWS + row) * C::VEC * C::LQ;
#pragma unroll
        for (int v = 0; v < C::VEC; v++) L[v * C::LQ + C::HL + c] = s.v[v];
        // the x-halo values: tile columns e < 0 or >= TXW
        if constexpr (C::XS) {
            if (c == 0 || c == C::LPR - 1) {
#pragma unroll
                for (int i = 0; i < C::NH; i++) {
                    const int e = c == 0 ? i - C::H : C::TXW + i;
                    const int lq = floor_div(e, C::VEC);
                    L[(e - lq * C::VEC) * C::LQ + C::HL + lq] = s.h[i];
                }
            }
        } else {
            const bool lo_h = c < C::H, hi_h = c >= C::LPR - C::H;
            if (lo_h || hi_h) {
                const int e = lo_h ? c - C::H : C::TXW + (c - (C::LPR - C::H));
                const int lq = floor_div(e, C::VEC);
                L[(e - lq * C::VEC) * C::LQ + C::HL + lq] = s.h[0];
            }
        }
    };

    bool xmask[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) xmask[v] = x + v >= LO && x + v < nx - HI;
    march<C>(k0, k1, load, load_aux, stage, [&](auto phase, int k, const auto& ring, const auto& aux) {
        constexpr int P = decltype(phase)::value;
        if (comp_row) {
            View<C, P> vw{lds, ring, aux, k, row, c, x < sy};
            op.run(vw, (long long)k * sz + (long long)j * sy, x, xmask);
        }
    });
}

// ------------------------------------------------------------ packed rows
// A tile of whole padded rows, for grids whose rows are narrower than a wave:
// the workgroup's 64*W lanes cover RT = 64*W / (sy/VEC) consecutive rows of
// the plane back to back, sy/VEC lanes per row, rows y0-H .. y0-H+RT-1, of
// which the middle RT-2H are computed (the LPR-shaped tiles would leave the
// lanes past a short row's end idle).  All neighbours along x lie in the
// staged rows, so no x-halo is loaded.  Requires sy % VEC == 0 and RT > 2H
// (pack_rows() below).
template <class C>
struct Packed {
    static constexpr int NL = 64 * C::W + 2 * C::HL;         // lane slots per component
    static constexpr size_t LDS_BYTES = sizeof(typename C::F) * C::NLS * C::VEC * NL;
};

// Rows per packed tile (halo rows included) for row pitch sy; 0 if the shape
// does not fit.
template <class C>
__host__ __device__ inline int pack_rows(long long sy)
{
    if (sy <= 0 || sy % C::VEC) return 0;
    const long long rt = (64LL * C::W) / (sy / C::VEC);
    return rt > 2 * C::H ? (int)rt : 0;
}

// LDS slot layout [NLS][VEC][NL], component-major and lane-contiguous as in
// the LPR-shaped tiles: the lanes of a packed tile hold consecutive VEC-point
// pieces of the row-major plane, so a neighbour (v+dx, dy) of lane L is
// component (v+dx) mod VEC of lane L + floor((v+dx)/VEC) + dy*(sy/VEC).
template <class C, int P = 0>
struct PView {
    using F = typename C::F;
    const F* lds;                              // guard lanes skipped
    const Slot<F, C::VEC, C::NH>* ring;        // [NP]
    const ASlot<F, C::VEC, C::NA>* aux;        // [NP]
    int k, lane, cpr;                          // plane, lane, lanes per row
    bool wr;                                   // always true: packed lanes lie in their row
    __device__ __forceinline__ F at(int v, int dx, int dy, int dz) const
    {
        const int lq = floor_div(v + dx, C::VEC), vq = v + dx - lq * C::VEC;
        return lds[(((k + dz) % C::NLS) * C::VEC + vq) * Packed<C>::NL + lane + lq + dy * cpr];
    }
    template <class S>
    __device__ __forceinline__ void put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                        const bool (&xm)[C::VEC]) const;
    __device__ __forceinline__ F col(int v, int dz) const
    {
        return ring[(P + C::ZLO + dz) % C::NP].v[v];
    }
    __device__ __forceinline__ F side(int a, int v) const { return aux[P % C::NP].a[a][v]; }
};

// Tile (by, bz) of a grid in packed rows; the same contract as plane_tile.
template <class C, int LO, int HI, int ZB0, int ZB1, class Op>
__device__ __forceinline__ void packed_tile(typename C::F* lds, int nx, int ny, int ns, long long sy,
                                            long long sz, int by, int bz, int kc, const Op& op)
{
    using S = typename C::S;
    using F = typename C::F;
    using T = typename C::T;
    constexpr int NL = Packed<C>::NL;
    F* L0 = lds + C::HL;
    const int lane = (int)threadIdx.x;
    const int cpr = (int)(sy / C::VEC);              // lanes per row
    const int RT = pack_rows<C>(sy);
    const int row = lane / cpr;
    const int x = (lane - row * cpr) * C::VEC;
    const bool lane_ok = row < RT;
    const int y0 = LO + by * (RT - 2 * C::H);
    const int k0 = ZB0 + bz * kc;
    const int k1 = min(k0 + kc, ns - ZB1);
    const int j = y0 - C::H + row;
    const bool row_ok = lane_ok && j < ny;           // j >= 0: y0 >= LO >= H
    const bool comp_row = lane_ok && row >= C::H && row < RT - C::H && j < ny - HI;
    const unsigned plane_bytes = (unsigned)(sz * (long long)sizeof(S));
    const unsigned off_v = row_ok ? (unsigned)(((long long)j * sy + x) * (long long)sizeof(S)) : KG_OOB;
    const unsigned off_a = comp_row ? off_v : KG_OOB;
    auto plane_rsrc = [&](const S* base, int k, bool want) {
        const bool ok = want && k >= 0 && k < ns;
        return make_rsrc(base + (long long)(ok ? k : 0) * sz, ok ? plane_bytes : 0u);
    };
    auto load = [&](int k, Slot<F, C::VEC, C::NH>& s, bool want) {
        bload<T, C::VEC>(plane_rsrc(op.src, k, want), off_v, s.v);
    };
    auto load_aux = [&](int k, ASlot<F, C::VEC, C::NA>& s, bool want) {
#pragma unroll
        for (int a = 0; a < C::NAUX; a++) bload<T, C::VEC>(plane_rsrc(op.side[a], k, want), off_a, s.a[a]);
    };
    auto stage = [&](int k, const Slot<F, C::VEC, C::NH>& s) {
        if (lane_ok) {
            F* L = L0 + (k % C::NLS) * C::VEC * NL + lane;
#pragma unroll
            for (int v = 0; v < C::VEC; v++) L[v * NL] = s.v[v];
        }
    };
    bool xmask[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) xmask[v] = x + v >= LO && x + v < nx - HI;
    march<C>(k0, k1, load, load_aux, stage, [&](auto phase, int k, const auto& ring, const auto& aux) {
        constexpr int P = decltype(phase)::value;
        if (comp_row) {
            PView<C, P> vw{L0, ring, aux, k, lane, cpr, true};
            op.run(vw, (long long)k * sz + (long long)j * sy, x, xmask);
        }
    });
}

// ------------------------------------------------------------ direct rows
// Single-plane stencils (the 2D kinds, applied to each xy-plane on its own)
// without LDS: a lane owns VEC points of PY consecutive rows of plane k and
// loads the rows it needs (three VEC-point pieces each, at x-VEC, x, x+VEC)
// through the caches, all before the first use; the workgroup's W waves are
// stacked in y.  A tile of the plane engine holds one plane here and would
// pay its pipeline fill for every tile.  Requires H <= VEC and VEC-aligned,
// padded rows (sy % VEC == 0).
template <class C, int PY_ = (C::H == 1 ? 4 : 2)>   // default fits 120 VGPRs
struct Rows2 {
    static constexpr int PY = PY_;                       // output rows per lane
    static constexpr int NR = PY + 2 * C::H;             // rows held per lane
    static constexpr int TXW = 64 * C::VEC;              // tile width
    static constexpr int TY = C::W * PY;                 // tile rows
    static_assert(C::H <= C::VEC, "x-neighbours within the adjacent pieces");
};

template <class C, int R, int NR>
struct RView {
    using F = typename C::F;
    const F (&win)[NR][3 * C::VEC];   // rows j0-H .., points x-VEC .. x+2VEC-1
    bool wr;
    __device__ __forceinline__ F at(int v, int dx, int dy, int) const
    {
        return win[R + C::H + dy][C::VEC + v + dx];
    }
    __device__ __forceinline__ F col(int v, int) const { return win[R + C::H][C::VEC + v]; }
    __device__ __forceinline__ F side(int, int) const { return F(0); }
    template <class S>
    __device__ __forceinline__ void put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                        const bool (&xm)[C::VEC]) const;
};

// Tile (bx, by) of plane k: xy-interior [LO, n-HI).  No barriers: lanes with
// nothing to do return at once.
template <class C, int LO, int HI, class Op, int PY = Rows2<C>::PY>
__device__ __forceinline__ void rows2_tile(int nx, int ny, long long sy, long long sz, int bx,
                                           int by, int k, const Op& op)
{
    using R2 = Rows2<C, PY>;
    using S = typename C::S;
    using F = typename C::F;
    using T = typename C::T;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const int x = bx * R2::TXW + lane * C::VEC;
    const int j0 = LO + (by * C::W + w) * R2::PY;
    if (x >= sy || j0 >= ny - HI) return;
    // pieces beside the row ends are clamped into the row: they then feed
    // only boundary points, which take the centre value
    const int xl = x >= C::VEC ? x - C::VEC : x;
    const int xr = x + C::VEC < sy ? x + C::VEC : x;
    const gmem<S>* base = as_global(op.src) + (long long)k * sz;
    F win[R2::NR][3 * C::VEC];
#pragma unroll
    for (int q = 0; q < R2::NR; q++) {
        const gmem<S>* row = base + (long long)min(j0 - C::H + q, ny - 1) * sy;
        F a[C::VEC], b[C::VEC], c[C::VEC];
        vload<T, C::VEC>(row + xl, a);
        vload<T, C::VEC>(row + x, b);
        vload<T, C::VEC>(row + xr, c);
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
            win[q][v] = a[v];
            win[q][C::VEC + v] = b[v];
            win[q][2 * C::VEC + v] = c[v];
        }
    }
    bool xmask[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) xmask[v] = x + v >= LO && x + v < nx - HI;
    kg_static_for<R2::PY>([&](auto rr) {
        constexpr int R = decltype(rr)::value;
        const int j = j0 + R;
        if (j < ny - HI)
            op.run(RView<C, R, R2::NR>{win, true}, (long long)k * sz + (long long)j * sy, x, xmask);
    });
}

// Single-grid launch of the direct row body over every plane of a dense
// grid: tiles enumerated x fastest, then y, then plane, in tile order ORD
// (kg_tile_order).
template <class C, int LO, int HI, class Op, int PY, int ORD = 0>
__global__ __launch_bounds__(64 * C::W)
void rows2_kernel(int nx, int ny, int ntx, int nty, Op op)
{
    const unsigned b = kg_tile_order<ORD>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int bx = (int)(b % (unsigned)ntx);
    const int by = (int)((b / (unsigned)ntx) % (unsigned)nty);
    const int k = (int)(b / ((unsigned)ntx * (unsigned)nty));
    rows2_tile<C, LO, HI, Op, PY>(nx, ny, nx, (long long)nx * ny, bx, by, k, op);
}

// ------------------------------------------------------------ column march
// For stencils that read only the xy-plane they write (gaussblur per plane,
// the 2D kinds).  A wave owns one row piece of 64*VEC points of one plane
// and marches it along y through a chunk of rows, keeping rows j-H .. j+H
// of its own columns (plus the D-row prefetch) in a register ring; the
// x-neighbours come from the adjacent lanes by wave-wide DPP shifts, lane 0
// holding the H points left of the piece and lane 63 the H points right of
// it.  Nothing is staged in LDS and there are no barriers, so the waves of a
// workgroup run independently.  C: x-shift mode, LPR = 64, LL = LR = 0,
// ZL = ZR = H (the ring's "planes" are rows).
template <class C, int P = 0>
struct CView {
    using F = typename C::F;
    const Slot<F, C::VEC, C::NH>* ring;   // [NP]
    const F (*sh)[2 * C::H];              // [NP]: the H points left and right of the piece
    bool wr;                              // the lane's points lie in the (padded) row
    // (x + v + dx, j + dy) of the lane's plane; |dx|, |dy| <= H
    __device__ __forceinline__ F at(int v, int dx, int dy, int) const
    {
        const int s = (P + C::ZLO + dy) % C::NP, e = v + dx;
        if (e < 0) return sh[s][C::H + e];
        if (e >= C::VEC) return sh[s][C::H + e - C::VEC];
        return ring[s].v[e];
    }
    __device__ __forceinline__ F col(int v, int) const { return ring[(P + C::ZLO) % C::NP].v[v]; }
    __device__ __forceinline__ F side(int, int) const { return F(0); }
    template <class S>
    __device__ __forceinline__ void put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                        const bool (&xm)[C::VEC]) const;
};

template <class C>
struct ColMarch {
    static_assert(C::XS && C::LPR == 64 && C::LL == 0 && C::LR == 0 && C::ZL == C::H &&
                  C::ZR == C::H && C::NAUX == 0 && C::H <= C::VEC, "column march configuration");
    static constexpr int TXW = 64 * C::VEC;   // piece width
};

// Tile (bx, by, bk) of the xy-interior [LO, n-HI) x all planes: the W
// waves form WZ columns of planes by W/WZ chunks of rows; wave w marches the
// kc rows of row chunk by*(W/WZ) + w/WZ in plane bk*WZ + w%WZ.  Every lane
// of a wave with work stays active to the end (the DPP shifts read all 64
// lanes); waves without rows return at once.
template <class C, int LO, int HI, class Op, int WZ = 1>
__device__ __forceinline__ void colmarch_tile(int nx, int ny, int ns, long long sy, long long sz,
                                              int bx, int by, int bk, int kc, const Op& op)
{
    static_assert(C::W % WZ == 0, "waves form WZ whole columns");
    using S = typename C::S;
    using F = typename C::F;
    using T = typename C::T;
    (void)ColMarch<C>{};
    const int lane = (int)(threadIdx.x & 63);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int k = bk * WZ + w % WZ;
    const int y0 = LO + (by * (C::W / WZ) + w / WZ) * kc;
    const int y1 = min(y0 + kc, ny - HI);
    if (k >= ns || y0 >= y1) return;
    const int x0 = bx * ColMarch<C>::TXW;
    const int x = x0 + lane * C::VEC;
    const unsigned off_v = x < nx ? (unsigned)((long long)x * (long long)sizeof(S)) : KG_OOB;
    // the H halo points of lane 0 (left of the piece) or lane 63 (right of
    // it) in one load: they lie in one row whenever the first does (nx and
    // x0 are multiples of VEC >= H), and start H*sizeof(S)-aligned
    const int hx = lane == 0 ? x0 - C::H : lane == 63 ? x0 + ColMarch<C>::TXW : -1;
    const unsigned off_h = hx >= 0 && hx < nx ? (unsigned)((long long)hx * (long long)sizeof(S)) : KG_OOB;
    const S* plane = op.src + (long long)k * sz;
    const unsigned row_bytes = (unsigned)(sy * (long long)sizeof(S));
    // row jr of the plane through its own resource; empty unless want and
    // 0 <= jr < ny
    auto load = [&](int jr, Slot<F, C::VEC, C::NH>& s, bool want) {
        const bool ok = want && jr >= 0 && jr < ny;
        const __amdgpu_buffer_rsrc_t r = make_rsrc(plane + (long long)(ok ? jr : 0) * sy,
                                                   ok ? row_bytes : 0u);
        bload<T, C::VEC>(r, off_v, s.v);
        bload<T, C::NH>(r, off_h, s.h);
    };
    auto no_aux = [&](int, ASlot<F, C::VEC, C::NA>&, bool) {};
    auto no_stage = [&](int, const Slot<F, C::VEC, C::NH>&) {};
    bool xmask[C::VEC];
#pragma unroll
    for (int v = 0; v < C::VEC; v++) xmask[v] = x + v >= LO && x + v < nx - HI;
    // The points beside the piece of each row in the window (rows j-H ..
    // j+H), taken from the neighbouring lanes once per row, when the row
    // enters the window, and kept with the row's ring slot.
    F sh[C::NP][2 * C::H];
    auto shift_in = [&](const Slot<F, C::VEC, C::NH>& r, F (&out)[2 * C::H]) {
#pragma unroll
        for (int i = 0; i < C::H; i++) {
            out[i] = dpp_shift<KG_DPP_WAVE_SHR1>(r.h[i], r.v[C::VEC - C::H + i]);
            out[C::H + i] = dpp_shift<KG_DPP_WAVE_SHL1>(r.h[i], r.v[i]);
        }
    };
    march<C, false>(y0, y1, load, no_aux, no_stage,
                    [&](auto phase, int j, const auto& ring, const auto&) {
        constexpr int P = decltype(phase)::value;
        if (j == y0) {   // the first row: the rows below and above it enter together
#pragma unroll
            for (int r = 0; r < C::ZLO + C::ZR; r++) shift_in(ring[(P + r) % C::NP], sh[(P + r) % C::NP]);
        }
        shift_in(ring[(P + C::ZLO + C::ZR) % C::NP], sh[(P + C::ZLO + C::ZR) % C::NP]);
        CView<C, P> vw{ring, (const F(*)[2 * C::H])sh, x < sy};
        op.run(vw, (long long)k * sz + (long long)j * sy, x, xmask);
    });
}

// Single-grid launch of the column march over every plane of a dense grid:
// tiles (bx, by, k), x fastest, then y, then plane, in tile order ORD
// (kg_tile_order).
template <class C, int LO, int HI, class Op, int ORD = 0>
__global__ __launch_bounds__(64 * C::W)
void colmarch_kernel(int nx, int ny, int ns, int ntx, int nty, int kc, Op op)
{
    const unsigned b = kg_tile_order<ORD>(blockIdx.x, gridDim.x, (unsigned)ntx);
    const int bx = (int)(b % (unsigned)ntx);
    const int by = (int)((b / (unsigned)ntx) % (unsigned)nty);
    const int k = (int)(b / ((unsigned)ntx * (unsigned)nty));
    colmarch_tile<C, LO, HI, Op>(nx, ny, ns, nx, (long long)nx * ny, bx, by, k, kc, op);
}

// Single-grid launch in the dense layout: one workgroup per tile of kc
// planes, tiles enumerated x fastest, then y, then z-chunk, in tile order ORD
// (kg_tile_order: hardware, XCD-contiguous or chunks of ORD tile rows).
// z-interior [ZB0, ns-ZB1), by default the same as the xy-interior.
// ORD < 0: tiles enumerated z-chunk fastest, then x, then y, and dealt to
// the XCDs in runs of -ORD whole z-columns (kg_xcd_chunk): the z-chunks of a
// tile column, which share the planes at their ends, run on one XCD.
// ORD == KG_ORD_REGION: each XCD owns a fixed region of the xy-tiles, an
// eighth of them in x-fastest order, and walks it z-chunk by z-chunk; the
// x-, y- and z-neighbours of a tile (which read its halo lines and end
// planes) then run on the same XCD, most of them at the same time.  The
// launch has 8 * region_tiles(ntx*nty) * nkc workgroups (region_blocks);
// slots past the last tile of a region return at once.
constexpr int KG_ORD_REGION = 100;
__host__ __device__ inline unsigned region_tiles(unsigned tiles_per_layer)
{
    return (tiles_per_layer + KG_NUM_XCD - 1) / KG_NUM_XCD;
}
__host__ __device__ inline unsigned region_blocks(unsigned tiles_per_layer, unsigned nkc)
{
    return KG_NUM_XCD * region_tiles(tiles_per_layer) * nkc;
}

template <class C, int LO, int HI, class Op, int ZB0 = LO, int ZB1 = HI, int ORD = 1>
__global__ __launch_bounds__(64 * C::W)
void plane_kernel(int nx, int ny, int ns, int ntx, int nty, int kc, Op op)
{
    __shared__ typename C::F lds[C::LDS_BYTES / sizeof(typename C::F)];
    int bx, by, bz;
    if constexpr (ORD == KG_ORD_REGION) {
        const unsigned L = (unsigned)ntx * (unsigned)nty, R = region_tiles(L);
        const unsigned x = blockIdx.x % KG_NUM_XCD, s = blockIdx.x / KG_NUM_XCD;
        const unsigned t = x * R + s % R;   // tile of the layer
        if (t >= L) return;                 // past the last tile: the whole workgroup leaves
        bz = (int)(s / R);
        bx = (int)(t % (unsigned)ntx);
        by = (int)(t / (unsigned)ntx);
    } else if constexpr (ORD < 0) {
        const unsigned nkc = gridDim.x / ((unsigned)ntx * (unsigned)nty);
        unsigned b = kg_xcd_chunk(blockIdx.x, gridDim.x, nkc * (unsigned)(-ORD));
        bz = (int)(b % nkc);
        b /= nkc;
        bx = (int)(b % (unsigned)ntx);
        by = (int)(b / (unsigned)ntx);
    } else {
        unsigned b = kg_tile_order<ORD>(blockIdx.x, gridDim.x, (unsigned)ntx);
        bx = (int)(b % (unsigned)ntx);
        b /= (unsigned)ntx;
        by = (int)(b % (unsigned)nty);
        bz = (int)(b / (unsigned)nty);
    }
    plane_tile<C, LO, HI, ZB0, ZB1, Op>(lds, nx, ny, ns, nx, (long long)nx * ny, bx, by, bz,
                                        kc, op);
}

// Store the VEC results of a lane: one vector non-temporal store when all
// points are interior, else per point.  Values are rounded to the storage type.
template <class T, int VEC>
__device__ __forceinline__ void store(typename T::S* dst, long long ob, int x,
                                      const typename T::F (&r)[VEC], const bool (&xmask)[VEC])
{
    using S = typename T::S;
    gmem<S>* d = as_global(dst);
    bool all = true;
#pragma unroll
    for (int v = 0; v < VEC; v++) all = all && xmask[v];
    if (VEC > 1 && all) {
        using R = typename Raw<VEC * (int)sizeof(S)>::t;
        S s[VEC];
#pragma unroll
        for (int v = 0; v < VEC; v++) s[v] = st_cvt<T>(r[v]);
        R raw;
        __builtin_memcpy(&raw, s, sizeof raw);
        __builtin_nontemporal_store(raw, (gmem<R>*)(d + ob + x));
    } else {
#pragma unroll
        for (int v = 0; v < VEC; v++)
            if (xmask[v]) __builtin_nontemporal_store(st_cvt<T>(r[v]), d + ob + x + v);
    }
}
template <int VEC>
__device__ __forceinline__ void store(double* dst, long long ob, int x, const double (&r)[VEC],
                                      const bool (&xmask)[VEC])
{
    store<F64, VEC>(dst, ob, x, r, xmask);
}

// The views' store: interior points only, or with Cfg::FULL the lane's whole
// vector, its non-interior points taking the input level's centre values.
template <class C, class V, class S>
__device__ __forceinline__ void put_row(const V& p, S* dst, long long ob, int x,
                                        typename C::F (&r)[C::VEC], const bool (&xm)[C::VEC])
{
    if constexpr (C::FULL) {
        if (!p.wr) return;
        bool all[C::VEC];
#pragma unroll
        for (int v = 0; v < C::VEC; v++) {
            if (!xm[v]) r[v] = p.col(v, 0);
            all[v] = true;
        }
        store<typename C::T, C::VEC>(dst, ob, x, r, all);
    } else {
        store<typename C::T, C::VEC>(dst, ob, x, r, xm);
    }
}
template <class C, int P>
template <class S>
__device__ __forceinline__ void View<C, P>::put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                                const bool (&xm)[C::VEC]) const
{
    put_row<C>(*this, dst, ob, x, r, xm);
}
template <class C, int R, int NR>
template <class S>
__device__ __forceinline__ void RView<C, R, NR>::put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                                     const bool (&xm)[C::VEC]) const
{
    put_row<C>(*this, dst, ob, x, r, xm);
}
template <class C, int P>
template <class S>
__device__ __forceinline__ void CView<C, P>::put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                                 const bool (&xm)[C::VEC]) const
{
    put_row<C>(*this, dst, ob, x, r, xm);
}
template <class C, int P>
template <class S>
__device__ __forceinline__ void PView<C, P>::put(S* dst, long long ob, int x, F (&r)[C::VEC],
                                                 const bool (&xm)[C::VEC]) const
{
    put_row<C>(*this, dst, ob, x, r, xm);
}

}  // namespace kgp
