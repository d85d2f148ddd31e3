// kg_batch.hip -- many independent grids stepped by one persistent launch.
//
// A batch owns a set of 3D grids of mixed kind (laplacian, wave13pt,
// gaussblur per xy-plane), storage type (FP64, FP32) and size.  Their levels
// live in HBM for the life of the batch (rows padded to 16 bytes) and are
// carved from an arena, so adding and removing grids never calls hipFree on
// the hot path.  One step of every grid is ONE launch of batch_k:
//
//   * the work list is every (grid, tile) pair, tiles cut per grid with the
//     tile shape of its size class (lanes per row 16/32/64 and planes per
//     tile), largest grids first;
//   * each resident workgroup pulls the next item from an atomic counter and
//     runs the plane engine's tile body for that grid's kind and type -- the
//     same bodies as the single-grid kernels, so a grid's result is bit for
//     bit what the single-grid entry point gives;
//   * the level a grid reads and writes follows from a device-resident step
//     counter, which the last workgroup to finish advances (and the work
//     counter is reset there), so a captured graph of K launches replays
//     correctly at any position of the time loop.
//
// The host API is C (kg_batch_* in include/kg_stencil.h).  kg_batch_tune
// times the candidate tile shapes of every size class present on the batch's
// own grids and keeps the fastest.
#include "kg_ops3d.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <new>
#include <vector>

namespace {

using namespace kgops;

// ------------------------------------------------------------ variants
#ifndef KG_BATCH_WAVES
#define KG_BATCH_WAVES 16
#endif
#ifndef KG_BATCH_DEPTH
#define KG_BATCH_DEPTH 2
#endif
#ifndef KG_BATCH_WPE
#define KG_BATCH_WPE 4
#endif
#ifndef KG_BATCH_VEC64
#define KG_BATCH_VEC64 2
#endif
#ifndef KG_BATCH_VEC32
#define KG_BATCH_VEC32 4
#endif
constexpr int BW = KG_BATCH_WAVES;   // waves per workgroup
constexpr int BD = KG_BATCH_DEPTH;   // prefetch depth (planes)
constexpr int NKIND = 5;        // KG_KIND_LAPLACIAN, _WAVE13PT, _GAUSSBLUR3, _JACOBI2, _LIFE2
constexpr int NTYPE = 2;        // KG_F64, KG_F32
constexpr int NLPR = 3;         // plane-engine shapes: lanes per row 64, 32, 16
constexpr int FLAT = NLPR;      // shape index of the flat (cache-served) body
constexpr int PACK = NLPR + 1;  // shape index of packed whole rows
constexpr int ROW2 = NLPR + 2;  // shape index of the direct row body (planar kinds only)
constexpr int COL = NLPR + 3;   // shape index of the column march (planar kinds only)
constexpr int NSHAPE = NLPR + 4;
constexpr int NVAR = NKIND * NTYPE * NSHAPE;
// API codes: 0 flat, -1 packed rows, -2 direct rows, -3 column march
constexpr int LPRS[NSHAPE] = {64, 32, 16, 0, -1, -2, -3};
// Kinds whose stencil reads only the plane it writes (gaussblur3 and the 2D
// kinds, applied to every xy-plane): they can run the direct row body.
constexpr bool is_planar(int kind)
{
    return kind == KG_KIND_GAUSSBLUR3 || kind == KG_KIND_JACOBI2 || kind == KG_KIND_LIFE2;
}
constexpr int KCS[] = {8, 16, 32, 64};
constexpr int NKC = sizeof(KCS) / sizeof(KCS[0]);

template <int K> struct Kind;
template <> struct Kind<KG_KIND_LAPLACIAN> {
    using B = Bounds<LaplacianOp>;
    using Sh = Shape<LaplacianOp>;
    template <class C> using Op = LaplacianOp<C>;
};
template <> struct Kind<KG_KIND_WAVE13PT> {
    using B = Bounds<Wave13ptOp>;
    using Sh = Shape<Wave13ptOp>;
    template <class C> using Op = Wave13ptOp<C>;
};
template <> struct Kind<KG_KIND_GAUSSBLUR3> {
    using B = Bounds<Gaussblur3Op>;
    using Sh = Shape<Gaussblur3Op>;
    template <class C> using Op = Gaussblur3Op<C>;
};

template <> struct Kind<KG_KIND_JACOBI2> {
    using B = Bounds<Jacobi2Op>;
    using Sh = Shape<Jacobi2Op>;
    template <class C> using Op = Jacobi2Op<C>;
};
template <> struct Kind<KG_KIND_LIFE2> {
    using B = Bounds<Life2Op>;
    using Sh = Shape<Life2Op>;
    template <class C> using Op = Life2Op<C>;
};

template <int DT> struct TypeOf;
template <> struct TypeOf<KG_F64> { using T = kgp::F64; static constexpr int VEC = KG_BATCH_VEC64; };
template <> struct TypeOf<KG_F32> { using T = kgp::F32; static constexpr int VEC = KG_BATCH_VEC32; };

// Prefetch depth (planes) per kind and type.  With one 16-wave workgroup
// per CU the bytes in flight per CU are about depth x tile rows x row bytes,
// and the achieved bandwidth follows them; each body takes the deepest
// prefetch that fits in 128 VGPRs without spilling (the ring holds
// ZL + ZHI + depth + 1 planes of VEC + 1 values).
#ifndef KG_BATCH_DEEP
#define KG_BATCH_DEEP 0
#endif
constexpr int depth_of(int kind, int dt)
{
    if (KG_BATCH_DEEP > 0) return KG_BATCH_DEEP;
    if (kind == KG_KIND_WAVE13PT) return 3;
    if (kind == KG_KIND_GAUSSBLUR3) return dt == KG_F64 ? 2 : 3;
    return dt == KG_F64 ? 3 : 4;
}

// Points per lane: 16-byte lanes, except wave13pt (13 neighbours, 3 levels),
// whose 16-byte FP64/FP32 bodies do not fit 128 VGPRs with a useful prefetch.
#ifndef KG_BATCH_WVEC64
#define KG_BATCH_WVEC64 1
#endif
#ifndef KG_BATCH_WVEC32
#define KG_BATCH_WVEC32 2
#endif
constexpr int vec_of(int kind, int dt)
{
    if (kind == KG_KIND_WAVE13PT) return dt == KG_F64 ? KG_BATCH_WVEC64 : KG_BATCH_WVEC32;
    return dt == KG_F64 ? TypeOf<KG_F64>::VEC : TypeOf<KG_F32>::VEC;
}

// Whole rows are written (Cfg::FULL): a batch grid's levels share their
// boundary values, and a row whose edge cache lines were left partly
// unwritten measured up to 2.3x slower to store (64-point FP32 rows).
template <int K, int DT, int LPR>
using VarCfg = kgp::Cfg<vec_of(K, DT), BW, Kind<K>::Sh::H, Kind<K>::Sh::LL, Kind<K>::Sh::LR,
                        Kind<K>::Sh::ZL, Kind<K>::Sh::ZR, depth_of(K, DT), Kind<K>::Sh::NAUX, 32,
                        typename TypeOf<DT>::T, LPR, true>;

constexpr int var_id(int kind, int dt, int li) { return (kind * NTYPE + dt) * NSHAPE + li; }

// Column march (kgp::colmarch_tile, planar kinds): the 16 waves of a tile
// form COLZ columns of planes by 16/COLZ chunks of kc rows; each wave keeps
// rows j-H .. j+H plus COLD prefetched rows of its columns in registers.
#ifndef KG_BATCH_COLZ
#define KG_BATCH_COLZ 4
#endif
#ifndef KG_BATCH_COLD64
#define KG_BATCH_COLD64 1
#endif
#ifndef KG_BATCH_COLD32
#define KG_BATCH_COLD32 2
#endif
constexpr int COLZ = KG_BATCH_COLZ;
template <int K, int DT>
using ColCfg = kgp::Cfg<TypeOf<DT>::VEC, BW, Kind<K>::Sh::H, 0, 0, Kind<K>::Sh::H, Kind<K>::Sh::H,
                        DT == KG_F64 ? KG_BATCH_COLD64 : KG_BATCH_COLD32, 0, 32,
                        typename TypeOf<DT>::T, 64, true, true>;

// Flat body configuration: one point per lane, operands from global memory.
template <int K, int DT>
using FlatCfg = kgp::Cfg<1, BW, Kind<K>::Sh::H, Kind<K>::Sh::LL, Kind<K>::Sh::LR, Kind<K>::Sh::ZL,
                         Kind<K>::Sh::ZR, BD, Kind<K>::Sh::NAUX, 32, typename TypeOf<DT>::T, 64>;

// Shape facts of every variant, for the host (tile counts) and the LDS size.
struct VarInfo {
    int txw, ty, lo, hi, zlo, zhi, vec, esize, h;
    size_t lds;
};

template <int K, int DT, int LI>
constexpr VarInfo var_info()
{
    using B = typename Kind<K>::B;
    if constexpr (LI == FLAT) {
        using C = FlatCfg<K, DT>;
        return VarInfo{0, 0, B::LO, B::HI, B::ZLO, B::ZHI, 1, (int)sizeof(typename C::S), C::H, 0};
    } else if constexpr (LI == PACK) {
        using C = VarCfg<K, DT, 64>;
        return VarInfo{0, 0, B::LO, B::HI, B::ZLO, B::ZHI, C::VEC, (int)sizeof(typename C::S),
                       C::H, kgp::Packed<C>::LDS_BYTES};
    } else if constexpr (LI == ROW2) {
        if constexpr (is_planar(K)) {
            using C = VarCfg<K, DT, 64>;
            return VarInfo{kgp::Rows2<C>::TXW, kgp::Rows2<C>::TY, B::LO, B::HI, B::ZLO, B::ZHI,
                           C::VEC, (int)sizeof(typename C::S), C::H, 0};
        } else {
            return var_info<K, DT, 0>();   // never used: build_work maps it to shape 0
        }
    } else if constexpr (LI == COL) {
        if constexpr (is_planar(K)) {
            using C = ColCfg<K, DT>;
            return VarInfo{kgp::ColMarch<C>::TXW, 0, B::LO, B::HI, B::ZLO, B::ZHI, C::VEC,
                           (int)sizeof(typename C::S), C::H, 0};
        } else {
            return var_info<K, DT, 0>();   // never used: build_work maps it to shape 0
        }
    } else {
        using C = VarCfg<K, DT, LPRS[LI]>;
        return VarInfo{C::TXW, C::TY, B::LO, B::HI, B::ZLO, B::ZHI, C::VEC,
                       (int)sizeof(typename C::S), C::H, C::LDS_BYTES};
    }
}

#define KG_VAR_LIST(X)                                                       \
    X(0, 0, 0) X(0, 0, 1) X(0, 0, 2) X(0, 1, 0) X(0, 1, 1) X(0, 1, 2)        \
    X(1, 0, 0) X(1, 0, 1) X(1, 0, 2) X(1, 1, 0) X(1, 1, 1) X(1, 1, 2)        \
    X(2, 0, 0) X(2, 0, 1) X(2, 0, 2) X(2, 1, 0) X(2, 1, 1) X(2, 1, 2)        \
    X(3, 0, 0) X(3, 0, 1) X(3, 0, 2) X(3, 1, 0) X(3, 1, 1) X(3, 1, 2)        \
    X(4, 0, 0) X(4, 0, 1) X(4, 0, 2) X(4, 1, 0) X(4, 1, 1) X(4, 1, 2)

#define KG_FLAT_LIST(X) \
    X(0, 0) X(0, 1) X(1, 0) X(1, 1) X(2, 0) X(2, 1) X(3, 0) X(3, 1) X(4, 0) X(4, 1)

#define KG_VAR_INFO6(K, DT) var_info<K, DT, 0>(), var_info<K, DT, 1>(), var_info<K, DT, 2>(), \
                            var_info<K, DT, FLAT>(), var_info<K, DT, PACK>(), var_info<K, DT, ROW2>(), \
                            var_info<K, DT, COL>(),
constexpr VarInfo VARS[NVAR] = {KG_FLAT_LIST(KG_VAR_INFO6)};
#define KG_ROW2_LIST(X) X(2, 0) X(2, 1) X(3, 0) X(3, 1) X(4, 0) X(4, 1)

// Rows of a packed tile (halo rows included) for a grid of variant v with
// row pitch sy; 0 when the grid's rows do not pack (see kgp::pack_rows).
inline int pack_rows_of(const VarInfo& v, long long sy)
{
    if (sy <= 0 || sy % v.vec) return 0;
    const long long rt = (64LL * BW) / (sy / v.vec);
    return rt > 2 * v.h ? (int)rt : 0;
}

constexpr size_t max_lds()
{
    size_t m = 0;
    for (int v = 0; v < NVAR; v++) m = VARS[v].lds > m ? VARS[v].lds : m;
    return m;
}
constexpr size_t BATCH_LDS = max_lds();

// ------------------------------------------------------- device structures
struct GridDesc {
    int var;              // variant id
    int nx, ny, ns;
    int kc;               // planes per tile
    int nlev;             // 2 (laplacian, gaussblur) or 3 (wave13pt)
    long long sy, sz;     // row and plane pitch (elements)
    long long step0;      // device step at which this grid's local step 0 runs
    void* lev[3];
    double c[7];          // kind coefficients (gaussblur: f, s0, s1, s2, s4, s5, s8)
};

struct Item {
    int grid, bx, by, bz;
    int var;              // the grid's variant id (read with the item)
};

struct BatchState {
    unsigned long long step;          // steps completed by the batch
    unsigned counter[KG_NUM_XCD];     // next item of each XCD's part of the list
    unsigned done;                    // workgroups finished this launch
};

// Not inlined: each variant then gets its own register allocation and the
// kernel needs the maximum of them (up to 116), not their union (which spills
// at the 128 registers a 1024-thread workgroup allows).
// Tile staging memory of batch_k, sized for the largest variant.  At file
// scope so that the non-inlined tile bodies address it as LDS.
__shared__ __attribute__((aligned(16))) char batch_lds[BATCH_LDS];

// Arguments of a non-inlined device function arrive in vector registers;
// these mark workgroup-uniform ones as such, so that the descriptor is read
// with scalar loads into scalar registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long uni(unsigned long long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ long long uni(long long v)
{
    return (long long)uni((unsigned long long)v);
}
template <class T>
__device__ __forceinline__ T* uni(T* p)
{
    return (T*)uni((unsigned long long)(uintptr_t)p);
}

// The op of grid g at batch step `step`: levels chosen from the grid's local
// step, coefficients converted to the arithmetic type.
template <int K, class C>
__device__ __forceinline__ typename Kind<K>::template Op<C> make_op(const GridDesc& g,
                                                                      unsigned long long step)
{
    using Op = typename Kind<K>::template Op<C>;
    using S = typename C::S;
    using F = typename C::F;
    // Level pointers read back into scalar registers: the tile bodies build
    // buffer resources from them, which must be workgroup-uniform (a resource
    // in vector registers makes the compiler loop over the lanes per load).
    const long long ls = uni((long long)step - g.step0);
    void* const lev[3] = {uni(g.lev[0]), uni(g.lev[1]), uni(g.lev[2])};
    Op op{};
    if constexpr (K == KG_KIND_WAVE13PT) {
        op.side[0] = (const S*)lev[ls % 3];
        op.src = (const S*)lev[(ls + 1) % 3];
        op.dst = (S*)lev[(ls + 2) % 3];
        op.m0 = (F)g.c[0];
        op.m1 = (F)g.c[1];
        op.m2 = (F)g.c[2];
    } else if constexpr (K == KG_KIND_LAPLACIAN) {
        op.src = (const S*)lev[ls & 1];
        op.dst = (S*)lev[(ls + 1) & 1];
        op.alpha = (F)g.c[0];
        op.beta = (F)g.c[1];
    } else if constexpr (K == KG_KIND_JACOBI2) {
        op.src = (const S*)lev[ls & 1];
        op.dst = (S*)lev[(ls + 1) & 1];
        op.c0 = (F)g.c[0];
        op.c1 = (F)g.c[1];
        op.c2 = (F)g.c[2];
    } else if constexpr (K == KG_KIND_LIFE2) {
        op.src = (const S*)lev[ls & 1];
        op.dst = (S*)lev[(ls + 1) & 1];
    } else {
        op.src = (const S*)lev[ls & 1];
        op.dst = (S*)lev[(ls + 1) & 1];
        op.f = (F)g.c[0];
        op.s0 = (F)g.c[1];
        op.s1 = (F)g.c[2];
        op.s2 = (F)g.c[3];
        op.s4 = (F)g.c[4];
        op.s5 = (F)g.c[5];
        op.s8 = (F)g.c[6];
    }
    return op;
}

template <int K, int DT, int LPR>
__device__ __attribute__((noinline)) void run_tile(const GridDesc* __restrict__ gp_, int bx_,
                                                   int by_, int bz_, unsigned long long step_)
{
    const GridDesc& g = *uni(gp_);
    const int bx = uni(bx_), by = uni(by_), bz = uni(bz_);
    const unsigned long long step = uni(step_);
    using C = VarCfg<K, DT, LPR>;
    using Op = typename Kind<K>::template Op<C>;
    using F = typename C::F;
    using B = typename Kind<K>::B;
    const Op op = make_op<K, C>(g, step);
    kgp::plane_tile<C, B::LO, B::HI, B::ZLO, B::ZHI, Op>(
        (F*)batch_lds, uni(g.nx), uni(g.ny), uni(g.ns), uni(g.sy), uni(g.sz), bx,
        by, bz, uni(g.kc), op);
}

template <int K, int DT>
__device__ __attribute__((noinline)) void run_pack(const GridDesc* __restrict__ gp_, int by_,
                                                   int bz_, unsigned long long step_)
{
    const GridDesc& g = *uni(gp_);
    const int by = uni(by_), bz = uni(bz_);
    const unsigned long long step = uni(step_);
    using C = VarCfg<K, DT, 64>;
    using Op = typename Kind<K>::template Op<C>;
    using F = typename C::F;
    using B = typename Kind<K>::B;
    const Op op = make_op<K, C>(g, step);
    kgp::packed_tile<C, B::LO, B::HI, B::ZLO, B::ZHI, Op>((F*)batch_lds, uni(g.nx), uni(g.ny),
                                                          uni(g.ns), uni(g.sy), uni(g.sz), by, bz,
                                                          uni(g.kc), op);
}

template <int K, int DT>
__device__ __attribute__((noinline)) void run_rows2(const GridDesc* __restrict__ gp_, int bx_,
                                                    int by_, int bz_, unsigned long long step_)
{
    static_assert(is_planar(K), "direct rows: kinds that read only their own plane");
    const GridDesc& g = *uni(gp_);
    const int bx = uni(bx_), by = uni(by_), bz = uni(bz_);
    const unsigned long long step = uni(step_);
    using C = VarCfg<K, DT, 64>;
    using Op = typename Kind<K>::template Op<C>;
    using B = typename Kind<K>::B;
    const Op op = make_op<K, C>(g, step);
    kgp::rows2_tile<C, B::LO, B::HI, Op>(uni(g.nx), uni(g.ny), uni(g.sy), uni(g.sz), bx, by,
                                         B::ZLO + bz, op);
}

template <int K, int DT>
__device__ __attribute__((noinline)) void run_col(const GridDesc* __restrict__ gp_, int bx_,
                                                  int by_, int bk_, unsigned long long step_)
{
    static_assert(is_planar(K), "column march: kinds that read only their own plane");
    const GridDesc& g = *uni(gp_);
    const int bx = uni(bx_), by = uni(by_), bk = uni(bk_);
    const unsigned long long step = uni(step_);
    using C = ColCfg<K, DT>;
    using Op = typename Kind<K>::template Op<C>;
    using B = typename Kind<K>::B;
    const Op op = make_op<K, C>(g, step);
    kgp::colmarch_tile<C, B::LO, B::HI, Op, COLZ>(uni(g.nx), uni(g.ny), uni(g.ns), uni(g.sy),
                                                  uni(g.sz), bx, by, bk, uni(g.kc), op);
}

// Flat body: the workgroup sweeps the interior points of planes
// ZLO + bz*kc .. +kc, one point per lane per round, x fastest; operands come
// from global memory, where a small grid's planes stay cached (L1/L2).
template <int K, int DT>
__device__ __attribute__((noinline)) void run_flat(const GridDesc* __restrict__ gp_, int bz_,
                                                   unsigned long long step_)
{
    const GridDesc& g = *uni(gp_);
    const int bz = uni(bz_);
    const unsigned long long step = uni(step_);
    using C = FlatCfg<K, DT>;
    using B = typename Kind<K>::B;
    const auto op = make_op<K, C>(g, step);
    const unsigned nxi = (unsigned)(g.nx - B::LO - B::HI), nyi = (unsigned)(g.ny - B::LO - B::HI);
    const int k0 = B::ZLO + bz * g.kc;
    const int k1 = min(k0 + g.kc, g.ns - B::ZHI);
    const unsigned plane = nxi * nyi;
    const unsigned npts = plane * (unsigned)(k1 - k0);
    FlatView<C> v;
    v.src = kgp::as_global(op.src);
    v.aux[0] = C::NAUX ? kgp::as_global(op.side[0]) : nullptr;
    v.sy = g.sy;
    v.sz = g.sz;
    const bool xm[1] = {true};
    for (unsigned p = threadIdx.x; p < npts; p += blockDim.x) {
        const unsigned kk = p / plane, r = p - kk * plane;
        const unsigned jj = r / nxi, ii = r - jj * nxi;
        const int i = B::LO + (int)ii;
        const long long ob = (long long)(k0 + (int)kk) * g.sz + (long long)(B::LO + (int)jj) * g.sy;
        v.o = ob + i;
        op.run(v, ob, i, xm);
    }
}

// 4 waves per SIMD: one 16-wave or two 8-wave workgroups per CU, 128 VGPRs.
// TRIAL only names the autotuner's timing launches apart from real steps in
// traces (batch_k<true> / batch_k<false>); the body is the same.
template <bool TRIAL>
__global__ __launch_bounds__(64 * BW, KG_BATCH_WPE)
void batch_k(const GridDesc* __restrict__ grids, const Item* __restrict__ items, int nitems,
             BatchState* st)
{
    __shared__ int next;
    const unsigned long long step = __hip_atomic_load(&st->step, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
    // The list is cut into KG_NUM_XCD contiguous parts.  Workgroups are
    // dispatched round-robin over the XCDs, so workgroup b runs on XCD b % 8
    // and takes items from part b % 8 first: neighbouring tiles, which share
    // halo rows and planes, then run on one XCD and share its L2.  A
    // workgroup whose part is empty moves on to the next part.
    // Taking runs of several items per atomic (fewer atomics, neighbouring
    // tiles in sequence on one workgroup) measured slower: 2, 4, 8 and 32
    // items per take cost 4 %, 7 %, 10 % and 17 % (FP32).  So did taking the
    // next item while the current one runs (5 %: the tile's first wait for
    // its loads then also waits for the atomic) and running a workgroup's
    // consecutive items of one variant in a loop inside the tile function
    // (3 %), which saves the callee-saved registers once per run of items
    // instead of once per item.
    const int home = (int)(blockIdx.x % KG_NUM_XCD);
    int part = home, lo = 0, hi = 0, parts_left = KG_NUM_XCD;
    lo = (int)((long long)part * nitems / KG_NUM_XCD);
    hi = (int)((long long)(part + 1) * nitems / KG_NUM_XCD);
    for (;;) {
        if (threadIdx.x == 0) next = lo + (int)atomicAdd(&st->counter[part], 1u);
        __syncthreads();
        const int i = next;
        __syncthreads();
        if (i >= hi) {
            if (--parts_left == 0) break;
            part = (part + 1) % KG_NUM_XCD;
            lo = (int)((long long)part * nitems / KG_NUM_XCD);
            hi = (int)((long long)(part + 1) * nitems / KG_NUM_XCD);
            continue;
        }
        const Item it = items[i];
        const GridDesc* g = grids + it.grid;
        switch (it.var) {
#define KG_VAR_CASE(K, DT, LI) \
    case var_id(K, DT, LI): run_tile<K, DT, LPRS[LI]>(g, it.bx, it.by, it.bz, step); break;
            KG_VAR_LIST(KG_VAR_CASE)
#undef KG_VAR_CASE
#define KG_FLAT_CASE(K, DT) \
    case var_id(K, DT, FLAT): run_flat<K, DT>(g, it.bz, step); break;                 \
    case var_id(K, DT, PACK): run_pack<K, DT>(g, it.by, it.bz, step); break;
            KG_FLAT_LIST(KG_FLAT_CASE)
#undef KG_FLAT_CASE
#define KG_ROW2_CASE(K, DT) \
    case var_id(K, DT, ROW2): run_rows2<K, DT>(g, it.bx, it.by, it.bz, step); break; \
    case var_id(K, DT, COL): run_col<K, DT>(g, it.bx, it.by, it.bz, step); break;
            KG_ROW2_LIST(KG_ROW2_CASE)
#undef KG_ROW2_CASE
        default: break;
        }
        __syncthreads();   // the next tile restages LDS
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&st->done, 1u) == gridDim.x - 1) {
            // every workgroup has taken its last item and read `step`
            for (int x = 0; x < KG_NUM_XCD; x++) atomicExch(&st->counter[x], 0u);
            atomicExch(&st->done, 0u);
            atomicExch(&st->step, step + 1);
            __threadfence();
        }
    }
}

// One tile per workgroup of a single grid (the sequential comparison path:
// one launch per grid and step, same tile bodies as batch_k).
template <int K, int DT, int LPR>
__global__ __launch_bounds__(64 * BW, KG_BATCH_WPE)
void seq_k(const GridDesc* __restrict__ gp, int ntx, int nty, unsigned long long step)
{
    const unsigned b = kg_xcd_swizzle(blockIdx.x, gridDim.x);
    const int bx = (int)(b % (unsigned)ntx);
    const int by = (int)((b / (unsigned)ntx) % (unsigned)nty);
    const int bz = (int)(b / ((unsigned)ntx * (unsigned)nty));
    run_tile<K, DT, LPR>(gp, bx, by, bz, step);
}

template <int K, int DT>
__global__ __launch_bounds__(64 * BW, KG_BATCH_WPE)
void seq_flat_k(const GridDesc* __restrict__ gp, unsigned long long step)
{
    run_flat<K, DT>(gp, (int)blockIdx.x, step);
}

template <int K, int DT>
__global__ __launch_bounds__(64 * BW, KG_BATCH_WPE)
void seq_rows2_k(const GridDesc* __restrict__ gp, int ntx, int nty, unsigned long long step)
{
    const unsigned b = blockIdx.x;
    run_rows2<K, DT>(gp, (int)(b % (unsigned)ntx), (int)((b / (unsigned)ntx) % (unsigned)nty),
                     (int)(b / ((unsigned)ntx * (unsigned)nty)), step);
}

template <int K, int DT>
__global__ __launch_bounds__(64 * BW, KG_BATCH_WPE)
void seq_col_k(const GridDesc* __restrict__ gp, int ntx, int nty, unsigned long long step)
{
    const unsigned b = blockIdx.x;
    run_col<K, DT>(gp, (int)(b % (unsigned)ntx), (int)((b / (unsigned)ntx) % (unsigned)nty),
                   (int)(b / ((unsigned)ntx * (unsigned)nty)), step);
}

template <int K, int DT>
__global__ __launch_bounds__(64 * BW, KG_BATCH_WPE)
void seq_pack_k(const GridDesc* __restrict__ gp, int nty, unsigned long long step)
{
    const unsigned b = kg_xcd_swizzle(blockIdx.x, gridDim.x);
    run_pack<K, DT>(gp, (int)(b % (unsigned)nty), (int)(b / (unsigned)nty), step);
}

// Uniform [-1, 1) values from a counter-based hash of (seed, element index):
// the same seed gives the same field whatever the launch shape.
__device__ __forceinline__ double hash_uniform(unsigned long long seed, unsigned long long i)
{
    unsigned long long h = (i + 1) * 0x9E3779B97F4A7C15ull ^ seed * 0xD1B54A32D192ED03ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    h *= 0x94D049BB133111EBull;
    h ^= h >> 32;
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Fill an nx x ny x ns field with row pitch sy (>= nx) and plane pitch
// sz = sy*ny: element (i,j,k) inside the box [lo, n-hi) x [lo, n-hi) x
// [zlo, ns-zhi) gets hash_uniform(seed, i + nx*(j + ny*k)), elements outside
// it hash_uniform(bseed, ...) (so the levels of a grid, filled with one bseed,
// share their boundary values), the row padding i >= nx gets 0 (so sums over
// the padded level are defined).
struct FillBox {
    int lo, hi, zlo, zhi;
};
template <class S>
__global__ __launch_bounds__(256)
void fill_k(S* __restrict__ p, int nx, int ny, int ns, long long sy, long long sz,
            unsigned long long seed, unsigned long long bseed, FillBox box)
{
    const long long n = sy * ny * ns;
    const long long stride = (long long)gridDim.x * 256;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += stride) {
        const long long i = e % sy, r = e / sy;
        const long long j = r % ny, k = r / ny;
        const bool in = i >= box.lo && i < nx - box.hi && j >= box.lo && j < ny - box.hi &&
                        k >= box.zlo && k < ns - box.zhi;
        const S v = i < nx ? (S)hash_uniform(in ? seed : bseed,
                                             (unsigned long long)(i + nx * (j + ny * k)))
                           : (S)0;
        kg_st(p + k * sz + j * sy + i, v);
    }
}

template <class S>
void launch_fill(hipStream_t st, void* p, int nx, int ny, int ns, long long sy, long long sz,
                 unsigned long long seed, unsigned long long bseed = 0, FillBox box = {0, 0, 0, 0})
{
    const long long n = sy * ny * ns;
    const long long blocks = std::min<long long>((n + 255) / 256, 256 * 32);
    if (n > 0)
        hipLaunchKernelGGL(fill_k<S>, dim3((unsigned)blocks), dim3(256), 0, st, (S*)p, nx, ny, ns,
                           sy, sz, seed, bseed, box);
}

__global__ void state_set_k(BatchState* st, unsigned long long step)
{
    if (threadIdx.x == 0) {
        atomicExch(&st->step, step);
        for (int x = 0; x < KG_NUM_XCD; x++) atomicExch(&st->counter[x], 0u);
        atomicExch(&st->done, 0u);
    }
}

// ------------------------------------------------------------ host side
struct Arena {
    // First-fit suballocator over hipMalloc'd chunks; blocks are 256-byte
    // aligned.  free() coalesces neighbours; chunks are released at the end.
    struct Chunk {
        char* base;
        size_t size;
        std::map<size_t, size_t> free_;   // offset -> size
    };
    std::vector<Chunk> chunks;
    size_t chunk_size = (size_t)1 << 30;
    size_t reserved = 0;

    static size_t up(size_t n) { return (n + 255) & ~(size_t)255; }

    int alloc(size_t n, void** out)
    {
        n = up(n);
        for (auto& c : chunks)
            for (auto itf = c.free_.begin(); itf != c.free_.end(); ++itf)
                if (itf->second >= n) {
                    const size_t off = itf->first, sz = itf->second;
                    c.free_.erase(itf);
                    if (sz > n) c.free_[off + n] = sz - n;
                    *out = c.base + off;
                    return KG_OK;
                }
        const size_t cs = std::max(chunk_size, n);
        char* p = nullptr;
        hipError_t e = hipMalloc((void**)&p, cs);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return kg_set_error(KG_ERR_NOMEM, "kg_batch: device allocation of %zu bytes failed: %s",
                                cs, hipGetErrorString(e));
        }
        reserved += cs;
        chunks.push_back(Chunk{p, cs, {}});
        if (cs > n) chunks.back().free_[n] = cs - n;
        *out = p;
        return KG_OK;
    }
    void release(void* ptr, size_t n)
    {
        n = up(n);
        for (auto& c : chunks) {
            char* p = (char*)ptr;
            if (p < c.base || p >= c.base + c.size) continue;
            size_t off = (size_t)(p - c.base);
            auto nx = c.free_.lower_bound(off);
            if (nx != c.free_.end() && off + n == nx->first) {
                n += nx->second;
                c.free_.erase(nx);
            }
            auto pv = c.free_.lower_bound(off);
            if (pv != c.free_.begin()) {
                --pv;
                if (pv->first + pv->second == off) {
                    off = pv->first;
                    n += pv->second;
                    c.free_.erase(pv);
                }
            }
            c.free_[off] = n;
            return;
        }
    }
    void destroy()
    {
        for (auto& c : chunks) (void)hipFree(c.base);
        chunks.clear();
        reserved = 0;
    }
};

struct HostGrid {
    kg_grid_spec spec;
    bool alive;
    int li, kc;            // tile shape: LPRS[li], planes per tile
    long long sy, sz;      // pitches (elements)
    size_t level_bytes;
    long long step0;
    void* lev[3];
};

int nlev_of(int kind) { return kind == KG_KIND_WAVE13PT ? 3 : 2; }

int size_class(const kg_grid_spec& s)
{
    auto bucket = [](int n) { return n <= 24 ? 0 : n <= 48 ? 1 : n <= 96 ? 2 : n <= 192 ? 3 : 4; };
    return ((s.kind * NTYPE + s.dtype) * 5 + bucket(s.nx)) * 5 + bucket(s.ny);
}

// Shape of a grid whose size class has not been tuned: the direct row body
// for the 2D kinds (for gaussblur3 it measured slower than the plane engine:
// serving capacity 196 against 210 jobs/s); packed rows when a row fits in a
// wave (build_work falls back to 64 lanes per row where rows do not pack),
// 64 lanes per row otherwise.
void default_shape(const kg_grid_spec& s, int& li, int& kc)
{
    const int vec = vec_of(s.kind, s.dtype);
    const bool two_d = s.kind == KG_KIND_JACOBI2 || s.kind == KG_KIND_LIFE2;
    li = two_d ? ROW2 : s.nx <= 64 * vec ? PACK : 0;
    kc = s.ns <= 24 ? 8 : s.ns <= 48 ? 16 : 32;
}

}  // namespace

struct kg_batch {
    kg_ctx* ctx;
    std::vector<HostGrid> grids;
    Arena arena;
    std::map<int, std::pair<int, int>> tuned;   // size class -> (li, kc)
    GridDesc* d_desc = nullptr;
    int desc_cap = 0;
    Item* d_items = nullptr;
    int item_cap = 0;
    int nitems = 0;
    BatchState* d_state = nullptr;
    bool dirty = true;
    long long steps = 0;                       // steps issued
    int nwg = 0;                               // persistent workgroups
    std::map<int, hipGraphExec_t> graphs;      // steps per replay -> graph
    double pts_per_step = 0, bytes_per_step = 0;
    // pinned staging of the work list's uploads; staged marks the end of the
    // last upload on the stream (the staging memory is rewritten only after it)
    GridDesc* h_desc = nullptr;
    size_t h_desc_cap = 0;
    Item* h_items = nullptr;
    size_t h_item_cap = 0;
    hipEvent_t staged = nullptr;
};

namespace {

void drop_graphs(kg_batch* b)
{
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

// Tiles of grid g with shape (li, kc), appended to out.
void cut_tiles(const HostGrid& g, int gi, int li, int kc, std::vector<Item>& out)
{
    const int var = var_id(g.spec.kind, g.spec.dtype, li);
    const VarInfo& v = VARS[var];
    const int nx = g.spec.nx, ny = g.spec.ny, ns = g.spec.ns;
    const int nxi = nx - v.lo - v.hi, nyi = ny - v.lo - v.hi, nzi = ns - v.zlo - v.zhi;
    if (nxi <= 0 || nyi <= 0 || nzi <= 0) return;
    if (li == FLAT) {
        for (int bz = 0; bz < (nzi + kc - 1) / kc; bz++) out.push_back(Item{gi, 0, 0, bz, var});
        return;
    }
    if (li == ROW2) {   // one plane per item
        const int ntx = (int)((g.sy + v.txw - 1) / v.txw), nty = (nyi + v.ty - 1) / v.ty;
        for (int bz = 0; bz < nzi; bz++)
            for (int by = 0; by < nty; by++)
                for (int bx = 0; bx < ntx; bx++) out.push_back(Item{gi, bx, by, bz, var});
        return;
    }
    if (li == COL) {   // COLZ planes by 16/COLZ row chunks of kc rows per item
        const int ntx = (int)((g.sy + v.txw - 1) / v.txw);
        const int rows = (BW / COLZ) * kc;
        const int nty = (nyi + rows - 1) / rows, nkb = (ns + COLZ - 1) / COLZ;
        for (int bk = 0; bk < nkb; bk++)
            for (int by = 0; by < nty; by++)
                for (int bx = 0; bx < ntx; bx++) out.push_back(Item{gi, bx, by, bk, var});
        return;
    }
    if (li == PACK) {
        const int ty = pack_rows_of(v, g.sy) - 2 * v.h;   // > 0: checked by build_work
        for (int bz = 0; bz < (nzi + kc - 1) / kc; bz++)
            for (int by = 0; by < (nyi + ty - 1) / ty; by++) out.push_back(Item{gi, 0, by, bz, var});
        return;
    }
    const int ntx = (nx + v.txw - 1) / v.txw;
    const int nty = (nyi + v.ty - 1) / v.ty;
    const int nkc = (nzi + kc - 1) / kc;
    for (int bz = 0; bz < nkc; bz++)
        for (int by = 0; by < nty; by++)
            for (int bx = 0; bx < ntx; bx++) out.push_back(Item{gi, bx, by, bz, var});
}

GridDesc make_desc(const HostGrid& g, int li, int kc)
{
    GridDesc d{};
    d.var = var_id(g.spec.kind, g.spec.dtype, li);
    d.nx = g.spec.nx;
    d.ny = g.spec.ny;
    d.ns = g.spec.ns;
    d.kc = kc;
    d.nlev = nlev_of(g.spec.kind);
    d.sy = g.sy;
    d.sz = g.sz;
    d.step0 = g.step0;
    for (int l = 0; l < 3; l++) d.lev[l] = g.lev[l];
    const double* c = g.spec.coef;
    if (g.spec.kind == KG_KIND_GAUSSBLUR3) {
        // f = 1/(s0 + 4(s1+s2+s4+s8) + 8 s5) in the grid's arithmetic type
        // (gaussblur/gaussblur.c:65)
        if (g.spec.dtype == KG_F64) d.c[0] = 1. / (c[0] + 4 * (c[1] + c[2] + c[3] + c[5]) + 8 * c[4]);
        else {
            const float s0 = (float)c[0], s1 = (float)c[1], s2 = (float)c[2], s4 = (float)c[3],
                        s5 = (float)c[4], s8 = (float)c[5];
            d.c[0] = (double)(1.f / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5));
        }
        for (int q = 0; q < 6; q++) d.c[1 + q] = c[q];
    } else {
        for (int q = 0; q < 6; q++) d.c[q] = c[q];
    }
    return d;
}

int ensure_device_arrays(kg_batch* b, int ndesc, int nitems)
{
    if (ndesc > b->desc_cap) {
        if (b->d_desc) KG_HIP(hipFree(b->d_desc));
        b->d_desc = nullptr;
        const int cap = std::max(ndesc, 2 * b->desc_cap);
        KG_HIP(hipMalloc((void**)&b->d_desc, (size_t)cap * sizeof(GridDesc)));
        b->desc_cap = cap;
    }
    if (nitems > b->item_cap) {
        if (b->d_items) KG_HIP(hipFree(b->d_items));
        b->d_items = nullptr;
        const int cap = std::max(nitems, 2 * b->item_cap);
        KG_HIP(hipMalloc((void**)&b->d_items, (size_t)cap * sizeof(Item)));
        b->item_cap = cap;
    }
    return KG_OK;
}

// Shape of grid g: the tuned one of its size class, else the default.
void shape_of(const kg_batch* b, const HostGrid& g, int& li, int& kc)
{
    auto it = b->tuned.find(size_class(g.spec));
    if (it != b->tuned.end()) {
        li = it->second.first;
        kc = it->second.second;
    } else {
        default_shape(g.spec, li, kc);
    }
}

// Upload descriptors and the work list (largest grids' tiles first).  With
// only_class >= 0, only that size class is listed, with shape (li, kc).
int build_work(kg_batch* b, int only_class = -1, int li_force = 0, int kc_force = 0)
{
    std::vector<GridDesc> descs(b->grids.size());
    std::vector<int> order;
    double pts = 0, bytes = 0;
    for (size_t gi = 0; gi < b->grids.size(); gi++) {
        HostGrid& g = b->grids[gi];
        if (!g.alive) continue;
        int li, kc;
        if (only_class >= 0) {
            if (size_class(g.spec) != only_class) continue;
            li = li_force;
            kc = kc_force;
        } else {
            shape_of(b, g, li, kc);
        }
        // grids whose rows do not pack take the 64-lane rows instead; the
        // direct row body is for the planar kinds only
        if (li == PACK && !pack_rows_of(VARS[var_id(g.spec.kind, g.spec.dtype, PACK)], g.sy)) li = 0;
        if ((li == ROW2 || li == COL) && !is_planar(g.spec.kind)) li = 0;
        g.li = li;
        g.kc = kc;
        descs[gi] = make_desc(g, li, kc);
        order.push_back((int)gi);
        const double n = (double)g.spec.nx * g.spec.ny * g.spec.ns;
        pts += n;
        bytes += n * (g.spec.dtype == KG_F64 ? 8 : 4) * (g.spec.kind == KG_KIND_WAVE13PT ? 3 : 2);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        const kg_grid_spec &x = b->grids[a].spec, &y = b->grids[c].spec;
        return (double)x.nx * x.ny * x.ns > (double)y.nx * y.ny * y.ns;
    });
    std::vector<Item> items;
    for (int gi : order) cut_tiles(b->grids[gi], gi, b->grids[gi].li, b->grids[gi].kc, items);
    if (int rc = ensure_device_arrays(b, (int)std::max<size_t>(descs.size(), 1),
                                      (int)std::max<size_t>(items.size(), 1)))
        return rc;
    // Uploads go through pinned staging buffers and stay asynchronous: the
    // work list can change between steps (grids joining and leaving a
    // serving batch) without draining the stream.  Only the previous upload
    // must have left the staging memory before it is rewritten.
    if (b->staged) KG_HIP(hipEventSynchronize(b->staged));
    else KG_HIP(hipEventCreateWithFlags(&b->staged, hipEventDisableTiming));
    if (descs.size() > b->h_desc_cap) {
        if (b->h_desc) KG_HIP(hipHostFree(b->h_desc));
        b->h_desc = nullptr;
        const size_t cap = std::max(descs.size(), 2 * b->h_desc_cap);
        KG_HIP(hipHostMalloc((void**)&b->h_desc, cap * sizeof(GridDesc), hipHostMallocDefault));
        b->h_desc_cap = cap;
    }
    if (items.size() > b->h_item_cap) {
        if (b->h_items) KG_HIP(hipHostFree(b->h_items));
        b->h_items = nullptr;
        const size_t cap = std::max(items.size(), 2 * b->h_item_cap);
        KG_HIP(hipHostMalloc((void**)&b->h_items, cap * sizeof(Item), hipHostMallocDefault));
        b->h_item_cap = cap;
    }
    if (!descs.empty()) {
        std::memcpy(b->h_desc, descs.data(), descs.size() * sizeof(GridDesc));
        KG_HIP(hipMemcpyAsync(b->d_desc, b->h_desc, descs.size() * sizeof(GridDesc),
                              hipMemcpyHostToDevice, b->ctx->stream));
    }
    if (!items.empty()) {
        std::memcpy(b->h_items, items.data(), items.size() * sizeof(Item));
        KG_HIP(hipMemcpyAsync(b->d_items, b->h_items, items.size() * sizeof(Item),
                              hipMemcpyHostToDevice, b->ctx->stream));
    }
    KG_HIP(hipEventRecord(b->staged, b->ctx->stream));
    b->nitems = (int)items.size();
    if (only_class < 0) {
        b->pts_per_step = pts;
        b->bytes_per_step = bytes;
        b->dirty = false;
    } else {
        b->dirty = true;   // the full list must be rebuilt before the next real step
    }
    return KG_OK;
}

int launch_steps(kg_batch* b, int n, bool trial = false)
{
    for (int s = 0; s < n; s++) {
        if (trial)
            hipLaunchKernelGGL(batch_k<true>, dim3(b->nwg), dim3(64 * BW), 0, b->ctx->stream,
                               b->d_desc, b->d_items, b->nitems, b->d_state);
        else
            hipLaunchKernelGGL(batch_k<false>, dim3(b->nwg), dim3(64 * BW), 0, b->ctx->stream,
                               b->d_desc, b->d_items, b->nitems, b->d_state);
    }
    KG_HIP(hipGetLastError());
    return KG_OK;
}

int check_batch(kg_batch* b)
{
    if (!b) return kg_set_error(KG_ERR_INVALID, "kg_batch: null batch");
    return kg_check_ctx(b->ctx);
}

int check_spec(const kg_grid_spec& s)
{
    if (s.kind < 0 || s.kind >= NKIND)
        return kg_set_error(KG_ERR_INVALID, "kg_batch: unknown kind %d", s.kind);
    if (s.dtype != KG_F64 && s.dtype != KG_F32)
        return kg_set_error(KG_ERR_INVALID, "kg_batch: dtype %d not batched (F64 or F32)", s.dtype);
    if (s.nx < 1 || s.ny < 1 || s.ns < 1 || (long long)s.nx * s.ny > (1LL << 31))
        return kg_set_error(KG_ERR_INVALID, "kg_batch: bad grid %dx%dx%d", s.nx, s.ny, s.ns);
    // padded rows (up to 3 extra values) of 8 or 4 bytes
    return kg_check_plane_bytes((long long)(s.nx + 3) * s.ny, s.dtype == KG_F64 ? 8 : 4);
}

}  // namespace

// ================================================================ C API

extern "C" int kg_batch_create(kg_ctx* ctx, kg_batch** out)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (!out) return kg_set_error(KG_ERR_INVALID, "kg_batch_create: null argument");
    *out = nullptr;
    kg_batch* b = new (std::nothrow) kg_batch();
    if (!b) return kg_set_error(KG_ERR_INVALID, "kg_batch_create: host allocation failed");
    b->ctx = ctx;
    KG_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc((void**)&b->d_state, sizeof(BatchState));
    if (e != hipSuccess) {
        delete b;
        return kg_set_error(KG_ERR_NOMEM, "kg_batch_create: %s", hipGetErrorString(e));
    }
    hipLaunchKernelGGL(state_set_k, dim3(1), dim3(64), 0, ctx->stream, b->d_state, 0ull);
    int per_cu = 0, cus = 0;
    KG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_k<false>, 64 * BW, 0));
    KG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    b->nwg = std::max(1, per_cu) * std::max(1, cus);
    KG_HIP(hipStreamSynchronize(ctx->stream));
    *out = b;
    return KG_OK;
}

extern "C" int kg_batch_add(kg_batch* b, const kg_grid_spec* spec, int* id)
{
    if (int rc = check_batch(b)) return rc;
    if (!spec) return kg_set_error(KG_ERR_INVALID, "kg_batch_add: null spec");
    if (int rc = check_spec(*spec)) return rc;
    HostGrid g{};
    g.spec = *spec;
    g.alive = true;
    const int vec = spec->dtype == KG_F64 ? TypeOf<KG_F64>::VEC : TypeOf<KG_F32>::VEC;
    const int es = spec->dtype == KG_F64 ? 8 : 4;
    g.sy = (spec->nx + vec - 1) / vec * vec;
    g.sz = g.sy * spec->ny;
    g.level_bytes = (size_t)g.sz * spec->ns * es;
    g.step0 = b->steps;
    for (int l = 0; l < nlev_of(spec->kind); l++) {
        if (int rc = b->arena.alloc(g.level_bytes, &g.lev[l])) {
            for (int q = 0; q < l; q++) b->arena.release(g.lev[q], g.level_bytes);
            return rc;
        }
    }
    b->grids.push_back(g);
    b->dirty = true;
    drop_graphs(b);
    if (id) *id = (int)b->grids.size() - 1;
    return KG_OK;
}

extern "C" int kg_batch_remove(kg_batch* b, int id)
{
    if (int rc = check_batch(b)) return rc;
    if (id < 0 || id >= (int)b->grids.size() || !b->grids[id].alive)
        return kg_set_error(KG_ERR_INVALID, "kg_batch_remove: no grid %d", id);
    // Work already queued on the stream may still read the levels.
    KG_HIP(hipStreamSynchronize(b->ctx->stream));
    HostGrid& g = b->grids[id];
    for (int l = 0; l < nlev_of(g.spec.kind); l++) b->arena.release(g.lev[l], g.level_bytes);
    g.alive = false;
    b->dirty = true;
    drop_graphs(b);
    return KG_OK;
}

static int level_copy(kg_batch* b, int id, int level, void* host, bool up)
{
    if (int rc = check_batch(b)) return rc;
    if (id < 0 || id >= (int)b->grids.size() || !b->grids[id].alive)
        return kg_set_error(KG_ERR_INVALID, "kg_batch: no grid %d", id);
    HostGrid& g = b->grids[id];
    if (level < 0 || level >= nlev_of(g.spec.kind))
        return kg_set_error(KG_ERR_INVALID, "kg_batch: grid %d has no level %d", id, level);
    if (!host) return kg_set_error(KG_ERR_INVALID, "kg_batch: null host buffer");
    const size_t es = g.spec.dtype == KG_F64 ? 8 : 4;
    const size_t w = (size_t)g.spec.nx * es, dp = (size_t)g.sy * es;
    const size_t h = (size_t)g.spec.ny * g.spec.ns;
    if (up)
        KG_HIP(hipMemcpy2DAsync(g.lev[level], dp, host, w, w, h, hipMemcpyHostToDevice,
                                b->ctx->stream));
    else
        KG_HIP(hipMemcpy2DAsync(host, w, g.lev[level], dp, w, h, hipMemcpyDeviceToHost,
                                b->ctx->stream));
    KG_HIP(hipStreamSynchronize(b->ctx->stream));
    return KG_OK;
}

extern "C" int kg_batch_upload(kg_batch* b, int id, int level, const void* host)
{
    return level_copy(b, id, level, (void*)host, true);
}

extern "C" int kg_batch_download(kg_batch* b, int id, int level, void* host)
{
    return level_copy(b, id, level, host, false);
}

extern "C" int kg_batch_newest_level(kg_batch* b, int id, int* level)
{
    if (int rc = check_batch(b)) return rc;
    if (id < 0 || id >= (int)b->grids.size() || !b->grids[id].alive || !level)
        return kg_set_error(KG_ERR_INVALID, "kg_batch_newest_level: bad grid %d", id);
    const HostGrid& g = b->grids[id];
    const long long ls = b->steps - g.step0;
    // laplacian/gaussblur: step s writes level (s+1)%2; wave13pt: level (s+2)%3
    *level = g.spec.kind == KG_KIND_WAVE13PT ? (int)((ls + 1) % 3) : (int)(ls % 2);
    return KG_OK;
}

extern "C" int kg_batch_steps_done(kg_batch* b, int id, long long* steps)
{
    if (int rc = check_batch(b)) return rc;
    if (id < 0 || id >= (int)b->grids.size() || !b->grids[id].alive || !steps)
        return kg_set_error(KG_ERR_INVALID, "kg_batch_steps_done: bad grid %d", id);
    *steps = b->steps - b->grids[id].step0;
    return KG_OK;
}

extern "C" int kg_batch_step(kg_batch* b, int nsteps, int use_graph)
{
    if (int rc = check_batch(b)) return rc;
    if (nsteps < 0) return kg_set_error(KG_ERR_INVALID, "kg_batch_step: negative step count");
    if (nsteps == 0) return KG_OK;
    if (b->dirty) {
        if (int rc = build_work(b)) return rc;
        drop_graphs(b);
    }
    kg_register_kernel("batch", (const void*)batch_k<false>);
    if (int rc = kg_launch_begin(b->ctx, "batch")) return rc;
    if (b->nitems > 0) {
        if (use_graph) {
            auto it = b->graphs.find(nsteps);
            if (it == b->graphs.end()) {
                hipGraph_t graph;
                KG_HIP(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
                int rc = launch_steps(b, nsteps);
                hipError_t e = hipStreamEndCapture(b->ctx->stream, &graph);
                if (rc) return rc;
                KG_HIP(e);
                hipGraphExec_t exec;
                e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                KG_HIP(e);
                it = b->graphs.emplace(nsteps, exec).first;
            }
            KG_HIP(hipGraphLaunch(it->second, b->ctx->stream));
        } else if (int rc = launch_steps(b, nsteps)) {
            return rc;
        }
    } else {
        // nothing to compute: the step counter still advances
        hipLaunchKernelGGL(state_set_k, dim3(1), dim3(64), 0, b->ctx->stream, b->d_state,
                           (unsigned long long)(b->steps + nsteps));
    }
    b->steps += nsteps;
    return kg_launch_end(b->ctx);
}

extern "C" int kg_batch_info(kg_batch* b, double* points_per_step, double* bytes_per_step,
                             int* work_items, int* workgroups)
{
    if (int rc = check_batch(b)) return rc;
    if (b->dirty)
        if (int rc = build_work(b)) return rc;
    if (points_per_step) *points_per_step = b->pts_per_step;
    if (bytes_per_step) *bytes_per_step = b->bytes_per_step;
    if (work_items) *work_items = b->nitems;
    if (workgroups) *workgroups = b->nwg;
    return KG_OK;
}

extern "C" int kg_batch_tune(kg_batch* b, int reps, int* classes_tuned)
{
    if (int rc = check_batch(b)) return rc;
    if (reps < 1) reps = 1;
    KG_HIP(hipStreamSynchronize(b->ctx->stream));
    std::map<int, int> classes;
    for (const auto& g : b->grids)
        if (g.alive) classes[size_class(g.spec)]++;
    // Tuning runs steps on a scratch step counter equal to the real one: each
    // trial rewrites the level the next real step writes anyway.
    BatchState* real = b->d_state;
    BatchState* scratch = nullptr;
    KG_HIP(hipMalloc((void**)&scratch, sizeof(BatchState)));
    hipEvent_t e0, e1;
    KG_HIP(hipEventCreate(&e0));
    KG_HIP(hipEventCreate(&e1));
    int rc = KG_OK;
    int ntuned = 0;
    for (const auto& cls : classes) {
        const int c = cls.first;
        float best = 1e30f;
        std::pair<int, int> pick{-1, -1};
        for (int li = 0; li < NSHAPE && !rc; li++)
            for (int ki = 0; ki < NKC && !rc; ki++) {
                if ((rc = build_work(b, c, li, KCS[ki]))) break;
                b->d_state = scratch;
                hipLaunchKernelGGL(state_set_k, dim3(1), dim3(64), 0, b->ctx->stream, scratch,
                                   (unsigned long long)b->steps);
                float ms = 0;
                rc = launch_steps(b, 1, true);   // warm-up
                if (!rc) {
                    hipLaunchKernelGGL(state_set_k, dim3(1), dim3(64), 0, b->ctx->stream, scratch,
                                       (unsigned long long)b->steps);
                    (void)hipEventRecord(e0, b->ctx->stream);
                    for (int r = 0; r < reps && !rc; r++) {
                        hipLaunchKernelGGL(state_set_k, dim3(1), dim3(64), 0, b->ctx->stream,
                                           scratch, (unsigned long long)b->steps);
                        rc = launch_steps(b, 1, true);
                    }
                    (void)hipEventRecord(e1, b->ctx->stream);
                    if (hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                        rc = kg_set_error(KG_ERR_HIP, "kg_batch_tune: event timing failed");
                }
                b->d_state = real;
                if (!rc && ms < best) {
                    best = ms;
                    pick = {li, KCS[ki]};
                }
            }
        if (rc) break;
        b->tuned[c] = pick;
        ntuned++;
    }
    b->d_state = real;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    b->dirty = true;
    drop_graphs(b);
    if (classes_tuned) *classes_tuned = ntuned;
    return rc;
}

extern "C" int kg_batch_grid_shape(kg_batch* b, int id, int* lanes_per_row, int* planes_per_tile)
{
    if (int rc = check_batch(b)) return rc;
    if (id < 0 || id >= (int)b->grids.size() || !b->grids[id].alive)
        return kg_set_error(KG_ERR_INVALID, "kg_batch_grid_shape: no grid %d", id);
    if (b->dirty)
        if (int rc = build_work(b)) return rc;
    if (lanes_per_row) *lanes_per_row = LPRS[b->grids[id].li];
    if (planes_per_tile) *planes_per_tile = b->grids[id].kc;
    return KG_OK;
}

extern "C" int kg_batch_set_shape(kg_batch* b, int id, int lanes_per_row, int planes_per_tile)
{
    if (int rc = check_batch(b)) return rc;
    if (id < 0 || id >= (int)b->grids.size() || !b->grids[id].alive)
        return kg_set_error(KG_ERR_INVALID, "kg_batch_set_shape: no grid %d", id);
    int li = -1;
    for (int q = 0; q < NSHAPE; q++)
        if (LPRS[q] == lanes_per_row) li = q;
    if (li < 0 || planes_per_tile < 1)
        return kg_set_error(KG_ERR_INVALID, "kg_batch_set_shape: lanes per row must be 16, 32, "
                                            "64, 0 (flat), -1 (packed rows) or -2 (direct rows, "
                                            "planar kinds) and planes per tile positive");
    b->tuned[size_class(b->grids[id].spec)] = {li, planes_per_tile};
    b->dirty = true;
    drop_graphs(b);
    return KG_OK;
}

extern "C" int kg_batch_memory(kg_batch* b, size_t* reserved, size_t* in_use)
{
    if (int rc = check_batch(b)) return rc;
    size_t used = 0;
    for (const auto& g : b->grids)
        if (g.alive) used += Arena::up(g.level_bytes) * nlev_of(g.spec.kind);
    if (reserved) *reserved = b->arena.reserved;
    if (in_use) *in_use = used;
    return KG_OK;
}

// Level l of grid g from seed: the interior from a hash of (seed, l,
// element), the boundary (every element outside the kind's interior box) from
// a hash of (seed, element), the same in every level; asynchronous.  jacobi2
// grids get the same values in both levels throughout (a relaxation towards a
// fixed boundary; the tolerance test compares the two levels).
static void fill_grid(kg_batch* b, const HostGrid& g, unsigned long long seed)
{
    const VarInfo& v = VARS[var_id(g.spec.kind, g.spec.dtype, 0)];
    const FillBox box{v.lo, v.hi, v.zlo, v.zhi};
    const unsigned long long bs = seed * 1000003ull;
    for (int l = 0; l < nlev_of(g.spec.kind); l++) {
        const unsigned long long s =
            bs + (g.spec.kind == KG_KIND_JACOBI2 ? 0ull : (unsigned long long)l);
        if (g.spec.dtype == KG_F64)
            launch_fill<double>(b->ctx->stream, g.lev[l], g.spec.nx, g.spec.ny, g.spec.ns, g.sy,
                                g.sz, s, bs, box);
        else
            launch_fill<float>(b->ctx->stream, g.lev[l], g.spec.nx, g.spec.ny, g.spec.ns, g.sy,
                               g.sz, s, bs, box);
    }
}

extern "C" int kg_batch_fill_grid(kg_batch* b, int id, unsigned long long seed)
{
    if (int rc = check_batch(b)) return rc;
    if (id < 0 || id >= (int)b->grids.size() || !b->grids[id].alive)
        return kg_set_error(KG_ERR_INVALID, "kg_batch_fill_grid: no grid %d", id);
    fill_grid(b, b->grids[id], seed);
    KG_HIP(hipGetLastError());
    return KG_OK;
}

extern "C" int kg_batch_fill(kg_batch* b, unsigned long long seed)
{
    if (int rc = check_batch(b)) return rc;
    for (size_t gi = 0; gi < b->grids.size(); gi++)
        if (b->grids[gi].alive) fill_grid(b, b->grids[gi], seed * 7919ull + gi);
    KG_HIP(hipGetLastError());
    KG_HIP(hipStreamSynchronize(b->ctx->stream));
    return KG_OK;
}

extern "C" int kg_batch_step_sequential(kg_batch* b, int nsteps)
{
    if (int rc = check_batch(b)) return rc;
    if (nsteps < 0) return kg_set_error(KG_ERR_INVALID, "kg_batch_step_sequential: negative count");
    if (b->dirty) {
        if (int rc = build_work(b)) return rc;
        drop_graphs(b);
    }
    kg_register_kernel("batch_sequential", (const void*)seq_k<0, 0, 64>);
    if (int rc = kg_launch_begin(b->ctx, "batch_sequential")) return rc;
    for (int s = 0; s < nsteps; s++) {
        const unsigned long long step = (unsigned long long)(b->steps + s);
        for (size_t gi = 0; gi < b->grids.size(); gi++) {
            const HostGrid& g = b->grids[gi];
            if (!g.alive) continue;
            const VarInfo& v = VARS[var_id(g.spec.kind, g.spec.dtype, g.li)];
            const int nxi = g.spec.nx - v.lo - v.hi, nyi = g.spec.ny - v.lo - v.hi;
            const int nzi = g.spec.ns - v.zlo - v.zhi;
            if (nxi <= 0 || nyi <= 0 || nzi <= 0) continue;
            if (g.li == FLAT) {
                const GridDesc* gp = b->d_desc + gi;
                const unsigned nkc = (unsigned)((nzi + g.kc - 1) / g.kc);
                switch (g.spec.kind * NTYPE + g.spec.dtype) {
#define KG_SEQ_FLAT(K, DT)                                                                      \
    case K * NTYPE + DT:                                                                        \
        hipLaunchKernelGGL((seq_flat_k<K, DT>), dim3(nkc), dim3(64 * BW), 0, b->ctx->stream, gp, \
                           step);                                                               \
        break;
                    KG_FLAT_LIST(KG_SEQ_FLAT)
#undef KG_SEQ_FLAT
                default: break;
                }
                continue;
            }
            if (g.li == ROW2) {
                const GridDesc* gp = b->d_desc + gi;
                const int ntx = (int)((g.sy + v.txw - 1) / v.txw), nty = (nyi + v.ty - 1) / v.ty;
                switch (g.spec.kind * NTYPE + g.spec.dtype) {
#define KG_SEQ_ROW2(K, DT)                                                                      \
    case K * NTYPE + DT:                                                                        \
        hipLaunchKernelGGL((seq_rows2_k<K, DT>), dim3((unsigned)(ntx * nty * nzi)), dim3(64 * BW), \
                           0, b->ctx->stream, gp, ntx, nty, step);                              \
        break;
                    KG_ROW2_LIST(KG_SEQ_ROW2)
#undef KG_SEQ_ROW2
                default: break;
                }
                continue;
            }
            if (g.li == COL) {
                const GridDesc* gp = b->d_desc + gi;
                const int rows = (BW / COLZ) * g.kc;
                const int ntx = (int)((g.sy + v.txw - 1) / v.txw), nty = (nyi + rows - 1) / rows;
                const int nkb = (g.spec.ns + COLZ - 1) / COLZ;
                switch (g.spec.kind * NTYPE + g.spec.dtype) {
#define KG_SEQ_COL(K, DT)                                                                       \
    case K * NTYPE + DT:                                                                        \
        hipLaunchKernelGGL((seq_col_k<K, DT>), dim3((unsigned)(ntx * nty * nkb)), dim3(64 * BW), \
                           0, b->ctx->stream, gp, ntx, nty, step);                              \
        break;
                    KG_ROW2_LIST(KG_SEQ_COL)
#undef KG_SEQ_COL
                default: break;
                }
                continue;
            }
            if (g.li == PACK) {
                const GridDesc* gp = b->d_desc + gi;
                const int ty = pack_rows_of(v, g.sy) - 2 * v.h;
                const int nty = (nyi + ty - 1) / ty, nkc = (nzi + g.kc - 1) / g.kc;
                switch (g.spec.kind * NTYPE + g.spec.dtype) {
#define KG_SEQ_PACK(K, DT)                                                                      \
    case K * NTYPE + DT:                                                                        \
        hipLaunchKernelGGL((seq_pack_k<K, DT>), dim3((unsigned)(nty * nkc)), dim3(64 * BW), 0,  \
                           b->ctx->stream, gp, nty, step);                                      \
        break;
                    KG_FLAT_LIST(KG_SEQ_PACK)
#undef KG_SEQ_PACK
                default: break;
                }
                continue;
            }
            const int ntx = (g.spec.nx + v.txw - 1) / v.txw, nty = (nyi + v.ty - 1) / v.ty;
            const int nkc = (nzi + g.kc - 1) / g.kc;
            const GridDesc* gp = b->d_desc + gi;
            switch (var_id(g.spec.kind, g.spec.dtype, g.li)) {
#define KG_SEQ_CASE(K, DT, LI)                                                                  \
    case var_id(K, DT, LI):                                                                     \
        hipLaunchKernelGGL((seq_k<K, DT, LPRS[LI]>), dim3((unsigned)(ntx * nty * nkc)),         \
                           dim3(64 * BW), 0, b->ctx->stream, gp, ntx, nty, step);               \
        break;
                KG_VAR_LIST(KG_SEQ_CASE)
#undef KG_SEQ_CASE
            default: break;
            }
        }
    }
    KG_HIP(hipGetLastError());
    // keep the device step counter in line with the host's
    b->steps += nsteps;
    hipLaunchKernelGGL(state_set_k, dim3(1), dim3(64), 0, b->ctx->stream, b->d_state,
                       (unsigned long long)b->steps);
    return kg_launch_end(b->ctx);
}

extern "C" int kg_fill_uniform(kg_ctx* ctx, int dtype, size_t n, void* p, unsigned long long seed)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (n && !p) return kg_set_error(KG_ERR_INVALID, "kg_fill_uniform: null buffer");
    if (n > (size_t)INT32_MAX * 64)
        return kg_set_error(KG_ERR_INVALID, "kg_fill_uniform: field too large");
    // as a 1 x n/64-row field when n is a multiple of 64, else n x 1
    const int nx = (n % 64 == 0) ? 64 : (int)n, ny = (n % 64 == 0) ? (int)(n / 64) : 1;
    if (n && n % 64 && n > (size_t)INT32_MAX)
        return kg_set_error(KG_ERR_INVALID, "kg_fill_uniform: field too large");
    if (dtype == KG_F64) launch_fill<double>(ctx->stream, p, nx, ny, 1, nx, (long long)nx * ny, seed);
    else if (dtype == KG_F32) launch_fill<float>(ctx->stream, p, nx, ny, 1, nx, (long long)nx * ny, seed);
    else return kg_set_error(KG_ERR_INVALID, "kg_fill_uniform: dtype %d (F64 or F32)", dtype);
    KG_HIP(hipGetLastError());
    return KG_OK;
}

extern "C" int kg_batch_destroy(kg_batch* b)
{
    if (!b) return KG_OK;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    drop_graphs(b);
    b->arena.destroy();
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_state) (void)hipFree(b->d_state);
    if (b->h_desc) (void)hipHostFree(b->h_desc);
    if (b->h_items) (void)hipHostFree(b->h_items);
    if (b->staged) (void)hipEventDestroy(b->staged);
    delete b;
    return KG_OK;
}

int kg_preload_batch() { return kg_preload_kernel((const void*)batch_k<false>); }

// ================================================================ serving
//
// kg_serve_run: a request loop over a batch.  Jobs (grids with a step count)
// arrive at given times; a job is admitted, in arrival order, as soon as the
// number of active jobs and the HBM budget allow (its levels are carved from
// the batch arena and filled on the device); every tick advances every active
// job by one step with one persistent launch, so jobs that arrive while others
// run join the next tick and share the device with them (continuous batching).
// A job leaves when its steps are done (jacobi2 jobs with tol > 0 also when
// the largest update max |w_new - w_old| falls below tol, checked every
// KG_SERVE_CHECK steps); its FP64 checksum and optionally its final field are
// taken before its memory is released.  Mode KG_SERVE_SEQUENTIAL admits one
// job at a time and steps it with one plain launch per step: the comparison
// path.
namespace {

constexpr int KG_SERVE_CHECK = 10;

double now_s(const std::chrono::steady_clock::time_point& t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

size_t job_bytes(const kg_job& j)
{
    const int vec = j.dtype == KG_F64 ? TypeOf<KG_F64>::VEC : TypeOf<KG_F32>::VEC;
    const int es = j.dtype == KG_F64 ? 8 : 4;
    const size_t sy = (size_t)((j.nx + vec - 1) / vec * vec);
    return Arena::up(sy * j.ny * j.ns * es) * nlev_of(j.kind);
}

struct Active {
    int job, id;
    long long done;
    long long check_at = -1;       // step of a convergence check in flight
    long long converged_at = -1;   // step at which the check found convergence
};

}  // namespace

extern "C" int kg_serve_run(kg_ctx* ctx, const kg_job* jobs, int njobs, size_t hbm_budget,
                            int max_active, int mode, kg_job_result* results,
                            kg_serve_stats* stats)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (njobs < 0 || (njobs && (!jobs || !results)))
        return kg_set_error(KG_ERR_INVALID, "kg_serve_run: bad job list");
    if (mode != KG_SERVE_CONTINUOUS && mode != KG_SERVE_SEQUENTIAL)
        return kg_set_error(KG_ERR_INVALID, "kg_serve_run: unknown mode %d", mode);
    if (max_active < 1) max_active = 1;
    if (mode == KG_SERVE_SEQUENTIAL) max_active = 1;
    std::vector<int> order(njobs);
    for (int i = 0; i < njobs; i++) {
        order[i] = i;
        kg_grid_spec s{};
        s.kind = jobs[i].kind;
        s.dtype = jobs[i].dtype;
        s.nx = jobs[i].nx;
        s.ny = jobs[i].ny;
        s.ns = jobs[i].ns;
        if (int rc = check_spec(s)) return rc;
        if (jobs[i].steps < 0) return kg_set_error(KG_ERR_INVALID, "kg_serve_run: negative steps");
        results[i] = kg_job_result{};
        results[i].status = KG_OK;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return jobs[a].arrival < jobs[c].arrival; });
    kg_batch* b = nullptr;
    if (int rc = kg_batch_create(ctx, &b)) return rc;
    std::deque<int> pending;
    std::vector<Active> active;
    size_t in_use = 0, peak = 0;
    int next = 0, finished = 0, peak_active = 0;
    double points = 0;
    int rc = KG_OK;
    const auto t0 = std::chrono::steady_clock::now();

    // A job ends with the field of step `upto` (<= a.done): a jacobi2 job
    // found converged one step late ends with its previous level, which that
    // step only read.
    auto finish = [&](const Active& a, long long upto) -> int {
        const kg_job& j = jobs[a.job];
        kg_job_result& r = results[a.job];
        int lev = 0;
        if (int e = kg_batch_newest_level(b, a.id, &lev)) return e;
        if (upto != a.done) {
            if (nlev_of(j.kind) != 2 || upto != a.done - 1)
                return kg_set_error(KG_ERR_INVALID, "kg_serve_run: no level for step %lld", upto);
            lev ^= 1;
        }
        const HostGrid& g = b->grids[a.id];
        double st[5];
        // checksum over the padded level: the padding columns are never
        // written after the fill, so equal fields give equal sums
        if (int e = kg_field_stats(ctx, j.dtype, (size_t)g.sz * j.ns, g.lev[lev], KG_F64, nullptr,
                                   st))
            return e;
        r.checksum[0] = st[0];
        r.checksum[1] = st[1];
        if (j.out)
            if (int e = kg_batch_download(b, a.id, lev, j.out)) return e;
        if (int e = kg_batch_remove(b, a.id)) return e;   // waits for the device
        r.finished = now_s(t0);
        r.steps_run = upto;
        in_use -= job_bytes(j);
        finished++;
        points += (double)j.nx * j.ny * j.ns * (double)upto;
        return KG_OK;
    };
    // Convergence checks (jacobi2 jobs with tol > 0, every KG_SERVE_CHECK
    // steps) are issued after a tick and read after the next tick has been
    // queued: waiting for them then leaves the device busy with that tick.
    double* check_res = nullptr;   // pinned: 5 statistics per job
    hipEvent_t checked = nullptr;
    bool checks_live = false;
    const double* stats_res = ctx->stats + kg_stats_scratch_bytes() / sizeof(double) - 5;
    if (hipHostMalloc((void**)&check_res, sizeof(double) * 5 * (size_t)std::max(njobs, 1),
                      hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&checked, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (check_res) (void)hipHostFree(check_res);
        kg_batch_destroy(b);
        return kg_set_error(KG_ERR_NOMEM, "kg_serve_run: pinned check buffer");
    }

    while (finished < njobs && !rc) {
        const double t = now_s(t0);
        while (next < njobs && jobs[order[next]].arrival <= t) pending.push_back(order[next++]);
        // admission in arrival order
        while (!pending.empty() && (int)active.size() < max_active) {
            const int ji = pending.front();
            const kg_job& j = jobs[ji];
            const size_t need = job_bytes(j);
            if (need > hbm_budget) {   // can never run
                pending.pop_front();
                results[ji].status = KG_ERR_NOMEM;
                results[ji].admitted = results[ji].finished = now_s(t0);
                finished++;
                continue;
            }
            if (in_use + need > hbm_budget) break;
            kg_grid_spec s{};
            s.kind = j.kind;
            s.dtype = j.dtype;
            s.nx = j.nx;
            s.ny = j.ny;
            s.ns = j.ns;
            for (int q = 0; q < 6; q++) s.coef[q] = j.coef[q];
            int id = -1;
            if ((rc = kg_batch_add(b, &s, &id))) break;
            fill_grid(b, b->grids[id], j.seed);
            pending.pop_front();
            in_use += need;
            peak = std::max(peak, in_use);
            results[ji].admitted = now_s(t0);
            active.push_back(Active{ji, id, 0});
        }
        if (rc) break;
        peak_active = std::max(peak_active, (int)active.size());
        // jobs with nothing to do leave at once
        for (size_t a = 0; a < active.size() && !rc;) {
            if (active[a].done >= jobs[active[a].job].steps) {
                rc = finish(active[a], active[a].done);
                active.erase(active.begin() + a);
            } else {
                a++;
            }
        }
        if (rc) break;
        if (active.empty()) {
            if (next < njobs && pending.empty()) {
                const double wait = jobs[order[next]].arrival - now_s(t0);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            }
            continue;
        }
        // one tick
        rc = mode == KG_SERVE_SEQUENTIAL ? kg_batch_step_sequential(b, 1) : kg_batch_step(b, 1, 0);
        if (rc) break;
        for (auto& a : active) a.done++;
        if (checks_live) {
            if (hipEventSynchronize(checked) != hipSuccess) {
                rc = kg_set_error(KG_ERR_HIP, "kg_serve_run: convergence check failed");
                break;
            }
            checks_live = false;
            for (auto& a : active)
                if (a.check_at >= 0) {
                    if (check_res[5 * a.job + 4] < jobs[a.job].tol) a.converged_at = a.check_at;
                    a.check_at = -1;
                }
        }
        bool issued = false;
        for (size_t a = 0; a < active.size() && !rc;) {
            Active& x = active[a];
            const kg_job& j = jobs[x.job];
            if (x.converged_at >= 0 || x.done >= j.steps) {
                rc = finish(x, x.converged_at >= 0 ? x.converged_at : x.done);
                active.erase(active.begin() + a);
                continue;
            }
            if (j.kind == KG_KIND_JACOBI2 && j.tol > 0 && x.done % KG_SERVE_CHECK == 0) {
                // largest update of this step: max |w_done - w_(done-1)|
                const HostGrid& g = b->grids[x.id];
                int lev = 0;
                if ((rc = kg_batch_newest_level(b, x.id, &lev))) break;
                if ((rc = kg_field_stats_launch(ctx->stream, j.dtype, (size_t)g.sz * j.ns, g.lev[lev],
                                                j.dtype, g.lev[lev ^ 1], ctx->stats)))
                    break;
                if (hipMemcpyAsync(check_res + 5 * x.job, stats_res, 5 * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream) != hipSuccess) {
                    rc = kg_set_error(KG_ERR_HIP, "kg_serve_run: convergence check copy failed");
                    break;
                }
                x.check_at = x.done;
                issued = true;
            }
            a++;
        }
        if (!rc && issued) {
            if (hipEventRecord(checked, ctx->stream) != hipSuccess)
                rc = kg_set_error(KG_ERR_HIP, "kg_serve_run: event record failed");
            else
                checks_live = true;
        }
    }
    if (checks_live) (void)hipEventSynchronize(checked);
    (void)hipEventDestroy(checked);
    (void)hipHostFree(check_res);
    const double wall = now_s(t0);
    if (rc) {
        kg_batch_destroy(b);
        return rc;
    }
    kg_batch_destroy(b);
    if (stats) {
        std::vector<double> lat;
        for (int i = 0; i < njobs; i++)
            if (results[i].status == KG_OK) lat.push_back(results[i].finished - jobs[i].arrival);
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            if (lat.empty()) return 0.0;
            const double pos = p * (lat.size() - 1);
            const size_t lo = (size_t)pos;
            const size_t hi = std::min(lo + 1, lat.size() - 1);
            return lat[lo] + (pos - lo) * (lat[hi] - lat[lo]);
        };
        double mean = 0;
        for (double v : lat) mean += v;
        stats->wall = wall;
        stats->points = points;
        stats->p50 = pct(0.5);
        stats->p99 = pct(0.99);
        stats->mean_latency = lat.empty() ? 0 : mean / lat.size();
        stats->peak_active = peak_active;
        stats->peak_bytes = peak;
    }
    return KG_OK;
}
