// kg_typed.hip -- precision variants beside the FP64 drop-in kernels, and the
// FP64 field statistics every precision run reports.
//
// laplacian and wave13pt with FP32 storage and FP32 arithmetic, or BF16
// storage (bfloat16 bit patterns) and FP32 arithmetic; gaussblur applied per
// xy-plane of a 3D grid in FP64 and FP32.  Same plane engine, same stencil
// bodies (kg_ops3d.h) and same source order as the FP64 kernels; only the
// types change.  Bytes per point drop from 16/24 (FP64) to 8/12 (FP32) and
// 4/6 (BF16).
//
// kg_field_stats accumulates sums, sums of squares and maxima in FP64 over a
// fixed partition (KG_STATS_BLOCKS workgroups, then one workgroup over their
// partials), so a result does not depend on timing and repeats bit for bit.
#include "kg_ops3d.h"

#include <algorithm>
#include <type_traits>

// Tile orders (kg_tile_order) of the FP32/BF16 plane-engine kernels and of the
// per-plane direct row body, measured at 512^3 (12-launch means): FP32
// laplacian 223 us XCD-contiguous, 210 in chunks of 2 tile rows (213-211 in
// hardware order or chunks of 4, 8), FP32 wave13pt 337 -> 315 us, BF16 within
// 3 %; per-plane gaussblur 242 us (FP32) in hardware order, 251-280 otherwise.
// x-shift mode for the FP32/BF16 plane-engine kernels: a lane's 4 (FP32)
// or 8 (BF16) points take their same-row x-neighbours from its own
// registers and, for the end points, from the neighbouring lanes by DPP.
#ifndef KG_TYPED_XS
#define KG_TYPED_XS 0
#endif
#ifndef KG_ORD_TYPED
#define KG_ORD_TYPED 2
#endif
// BF16 wave13pt (64-plane z-chunks sharing 4 planes at each seam) enumerates
// its tiles z-chunk fastest and deals them to the XCDs in runs of 2 whole
// z-columns (plane_kernel's ORD < 0), so the chunks of a column run on one
// XCD: 170-177 us on the first ten launches at 512^3 (179-181 mean, three
// levels rotated) against 189-196 (193) in chunks of 2 tile rows.  FP32
// wave13pt lost that way (runs of 1, 2, 4 columns: 326-351 us against 309-311)
// and so did both laplacians (FP32 232-248 us against 201-203; BF16 158-168
// against 137-138 with 8-plane chunks, 160-173 with 16-64-plane chunks in
// z-column runs or chunked tile rows).
#ifndef KG_ORD_WF32
#define KG_ORD_WF32 KG_ORD_TYPED
#endif
#ifndef KG_ORD_WBF
#define KG_ORD_WBF -2
#endif
#ifndef KG_GB3_W64
#define KG_GB3_W64 4
#endif
#ifndef KG_GB3_D64
#define KG_GB3_D64 1
#endif
#ifndef KG_GB3_YC64
#define KG_GB3_YC64 16
#endif
#ifndef KG_GB3_W32
#define KG_GB3_W32 4
#endif
#ifndef KG_GB3_D32
#define KG_GB3_D32 2
#endif
#ifndef KG_GB3_YC32
#define KG_GB3_YC32 16
#endif
#ifndef KG_ORD_ROWS2
#define KG_ORD_ROWS2 0
#endif

namespace {

using namespace kgops;

#define KG_REQUIRE(cond, msg) \
    do { if (!(cond)) return kg_set_error(KG_ERR_INVALID, "%s", msg); } while (0)

// Single-grid launch of op Op in storage type T: 16-byte lanes (VEC = VMAX)
// when nx is a multiple of VMAX, else one point per lane.
template <template <class> class OpT, class T, int VMAX, int W, int D, int KC, int ORD = KG_ORD_TYPED,
          class Fill>
static int launch_typed(kg_ctx* ctx, const char* name, int nx, int ny, int ns, bool al, Fill fill)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    if (int rc = kg_check_plane_bytes((long long)nx * ny, (int)sizeof(typename T::S))) return rc;
    constexpr int LO = Bounds<OpT>::LO, HI = Bounds<OpT>::HI;
    constexpr int ZLO = Bounds<OpT>::ZLO, ZHI = Bounds<OpT>::ZHI;
    auto go = [&](auto vtag) -> int {
        constexpr int V = decltype(vtag)::value;
        using C = OpCfg<OpT, T, V, W, D, KC, 64, (V > 1) && KG_TYPED_XS>;
        using Op = OpT<C>;
        Op op{};
        fill(op);
        const kgp::Grid g = kgp::make_grid<C>(nx, ny, ns, LO, HI, ZLO, ZHI, KC);
        auto kern = kgp::plane_kernel<C, LO, HI, Op, ZLO, ZHI, ORD>;
        kg_register_kernel(name, (const void*)kern);
        if (int rc = kg_launch_begin(ctx, name)) return rc;
        if (!g.empty)
            hipLaunchKernelGGL(kern, dim3(g.blocks), dim3(64 * C::W), 0, ctx->stream,
                               nx, ny, ns, g.ntx, g.nty, KC, op);
        return kg_launch_end(ctx);
    };
    if (nx % VMAX == 0 && al) return go(std::integral_constant<int, VMAX>{});
    return go(std::integral_constant<int, 1>{});
}

// Per-plane stencils (gaussblur3) on the direct row body when nx % VEC == 0
// and the fields are 16-byte aligned, else on the plane engine.  W waves
// stacked in y per tile, PY output rows per lane.  Measured at 512^3 (cold):
// FP64 W=8/PY=4 415 us (plane engine 492, PY=8 426, W=16 997); FP32
// W=8/PY=8 232 us (plane engine 257, PY=4 267, PY=16 239).
template <template <class> class OpT, class T, int VEC, int W, int PY, class Fill>
static int launch_rows2(kg_ctx* ctx, const char* name, int nx, int ny, int ns, bool al, Fill fill)
{
    if (nx % VEC != 0 || !al) return launch_typed<OpT, T, VEC, 8, 2, 16>(ctx, name, nx, ny, ns, al, fill);
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    constexpr int LO = Bounds<OpT>::LO, HI = Bounds<OpT>::HI;
    static_assert(Bounds<OpT>::ZLO == 0 && Bounds<OpT>::ZHI == 0, "every plane computed");
    using C = OpCfg<OpT, T, VEC, W, 2, 16>;
    using R2 = kgp::Rows2<C, PY>;
    using Op = OpT<C>;
    Op op{};
    fill(op);
    const int nxi = nx - LO - HI, nyi = ny - LO - HI;
    const int ntx = (nx + R2::TXW - 1) / R2::TXW, nty = nyi > 0 ? (nyi + R2::TY - 1) / R2::TY : 0;
    auto kern = kgp::rows2_kernel<C, LO, HI, Op, PY, KG_ORD_ROWS2>;
    kg_register_kernel(name, (const void*)kern);
    if (int rc = kg_launch_begin(ctx, name)) return rc;
    if (nxi > 0 && nyi > 0 && ns > 0)
        hipLaunchKernelGGL(kern, dim3((unsigned)ntx * (unsigned)nty * (unsigned)ns), dim3(64 * C::W),
                           0, ctx->stream, nx, ny, ntx, nty, op);
    return kg_launch_end(ctx);
}

// Per-plane stencils on the column march (kgp::colmarch_tile): W waves per
// workgroup, each marching YC rows of one plane with a D-row prefetch.  Same
// requirements as the direct row body (nx % VEC == 0, 16-byte aligned
// fields), else the plane engine.  gaussblur3 at 512^3, levels rotated,
// first ten launches: FP32 W=4/D=4/YC=16 214 us against 255-264 on the
// direct row body (YC=8 214, 12 225, 32 238, 64 256: fewer workgroups leave
// a tail); FP64 W=4/D=2/YC=16 418 us against 451-472.  Keeping each row's
// shifted neighbour points with its ring slot (one set of DPP moves per row
// instead of one per row and use) then gave FP32 207-209 us at D=2 (D=3
// 214, D=4 210) and FP64 404-412 us (D=1 404, D=2 408-412).  A resident grid
// looping over the tiles (no tail round) was slower at 16-64 rows per wave:
// FP32 220-251 us, FP64 431-447 us; so was marching odd row chunks
// downwards, which makes two neighbouring chunks read their 4 shared rows at
// the same time (FP32 212-213 us, FP64 unchanged), and FP32 32-byte lanes
// (two 16-byte loads; a wave's piece spans a whole 512-point row, so no
// x-halo lines are read twice): 265-292 us over D=1-3, 8 rows per wave and
// 8 waves per workgroup.  Tile order (ORD, kg_tile_order), cold caches:
// chunks of 4 tile rows per XCD 202.6-202.7 us (FP32) and 393-397 us (FP64)
// against 206.9-207.4 and 404-407 in hardware order; chunks of 2 or 8 rows
// alike, 16 rows FP64 412-417 us, XCD-contiguous eighths FP32 223-225 us.
// With two levels ping-ponged (the input just written by the previous
// launch) the chunked order keeps its gain in FP64 (390 against 405 us) but
// loses in FP32 (227-250 against 213 us), so FP32 keeps the hardware order.
#ifndef KG_ORD_CM64
#define KG_ORD_CM64 4
#endif
#ifndef KG_ORD_CM32
#define KG_ORD_CM32 0
#endif
#ifndef KG_GB3_ENGINE
#define KG_GB3_ENGINE 1
#endif
template <template <class> class OpT, class T, int VEC, int W, int D, int YC, int ORD, class Fill>
static int launch_colmarch(kg_ctx* ctx, const char* name, int nx, int ny, int ns, bool al, Fill fill)
{
    if (nx % VEC != 0 || !al) return launch_typed<OpT, T, VEC, 8, 2, 16>(ctx, name, nx, ny, ns, al, fill);
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    if (int rc = kg_check_plane_bytes((long long)nx * ny, (int)sizeof(typename T::S))) return rc;
    constexpr int LO = Bounds<OpT>::LO, HI = Bounds<OpT>::HI;
    static_assert(Bounds<OpT>::ZLO == 0 && Bounds<OpT>::ZHI == 0, "every plane computed");
    constexpr int H = Shape<OpT>::H;
    using C = kgp::Cfg<VEC, W, H, 0, 0, H, H, D, 0, YC, T, 64, false, true>;
    using Op = OpT<C>;
    Op op{};
    fill(op);
    const int nxi = nx - LO - HI, nyi = ny - LO - HI;
    const int ntx = (nx + kgp::ColMarch<C>::TXW - 1) / kgp::ColMarch<C>::TXW;
    const int nty = nyi > 0 ? (nyi + W * YC - 1) / (W * YC) : 0;
    auto kern = kgp::colmarch_kernel<C, LO, HI, Op, ORD>;
    kg_register_kernel(name, (const void*)kern);
    if (int rc = kg_launch_begin(ctx, name)) return rc;
    if (nxi > 0 && nyi > 0 && ns > 0)
        hipLaunchKernelGGL(kern, dim3((unsigned)ntx * (unsigned)nty * (unsigned)ns), dim3(64 * W),
                           0, ctx->stream, nx, ny, ns, ntx, nty, YC, op);
    return kg_launch_end(ctx);
}

// ------------------------------------------------------------ statistics
constexpr int KG_STATS_BLOCKS = 1024;
constexpr int KG_STATS_THREADS = 256;

template <int DT>
__device__ __forceinline__ double ld_any(const void* p, size_t i)
{
    if constexpr (DT == KG_F64) return ((const double*)p)[i];
    else if constexpr (DT == KG_F32) return (double)((const float*)p)[i];
    else return (double)kgp::bf16_in(((const uint16_t*)p)[i]);
}

// Per-block partials: [0] sum x, [1] sum x^2, [2] max |x|, [3] sum (x-r)^2,
// [4] max |x-r|, written to part[block*5 + q].
template <int DX, int DR>
__global__ __launch_bounds__(KG_STATS_THREADS)
void stats_k(size_t n, const void* __restrict__ x, const void* __restrict__ r,
             double* __restrict__ part)
{
    __shared__ double red[5][KG_STATS_THREADS];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    const size_t stride = (size_t)gridDim.x * KG_STATS_THREADS;
    for (size_t i = (size_t)blockIdx.x * KG_STATS_THREADS + threadIdx.x; i < n; i += stride) {
        const double v = ld_any<DX>(x, i);
        a0 += v;
        a1 += v * v;
        a2 = fmax(a2, fabs(v));
        if (DR >= 0) {
            const double d = v - ld_any<(DR < 0 ? 0 : DR)>(r, i);
            a3 += d * d;
            a4 = fmax(a4, fabs(d));
        }
    }
    const int t = threadIdx.x;
    red[0][t] = a0;
    red[1][t] = a1;
    red[2][t] = a2;
    red[3][t] = a3;
    red[4][t] = a4;
    __syncthreads();
    for (int h = KG_STATS_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) {
            red[0][t] += red[0][t + h];
            red[1][t] += red[1][t + h];
            red[2][t] = fmax(red[2][t], red[2][t + h]);
            red[3][t] += red[3][t + h];
            red[4][t] = fmax(red[4][t], red[4][t + h]);
        }
        __syncthreads();
    }
    if (t < 5) part[blockIdx.x * 5 + t] = red[t][0];
}

__global__ __launch_bounds__(KG_STATS_THREADS)
void stats_final_k(int nblocks, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double red[5][KG_STATS_THREADS];
    const int t = threadIdx.x;
    double a[5] = {0, 0, 0, 0, 0};
    for (int b = t; b < nblocks; b += KG_STATS_THREADS) {
        a[0] += part[b * 5 + 0];
        a[1] += part[b * 5 + 1];
        a[2] = fmax(a[2], part[b * 5 + 2]);
        a[3] += part[b * 5 + 3];
        a[4] = fmax(a[4], part[b * 5 + 4]);
    }
#pragma unroll
    for (int q = 0; q < 5; q++) red[q][t] = a[q];
    __syncthreads();
    for (int h = KG_STATS_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) {
            red[0][t] += red[0][t + h];
            red[1][t] += red[1][t + h];
            red[2][t] = fmax(red[2][t], red[2][t + h]);
            red[3][t] += red[3][t + h];
            red[4][t] = fmax(red[4][t], red[4][t + h]);
        }
        __syncthreads();
    }
    if (t < 5) out[t] = red[t][0];
}

template <int DX, int DR>
static void launch_stats(hipStream_t st, size_t n, const void* x, const void* r, double* part,
                         double* out)
{
    hipLaunchKernelGGL((stats_k<DX, DR>), dim3(KG_STATS_BLOCKS), dim3(KG_STATS_THREADS), 0, st, n,
                       x, r, part);
    hipLaunchKernelGGL(stats_final_k, dim3(1), dim3(KG_STATS_THREADS), 0, st, KG_STATS_BLOCKS,
                       part, out);
}

template <int DX>
static int stats_dispatch_ref(hipStream_t st, int rdt, size_t n, const void* x, const void* r,
                              double* part, double* out)
{
    if (!r) launch_stats<DX, -1>(st, n, x, r, part, out);
    else if (rdt == KG_F64) launch_stats<DX, KG_F64>(st, n, x, r, part, out);
    else if (rdt == KG_F32) launch_stats<DX, KG_F32>(st, n, x, r, part, out);
    else if (rdt == KG_BF16) launch_stats<DX, KG_BF16>(st, n, x, r, part, out);
    else return kg_set_error(KG_ERR_INVALID, "kg_field_stats: unknown reference dtype %d", rdt);
    return KG_OK;
}

// ------------------------------------------------------------ conversion
template <int DS, int DD>
__global__ __launch_bounds__(256)
void convert_k(size_t n, const void* __restrict__ src, void* __restrict__ dst)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const double v = ld_any<DS>(src, i);
        if constexpr (DD == KG_F64) ((double*)dst)[i] = v;
        else if constexpr (DD == KG_F32) ((float*)dst)[i] = (float)v;
        // double -> float -> bfloat16: exact except where the float rounding
        // decides a bfloat16 tie (not for values that came from bfloat16/float)
        else ((uint16_t*)dst)[i] = kgp::bf16_out((float)v);
    }
}

template <int DS>
static int convert_to(hipStream_t st, int dd, size_t n, const void* src, void* dst)
{
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 32);
    if (n == 0) return KG_OK;
    if (dd == KG_F64) hipLaunchKernelGGL((convert_k<DS, KG_F64>), dim3(blocks), dim3(256), 0, st, n, src, dst);
    else if (dd == KG_F32) hipLaunchKernelGGL((convert_k<DS, KG_F32>), dim3(blocks), dim3(256), 0, st, n, src, dst);
    else if (dd == KG_BF16) hipLaunchKernelGGL((convert_k<DS, KG_BF16>), dim3(blocks), dim3(256), 0, st, n, src, dst);
    else return kg_set_error(KG_ERR_INVALID, "kg_convert: unknown dtype %d", dd);
    KG_HIP(hipGetLastError());
    return KG_OK;
}

}  // namespace

extern "C" int kg_convert(kg_ctx* ctx, int dst_dtype, void* dst, int src_dtype, const void* src,
                          size_t n)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (n && (!dst || !src)) return kg_set_error(KG_ERR_INVALID, "kg_convert: null buffer");
    if (src_dtype == KG_F64) return convert_to<KG_F64>(ctx->stream, dst_dtype, n, src, dst);
    if (src_dtype == KG_F32) return convert_to<KG_F32>(ctx->stream, dst_dtype, n, src, dst);
    if (src_dtype == KG_BF16) return convert_to<KG_BF16>(ctx->stream, dst_dtype, n, src, dst);
    return kg_set_error(KG_ERR_INVALID, "kg_convert: unknown dtype %d", src_dtype);
}

// ======================================================= precision variants
// Tile shapes (waves, prefetch planes, planes per workgroup) of the FP32
// kernels, from sweeps at 512^3 (cold caches) in chunked tile order:
// laplacian 12/3/8 201 us (12/3/16 209, 12/2/16 210, 8/3/16 214, 16/3/16
// 227, 12/3/32 221); wave13pt 16/3/64 312 us (12/3/64 309, 16/2/64 315,
// 16/3/32 320).
#ifndef KG_LF32_W
#define KG_LF32_W 12
#endif
#ifndef KG_LF32_D
#define KG_LF32_D 3
#endif
#ifndef KG_LF32_KC
#define KG_LF32_KC 8
#endif
#ifndef KG_WF32_W
#define KG_WF32_W 16
#endif
#ifndef KG_WF32_D
#define KG_WF32_D 3
#endif
#ifndef KG_WF32_KC
#define KG_WF32_KC 64
#endif
// BF16 shapes measured at 512^3 with the hardware bf16 conversion:
// laplacian 16-byte lanes, W=8/D=2/KC=8 110 us (KC=16 113, KC=4 123, W=12
// 128); wave13pt 16-byte lanes, W=12/D=2/KC=64 200 us (KC=32 204-221, D=3
// 215; 8-byte lanes W=16/KC=32 232-244).
#ifndef KG_LBF_V
#define KG_LBF_V 8
#endif
#ifndef KG_LBF_W
#define KG_LBF_W 8
#endif
#ifndef KG_LBF_D
#define KG_LBF_D 2
#endif
#ifndef KG_LBF_KC
#define KG_LBF_KC 8
#endif
#ifndef KG_WBF_V
#define KG_WBF_V 8
#endif
#ifndef KG_WBF_W
#define KG_WBF_W 12
#endif
#ifndef KG_WBF_D
#define KG_WBF_D 2
#endif
#ifndef KG_WBF_KC
#define KG_WBF_KC 64
#endif

extern "C" int kg_laplacian_f32(kg_ctx* ctx, int nx, int ny, int ns, float alpha, float beta,
                                const float* w0, float* w1)
{
    KG_REQUIRE(w0 && w1, "kg_laplacian_f32: null buffer");
    KG_REQUIRE((const void*)w0 != (const void*)w1, "kg_laplacian_f32: w0 and w1 must be distinct");
    return launch_typed<LaplacianOp, kgp::F32, 4, KG_LF32_W, KG_LF32_D, KG_LF32_KC>(ctx, "laplacian_f32", nx, ny, ns,
                                                            kg_aligned16(w0, w1), [&](auto& op) {
        op.src = w0;
        op.dst = w1;
        op.alpha = alpha;
        op.beta = beta;
    });
}

extern "C" int kg_laplacian_bf16(kg_ctx* ctx, int nx, int ny, int ns, float alpha, float beta,
                                 const uint16_t* w0, uint16_t* w1)
{
    KG_REQUIRE(w0 && w1, "kg_laplacian_bf16: null buffer");
    KG_REQUIRE((const void*)w0 != (const void*)w1, "kg_laplacian_bf16: w0 and w1 must be distinct");
    return launch_typed<LaplacianOp, kgp::BF16, KG_LBF_V, KG_LBF_W, KG_LBF_D, KG_LBF_KC>(ctx, "laplacian_bf16", nx, ny, ns,
                                                             kg_aligned16(w0, w1), [&](auto& op) {
        op.src = w0;
        op.dst = w1;
        op.alpha = alpha;
        op.beta = beta;
    });
}

extern "C" int kg_wave13pt_f32(kg_ctx* ctx, int nx, int ny, int ns, float m0, float m1, float m2,
                               const float* w0, const float* w1, float* w2)
{
    KG_REQUIRE(w0 && w1 && w2, "kg_wave13pt_f32: null buffer");
    KG_REQUIRE(w2 != w0 && w2 != w1, "kg_wave13pt_f32: w2 aliases an input");
    return launch_typed<Wave13ptOp, kgp::F32, 4, KG_WF32_W, KG_WF32_D, KG_WF32_KC, KG_ORD_WF32>(ctx, "wave13pt_f32", nx, ny, ns,
                                                            kg_aligned16(w0, w1, w2), [&](auto& op) {
        op.src = w1;
        op.side[0] = w0;
        op.dst = w2;
        op.m0 = m0;
        op.m1 = m1;
        op.m2 = m2;
    });
}

extern "C" int kg_wave13pt_bf16(kg_ctx* ctx, int nx, int ny, int ns, float m0, float m1, float m2,
                                const uint16_t* w0, const uint16_t* w1, uint16_t* w2)
{
    KG_REQUIRE(w0 && w1 && w2, "kg_wave13pt_bf16: null buffer");
    KG_REQUIRE(w2 != w0 && w2 != w1, "kg_wave13pt_bf16: w2 aliases an input");
    // 16-byte lanes (8 points) in 12-wave workgroups (768 threads, so up to
    // 168 registers a lane: the 16-wave tile would spill at 128), 64 planes
    // per workgroup, 2-plane prefetch; the measurements are listed with the
    // KG_WBF_* defaults above.
    return launch_typed<Wave13ptOp, kgp::BF16, KG_WBF_V, KG_WBF_W, KG_WBF_D, KG_WBF_KC, KG_ORD_WBF>(ctx, "wave13pt_bf16", nx, ny, ns,
                                                             kg_aligned16(w0, w1, w2), [&](auto& op) {
        op.src = w1;
        op.side[0] = w0;
        op.dst = w2;
        op.m0 = m0;
        op.m1 = m1;
        op.m2 = m2;
    });
}

extern "C" int kg_gaussblur3_f64(kg_ctx* ctx, int nx, int ny, int ns, double s0, double s1,
                                 double s2, double s4, double s5, double s8, const double* w0,
                                 double* w1)
{
    KG_REQUIRE(w0 && w1, "kg_gaussblur3_f64: null buffer");
    KG_REQUIRE(w0 != w1, "kg_gaussblur3_f64: w0 and w1 must be distinct");
    const double f = 1. / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5);   // gaussblur/gaussblur.c:65
#if KG_GB3_ENGINE == 1
    return launch_colmarch<Gaussblur3Op, kgp::F64, 2, KG_GB3_W64, KG_GB3_D64, KG_GB3_YC64, KG_ORD_CM64>(ctx, "gaussblur3_f64", nx, ny, ns,
#else
    return launch_rows2<Gaussblur3Op, kgp::F64, 2, 8, 4>(ctx, "gaussblur3_f64", nx, ny, ns,
#endif
                                                             kg_aligned16(w0, w1), [&](auto& op) {
        op.src = w0;
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

extern "C" int kg_gaussblur3_f32(kg_ctx* ctx, int nx, int ny, int ns, float s0, float s1,
                                 float s2, float s4, float s5, float s8, const float* w0, float* w1)
{
    KG_REQUIRE(w0 && w1, "kg_gaussblur3_f32: null buffer");
    KG_REQUIRE(w0 != w1, "kg_gaussblur3_f32: w0 and w1 must be distinct");
    const float f = 1.f / (s0 + 4 * (s1 + s2 + s4 + s8) + 8 * s5);
#if KG_GB3_ENGINE == 1
    return launch_colmarch<Gaussblur3Op, kgp::F32, 4, KG_GB3_W32, KG_GB3_D32, KG_GB3_YC32, KG_ORD_CM32>(ctx, "gaussblur3_f32", nx, ny, ns,
#else
    return launch_rows2<Gaussblur3Op, kgp::F32, 4, 8, 8>(ctx, "gaussblur3_f32", nx, ny, ns,
#endif
                                                             kg_aligned16(w0, w1), [&](auto& op) {
        op.src = w0;
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

// ================================================================ statistics

extern "C" int kg_field_stats(kg_ctx* ctx, int dtype, size_t n, const void* x, int ref_dtype,
                              const void* ref, double* out)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    KG_REQUIRE(out, "kg_field_stats: null output");
    KG_REQUIRE(n == 0 || x, "kg_field_stats: null field");
    KG_REQUIRE(dtype == KG_F64 || dtype == KG_F32 || dtype == KG_BF16,
               "kg_field_stats: unknown dtype");
    KG_REQUIRE(!ref || ref_dtype == KG_F64 || ref_dtype == KG_F32 || ref_dtype == KG_BF16,
               "kg_field_stats: unknown reference dtype");
    int rc = KG_OK;
    double* dev = ctx->stats;   // KG_STATS_BLOCKS*5 partials followed by 5 results
    kg_register_kernel("field_stats", (const void*)stats_final_k);
    if ((rc = kg_launch_begin(ctx, "field_stats"))) return rc;
    if ((rc = kg_field_stats_launch(ctx->stream, dtype, n, x, ref_dtype, ref, dev))) return rc;
    if ((rc = kg_launch_end(ctx))) return rc;
    KG_HIP(hipMemcpyAsync(out, dev + KG_STATS_BLOCKS * 5, 5 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    KG_HIP(hipStreamSynchronize(ctx->stream));
    return KG_OK;
}

int kg_field_stats_launch(hipStream_t st, int dtype, size_t n, const void* x, int ref_dtype,
                          const void* ref, double* scratch)
{
    double* res = scratch + KG_STATS_BLOCKS * 5;
    switch (dtype) {
    case KG_F64: return stats_dispatch_ref<KG_F64>(st, ref_dtype, n, x, ref, scratch, res);
    case KG_F32: return stats_dispatch_ref<KG_F32>(st, ref_dtype, n, x, ref, scratch, res);
    case KG_BF16: return stats_dispatch_ref<KG_BF16>(st, ref_dtype, n, x, ref, scratch, res);
    default: return kg_set_error(KG_ERR_INVALID, "kg_field_stats: unknown dtype %d", dtype);
    }
}

size_t kg_stats_scratch_bytes() { return (KG_STATS_BLOCKS * 5 + 5) * sizeof(double); }

int kg_preload_typed()
{
    return kg_preload_kernel((const void*)stats_final_k);
}
