// kg_internal.h -- shared pieces of libkgstencil.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "../../include/kg_stencil.h"

struct kg_ctx {
    int device;
    hipStream_t stream;
    int timing;                       // per-launch event pairs on/off
    hipEvent_t* ev_start;             // ring of KG_TIMING_RING pairs
    hipEvent_t* ev_stop;
    long long launches;               // launches recorded so far
    long long drained;                // launches already returned by kg_timing_drain
    const char* last_kernel;          // name of the most recent launch
    double* stats;                    // device partials of kg_field_stats
    int cus;                          // compute units of the device
    int lds_cu;                       // LDS bytes per compute unit
};

// Bytes of the context's kg_field_stats scratch.
size_t kg_stats_scratch_bytes();
// The kernels of kg_field_stats on stream st, with their partials in
// scratch (kg_stats_scratch_bytes() bytes); the 5 results land in the last
// 5 doubles of scratch.  No timing, no copy, no synchronisation.
int kg_field_stats_launch(hipStream_t st, int dtype, size_t n, const void* x, int ref_dtype,
                          const void* ref, double* scratch);
// Per-device set-up of the Poisson solver's static workspace (kg_init).
int kg_poisson_static_setup(int device);

// Records the most recent error message (thread-local) and returns the code.
int kg_set_error(int code, const char* fmt, ...);

// Wrap a HIP call: on failure record the message and return KG_ERR_HIP.
#define KG_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t kg_e_ = (call);                                                         \
        if (kg_e_ != hipSuccess)                                                           \
            return kg_set_error(KG_ERR_HIP, "%s failed: %s (%s:%d)", #call,                \
                                hipGetErrorString(kg_e_), __FILE__, __LINE__);             \
    } while (0)

// Launch bookkeeping used by every entry point: kg_launch_begin() records the
// start event of the next ring slot, kg_launch_end() checks the launch and
// records the stop event.
int kg_launch_begin(kg_ctx* ctx, const char* name);
int kg_launch_end(kg_ctx* ctx);

// Registry of entry-point name -> kernel symbol, for kg_kernel_regcount.
struct kg_kernel_entry {
    const char* name;
    const void* func;
};
void kg_register_kernel(const char* name, const void* func);

// Device code is loaded lazily, one code object per source file, at the first
// launch of any of its kernels (about 1 ms).  kg_init calls these so that the
// load happens in the init phase, not inside the first timed launch.
int kg_preload_stencil3d();
int kg_preload_tricubic();
int kg_preload_stencil2d();
int kg_preload_matmul();
int kg_preload_util();
int kg_preload_life();
int kg_preload_typed();
int kg_preload_batch();
int kg_preload_poisson();
int kg_preload_vecmath();
int kg_preload_kernel(const void* func);

// Validation helpers shared by the entry points.
int kg_check_ctx(kg_ctx* ctx);
int kg_check_dims3(int nx, int ny, int ns);
int kg_check_dims2(int nx, int ny);
int kg_check_plane_bytes(long long points, int elem_bytes);
// True when every pointer is 16-byte aligned: the kernels' 16-byte lanes
// (two doubles, four floats) are taken only then, 8-byte or narrower ones
// otherwise.
template <class... P>
inline bool kg_aligned16(const P*... p)
{
    return !((0u | ... | (unsigned)((uintptr_t)p & 15u)));
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; each XCD
// has its own L2.  kg_xcd_swizzle maps the hardware workgroup id to a logical
// tile id so that XCD x receives the contiguous logical range
// [x*n/8, (x+1)*n/8): neighbouring tiles (which share halo cache lines) then
// sit in the same L2.  Ids past the last multiple of 8 map to themselves.
#ifndef KG_XCD_SWIZZLE
#define KG_XCD_SWIZZLE 1
#endif
__device__ __forceinline__ unsigned kg_xcd_swizzle(unsigned b, unsigned n)
{
    if (!KG_XCD_SWIZZLE) return b;
    const unsigned per = n >> 3;
    if (b < (per << 3)) return (b & 7u) * per + (b >> 3);
    return b;
}

#define KG_NUM_XCD 8

// Dynamic LDS to reserve (unused) so that at most occ workgroups, each with
// static_lds bytes of its own, are resident on one compute unit: occ+1
// workgroups of lds_cu/occ - 256 bytes do not fit in lds_cu.  0 when occ <= 0
// or the workgroup's own LDS already allows no more.
inline size_t kg_occupancy_lds(const kg_ctx* ctx, int occ, size_t static_lds)
{
    if (occ <= 0) return 0;
    const size_t per = (size_t)ctx->lds_cu / (size_t)occ;
    return per > static_lds + 256 ? per - static_lds - 256 : 0;
}

// kg_xcd_chunk deals runs of `chunk` consecutive logical tiles to the XCDs in
// turn: XCD x receives logical chunks x, x+8, x+16, ..., so at any moment
// the 8 XCDs work on 8 consecutive chunks (the chip streams one window) and
// the tiles of a chunk share one L2.  Ids past the last whole round of 8
// chunks map to themselves.
__device__ __forceinline__ unsigned kg_xcd_chunk(unsigned b, unsigned n, unsigned chunk)
{
    if (!KG_XCD_SWIZZLE) return b;
    const unsigned round = chunk * KG_NUM_XCD;
    if (b >= n / round * round) return b;
    const unsigned x = b % KG_NUM_XCD, s = b / KG_NUM_XCD;   // XCD, its sequence number
    return ((s / chunk) * KG_NUM_XCD + x) * chunk + s % chunk;
}

// Logical tile id of workgroup b of n, for tiles numbered x fastest with ntx
// tiles per tile row.  ORD 0: hardware order (consecutive tiles on different
// XCDs; the chip streams one window).  ORD 1: each XCD takes a contiguous
// eighth (kg_xcd_swizzle).  ORD >= 2: chunks of ORD tile rows per XCD
// (kg_xcd_chunk): a tile's x-neighbours and most of its y-neighbours share
// its L2, and the chip still streams one window of 8 * ORD tile rows.
template <int ORD>
__device__ __forceinline__ unsigned kg_tile_order(unsigned b, unsigned n, unsigned ntx)
{
    static_assert(ORD >= 0, "tile order");
    if constexpr (ORD == 0) return b;
    else if constexpr (ORD == 1) return kg_xcd_swizzle(b, n);
    else return kg_xcd_chunk(b, n, ntx * (unsigned)ORD);
}

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N-1>{}):
// a loop unrolled at compile time with the index usable as a constant.
template <class Fn, int... I>
__device__ __forceinline__ void kg_static_for_impl(Fn&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class Fn>
__device__ __forceinline__ void kg_static_for(Fn&& f)
{
    kg_static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Streaming store for stencil outputs.  Output grids are written once per step
// and not read again by the same kernel; marking the store non-temporal keeps
// it from displacing, in L2, the input lines that neighbouring rows and planes
// are about to re-read.  Measured on MI355X (512x256x256 laplacian): 153 us
// with ordinary stores, 87 us with non-temporal ones.
template <typename T>
__device__ __forceinline__ void kg_st(T* p, T v)
{
    __builtin_nontemporal_store(v, p);
}
