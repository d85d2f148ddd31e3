// kg_vecmath.hip -- the suite's three non-stencil tests: vecadd, matvec and
// sincos.  The reference builds them with real = double (makefile.in:12
// overrides the `real ?= float` of each gcc makefile; sincos's Fortran gets
// -fdefault-real-8), so the entry points kg_vecadd, kg_matvec and kg_sincos
// compute in FP64; the _f32 entry points are single-precision forms beside
// them.  They sit outside the stencil path and complete the benchmark table.
//
//   vecadd  w2 = w0 + w1 over every point (vecadd/vecadd.c:38-84; the host
//           loop rotates the three levels)
//   matvec  y[j] = sum_i A[i + nx*j] * x[i], i ascending (matvec/matvec.c:38-70)
//   sincos  xy = sin(x) + cos(y) over every point (sincos/sincos.F90:14-74)
//
// vecadd and sincos are one-read-per-input streams: each thread takes one
// 16-byte piece of every array (two doubles or four floats), workgroups in
// address order.  matvec keeps the reference's sequential sum per row
// (bit-identical to an IEEE build), one lane per row: for 16-byte aligned
// rows the lane streams its own row (matvec_rows_k: 512 x 65536 FP64 in
// 63 us against 251 us for the LDS-staged form); otherwise a workgroup owns
// 128 rows, stages column chunks of A in LDS with coalesced loads, and each
// lane sums its own row from LDS in column order.
#include "kg_plane.h"

#include <algorithm>

namespace {

template <class T> struct Vec16;
template <> struct Vec16<double> { typedef double t __attribute__((ext_vector_type(2))); };
template <> struct Vec16<float> { typedef float t __attribute__((ext_vector_type(4))); };
template <class T> using v16 = typename Vec16<T>::t;
template <class T> constexpr int per16() { return (int)(16 / sizeof(T)); }

// One 16-byte piece per thread (pieces [0, npc)); the elements past the last
// whole piece by the first threads of workgroup 0.
template <class T>
__global__ __launch_bounds__(256)
void vecadd_k(long long npc, long long n, const T* __restrict__ w0, const T* __restrict__ w1,
              T* __restrict__ w2)
{
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q < npc) {
        const v16<T> a = ((const v16<T>*)w0)[q], b = ((const v16<T>*)w1)[q];
        kg_st(&((v16<T>*)w2)[q], a + b);
    }
    const long long tail = n - npc * per16<T>();
    if (blockIdx.x == 0 && (long long)threadIdx.x < tail) {
        const long long e = npc * per16<T>() + threadIdx.x;
        w2[e] = w0[e] + w1[e];
    }
}

// Element-wise form for buffers that are not 16-byte aligned.
template <class T>
__global__ __launch_bounds__(256)
void vecadd1_k(long long n, const T* __restrict__ w0, const T* __restrict__ w1, T* __restrict__ w2)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) w2[e] = w0[e] + w1[e];
}

__device__ __forceinline__ double sincos_one(double x, double y) { return sin(x) + cos(y); }
__device__ __forceinline__ float sincos_one(float x, float y) { return sinf(x) + cosf(y); }

template <class T>
__global__ __launch_bounds__(256)
void sincos_k(long long npc, long long n, const T* __restrict__ x, const T* __restrict__ y,
              T* __restrict__ xy)
{
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q < npc) {
        const v16<T> a = ((const v16<T>*)x)[q], b = ((const v16<T>*)y)[q];
        v16<T> o;
#pragma unroll
        for (int v = 0; v < per16<T>(); v++) o[v] = sincos_one(a[v], b[v]);
        kg_st(&((v16<T>*)xy)[q], o);
    }
    const long long tail = n - npc * per16<T>();
    if (blockIdx.x == 0 && (long long)threadIdx.x < tail) {
        const long long e = npc * per16<T>() + threadIdx.x;
        xy[e] = sincos_one(x[e], y[e]);
    }
}

template <class T>
__global__ __launch_bounds__(256)
void sincos1_k(long long n, const T* __restrict__ x, const T* __restrict__ y, T* __restrict__ xy)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) xy[e] = sincos_one(x[e], y[e]);
}

constexpr int MV_ROWS = 128;   // rows (lanes) per workgroup

// Columns per LDS chunk: 128 rows x (MV_COLS + 1) values, about 33 KB for
// either type (the extra column keeps a lane's row reads off its
// neighbours' banks).
template <class T> constexpr int mv_cols() { return sizeof(T) == 8 ? 32 : 64; }

template <class T>
__global__ __launch_bounds__(MV_ROWS)
void matvec_k(int nx, int ny, const T* __restrict__ A, const T* __restrict__ x, T* __restrict__ y)
{
    constexpr int COLS = mv_cols<T>();
    __shared__ T tile[MV_ROWS][COLS + 1];
    __shared__ T xs[COLS];
    const int j0 = blockIdx.x * MV_ROWS;
    const int t = threadIdx.x;
    const int j = j0 + t;
    T acc = 0;
    for (int c0 = 0; c0 < nx; c0 += COLS) {
        const int nc = min(COLS, nx - c0);
        // coalesced: MV_ROWS / COLS rows x COLS columns per pass of the lanes
        for (int r = t / COLS; r < MV_ROWS; r += MV_ROWS / COLS) {
            const int c = t % COLS;
            const int jj = j0 + r;
            tile[r][c] = (jj < ny && c < nc) ? A[(long long)jj * nx + c0 + c] : T(0);
        }
        if (t < COLS) xs[t] = t < nc ? x[c0 + t] : T(0);
        __syncthreads();
        for (int c = 0; c < nc; c++) acc += tile[t][c] * xs[c];
        __syncthreads();
    }
    if (j < ny) y[j] = acc;
}

// Row-streaming form for rows of a multiple of per16 values on 16-byte
// aligned A (every row starts on a 16-byte boundary): lane j sums row j
// itself, reading it with 16-byte loads CH values at a time, the next chunk
// in flight while the current one is summed (two register buffers, named at
// compile time so no copy waits for a load).  x[i] is the same for every
// lane and comes through the scalar cache.  One lane per row leaves few
// waves (65536 rows: 4 per compute unit), so the rows' bytes in flight, not
// the lane count, decide the speed.  The sum is the reference's: i
// ascending, product then add.  FP64 chunks of 16, 48 or 64 values (instead
// of 32) measured the same 61-64 us at 512 x 65536: with every lane on its
// own 4 KB row, the reads are scattered over the whole matrix at once.  A
// row-block form that reads 8 whole rows per workgroup as one contiguous
// block into LDS and sums them there (8 lanes per workgroup) was
// bit-identical and slower, 115 us: the serial sums then dominate.
template <class T> constexpr int mv_chunk() { return sizeof(T) == 8 ? 32 : 64; }

template <class T>
__device__ __forceinline__ void mv_load(__amdgpu_buffer_rsrc_t r, unsigned off, T (&buf)[mv_chunk<T>()])
{
    constexpr int P = per16<T>();
#pragma unroll
    for (int q = 0; q < mv_chunk<T>() / P; q++) {
        const auto raw = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(off + 16u * q), 0, 0);
        const v16<T> v = __builtin_bit_cast(v16<T>, raw);
#pragma unroll
        for (int e = 0; e < P; e++) buf[q * P + e] = v[e];
    }
}

// FULL: the whole chunk lies in the row (no per-value test, so the uniform
// x loads merge into wide scalar loads).
template <class T, bool FULL>
__device__ __forceinline__ T mv_sum(T acc, const T (&buf)[mv_chunk<T>()], const T* __restrict__ x,
                                    int i0, int nx)
{
#pragma unroll
    for (int c = 0; c < mv_chunk<T>(); c++)
        if (FULL || i0 + c < nx) acc = acc + buf[c] * x[i0 + c];
    return acc;
}

template <class T>
__global__ __launch_bounds__(64)
void matvec_rows_k(int nx, int ny, const T* __restrict__ A, const T* __restrict__ x, T* __restrict__ y,
                   unsigned a_bytes)
{
    constexpr int CH = mv_chunk<T>();
    const int j = (int)(blockIdx.x * 64 + threadIdx.x);
    const __amdgpu_buffer_rsrc_t r = kgp::make_rsrc(A, a_bytes);
    // lanes past the last row read nothing (offset beyond the resource)
    const unsigned row = j < ny ? (unsigned)((long long)j * nx * (long long)sizeof(T)) : kgp::KG_OOB;
    T b0[CH], b1[CH];
    T acc = 0;
    mv_load<T>(r, row, b0);
    int i0 = 0;
    for (; i0 + 2 * CH <= nx; i0 += 2 * CH) {
        mv_load<T>(r, row + (unsigned)((i0 + CH) * sizeof(T)), b1);
        acc = mv_sum<T, true>(acc, b0, x, i0, nx);
        if (i0 + 2 * CH < nx) mv_load<T>(r, row + (unsigned)((i0 + 2 * CH) * sizeof(T)), b0);
        acc = mv_sum<T, true>(acc, b1, x, i0 + CH, nx);
    }
    if (i0 < nx) {   // a last chunk, in b0 (whole or in part)
        acc = mv_sum<T, false>(acc, b0, x, i0, nx);
        if (i0 + CH < nx) {
            mv_load<T>(r, row + (unsigned)((i0 + CH) * sizeof(T)), b1);
            acc = mv_sum<T, false>(acc, b1, x, i0 + CH, nx);
        }
    }
    if (j < ny) y[j] = acc;
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

// A launch holds fewer than 2^32 threads: one per 16-byte piece (aligned
// buffers) or per element (otherwise).
int check_len(const char* who, long long n, int per_thread)
{
    if (n / per_thread > (1LL << 32) - 512)
        return kg_set_error(KG_ERR_INVALID, "%s: %lld points exceed one launch", who, n);
    return KG_OK;
}

template <class T>
int vecadd(kg_ctx* ctx, const char* who, const char* name, int nx, int ny, int ns, const T* w0,
           const T* w1, T* w2)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    if (!w0 || !w1 || !w2) return kg_set_error(KG_ERR_INVALID, "%s: null buffer", who);
    if (w2 == w0 || w2 == w1) return kg_set_error(KG_ERR_INVALID, "%s: w2 aliases an input", who);
    const long long n = (long long)nx * ny * ns;
    const bool al = kg_aligned16(w0, w1, w2);
    if (int rc = check_len(who, n, al ? per16<T>() : 1)) return rc;
    const void* k = al ? (const void*)vecadd_k<T> : (const void*)vecadd1_k<T>;
    kg_register_kernel(name, k);
    if (int rc = kg_launch_begin(ctx, name)) return rc;
    if (n > 0) {
        if (al) {
            const long long npc = n / per16<T>();
            hipLaunchKernelGGL(vecadd_k<T>, dim3(blocks_for(std::max(npc, 1LL))), dim3(256), 0,
                               ctx->stream, npc, n, w0, w1, w2);
        } else {
            hipLaunchKernelGGL(vecadd1_k<T>, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, n, w0,
                               w1, w2);
        }
    }
    return kg_launch_end(ctx);
}

template <class T>
int matvec(kg_ctx* ctx, const char* who, const char* name, int nx, int ny, const T* A, const T* x,
           T* y)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims2(nx, ny)) return rc;
    if (!A || !x || !y) return kg_set_error(KG_ERR_INVALID, "%s: null buffer", who);
    if (y == A || y == x) return kg_set_error(KG_ERR_INVALID, "%s: y aliases an input", who);
    const long long a_bytes = (long long)nx * ny * (long long)sizeof(T);
    const bool rows16 = nx % per16<T>() == 0 && !((uintptr_t)A & 15) && a_bytes < (1LL << 31);
    kg_register_kernel(name, rows16 ? (const void*)matvec_rows_k<T> : (const void*)matvec_k<T>);
    if (int rc = kg_launch_begin(ctx, name)) return rc;
    if (ny > 0 && rows16 && nx > 0)
        hipLaunchKernelGGL(matvec_rows_k<T>, dim3((ny + 63) / 64), dim3(64), 0, ctx->stream, nx, ny,
                           A, x, y, (unsigned)a_bytes);
    else if (ny > 0)
        hipLaunchKernelGGL(matvec_k<T>, dim3((ny + MV_ROWS - 1) / MV_ROWS), dim3(MV_ROWS), 0,
                           ctx->stream, nx, ny, A, x, y);
    return kg_launch_end(ctx);
}

template <class T>
int sincos_(kg_ctx* ctx, const char* who, const char* name, int nx, int ny, int ns, const T* x,
            const T* y, T* xy)
{
    if (int rc = kg_check_ctx(ctx)) return rc;
    if (int rc = kg_check_dims3(nx, ny, ns)) return rc;
    if (!x || !y || !xy) return kg_set_error(KG_ERR_INVALID, "%s: null buffer", who);
    if (xy == x || xy == y) return kg_set_error(KG_ERR_INVALID, "%s: xy aliases an input", who);
    const long long n = (long long)nx * ny * ns;
    const bool al = kg_aligned16(x, y, xy);
    if (int rc = check_len(who, n, al ? per16<T>() : 1)) return rc;
    const void* k = al ? (const void*)sincos_k<T> : (const void*)sincos1_k<T>;
    kg_register_kernel(name, k);
    if (int rc = kg_launch_begin(ctx, name)) return rc;
    if (n > 0) {
        if (al) {
            const long long npc = n / per16<T>();
            hipLaunchKernelGGL(sincos_k<T>, dim3(blocks_for(std::max(npc, 1LL))), dim3(256), 0,
                               ctx->stream, npc, n, x, y, xy);
        } else {
            hipLaunchKernelGGL(sincos1_k<T>, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, n, x,
                               y, xy);
        }
    }
    return kg_launch_end(ctx);
}

}  // namespace

extern "C" int kg_vecadd(kg_ctx* ctx, int nx, int ny, int ns, const double* w0, const double* w1,
                         double* w2)
{
    return vecadd<double>(ctx, "kg_vecadd", "vecadd", nx, ny, ns, w0, w1, w2);
}

extern "C" int kg_matvec(kg_ctx* ctx, int nx, int ny, const double* A, const double* x, double* y)
{
    return matvec<double>(ctx, "kg_matvec", "matvec", nx, ny, A, x, y);
}

extern "C" int kg_sincos(kg_ctx* ctx, int nx, int ny, int ns, const double* x, const double* y,
                         double* xy)
{
    return sincos_<double>(ctx, "kg_sincos", "sincos", nx, ny, ns, x, y, xy);
}

extern "C" int kg_vecadd_f32(kg_ctx* ctx, int nx, int ny, int ns, const float* w0, const float* w1,
                             float* w2)
{
    return vecadd<float>(ctx, "kg_vecadd_f32", "vecadd_f32", nx, ny, ns, w0, w1, w2);
}

extern "C" int kg_matvec_f32(kg_ctx* ctx, int nx, int ny, const float* A, const float* x, float* y)
{
    return matvec<float>(ctx, "kg_matvec_f32", "matvec_f32", nx, ny, A, x, y);
}

extern "C" int kg_sincos_f32(kg_ctx* ctx, int nx, int ny, int ns, const float* x, const float* y,
                             float* xy)
{
    return sincos_<float>(ctx, "kg_sincos_f32", "sincos_f32", nx, ny, ns, x, y, xy);
}

int kg_preload_vecmath() { return kg_preload_kernel((const void*)vecadd_k<double>); }
