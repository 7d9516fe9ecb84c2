// mgs_device.h -- launch geometry and device helpers shared by the kernel
// files of libmgsolve.
#pragma once

#include "mgs_common.h"

namespace mgs {

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
struct Launch {
    dim3 grid, block;
    int TX, TY, XB;
    unsigned NB;
};

// blocks of (TX, TY) threads over `rows` rows of `width` cells each
inline Launch make_launch(int width, long rows)
{
    Launch L;
    L.TX = width < 256 ? width : 256;
    if (L.TX < 1) L.TX = 1;
    int ty = 256 / L.TX;
    if (ty < 1) ty = 1;
    if (ty > rows) ty = (int)rows;
    L.TY = ty;
    L.XB = (width + L.TX - 1) / L.TX;
    const long rb = (rows + L.TY - 1) / L.TY;
    L.NB = (unsigned)(rb * L.XB);
    const unsigned per = (L.NB + 7) / 8;
    L.grid = dim3(per * 8);
    L.block = dim3(L.TX, L.TY);
    return L;
}

// XCD-aware mapping: the hardware hands consecutive block ids round-robin to
// the 8 XCDs; remap so each XCD gets a contiguous run of logical blocks.
__device__ inline bool map_block(unsigned NB, unsigned& lb)
{
    const unsigned b = blockIdx.x;
    const unsigned per = (NB + 7) >> 3;
    lb = (b & 7u) * per + (b >> 3);
    return lb < NB;
}

// logical (x index, row index) of this thread
// (rows < 2^31 and 32-bit index math throughout: n^3 <= 2^30 for n <= 1024)
__device__ inline bool map_xy(int width, long rows, int TX, int TY, int XB, unsigned NB, int& x, unsigned& row)
{
    unsigned lb;
    if (!map_block(NB, lb)) return false;
    unsigned xb, rb;
    if (XB == 1) {
        xb = 0;
        rb = lb;
    } else {
        xb = lb % (unsigned)XB;
        rb = lb / (unsigned)XB;
    }
    x = (int)(xb * TX + threadIdx.x);
    row = rb * (unsigned)TY + threadIdx.y;
    return x < width && row < (unsigned)rows;
}

// write one block partial; idle blocks of the rounded-up grid write nothing
__device__ inline void store_partial(double* partials, unsigned NB, double t)
{
    unsigned lb;
    if (threadIdx.x == 0 && threadIdx.y == 0 && partials && map_block(NB, lb)) partials[lb] = t;
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Element i of a device array addressed by a 32-bit byte offset: the access
// then takes the (uniform) array base from scalar registers and one 32-bit
// vector offset, instead of a 64-bit vector address per access.  Requires
// i * sizeof(T) < 2^32: a level holds at most 2^30 cells, so any float index
// of a level, or any index within one colour half of a double level.
template <typename T>
__device__ __forceinline__ const T& elem(const T* a, unsigned i)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a) + i * (unsigned)sizeof(T));
}

template <typename T>
__device__ __forceinline__ T& elem(T* a, unsigned i)
{
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(a) + i * (unsigned)sizeof(T));
}

// Loads and stores of values a pass reads or writes once (a cell's own
// data, its result; never the field whose window neighbours re-read).  With
// NT they are non-temporal: on levels whose arrays exceed the last-level
// cache (NT_MIN_N), and in batched kernels whose one pass streams the level
// 0 of many requests, that keeps the L2 for the re-read field; on smaller
// levels the cache holds the arrays from one pass to the next, and plain
// accesses keep them there.
constexpr int NT_MIN_N = 512;

template <bool NT, typename T>
__device__ __forceinline__ T ld_once(const T& r)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(&r);
    else
        return r;
}

template <bool NT, typename T>
__device__ __forceinline__ void st_once(T& r, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, &r);
    else
        r = v;
}

// element i of a whole level's array (up to 2^30 cells): elem() where the
// byte offset fits 32 bits for any index (4-byte elements), a[i] otherwise
template <typename T>
__device__ __forceinline__ const T& lvl(const T* a, unsigned i)
{
    if constexpr (sizeof(T) <= 4)
        return elem(a, i);
    else
        return a[i];
}

template <typename T>
__device__ __forceinline__ T& lvl(T* a, unsigned i)
{
    if constexpr (sizeof(T) <= 4)
        return elem(a, i);
    else
        return a[i];
}
template <int FRN, typename T>
__device__ __forceinline__ T wpow(T x)
{
    if constexpr (FRN == 0)
        return x;
    else if constexpr (FRN == 1)
        return x * x;
    else
        return x * x * x;
}

// the six colour-split neighbours of a colour-c cell at (m,j,k), as 32-bit
// offsets (a level holds at most 2^30 cells)
struct Nb {
    unsigned xp, xm, yp, ym, zp, zm, me;
};

__device__ __forceinline__ Nb neighbours(const Geo& g, int c, int m, int j, int k)
{
    Nb r;
    const int s = (c + j + k) & 1;
    int mxm = m + s - 1;
    if (mxm < 0) mxm += g.h;
    int mxp = m + s;
    if (mxp >= g.h) mxp -= g.h;
    const unsigned n = (unsigned)g.n, h = (unsigned)g.h;
    const unsigned o = (unsigned)(1 - c) * (unsigned)g.half;
    const unsigned r0 = ((unsigned)k * n + (unsigned)j) * h;
    r.xp = o + r0 + mxp;
    r.xm = o + r0 + mxm;
    r.yp = o + ((unsigned)k * n + (unsigned)wrap_inc(j, g.n)) * h + m;
    r.ym = o + ((unsigned)k * n + (unsigned)wrap_dec(j, g.n)) * h + m;
    r.zp = o + ((unsigned)wrap_inc(k, g.n) * n + (unsigned)j) * h + m;
    r.zm = o + ((unsigned)wrap_dec(k, g.n) * n + (unsigned)j) * h + m;
    r.me = (unsigned)c * (unsigned)g.half + r0 + m;
    return r;
}

// block sum of one double per thread; thread 0 of the block gets the total
__device__ inline double block_sum(double v)
{
    __shared__ double wsum[64];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int nthr = blockDim.x * blockDim.y;
    if (nthr < 64) {
        // a partial wave: lanes beyond the block do not exist, so no shuffles
        wsum[tid] = v;
        __syncthreads();
        double t = 0.0;
        if (tid == 0)
            for (int w = 0; w < nthr; w++) t += wsum[w];
        return t;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) wsum[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (tid == 0) {
        const int nw = (nthr + 63) >> 6;
        for (int w = 0; w < nw; w++) t += wsum[w];
    }
    return t;
}

// sum of p[lo, hi) by one block: thread t adds p[lo + t], p[lo + t + B], ...
// (B = blockDim.x, a power of two <= 1024), then a tree over the block in
// s[0..B); the result is s[0].  The order depends only on (lo, hi, B).
__device__ inline void block_chunk_sum(const double* __restrict__ p, int lo, int hi, double* s)
{
    double v = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += blockDim.x) v += p[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
}

}  // namespace mgs
