// mgs_common.h -- shared device/host definitions of libmgsolve.
//
// Grid layout on the device ("colour-split"): a level with n cells per side
// (n even) is stored as two half arrays, colour 0 ("red", i+j+k even) first,
// then colour 1 ("black").  Each half is indexed [k][j][m], m = i >> 1 in
// [0, n/2), so one colour pass of the red-black smoother reads and writes
// whole cache lines.  Cell (i,j,k) here is the reference's FI(M1,M2,M3) with
// (M1,M2,M3) = (i+1,j+1,k+1); the reference's red pass (redstep=.TRUE.,
// PMP2MGsolver_fR.f90:88-96) updates M1+M2+M3 odd, i.e. colour 0 here.
//
// Neighbours of a colour-c cell at (m,j,k) all have colour 1-c: the x
// neighbours sit at m+s-1 and m+s (s = (c+j+k)&1 = i&1, wrapped into
// [0,n/2)), the y and z neighbours at the same m in rows j+-1 / k+-1.
// The 2x2x2 children of coarse cell (I,J,K) are, in each of the four fine
// rows (2J+b, 2K+c), the entry m = I of both colours.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mgs {

struct Geo {
    int n;        // cells per side
    int h;        // n / 2
    size_t half;  // n * n * h: cells of one colour
};

__host__ __device__ inline Geo make_geo(int n)
{
    Geo g;
    g.n = n;
    g.h = n >> 1;
    g.half = (size_t)n * (size_t)n * (size_t)g.h;
    return g;
}

// flat offset of cell (i,j,k) in the colour-split array
__host__ __device__ inline size_t split_index(const Geo& g, int i, int j, int k)
{
    const int c = (i + j + k) & 1;
    return (size_t)c * g.half + ((size_t)k * g.n + j) * g.h + (i >> 1);
}

__host__ __device__ inline size_t row_base(const Geo& g, int j, int k)
{
    return ((size_t)k * g.n + j) * g.h;
}

__host__ __device__ inline int wrap_dec(int a, int n) { return a == 0 ? n - 1 : a - 1; }
__host__ __device__ inline int wrap_inc(int a, int n) { return a == n - 1 ? 0 : a + 1; }

// f(R) model constants of one level (computed on the host in double,
// mgs_fr_coefs in mgs_api.cpp), PMP2MGsolver_fR.f90:54-71, 333-346.
struct FrCoef {
    // smoother (closed form u^{n+2} + P u + Q = 0)
    double fct1s, fct2s, Q2, Q, Q2onethird, Q4, Delta11, Delta20;
    // residual / RHS
    double fct1r, fct2r, fct3r, a_fR_rt;
    // fused level-1 restriction (dx = 1)
    double fct3_1, fct4_1;
    double ubg;
    // normalisation of the local polynomial for the FP32 seed:
    // x = u / ubg solves x^{n+2} + (P inv_w) x + Q2 inv_w = 0, inv_w = ubg^-(n+1)
    double inv_w;
    double inv_ubg;  // 1 / ubg
    int reference_roots;  // 1: evaluate the reference's closed form literally
};

// Delta-form (mixed precision) constants of one level:
//   D(d) = A (sum_nb d - 6 d) + C (1/xb - 1/x),  x^{n+1} = xb^{n+1} + d
struct DfCoef {
    float A, C;       // A = ctilde^2 s ubg^{n+1} / dx^2, C = Rbar a^2 / 3
    float inv6A;      // 1 / (6 A)
    double A64, C64;  // same in double (coarsest-level FP64 solve)
};

}  // namespace mgs
