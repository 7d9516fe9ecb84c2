// mgs_fr_dev.h -- device-side f(R) operators shared by the standalone
// kernels (mgs_fr.hip) and the batched serving kernels (mgs_serve.hip):
// local roots, stencil ops for the z-marching drivers, and per-column /
// per-cell bodies of the transfer kernels.
#pragma once

#include <math.h>

#include "mgs_device.h"
#include "mgs_stencil.h"

namespace mgs {

// ---------------------------------------------------------------------------
// closed-form local roots of u^{n+2} + P u + Q = 0 as the reference writes
// them, PMP2MGsolver_fR.f90:105-112 (n=0), 159-179 (n=1), 240-256 (n=2)
// ---------------------------------------------------------------------------
template <int FRN>
__device__ __forceinline__ double fr_root64(double P, const FrCoef& c)
{
    const double ZERO = 1.0e-20;
    const double ONEOVERTHREE = 1.0 / 3.0;
    const double TWOOVERTHREE = 2.0 / 3.0;
    const double TWOPIOVERTHREE = 2.0 / 3.0 * 3.14159265358979323846;
    if constexpr (FRN == 0) {
        return 0.5 * (-P + sqrt(P * P - c.Q4));
    } else if constexpr (FRN == 1) {
        const double Delta10 = -3.0 * P;
        const double CC = c.Delta11 * c.Delta11 - 4.0 * Delta10 * Delta10 * Delta10;
        if (CC >= 0.0) {
            const double tmp1 = 0.5 * (sqrt(CC) + c.Delta11);
            const double tmp2 = tmp1 - c.Delta11;
            if (fabs(P) < ZERO) return c.Q2onethird;
            return -(copysign(pow(fabs(tmp1), ONEOVERTHREE), tmp1) - copysign(pow(fabs(tmp2), ONEOVERTHREE), tmp2)) *
                   ONEOVERTHREE;
        }
        const double sd = sqrt(Delta10);
        const double S = acos(0.5 * c.Delta11 / (sd * sd * sd));
        return -TWOOVERTHREE * sd * cos(S * ONEOVERTHREE + TWOPIOVERTHREE);
    } else {
        const double Delta21 = 27.0 * P * P;
        const double tmp1 = 0.5 * (sqrt(Delta21 * Delta21 - 4.0 * c.Delta20 * c.Delta20 * c.Delta20) + Delta21);
        const double cb = copysign(pow(fabs(tmp1), ONEOVERTHREE), tmp1);
        const double S = sqrt(cb + c.Delta20 / cb) / sqrt(12.0);
        if (P > ZERO) return -S + 0.5 * sqrt(-4.0 * S * S + P / S);
        if (P < -ZERO) return S + 0.5 * sqrt(-4.0 * S * S - P / S);
        return sqrt(sqrt(-c.Q));
    }
}

// ---------------------------------------------------------------------------
// Seed of the positive root of x^{n+2} + P x + Q = 0 (Q < 0, O(1) normalised
// coefficients) in FP32, written without the cancellations of the literal
// closed forms:
//   n=0: the quadratic root, as -2Q/(P + sqrt(P^2-4Q)) when P > 0;
//   n=1: Cardano with t1 = 54 P^3 / (sqrt(CC) + |27Q|) (= t1 of the literal
//        form) and, when P > 0, cbrt(t2) - cbrt(t1) = |27Q| / (a^2 + ab + b^2);
//        the trigonometric form when CC < 0;
//   n=2: Ferrari for P <= 0; for P > 0 the bound min(|Q|/P, |Q|^{1/4}), which
//        lies within a factor 2 above the root.
// Hardware transcendentals: the seed is refined by Newton's method.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fast_cbrt(float t)
{
    if (t == 0.0f) return 0.0f;
    const float r = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(fabsf(t)) * (1.0f / 3.0f));
    return copysignf(r, t);
}

// hardware square root and reciprocal (about 1 ulp): the seed only has to
// be close, the Newton steps that follow set the accuracy
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <int FRN>
__device__ __forceinline__ float seed_root(float P, float Q)
{
    if constexpr (FRN == 0) {
        const float sq = hw_sqrt(fmaxf(P * P - 4.0f * Q, 0.0f));
        return P > 0.0f ? (-2.0f * Q) * hw_rcp(P + sq) : 0.5f * (sq - P);
    } else if constexpr (FRN == 1) {
        const float aD1 = -27.0f * Q;
        const float CC = aD1 * aD1 + 108.0f * P * P * P;
        if (CC >= 0.0f) {
            const float sq = hw_sqrt(CC);
            const float t2 = 0.5f * (sq + aD1);
            const float t1 = 54.0f * P * P * P * hw_rcp(sq + aD1);
            const float a = fast_cbrt(t2), b = fast_cbrt(t1);
            if (t1 > 0.0f) return aD1 * hw_rcp(3.0f * (a * a + a * b + b * b));
            return (a - b) * (1.0f / 3.0f);
        }
        const float sd = hw_sqrt(-3.0f * P);
        const float arg = fminf(fmaxf(-0.5f * aD1 * hw_rcp(sd * sd * sd), -1.0f), 1.0f);
        return -(2.0f / 3.0f) * sd * __cosf(acosf(arg) * (1.0f / 3.0f) + 2.0943951023931955f);
    } else {
        const float q4 = hw_sqrt(hw_sqrt(-Q));
        if (P > 0.0f) return fminf(-Q * hw_rcp(P), q4);
        if (P == 0.0f) return q4;
        const float D20 = 12.0f * Q;
        const float D21 = 27.0f * P * P;
        const float t1 = 0.5f * (hw_sqrt(fmaxf(D21 * D21 - 4.0f * D20 * D20 * D20, 0.0f)) + D21);
        const float cb = fast_cbrt(t1);
        const float S = hw_sqrt(fmaxf(cb + D20 * hw_rcp(cb), 0.0f)) * 0.28867513459481287f;
        return S + 0.5f * hw_sqrt(fmaxf(-4.0f * S * S - P * hw_rcp(S), 0.0f));
    }
}

// 1/x in FP64 from the hardware estimate and one Newton-Raphson correction
// (relative error ~1e-15: enough for a Newton step's length)
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    const double e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Newton's method in FP64 on u^{n+2} + P u + Q = 0 from u.  With Q < 0 the
// function is convex for u > 0 and negative at 0, so its derivative at the
// positive root is at least -Q/u > 0 and the iteration converges
// quadratically with an O(1) constant in the normalised variable: once a
// step is below 1e-11 u the error left is ~1e-22 u, so it stops there.
template <int FRN>
__device__ __forceinline__ double newton_root64(double u, double P, double Q)
{
#pragma unroll 1
    for (int it = 0; it < 8; it++) {
        const double un1 = wpow<FRN>(u);
        const double f = un1 * u + P * u + Q;
        const double fp = (double)(FRN + 2) * un1 + P;
        if (!(fp > 0.0)) break;
        const double d = f * rcp64(fp);
        const double u1 = u - d;
        if (!(u1 > 0.0)) break;
        u = u1;
        if (fabs(d) <= 1.0e-11 * u) break;
    }
    return u;
}

// 32-bit colour-split offset of cell (i,j,k)
__device__ __forceinline__ unsigned sidx(const Geo& g, int i, int j, int k)
{
    return (unsigned)((i + j + k) & 1) * (unsigned)g.half + ((unsigned)k * (unsigned)g.n + (unsigned)j) * (unsigned)g.h +
           (unsigned)(i >> 1);
}

// ===========================================================================
// Plain FP64 path
// ===========================================================================

// relaxation_iterations_fR, one colour, PMP2MGsolver_fR.f90:20-301.
// REFROOT: the reference's closed form evaluated literally; otherwise the
// same root from the FP32 seed refined by FP64 Newton steps.
template <int FRN, bool COARSE, bool REFROOT, bool NT = false>
struct FrSmoothOp {
    using T = double;
    struct Own {
        double rho, f;
    };
    double* u;
    const double* rho;
    const double* f;
    FrCoef c;
    __device__ double* field() const { return u; }
    __device__ __forceinline__ Own load(unsigned o, unsigned) const
    {
        Own w;
        w.rho = ld_once<NT>(rho[o]);
        w.f = COARSE ? ld_once<NT>(f[o]) : 0.0;
        return w;
    }
    __device__ __forceinline__ void update(unsigned o, unsigned, const Own& w, double, double xp, double xm,
                                           double yp, double ym, double zp, double zm) const
    {
        const double S =
            wpow<FRN>(xp) + wpow<FRN>(xm) + wpow<FRN>(yp) + wpow<FRN>(ym) + wpow<FRN>(zp) + wpow<FRN>(zm);
        double P;
        if constexpr (!COARSE)
            P = c.fct1s * w.rho - c.Q2 - S / 6.0;
        else
            P = c.fct1s * w.rho - c.Q2 + c.fct2s * w.f - S / 6.0;
        if constexpr (REFROOT) {
            st_once<NT>(u[o], fr_root64<FRN>(P, c));
        } else {
            float x = seed_root<FRN>((float)(P * c.inv_w), (float)(c.Q2 * c.inv_w));
            if (!(x > 0.0f) || !(x < 3.0e38f)) x = 1.0f;  // Newton from u_bg
            st_once<NT>(u[o], newton_root64<FRN>((double)x * c.ubg, P, c.Q));
        }
    }
};

// residual of one cell, calculate_residual_fR, PMP2MGsolver_fR.f90:434-458;
// COARSE subtracts f (:482-488).  Also the FAS right-hand side (:836-923)
// with (u, f, out) = (R u, R r, f_l).
template <int FRN, bool COARSE, bool NT = false>
struct FrResOp {
    using T = double;
    struct Own {
        double rho, f;
    };
    const double* u;
    const double* rho;
    const double* f;
    double* out;
    FrCoef c;
    __device__ const double* field() const { return u; }
    __device__ __forceinline__ Own load(unsigned o) const
    {
        Own w;
        w.rho = ld_once<NT>(rho[o]);
        w.f = COARSE ? ld_once<NT>(f[o]) : 0.0;
        return w;
    }
    __device__ __forceinline__ double cell(unsigned o, const Own& w, double uc, double xp, double xm, double yp,
                                           double ym, double zp, double zm) const
    {
        const double S =
            wpow<FRN>(xp) + wpow<FRN>(xm) + wpow<FRN>(yp) + wpow<FRN>(ym) + wpow<FRN>(zp) + wpow<FRN>(zm);
        double OP = S - wpow<FRN>(uc) * 6.0;
        OP = OP * c.fct3r;
        OP = OP + c.fct1r * w.rho;
        OP = OP + c.fct2r * (c.a_fR_rt / uc - 1.0);
        if constexpr (COARSE) OP = OP - w.f;
        if (out) st_once<NT>(out[o], OP);
        return OP;
    }
};

// Fine cells seen from coarse column (I, J) in one fine plane z: the 2x2
// children (x0,y0), (x1,y0), (x0,y1), (x1,y1) and the 8 face neighbours
// (xl,y0), (xl,y1), (xu,y0), (xu,y1), (x0,yl), (x1,yl), (x0,yu), (x1,yu),
// kept as colour parity and in-plane offset so that a plane loads with one
// add per value
struct Pts12 {
    unsigned par[12], row[12], plane, half;
    __device__ __forceinline__ void init(const Geo& gf, int I, int J)
    {
        const int N = gf.n;
        const int x0 = 2 * I, x1 = 2 * I + 1, y0 = 2 * J, y1 = 2 * J + 1;
        const int xl = (I == 0) ? N - 1 : 2 * I - 1, xu = (2 * I + 2 == N) ? 0 : 2 * I + 2;
        const int yl = (J == 0) ? N - 1 : 2 * J - 1, yu = (2 * J + 2 == N) ? 0 : 2 * J + 2;
        const int X[12] = {x0, x1, x0, x1, xl, xl, xu, xu, x0, x1, x0, x1};
        const int Y[12] = {y0, y0, y1, y1, y0, y1, y0, y1, yl, yl, yu, yu};
#pragma unroll
        for (int q = 0; q < 12; q++) {
            par[q] = (unsigned)((X[q] + Y[q]) & 1);
            row[q] = (unsigned)Y[q] * (unsigned)gf.h + (unsigned)(X[q] >> 1);
        }
        plane = (unsigned)gf.n * (unsigned)gf.h;
        half = (unsigned)gf.half;
    }
    __device__ __forceinline__ unsigned at(int q, int z) const
    {
        return ((par[q] + (unsigned)z) & 1u) * half + (unsigned)z * plane + row[q];
    }
};

// Coarse 3x3 block around column (I, J) of a coarse level: the colour
// parity and in-plane offset of each (x, y) of the block, so that a plane K
// of the block loads with two adds per value
struct Block9 {
    // x offsets (m) of columns I-1, I, I+1 and row offsets of rows J-1, J, J+1;
    // the colour of (x, y) is p0 ^ (x + y) & 1 (nc even: the wrap keeps parity)
    unsigned xo[3], yo[3], p0, plane, half;
    __device__ __forceinline__ void init(const Geo& gc, int I, int J)
    {
        const int nc = gc.n;
        xo[0] = (unsigned)(wrap_dec(I, nc) >> 1);
        xo[1] = (unsigned)(I >> 1);
        xo[2] = (unsigned)(wrap_inc(I, nc) >> 1);
        yo[0] = (unsigned)wrap_dec(J, nc) * (unsigned)gc.h;
        yo[1] = (unsigned)J * (unsigned)gc.h;
        yo[2] = (unsigned)wrap_inc(J, nc) * (unsigned)gc.h;
        p0 = (unsigned)((I + J) & 1);
        plane = (unsigned)gc.n * (unsigned)gc.h;
        half = (unsigned)gc.half;
    }
    // offset of point q = x + 3 y of the 3x3 block in coarse plane K
    __device__ __forceinline__ unsigned at(int q, int K) const
    {
        const int x = q % 3, y = q / 3;
        return ((p0 + (unsigned)(x + y) + (unsigned)K) & 1u) * half + (unsigned)K * plane + yo[y] + xo[x];
    }
};

// offset of son (2I+s1, 2J+s2, 2K+s3) on the fine level: colour (s1+s2+s3)&1,
// m = I
__device__ __forceinline__ unsigned son_index(const Geo& gf, int I, int J, int K, int s1, int s2, int s3)
{
    return (unsigned)((s1 + s2 + s3) & 1) * (unsigned)gf.half +
           ((unsigned)(2 * K + s3) * (unsigned)gf.n + (unsigned)(2 * J + s2)) * (unsigned)gf.h + (unsigned)I;
}

// son (2I+s1, 2J+s2, 2K+s3) of a double level: its colour half (the colour
// is known at compile time, so the half's base stays scalar) and an offset in
// it, which fits a 32-bit byte offset (a half holds at most 2^29 cells)
__device__ __forceinline__ double& son64(double* uf, const Geo& gf, int I, int J, int K, int s1, int s2, int s3)
{
    double* h = uf + (size_t)((s1 + s2 + s3) & 1) * gf.half;
    return elem(h, ((unsigned)(2 * K + s3) * (unsigned)gf.n + (unsigned)(2 * J + s2)) * (unsigned)gf.h + (unsigned)I);
}

// correct_solution, PMP2extradof.f90:1361-1578: u_h += P(u_H - R u_h) with the
// 27/9/3/1 (/64) weights; guard: keep the old value unless the new one is > 0.
// A thread owns coarse column (I, J); the sons in fine plane 2K take the 3x3
// coarse block of planes K-1 (hz, rz) and K (hk, rk), those in plane 2K+1
// planes K+1 and K.  The sum keeps the reference's order P = P + w u_H - w R u
// over the 8 contributors of a son.
//
// ONLY1: correct the colour-1 sons only.  Inside the f(R) cycles a
// correction is always followed by a colour-0 relaxation pass, whose new
// values depend on the colour-1 neighbours alone (the local root does not
// read the cell's old value), so corrected colour-0 values would be
// overwritten unread: skipping them halves the fine-level traffic and leaves
// every result bit-identical.
template <int S3, bool ONLY1 = false>
__device__ __forceinline__ bool son_corrected(int t)
{
    return !ONLY1 || (((t & 1) + (t >> 1) + S3) & 1) == 1;
}

// the old values of the sons of fine plane 2K+S3 that are corrected (loaded
// before any son of the coarse plane is stored: each son is read and written
// by its own thread only)
template <int S3, bool ONLY1, bool NT = false>
__device__ __forceinline__ void correct64_sons_load(double* __restrict__ uf, const Geo& gf, int I, int J, int K,
                                                    double* old)
{
#pragma unroll
    for (int t = 0; t < 4; t++)
        if (son_corrected<S3, ONLY1>(t)) old[t] = ld_once<NT>(son64(uf, gf, I, J, K, t & 1, t >> 1, S3));
}

template <int S3, bool ONLY1, bool GUARD, bool NT = false>
__device__ __forceinline__ void correct64_sons_apply(const double* hz, const double* rz, const double* hk,
                                                     const double* rk, double* __restrict__ uf, const Geo& gf, int I,
                                                     int J, int K, const double* old)
{
    const double a = 1.0 / 64.0, b = 3.0 * a, c9 = 9.0 * a, d = 27.0 * a;
    const double wt[8] = {a, b, b, c9, b, c9, c9, d};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (!son_corrected<S3, ONLY1>(t)) continue;
        const int s1 = t & 1, s2 = t >> 1;
        double P = 0.0;
#pragma unroll
        for (int ind = 0; ind < 8; ind++) {
            const int x = (ind & 1) ? 1 : (s1 ? 2 : 0);
            const int y = (ind & 2) ? 1 : (s2 ? 2 : 0);
            const int q = x + 3 * y;
            // contributor in plane K (ind & 4) or in the plane beside it
            const double vh = (ind & 4) ? hk[q] : hz[q];
            const double vr = (ind & 4) ? rk[q] : rz[q];
            P = P + wt[ind] * vh;
            P = P - wt[ind] * vr;
        }
        const double tmpt = old[t] + P;
        if (!GUARD || tmpt > 0.0) st_once<NT>(son64(uf, gf, I, J, K, s1, s2, S3), tmpt);
    }
}

// ===========================================================================
// Mixed precision: FP32 delta-form FAS
// ===========================================================================

// x from w = x^{n+1}
template <int FRN, typename T>
__device__ __forceinline__ T xroot(T w)
{
    if constexpr (FRN == 0)
        return w;
    else if constexpr (FRN == 1)
        return sqrt(w);
    else
        return cbrt(w);
}

// (x^{n+1} - xb^{n+1}) / (x - xb)
template <int FRN, typename T>
__device__ __forceinline__ T dden(T x, T xb)
{
    if constexpr (FRN == 0)
        return (T)1;
    else if constexpr (FRN == 1)
        return x + xb;
    else
        return x * x + x * xb + xb * xb;
}

// local term C (1/xb - 1/x) of D for increment d (x consistent with d)
template <int FRN, typename T>
__device__ __forceinline__ T dlocal(T d, T x, T xb, T C)
{
    return C * (d / dden<FRN>(x, xb)) / (x * xb);
}

// solve A (S - 6 d) + C (1/xb - 1/x(d)) = r for d
template <int FRN>
__device__ __forceinline__ float df_solve(float S, float xb, float r, const DfCoef& c)
{
    const float wb = wpow<FRN>(xb);
    const float rxb = __builtin_amdgcn_rcpf(xb);
    const float K = c.A * S + 6.0f * c.A * wb + c.C * rxb - r;
    float x = seed_root<FRN>(-K * c.inv6A, c.C * c.inv6A);
    if (!(x > 0.0f) || !isfinite(x)) x = xb;
    float d = wpow<FRN>(x) - wb;
    // One Newton step in d from the cancellation-free seed leaves an error
    // ~ (F''/F') (seed error)^2 ~ 1e-13 wb; the n=2 seed for P > 0 is only a
    // bound within a factor 2 of the root and takes four.
    constexpr int NIT = FRN == 2 ? 4 : 1;
#pragma unroll
    for (int it = 0; it < NIT; it++) {
        if (it > 0) x = xroot<FRN>(fmaxf(wb + d, 1.0e-30f));
        // one reciprocal q = 1 / (x dden(x, xb)) gives 1/x = dden q and
        // C (1/xb - 1/x) = C d / (dden(x, xb) x xb) = C d q / xb
        const float dd = dden<FRN>(x, xb);
        const float q = __builtin_amdgcn_rcpf(x * dd);
        const float rx = dd * q;
        const float loc = c.C * d * q * rxb;
        const float F = c.A * (S - 6.0f * d) + loc - r;
        // dF/dd = -6A + C / ((n+1) x^{n+2})
        const float Fp = -6.0f * c.A + c.C * (1.0f / (float)(FRN + 1)) * rx * wpow<FRN>(rx);
        d = d - F * __builtin_amdgcn_rcpf(Fp);
    }
    // stay inside the domain x > 0
    if (!(wb + d > 0.0f)) d = -0.5f * wb;
    return d;
}

template <int FRN, bool NT = false>
struct DfSmoothOp {
    using T = float;
    struct Own {
        float xb, rhs;
    };
    float* d;
    const float* xb;
    const float* rhs;
    DfCoef c;
    __device__ float* field() const { return d; }
    __device__ __forceinline__ Own load(unsigned o, unsigned) const
    {
        return Own{ld_once<NT>(elem(xb, o)), ld_once<NT>(elem(rhs, o))};
    }
    __device__ __forceinline__ void update(unsigned o, unsigned, const Own& w, float, float xp, float xm, float yp,
                                           float ym, float zp, float zm) const
    {
        const float S = xp + xm + yp + ym + zp + zm;
        st_once<NT>(elem(d, o), df_solve<FRN>(S, w.xb, w.rhs, c));
    }
};

// u + u_bg (x(d) - xb) at one cell, evaluated in FP64 as u_bg d / dden(x, xb)
// (u unchanged where xb^{n+1} + d is not positive)
template <int FRN>
__device__ __forceinline__ double outer_updated(double u, float d, float xb, double ubg)
{
    const double dl = d, xbc = xb;
    const double w = wpow<FRN>(xbc) + dl;
    if (!(w > 0.0)) return u;
    const double x = xroot<FRN>(w);
    return u + ubg * (dl / dden<FRN>(x, xbc));
}

// the last colour-1 pass of the inner cycle fused with the outer update: the
// new colour-1 increment goes straight into u (d itself is not needed
// afterwards) and every colour-0 cell, visited through the window, is
// updated from its final increment
template <int FRN, bool NT = false>
struct DfSmoothUpdateOp {
    using T = float;
    struct Own {
        float xb, rhs, xbo;
        double u, uo;
    };
    const float* d;
    const float* xb;
    const float* rhs;
    double* u;
    DfCoef c;
    double ubg;
    __device__ const float* field() const { return d; }
    __device__ __forceinline__ Own load(unsigned o, unsigned oo) const
    {
        Own w;
        w.xb = ld_once<NT>(elem(xb, o));
        w.rhs = ld_once<NT>(elem(rhs, o));
        w.u = ld_once<NT>(u[o]);
        w.xbo = ld_once<NT>(elem(xb, oo));
        w.uo = ld_once<NT>(u[oo]);
        return w;
    }
    __device__ __forceinline__ void update(unsigned o, unsigned oo, const Own& w, float dother, float xp, float xm,
                                           float yp, float ym, float zp, float zm) const
    {
        const float S = xp + xm + yp + ym + zp + zm;
        st_once<NT>(u[o], outer_updated<FRN>(w.u, df_solve<FRN>(S, w.xb, w.rhs, c), w.xb, ubg));
        st_once<NT>(u[oo], outer_updated<FRN>(w.uo, dother, w.xbo, ubg));
    }
};

// FP64 outer residual r = N(u) of the mixed loop (the plain path's residual),
// emitting the inner problem's data: g = -r and xb = u / u_bg in FP32, and
// the inner increment after its first colour-0 pass from d = 0 (where every
// neighbour sum vanishes): d = solve(0) on colour 0.  Colour 1 of d is not
// written: the inner cycle's next pass is the colour-1 pass, which sets it
// from colour 0 without reading it.  NORM_ONLY: the same residual and
// partial sums, without the inner problem's data (a solve's last residual).
template <int FRN, bool NORM_ONLY = false, bool NT = false>
struct FrOuterOp {
    using T = double;
    const double* u;
    const double* delta;
    float* g;
    float* xb;
    float* d;
    FrCoef c;
    DfCoef dc;
    unsigned half;
    struct Own {
        double delta;
    };
    __device__ const double* field() const { return u; }
    __device__ __forceinline__ Own load(unsigned o) const { return Own{ld_once<NT>(delta[o])}; }
    __device__ __forceinline__ double cell(unsigned o, const Own& w, double uc, double xp, double xm, double yp,
                                           double ym, double zp, double zm) const
    {
        const double S =
            wpow<FRN>(xp) + wpow<FRN>(xm) + wpow<FRN>(yp) + wpow<FRN>(ym) + wpow<FRN>(zp) + wpow<FRN>(zm);
        // the plain residual's operator, with 1/u from the hardware
        // reciprocal refined once (a few ulp: the inner problem only needs
        // g and xb in FP32; the norm is unaffected at any tolerance used)
        const double ru = rcp64(uc);
        double OP = S - wpow<FRN>(uc) * 6.0;
        OP = OP * c.fct3r;
        OP = OP + c.fct1r * w.delta;
        OP = OP + c.fct2r * (c.a_fR_rt * ru - 1.0);
        if constexpr (!NORM_ONLY) {
            const float gv = (float)(-OP), xv = (float)(uc * c.inv_ubg);
            st_once<NT>(elem(g, o), gv);
            st_once<NT>(elem(xb, o), xv);
            if (o < half) st_once<NT>(elem(d, o), df_solve<FRN>(0.0f, xv, gv, dc));
        }
        return OP;
    }
};

// D(d) - rhs at a cell; with (d, rhs, out) = (R d, R r, f) the inner FAS
// right-hand side f = D(R d) - R r
template <int FRN, bool NT = false>
struct DfResOp {
    using T = float;
    const float* d;
    const float* xb;
    const float* rhs;
    float* out;
    DfCoef c;
    struct Own {
        float xb, rhs;
    };
    __device__ const float* field() const { return d; }
    __device__ __forceinline__ Own load(unsigned o) const { return Own{ld_once<NT>(elem(xb, o)), ld_once<NT>(elem(rhs, o))}; }
    __device__ __forceinline__ double cell(unsigned o, const Own& w, float dc, float xp, float xm, float yp, float ym,
                                           float zp, float zm) const
    {
        const float S = xp + xm + yp + ym + zp + zm;
        const float xbc = w.xb;
        const float x = xroot<FRN>(fmaxf(wpow<FRN>(xbc) + dc, 1.0e-30f));
        const float r = c.A * (S - 6.0f * dc) + dlocal<FRN>(dc, x, xbc, c.C) - w.rhs;
        st_once<NT>(elem(out, o), r);
        return 0.0;
    }
};

// inner residual D(d) - rhs of a level averaged over the 8 children of each
// coarse cell (rhs = g on level 0, the FAS right-hand side below); the
// children's Laplacians collapse to -3 sum(children) + sum(24 outer faces).
// One coarse column (I, J) = (cc.m, cc.j) marching over its run of K, like
// k_fr_restrict_residual1_64.  The children's increments it reads also give
// R d, stored as the frozen copy rd1 and the coarse start value d1.
template <int FRN, bool SHFL = false>
__device__ __forceinline__ void df_restrict_residual_col(const float* __restrict__ d0, const float* __restrict__ xb0,
                                                         const float* __restrict__ g0, float* __restrict__ rres1,
                                                         float* __restrict__ rd1, float* __restrict__ d1,
                                                         const Geo& gf, const Geo& gc, const DfCoef& c, const ZCol& cc,
                                                         bool lok = false, bool rok = false)
{
    const int I = cc.m, J = cc.j, N = gf.n;
    Pts12 F;
    F.init(gf, I, J);
    // per fine plane: local term minus rhs and increment of the 4 children,
    // increments of the 8 face neighbours
    struct Pl {
        float l[4], d[4], f[8];
    };
    auto load = [&](int z, Pl& p) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const unsigned o = F.at(q, z);
            const float dc = elem(d0, o), xbc = elem(xb0, o);
            const float x = xroot<FRN>(fmaxf(wpow<FRN>(xbc) + dc, 1.0e-30f));
            p.l[q] = dlocal<FRN>(dc, x, xbc, c.C) - elem(g0, o);
            p.d[q] = dc;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            // SHFL: the x-faces 2I-1 and 2I+2 are children of columns I-1
            // and I+1, which the adjacent lanes load where lok / rok say so
            if (SHFL && q < 2 && lok) continue;
            if (SHFL && q >= 2 && q < 4 && rok) continue;
            p.f[q] = elem(d0, F.at(4 + q, z));
        }
        if constexpr (SHFL) {
            const float l1 = __shfl_up(p.d[1], 1u, 64), l3 = __shfl_up(p.d[3], 1u, 64);
            const float r0 = __shfl_down(p.d[0], 1u, 64), r2 = __shfl_down(p.d[2], 1u, 64);
            if (lok) {
                p.f[0] = l1;
                p.f[1] = l3;
            }
            if (rok) {
                p.f[2] = r0;
                p.f[3] = r2;
            }
        }
    };
    Pl p0, p1, p2;
    float dl[4];
    const int zs = 2 * cc.k0;
    load(zs, p0);
    {
        const int zl = zs == 0 ? N - 1 : zs - 1;
#pragma unroll
        for (int q = 0; q < 4; q++) dl[q] = elem(d0, F.at(q, zl));
    }
    for (int K = cc.k0; K < cc.k1; K++) {
        load(2 * K + 1, p1);
        load(2 * K + 2 == N ? 0 : 2 * K + 2, p2);
        // s8 adds the children in restrict8_cell's order (x fastest, from
        // the first child), so R d below is bit-identical to it
        float loc = 0.0f, s8 = p0.d[0];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            loc += p0.l[q];
            if (q > 0) s8 += p0.d[q];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            loc += p1.l[q];
            s8 += p1.d[q];
        }
        float outer = p0.f[0] + p0.f[1] + p1.f[0] + p1.f[1];
        outer += p0.f[2] + p0.f[3] + p1.f[2] + p1.f[3];
        outer += p0.f[4] + p0.f[5] + p1.f[4] + p1.f[5];
        outer += p0.f[6] + p0.f[7] + p1.f[6] + p1.f[7];
        outer += dl[0] + dl[1] + dl[2] + dl[3];
        outer += p2.d[0] + p2.d[1] + p2.d[2] + p2.d[3];
        const unsigned oc = sidx(gc, I, J, K);
        elem(rres1, oc) = (c.A * (outer - 3.0f * s8) + loc) * 0.125f;
        const float rdv = s8 / 8.0f;
        elem(rd1, oc) = rdv;
        elem(d1, oc) = rdv;
#pragma unroll
        for (int q = 0; q < 4; q++) dl[q] = p1.d[q];
        p0 = p2;
    }
}

// restricted base of coarse cell (I, J, K): xb_c = (mean of the children's
// xb^{n+1})^{1/(n+1)}
template <int FRN>
__device__ __forceinline__ void restrict_base_cell(const float* __restrict__ xf, float* __restrict__ xc, const Geo& gf,
                                                   const Geo& gc, int I, int J, int K)
{
    float w = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; q++)
        w += wpow<FRN>(elem(xf, sidx(gf, 2 * I + (q & 1), 2 * J + ((q >> 1) & 1), 2 * K + ((q >> 2) & 1))));
    elem(xc, sidx(gc, I, J, K)) = xroot<FRN>(w * 0.125f);
}

// inner coarse-grid correction d_h += P(d_H - R d_h), kept only where the
// corrected x stays positive (the delta-form image of the reference's guard);
// z-marching like k_correct64, on e = d_H - R d_h.  Colour-1 sons only: every
// inner correction is followed by a colour-0 pass (see son_corrected).
// df_sons_load loads the sons' old d and raw base xb (the weight xb^{n+1} is
// formed at use, so no arithmetic waits on a load in flight).
template <int S3, bool NT = false>
__device__ __forceinline__ void df_sons_load(const float* __restrict__ df, const float* __restrict__ xbf, const Geo& gf,
                                             int I, int J, int K, float* old, float* xb)
{
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (!son_corrected<S3, true>(t)) continue;
        const unsigned fi = son_index(gf, I, J, K, t & 1, t >> 1, S3);
        old[t] = ld_once<NT>(elem(df, fi));
        xb[t] = ld_once<NT>(elem(xbf, fi));
    }
}

// The raw coarse values of plane K that e = dH - rd over the 3x3 block needs.
// SHFL: each row's centre is loaded and its x-neighbours come from the
// adjacent lanes of the wave, which hold columns I-1 and I+1 of the same row
// and plane run where lok / rok say so (else they are loaded too): a third of
// the loads, same values.  Loading and forming e are separate steps so that
// the loads of the next plane can be in flight while this one is stored.
struct ERaw {
    float h[9], r[9];
};

template <bool SHFL>
__device__ __forceinline__ void e_load(const float* __restrict__ dH, const float* __restrict__ rd, const Block9& B,
                                       int K, bool lok, bool rok, ERaw& w)
{
#pragma unroll
    for (int y = 0; y < 3; y++) {
#pragma unroll
        for (int x = 0; x < 3; x++) {
            if (SHFL && x == 0 && lok) continue;
            if (SHFL && x == 2 && rok) continue;
            const unsigned o = B.at(x + 3 * y, K);
            w.h[x + 3 * y] = elem(dH, o);
            w.r[x + 3 * y] = elem(rd, o);
        }
    }
}

template <bool SHFL>
__device__ __forceinline__ void e_form(const ERaw& w, bool lok, bool rok, float* e)
{
#pragma unroll
    for (int y = 0; y < 3; y++) {
        const float c = w.h[1 + 3 * y] - w.r[1 + 3 * y];
        e[3 * y + 1] = c;
        if constexpr (SHFL) {
            const float l = __shfl_up(c, 1u, 64);
            const float r = __shfl_down(c, 1u, 64);
            e[3 * y] = lok ? l : w.h[3 * y] - w.r[3 * y];
            e[3 * y + 2] = rok ? r : w.h[3 * y + 2] - w.r[3 * y + 2];
        } else {
            e[3 * y] = w.h[3 * y] - w.r[3 * y];
            e[3 * y + 2] = w.h[3 * y + 2] - w.r[3 * y + 2];
        }
    }
}

// new values of the corrected sons of fine plane 2K + S3 (d + P e, P the
// trilinear weights 1, 3, 9, 27 / 64 over the coarse planes ez and ek) and
// whether each keeps x > 0
template <int FRN, int S3>
__device__ __forceinline__ void df_sons_new(const float* ez, const float* ek, const float* old, const float* xb,
                                            float* dn, bool* keep)
{
    const float a = 1.0f / 64.0f, b = 3.0f * a, c9 = 9.0f * a, dd = 27.0f * a;
    const float wt[8] = {a, b, b, c9, b, c9, c9, dd};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (!son_corrected<S3, true>(t)) continue;
        const int s1 = t & 1, s2 = t >> 1;
        float P = 0.0f;
#pragma unroll
        for (int ind = 0; ind < 8; ind++) {
            const int x = (ind & 1) ? 1 : (s1 ? 2 : 0);
            const int y = (ind & 2) ? 1 : (s2 ? 2 : 0);
            const int q = x + 3 * y;
            P += wt[ind] * ((ind & 4) ? ek[q] : ez[q]);
        }
        dn[t] = old[t] + P;
        keep[t] = wpow<FRN>(xb[t]) + dn[t] > 0.0f;
    }
}

template <int S3, bool NT = false>
__device__ __forceinline__ void df_sons_store(float* __restrict__ df, const Geo& gf, int I, int J, int K,
                                              const float* dn, const bool* keep)
{
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (!son_corrected<S3, true>(t)) continue;
        if (keep[t]) st_once<NT>(elem(df, son_index(gf, I, J, K, t & 1, t >> 1, S3)), dn[t]);
    }
}

// Software-pipelined over the column's coarse planes: plane K's sons are
// computed, then every load of plane K+1 (its sons' old d and base weight,
// coarse plane K+2) is issued, then plane K's sons are stored.  A son is read
// and written by its own thread only and the planes' sons are distinct, so
// the values are those of the plane-by-plane order.
template <int FRN, bool SHFL = false, bool NT = false>
__device__ __forceinline__ void df_correct_col(const float* __restrict__ dH, const float* __restrict__ rd,
                                               float* __restrict__ df, const float* __restrict__ xbf, const Geo& gf,
                                               const Geo& gc, const ZCol& cc, bool lok = false, bool rok = false)
{
    const int I = cc.m, J = cc.j;
    Block9 B;
    B.init(gc, I, J);
    ERaw wm, wk, wp;  // coarse planes K-1, K, K+1
    float o0[4], w0[4], o1[4], w1[4];  // old d and raw xb of the sons of fine planes 2K, 2K+1
    e_load<SHFL>(dH, rd, B, wrap_dec(cc.k0, gc.n), lok, rok, wm);
    e_load<SHFL>(dH, rd, B, cc.k0, lok, rok, wk);
    e_load<SHFL>(dH, rd, B, wrap_inc(cc.k0, gc.n), lok, rok, wp);
    df_sons_load<0, NT>(df, xbf, gf, I, J, cc.k0, o0, w0);
    df_sons_load<1, NT>(df, xbf, gf, I, J, cc.k0, o1, w1);
    float em[9], ek[9], ep[9];
    e_form<SHFL>(wm, lok, rok, em);
    e_form<SHFL>(wk, lok, rok, ek);
    for (int K = cc.k0; K < cc.k1; K++) {
        e_form<SHFL>(wp, lok, rok, ep);
        float dn0[4], dn1[4];
        bool k0[4], k1[4];
        df_sons_new<FRN, 0>(em, ek, o0, w0, dn0, k0);
        df_sons_new<FRN, 1>(ep, ek, o1, w1, dn1, k1);
        if (K + 1 < cc.k1) {
            df_sons_load<0, NT>(df, xbf, gf, I, J, K + 1, o0, w0);
            df_sons_load<1, NT>(df, xbf, gf, I, J, K + 1, o1, w1);
            e_load<SHFL>(dH, rd, B, wrap_inc(K + 1, gc.n), lok, rok, wp);
        }
        df_sons_store<0, NT>(df, gf, I, J, K, dn0, k0);
        df_sons_store<1, NT>(df, gf, I, J, K, dn1, k1);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            em[q] = ek[q];
            ek[q] = ep[q];
        }
    }
}

// coarsest level: red-black sweeps of the delta-form smoother in FP64 by one
// workgroup (the level holds at most a few thousand cells)
template <int FRN>
__device__ __forceinline__ void df_coarsest_block(float* __restrict__ d, const float* __restrict__ xb,
                                                  const float* __restrict__ rhs, const Geo& g, int sweeps,
                                                  const DfCoef& c)
{
    for (int sw = 0; sw < sweeps; sw++) {
        for (int colour = 0; colour < 2; colour++) {
            for (int t = threadIdx.x; t < (int)g.half; t += blockDim.x) {
                const int m = t % g.h;
                const int row = t / g.h;
                const int j = row % g.n, k = row / g.n;
                const Nb q = neighbours(g, colour, m, j, k);
                const double S = (double)d[q.xp] + d[q.xm] + d[q.yp] + d[q.ym] + d[q.zp] + d[q.zm];
                const double xbc = xb[q.me];
                const double wb = wpow<FRN>(xbc);
                const double r = rhs[q.me];
                // Newton on the increment in FP64 from the current value
                double dl = d[q.me];
                for (int it = 0; it < 30; it++) {
                    double x = xroot<FRN>(fmax(wb + dl, 1.0e-300));
                    const double F = c.A64 * (S - 6.0 * dl) + dlocal<FRN>(dl, x, xbc, c.C64) - r;
                    const double Fp = -6.0 * c.A64 + c.C64 / ((double)(FRN + 1) * x * wpow<FRN>(x));
                    double dn = dl - F / Fp;
                    if (!(wb + dn > 0.0)) dn = 0.5 * (dl - wb);  // halve the distance to x = 0
                    const bool done = fabs(dn - dl) <= 1.0e-15 * (fabs(dn) + wb);
                    dl = dn;
                    if (done) break;
                }
                d[q.me] = (float)dl;
            }
            __syncthreads();
        }
    }
}

// 8-child average (restrict_density, restrict_scalar_field,
// restrict_residual_fR at levels > 1, PMP2extradof.f90:1198-1247,
// 1293-1345, PMP2MGsolver_fR.f90:750-779) of coarse cell (I, J, K); out2:
// optional second copy
template <typename T>
__device__ __forceinline__ void restrict8_cell(const T* __restrict__ fine, T* __restrict__ out1, T* __restrict__ out2,
                                               const Geo& gf, const Geo& gc, int I, int J, int K)
{
    // children in the reference's order (x fastest), accumulated in T
    T P = lvl(fine, sidx(gf, 2 * I, 2 * J, 2 * K));
    P = P + lvl(fine, sidx(gf, 2 * I + 1, 2 * J, 2 * K));
    P = P + lvl(fine, sidx(gf, 2 * I, 2 * J + 1, 2 * K));
    P = P + lvl(fine, sidx(gf, 2 * I + 1, 2 * J + 1, 2 * K));
    P = P + lvl(fine, sidx(gf, 2 * I, 2 * J, 2 * K + 1));
    P = P + lvl(fine, sidx(gf, 2 * I + 1, 2 * J, 2 * K + 1));
    P = P + lvl(fine, sidx(gf, 2 * I, 2 * J + 1, 2 * K + 1));
    P = P + lvl(fine, sidx(gf, 2 * I + 1, 2 * J + 1, 2 * K + 1));
    const unsigned o = sidx(gc, I, J, K);
    const T v = P / (T)8.0;
    lvl(out1, o) = v;
    if (out2) lvl(out2, o) = v;
}

// ---------------------------------------------------------------------------
// Tail of the inner V-cycle: every level from l0 (at most 16^3 cells) down to
// the coarsest, run by one workgroup in one launch.  The coarse levels are
// too small to fill the GPU and a launch per operation would cost more than
// the work; here each operation is one pass of the workgroup over the
// level's cells, separated by barriers.  Per cell the arithmetic is that of
// the level kernels (a red-black pass, the restrictions and the
// prolongation do not depend on the order cells are visited in), so the
// results are identical to the launch-per-operation cycle.
// ---------------------------------------------------------------------------
struct TailLevel {
    Geo g;
    float *d, *xb, *f, *rres, *rd;
    DfCoef dc;
};

constexpr int TAIL_MAXL = 8;
constexpr int TAIL_MAX_N = 16;  // levels with n <= 16 run in the tail
// LDS for the tail's arrays: d, xb, g of its first level and d, xb, g, rres,
// rd of each coarser one (15208 floats for 16^3 down to 2^3)
constexpr int TAIL_LDS_FLOATS = 15232;

struct TailDesc {
    int nl;  // levels below the first (L[nl] is the coarsest)
    TailLevel L[TAIL_MAXL];
};

// cell t of a colour half -> the six neighbours' offsets and its own
__device__ __forceinline__ Nb tail_cell(const Geo& g, int colour, int t)
{
    const int m = t % g.h, row = t / g.h;
    return neighbours(g, colour, m, row % g.n, row / g.n);
}

template <class OP>
__device__ __forceinline__ void tail_update(const OP& op, const Geo& g, int colour, int t)
{
    const Nb q = tail_cell(g, colour, t);
    const auto* a = op.field();
    const unsigned oo = q.me + (colour ? 0u - (unsigned)g.half : (unsigned)g.half);
    op.update(q.me, oo, op.load(q.me, oo), a[oo], a[q.xp], a[q.xm], a[q.yp], a[q.ym], a[q.zp], a[q.zm]);
}

template <class OP>
__device__ __forceinline__ void tail_cell_op(const OP& op, const Geo& g, int t)
{
    const int colour = t >= (int)g.half ? 1 : 0;
    const Nb q = tail_cell(g, colour, t - colour * (int)g.half);
    const auto* a = op.field();
    (void)op.cell(q.me, op.load(q.me), a[q.me], a[q.xp], a[q.xm], a[q.yp], a[q.ym], a[q.zp], a[q.zm]);
}

__device__ __forceinline__ void coarse_ijk(const Geo& gc, int t, int& I, int& J, int& K)
{
    const unsigned n = (unsigned)gc.n;
    I = (int)((unsigned)t % n);
    J = (int)(((unsigned)t / n) % n);
    K = (int)((unsigned)t / (n * n));
}

// The tail runs out of LDS: the first level's d, xb and g and every coarser
// level's xb are copied in, the cycle works on the LDS copies (the arithmetic
// and its order are those of the level kernels), and the first level's d is
// copied back; coarser levels are internal to the cycle.  The launcher checks
// that the levels fit in TAIL_LDS_FLOATS.
struct TailArrays {
    float *d, *xb, *f, *rres, *rd;
};

// level i's arrays in the LDS block: d, xb, g (and rres, rd below the first
// level), levels one after the other
__device__ __forceinline__ TailArrays tail_lds_level(const TailDesc& G, int i, float* lds)
{
    int off = 0;
    for (int q = 0; q < i; q++) off += (q == 0 ? 3 : 5) * G.L[q].g.n * G.L[q].g.n * G.L[q].g.n;
    const int cells = G.L[i].g.n * G.L[i].g.n * G.L[i].g.n;
    float* p = lds + off;
    return TailArrays{p, p + cells, p + 2 * cells, i > 0 ? p + 3 * cells : nullptr, i > 0 ? p + 4 * cells : nullptr};
}

template <int FRN>
__device__ void df_tail_vcycle(const TailDesc& G)
{
    __shared__ float lds[TAIL_LDS_FLOATS];
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int i = 0; i <= G.nl; i++) {
        const TailArrays A = tail_lds_level(G, i, lds);
        const int cells = G.L[i].g.n * G.L[i].g.n * G.L[i].g.n;
        for (int t = tid; t < cells; t += nthr) A.xb[t] = G.L[i].xb[t];
        if (i == 0)
            for (int t = tid; t < cells; t += nthr) {
                A.d[t] = G.L[0].d[t];
                A.f[t] = G.L[0].f[t];
            }
    }
    __syncthreads();
    for (int i = 0; i < G.nl; i++) {
        const TailLevel &L = G.L[i], &C = G.L[i + 1];
        const TailArrays l = tail_lds_level(G, i, lds), c = tail_lds_level(G, i + 1, lds);
        for (int it = 0; it < 4; it++) {
            for (int t = tid; t < (int)L.g.half; t += nthr)
                tail_update(DfSmoothOp<FRN>{l.d, l.xb, l.f, L.dc}, L.g, it & 1, t);
            __syncthreads();
        }
        const int ncc = C.g.n * C.g.n * C.g.n;
        for (int t = tid; t < ncc; t += nthr) {
            int I, J, K;
            coarse_ijk(C.g, t, I, J, K);
            df_restrict_residual_col<FRN>(l.d, l.xb, l.f, c.rres, c.rd, c.d, L.g, C.g, L.dc, ZCol{I, J, K, K + 1});
        }
        __syncthreads();
        for (int t = tid; t < ncc; t += nthr) tail_cell_op(DfResOp<FRN>{c.rd, c.xb, c.rres, c.f, C.dc}, C.g, t);
        __syncthreads();
    }
    {
        const TailArrays b = tail_lds_level(G, G.nl, lds);
        df_coarsest_block<FRN>(b.d, b.xb, b.f, G.L[G.nl].g, 8, G.L[G.nl].dc);
    }
    for (int i = G.nl - 1; i >= 0; i--) {
        const TailLevel &L = G.L[i], &C = G.L[i + 1];
        const TailArrays l = tail_lds_level(G, i, lds), c = tail_lds_level(G, i + 1, lds);
        const int ncc = C.g.n * C.g.n * C.g.n;
        for (int t = tid; t < ncc; t += nthr) {
            int I, J, K;
            coarse_ijk(C.g, t, I, J, K);
            df_correct_col<FRN>(c.d, c.rd, l.d, l.xb, L.g, C.g, ZCol{I, J, K, K + 1});
        }
        __syncthreads();
        for (int it = 0; it < 4; it++) {
            for (int t = tid; t < (int)L.g.half; t += nthr)
                tail_update(DfSmoothOp<FRN>{l.d, l.xb, l.f, L.dc}, L.g, it & 1, t);
            __syncthreads();
        }
    }
    const TailArrays a0 = tail_lds_level(G, 0, lds);
    const int cells0 = G.L[0].g.n * G.L[0].g.n * G.L[0].g.n;
    for (int t = tid; t < cells0; t += nthr) G.L[0].d[t] = a0.d[t];
}

}  // namespace mgs
