// mgs_fr.hip -- f(R) multigrid kernels for gfx950 (CDNA4).
//
// Plain FP64 path: the reference's operators (PMP2MGsolver_fR.f90,
// PMP2extradof.f90:1168-1578) evaluated in FP64 with the reference's operand
// order, on the colour-split layout of mgs_common.h.  Built with
// -ffp-contract=off, so the residual, the restrictions, the right-hand side
// and the prolongation round every operation as the CPU oracle does and are
// bit-identical to it.  The smoother solves the same local polynomial
// u^{n+2} + P u + Q = 0; by default its root comes from a cancellation-free
// FP32 seed refined by FP64 Newton steps (MGS_FLAG_REFERENCE_ROOTS selects
// the reference's closed forms evaluated literally with the device libm).
//
// Mixed path: the FP32 delta-form FAS used inside the FP64 defect-correction
// loop (mgs_api.cpp, solve_mixed).  Unknown d = x^{n+1} - xb^{n+1} with
// x = u / u_bg and xb an FP32 copy of the current FP64 iterate.  The inner
// operator
//     D(d) = A (sum_nb d_nb - 6 d) + C (1/xb - 1/x)
// is N(xb + .) - N(xb) written so that every term is proportional to d: FP32
// rounding then scales with the correction and the outer FP64 residual
// converges to FP64 accuracy.  Each cell is solved from the FP32 closed-form
// seed followed by a Newton step on D(d) = rhs in d.
//
// Stencil kernels march over z (mgs_stencil.h); transfer kernels map one
// thread to one coarse cell; all are memory-bound, with 32-bit index math.

#include <math.h>

#include "mgs_device.h"
#include "mgs_kernels.h"
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
template <int FRN, bool COARSE, bool REFROOT>
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
        w.rho = rho[o];
        w.f = COARSE ? f[o] : 0.0;
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
            u[o] = fr_root64<FRN>(P, c);
        } else {
            float x = seed_root<FRN>((float)(P * c.inv_w), (float)(c.Q2 * c.inv_w));
            if (!(x > 0.0f) || !(x < 3.0e38f)) x = 1.0f;  // Newton from u_bg
            u[o] = newton_root64<FRN>((double)x * c.ubg, P, c.Q);
        }
    }
};

// residual of one cell, calculate_residual_fR, PMP2MGsolver_fR.f90:434-458;
// COARSE subtracts f (:482-488).  Also the FAS right-hand side (:836-923)
// with (u, f, out) = (R u, R r, f_l).
template <int FRN, bool COARSE>
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
        w.rho = rho[o];
        w.f = COARSE ? f[o] : 0.0;
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
        if (out) out[o] = OP;
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

// restrict_residual_fR(1): the 8 children's residuals averaged, recomputed
// from u and delta, PMP2MGsolver_fR.f90:607-749.  A thread owns coarse column
// (I, J) and marches over K; fine plane 2K+2 (the upper z faces of K) is the
// lower child plane of K+1 and plane 2K-1 is carried from K-1, so u and delta
// are read once.  The operation order is the reference's.
template <int FRN>
__global__ __launch_bounds__(256) void k_fr_restrict_residual1_64(const double* __restrict__ u0,
                                                                  const double* __restrict__ delta,
                                                                  double* __restrict__ rres1, Geo gf, Geo gc,
                                                                  FrCoef c, int TX, int TY, int XB, int JB, int KC,
                                                                  unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    const int I = cc.m, J = cc.j, N = gf.n;
    Pts12 F;
    F.init(gf, I, J);
    // per fine plane: u^{n+1}, fct1 delta and fct4/u of the 4 children, and
    // u^{n+1} of the 8 face neighbours
    struct Pl {
        double w[4], t1[4], t2[4], f[8];
    };
    auto load = [&](int z, Pl& p) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const unsigned o = F.at(q, z);
            const double u = u0[o];
            p.w[q] = wpow<FRN>(u);
            p.t1[q] = c.fct1r * delta[o];
            p.t2[q] = c.fct4_1 / u;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) p.f[q] = wpow<FRN>(u0[F.at(4 + q, z)]);
    };
    Pl p0, p1, p2;
    double wl[4];
    const int zs = 2 * cc.k0;
    load(zs, p0);
    {
        const int zl = zs == 0 ? N - 1 : zs - 1;
#pragma unroll
        for (int q = 0; q < 4; q++) wl[q] = wpow<FRN>(u0[F.at(q, zl)]);
    }
    for (int K = cc.k0; K < cc.k1; K++) {
        load(2 * K + 1, p1);
        load(2 * K + 2 == N ? 0 : 2 * K + 2, p2);
        double Q = -8.0 * c.fct2r;
#pragma unroll
        for (int q = 0; q < 4; q++) Q = Q + p0.t1[q] + p0.t2[q];
#pragma unroll
        for (int q = 0; q < 4; q++) Q = Q + p1.t1[q] + p1.t2[q];
        double s8 = p0.w[0];
        s8 = s8 + p0.w[1];
        s8 = s8 + p0.w[2];
        s8 = s8 + p0.w[3];
#pragma unroll
        for (int q = 0; q < 4; q++) s8 = s8 + p1.w[q];
        double P = -3.0 * s8;
        P = P + p0.f[0] + p0.f[1] + p1.f[0] + p1.f[1];  // x-1 faces
        P = P + p0.f[2] + p0.f[3] + p1.f[2] + p1.f[3];  // x+2 faces
        P = P + p0.f[4] + p0.f[5] + p1.f[4] + p1.f[5];  // y-1 faces
        P = P + p0.f[6] + p0.f[7] + p1.f[6] + p1.f[7];  // y+2 faces
        P = P + wl[0] + wl[1] + wl[2] + wl[3];          // z-1 faces
        P = P + p2.w[0] + p2.w[1] + p2.w[2] + p2.w[3];  // z+2 faces
        rres1[sidx(gc, I, J, K)] = (P * c.fct3_1 + Q) / 8.0;
#pragma unroll
        for (int q = 0; q < 4; q++) wl[q] = p1.w[q];
        p0 = p2;
    }
}

// 8-child average (restrict_density, restrict_scalar_field,
// restrict_residual_fR at levels > 1, PMP2extradof.f90:1198-1247,
// 1293-1345, PMP2MGsolver_fR.f90:750-779); out2: optional second copy
template <typename T>
__global__ __launch_bounds__(256) void k_restrict8(const T* __restrict__ fine, T* __restrict__ out1,
                                                   T* __restrict__ out2, Geo gf, Geo gc, int TX, int TY, int XB,
                                                   unsigned NB)
{
    int I;
    unsigned row;
    if (!map_xy(gc.n, (long)gc.n * gc.n, TX, TY, XB, NB, I, row)) return;
    const int J = (int)(row % (unsigned)gc.n), K = (int)(row / (unsigned)gc.n);
    // children in the reference's order (x fastest), accumulated in T
    T P = fine[sidx(gf, 2 * I, 2 * J, 2 * K)];
    P = P + fine[sidx(gf, 2 * I + 1, 2 * J, 2 * K)];
    P = P + fine[sidx(gf, 2 * I, 2 * J + 1, 2 * K)];
    P = P + fine[sidx(gf, 2 * I + 1, 2 * J + 1, 2 * K)];
    P = P + fine[sidx(gf, 2 * I, 2 * J, 2 * K + 1)];
    P = P + fine[sidx(gf, 2 * I + 1, 2 * J, 2 * K + 1)];
    P = P + fine[sidx(gf, 2 * I, 2 * J + 1, 2 * K + 1)];
    P = P + fine[sidx(gf, 2 * I + 1, 2 * J + 1, 2 * K + 1)];
    const unsigned o = sidx(gc, I, J, K);
    const T v = P / (T)8.0;
    out1[o] = v;
    if (out2) out2[o] = v;
}

// Coarse 3x3 block around column (I, J) of a coarse level: the colour
// parity and in-plane offset of each (x, y) of the block, so that a plane K
// of the block loads with one add per value
struct Block9 {
    unsigned par[9], row[9], plane, half;
    __device__ __forceinline__ void init(const Geo& gc, int I, int J)
    {
        const int nc = gc.n;
        const int Xs[3] = {wrap_dec(I, nc), I, wrap_inc(I, nc)};
        const int Ys[3] = {wrap_dec(J, nc), J, wrap_inc(J, nc)};
#pragma unroll
        for (int y = 0; y < 3; y++)
#pragma unroll
            for (int x = 0; x < 3; x++) {
                par[x + 3 * y] = (unsigned)((Xs[x] + Ys[y]) & 1);
                row[x + 3 * y] = (unsigned)Ys[y] * (unsigned)gc.h + (unsigned)(Xs[x] >> 1);
            }
        plane = (unsigned)gc.n * (unsigned)gc.h;
        half = (unsigned)gc.half;
    }
    __device__ __forceinline__ unsigned at(int q, int K) const
    {
        return ((par[q] + (unsigned)K) & 1u) * half + (unsigned)K * plane + row[q];
    }
};

// offset of son (2I+s1, 2J+s2, 2K+s3) on the fine level: colour (s1+s2+s3)&1,
// m = I
__device__ __forceinline__ unsigned son_index(const Geo& gf, int I, int J, int K, int s1, int s2, int s3)
{
    return (unsigned)((s1 + s2 + s3) & 1) * (unsigned)gf.half +
           ((unsigned)(2 * K + s3) * (unsigned)gf.n + (unsigned)(2 * J + s2)) * (unsigned)gf.h + (unsigned)I;
}

// correct_solution, PMP2extradof.f90:1361-1578: u_h += P(u_H - R u_h) with the
// 27/9/3/1 (/64) weights; guard: keep the old value unless the new one is > 0.
// A thread owns coarse column (I, J) and marches over K with two planes of
// the 3x3 coarse block in registers: the sons in fine plane 2K use planes
// K-1 and K, then plane K+1 replaces K-1 for the sons in plane 2K+1.  The sum
// keeps the reference's order P = P + w u_H - w R u over the 8 contributors
// of a son.
template <int S3>
__device__ __forceinline__ void correct64_sons(const double* hz, const double* rz, const double* hk,
                                               const double* rk, double* __restrict__ uf, const Geo& gf, int I, int J,
                                               int K, int guard)
{
    const double a = 1.0 / 64.0, b = 3.0 * a, c9 = 9.0 * a, d = 27.0 * a;
    const double wt[8] = {a, b, b, c9, b, c9, c9, d};
    double old[4];
#pragma unroll
    for (int t = 0; t < 4; t++) old[t] = uf[son_index(gf, I, J, K, t & 1, t >> 1, S3)];
#pragma unroll
    for (int t = 0; t < 4; t++) {
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
        if (!guard || tmpt > 0.0) uf[son_index(gf, I, J, K, s1, s2, S3)] = tmpt;
    }
}

__global__ __launch_bounds__(256) void k_correct64(const double* __restrict__ uH, const double* __restrict__ ru,
                                                   double* __restrict__ uf, Geo gf, Geo gc, int guard, int TX, int TY,
                                                   int XB, int JB, int KC, unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    const int I = cc.m, J = cc.j;
    Block9 B;
    B.init(gc, I, J);
    double hs[9], rs[9], hk[9], rk[9];  // plane beside K (K-1, then K+1) and plane K
    const int Km = wrap_dec(cc.k0, gc.n);
#pragma unroll
    for (int q = 0; q < 9; q++) {
        hs[q] = uH[B.at(q, Km)];
        rs[q] = ru[B.at(q, Km)];
        hk[q] = uH[B.at(q, cc.k0)];
        rk[q] = ru[B.at(q, cc.k0)];
    }
    for (int K = cc.k0; K < cc.k1; K++) {
        correct64_sons<0>(hs, rs, hk, rk, uf, gf, I, J, K, guard);
        const int Kp = wrap_inc(K, gc.n);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            hs[q] = uH[B.at(q, Kp)];
            rs[q] = ru[B.at(q, Kp)];
        }
        correct64_sons<1>(hs, rs, hk, rk, uf, gf, I, J, K, guard);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const double th = hk[q], tr = rk[q];
            hk[q] = hs[q];
            rk[q] = rs[q];
            hs[q] = th;
            rs[q] = tr;
        }
    }
}

// deterministic sum of block partials: stage 1 folds contiguous chunks, stage
// 2 (one block) the chunk sums; the order depends only on the count
__global__ __launch_bounds__(1024) void k_sum_chunks(const double* __restrict__ p, int count, int chunk,
                                                     double* __restrict__ out)
{
    __shared__ double s[1024];
    const int lo = blockIdx.x * chunk;
    const int hi = lo + chunk < count ? lo + chunk : count;
    double v = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += blockDim.x) v += p[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_pack(const TI* __restrict__ lex, TO* __restrict__ split, Geo g, int TX,
                                              int TY, int XB, unsigned NB)
{
    int i;
    unsigned row;
    if (!map_xy(g.n, (long)g.n * g.n, TX, TY, XB, NB, i, row)) return;
    const int j = (int)(row % (unsigned)g.n), k = (int)(row / (unsigned)g.n);
    split[sidx(g, i, j, k)] = (TO)lex[(size_t)row * g.n + i];
}

template <typename TO>
__global__ __launch_bounds__(256) void k_unpack(const double* __restrict__ split, TO* __restrict__ lex, Geo g,
                                                double scale, int TX, int TY, int XB, unsigned NB)
{
    int i;
    unsigned row;
    if (!map_xy(g.n, (long)g.n * g.n, TX, TY, XB, NB, i, row)) return;
    const int j = (int)(row % (unsigned)g.n), k = (int)(row / (unsigned)g.n);
    const double v = split[sidx(g, i, j, k)];
    lex[(size_t)row * g.n + i] = (TO)(scale == 1.0 ? v : v * scale);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, size_t count, T v)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy(const T* __restrict__ s, T* __restrict__ d, size_t count)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) d[i] = s[i];
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
        const float rx = __builtin_amdgcn_rcpf(x);
        // C (1/xb - 1/x) = C d / (dden(x, xb) x xb)
        const float loc = c.C * d * __builtin_amdgcn_rcpf(dden<FRN>(x, xb)) * rx * rxb;
        const float F = c.A * (S - 6.0f * d) + loc - r;
        const float Fp = -6.0f * c.A + c.C * (1.0f / (float)(FRN + 1)) * rx * __builtin_amdgcn_rcpf(wpow<FRN>(x));
        d = d - F * __builtin_amdgcn_rcpf(Fp);
    }
    // stay inside the domain x > 0
    if (!(wb + d > 0.0f)) d = -0.5f * wb;
    return d;
}

template <int FRN>
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
    __device__ __forceinline__ Own load(unsigned o, unsigned) const { return Own{xb[o], rhs[o]}; }
    __device__ __forceinline__ void update(unsigned o, unsigned, const Own& w, float, float xp, float xm, float yp,
                                           float ym, float zp, float zm) const
    {
        const float S = xp + xm + yp + ym + zp + zm;
        d[o] = df_solve<FRN>(S, w.xb, w.rhs, c);
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
template <int FRN>
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
        w.xb = xb[o];
        w.rhs = rhs[o];
        w.u = u[o];
        w.xbo = xb[oo];
        w.uo = u[oo];
        return w;
    }
    __device__ __forceinline__ void update(unsigned o, unsigned oo, const Own& w, float dother, float xp, float xm,
                                           float yp, float ym, float zp, float zm) const
    {
        const float S = xp + xm + yp + ym + zp + zm;
        u[o] = outer_updated<FRN>(w.u, df_solve<FRN>(S, w.xb, w.rhs, c), w.xb, ubg);
        u[oo] = outer_updated<FRN>(w.uo, dother, w.xbo, ubg);
    }
};

// FP64 outer residual r = N(u) of the mixed loop (the plain path's residual),
// emitting the inner problem's data: g = -r and xb = u / u_bg in FP32, and
// the inner increment after its first colour-0 pass from d = 0 (where every
// neighbour sum vanishes): d = solve(0) on colour 0, d = 0 on colour 1
template <int FRN>
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
    __device__ __forceinline__ Own load(unsigned o) const { return Own{delta[o]}; }
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
        const float gv = (float)(-OP), xv = (float)(uc * c.inv_ubg);
        g[o] = gv;
        xb[o] = xv;
        d[o] = o < half ? df_solve<FRN>(0.0f, xv, gv, dc) : 0.0f;
        return OP;
    }
};

// D(d) - rhs at a cell; with (d, rhs, out) = (R d, R r, f) the inner FAS
// right-hand side f = D(R d) - R r
template <int FRN>
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
    __device__ __forceinline__ Own load(unsigned o) const { return Own{xb[o], rhs[o]}; }
    __device__ __forceinline__ double cell(unsigned o, const Own& w, float dc, float xp, float xm, float yp, float ym,
                                           float zp, float zm) const
    {
        const float S = xp + xm + yp + ym + zp + zm;
        const float xbc = w.xb;
        const float x = xroot<FRN>(fmaxf(wpow<FRN>(xbc) + dc, 1.0e-30f));
        const float r = c.A * (S - 6.0f * dc) + dlocal<FRN>(dc, x, xbc, c.C) - w.rhs;
        out[o] = r;
        return 0.0;
    }
};

// inner residual D(d) - rhs of a level averaged over the 8 children of each
// coarse cell (rhs = g on level 0, the FAS right-hand side below); the
// children's Laplacians collapse to -3 sum(children) + sum(24 outer faces).
// z-marching like k_fr_restrict_residual1_64.
template <int FRN>
__global__ __launch_bounds__(256) void k_df_restrict_residual(const float* __restrict__ d0,
                                                              const float* __restrict__ xb0,
                                                              const float* __restrict__ g0, float* __restrict__ rres1,
                                                              Geo gf, Geo gc, DfCoef c, int TX, int TY, int XB, int JB,
                                                              int KC, unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
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
            const float dc = d0[o], xbc = xb0[o];
            const float x = xroot<FRN>(fmaxf(wpow<FRN>(xbc) + dc, 1.0e-30f));
            p.l[q] = dlocal<FRN>(dc, x, xbc, c.C) - g0[o];
            p.d[q] = dc;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) p.f[q] = d0[F.at(4 + q, z)];
    };
    Pl p0, p1, p2;
    float dl[4];
    const int zs = 2 * cc.k0;
    load(zs, p0);
    {
        const int zl = zs == 0 ? N - 1 : zs - 1;
#pragma unroll
        for (int q = 0; q < 4; q++) dl[q] = d0[F.at(q, zl)];
    }
    for (int K = cc.k0; K < cc.k1; K++) {
        load(2 * K + 1, p1);
        load(2 * K + 2 == N ? 0 : 2 * K + 2, p2);
        float loc = 0.0f, s8 = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            loc += p0.l[q];
            s8 += p0.d[q];
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
        rres1[sidx(gc, I, J, K)] = (c.A * (outer - 3.0f * s8) + loc) * 0.125f;
#pragma unroll
        for (int q = 0; q < 4; q++) dl[q] = p1.d[q];
        p0 = p2;
    }
}

// restricted base: xb_c = (mean of the children's xb^{n+1})^{1/(n+1)}
template <int FRN>
__global__ __launch_bounds__(256) void k_restrict_base(const float* __restrict__ xf, float* __restrict__ xc, Geo gf,
                                                       Geo gc, int TX, int TY, int XB, unsigned NB)
{
    int I;
    unsigned row;
    if (!map_xy(gc.n, (long)gc.n * gc.n, TX, TY, XB, NB, I, row)) return;
    const int J = (int)(row % (unsigned)gc.n), K = (int)(row / (unsigned)gc.n);
    float w = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; q++)
        w += wpow<FRN>(xf[sidx(gf, 2 * I + (q & 1), 2 * J + ((q >> 1) & 1), 2 * K + ((q >> 2) & 1))]);
    xc[sidx(gc, I, J, K)] = xroot<FRN>(w * 0.125f);
}

// inner coarse-grid correction d_h += P(d_H - R d_h), kept only where the
// corrected x stays positive (the delta-form image of the reference's guard);
// z-marching like k_correct64, on e = d_H - R d_h
template <int FRN, int S3>
__device__ __forceinline__ void df_correct_sons(const float* ez, const float* ek, float* __restrict__ df,
                                                const float* __restrict__ xbf, const Geo& gf, int I, int J, int K)
{
    const float a = 1.0f / 64.0f, b = 3.0f * a, c9 = 9.0f * a, dd = 27.0f * a;
    const float wt[8] = {a, b, b, c9, b, c9, c9, dd};
    float old[4], wb[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const unsigned fi = son_index(gf, I, J, K, t & 1, t >> 1, S3);
        old[t] = df[fi];
        wb[t] = wpow<FRN>(xbf[fi]);
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int s1 = t & 1, s2 = t >> 1;
        float P = 0.0f;
#pragma unroll
        for (int ind = 0; ind < 8; ind++) {
            const int x = (ind & 1) ? 1 : (s1 ? 2 : 0);
            const int y = (ind & 2) ? 1 : (s2 ? 2 : 0);
            const int q = x + 3 * y;
            P += wt[ind] * ((ind & 4) ? ek[q] : ez[q]);
        }
        const float dn = old[t] + P;
        if (wb[t] + dn > 0.0f) df[son_index(gf, I, J, K, s1, s2, S3)] = dn;
    }
}

template <int FRN>
__global__ __launch_bounds__(256) void k_df_correct(const float* __restrict__ dH, const float* __restrict__ rd,
                                                    float* __restrict__ df, const float* __restrict__ xbf, Geo gf,
                                                    Geo gc, int TX, int TY, int XB, int JB, int KC, unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    const int I = cc.m, J = cc.j;
    Block9 B;
    B.init(gc, I, J);
    float es[9], ek[9];  // plane beside K (K-1, then K+1) and plane K
    const int Km = wrap_dec(cc.k0, gc.n);
#pragma unroll
    for (int q = 0; q < 9; q++) {
        const unsigned om = B.at(q, Km), o0 = B.at(q, cc.k0);
        es[q] = dH[om] - rd[om];
        ek[q] = dH[o0] - rd[o0];
    }
    for (int K = cc.k0; K < cc.k1; K++) {
        df_correct_sons<FRN, 0>(es, ek, df, xbf, gf, I, J, K);
        const int Kp = wrap_inc(K, gc.n);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const unsigned op = B.at(q, Kp);
            es[q] = dH[op] - rd[op];
        }
        df_correct_sons<FRN, 1>(es, ek, df, xbf, gf, I, J, K);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const float t = ek[q];
            ek[q] = es[q];
            es[q] = t;
        }
    }
}

// coarsest level: red-black sweeps of the delta-form smoother in FP64 in one
// workgroup (the level holds at most a few thousand cells)
template <int FRN>
__global__ __launch_bounds__(1024) void k_df_coarsest(float* __restrict__ d, const float* __restrict__ xb,
                                                      const float* __restrict__ rhs, Geo g, int sweeps, DfCoef c)
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

// ===========================================================================
// launchers
// ===========================================================================
#define LAUNCH_ARGS L.TX, L.TY, L.XB, L.NB
#define ZARGS L.TX, L.TY, L.XB, L.JB, L.KC, L.NB

#define FRN_DISPATCH(frn, KERNEL, ...)                                                   \
    do {                                                                                 \
        if ((frn) == 0)                                                                  \
            hipLaunchKernelGGL((KERNEL<0>), L.grid, L.block, 0, s, __VA_ARGS__);         \
        else if ((frn) == 1)                                                             \
            hipLaunchKernelGGL((KERNEL<1>), L.grid, L.block, 0, s, __VA_ARGS__);         \
        else                                                                             \
            hipLaunchKernelGGL((KERNEL<2>), L.grid, L.block, 0, s, __VA_ARGS__);         \
    } while (0)

template <class OP>
static void colour_pass(const OP& op, const Geo& g, int colour, hipStream_t s)
{
    const ZLaunch L = make_zlaunch(g);
    hipLaunchKernelGGL((k_colour_pass<OP>), L.grid, L.block, 0, s, op, g, colour, ZARGS);
}

template <class OP>
static int both_colours(const OP& op, double* partials, const Geo& g, hipStream_t s)
{
    const ZLaunch L = make_zlaunch(g);
    hipLaunchKernelGGL((k_both_colours<OP>), L.grid, L.block, 0, s, op, partials, g, ZARGS);
    return (int)L.NB;
}

void launch_pack_f64_from_f64(const double* lex, double* split, int n, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_pack<double, double>), L.grid, L.block, 0, s, lex, split, g, LAUNCH_ARGS);
}

void launch_pack_f64_from_f32(const float* lex, double* split, int n, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_pack<float, double>), L.grid, L.block, 0, s, lex, split, g, LAUNCH_ARGS);
}

void launch_unpack_f64_to_f64(const double* split, double* lex, int n, double scale, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_unpack<double>), L.grid, L.block, 0, s, split, lex, g, scale, LAUNCH_ARGS);
}

void launch_unpack_f64_to_f32(const double* split, float* lex, int n, double scale, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_unpack<float>), L.grid, L.block, 0, s, split, lex, g, scale, LAUNCH_ARGS);
}

static dim3 flat_grid(size_t count)
{
    size_t b = (count + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

void launch_fill_f64(double* p, size_t count, double v, hipStream_t s)
{
    hipLaunchKernelGGL((k_fill<double>), flat_grid(count), dim3(256), 0, s, p, count, v);
}

void launch_fill_f32(float* p, size_t count, float v, hipStream_t s)
{
    hipLaunchKernelGGL((k_fill<float>), flat_grid(count), dim3(256), 0, s, p, count, v);
}

void launch_copy_f64(const double* src, double* dst, size_t count, hipStream_t s)
{
    hipLaunchKernelGGL((k_copy<double>), flat_grid(count), dim3(256), 0, s, src, dst, count);
}

void launch_fr_smooth64(int frn, int level, double* u, const double* rho, const double* f, int n, int colour,
                        const FrCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
#define SM(FRN, CO, RR) colour_pass(FrSmoothOp<FRN, CO, RR>{u, rho, f, c}, g, colour, s)
#define SM2(FRN, CO)            \
    do {                        \
        if (c.reference_roots)  \
            SM(FRN, CO, true);  \
        else                    \
            SM(FRN, CO, false); \
    } while (0)
    if (level == 0) {
        if (frn == 0) SM2(0, false);
        else if (frn == 1) SM2(1, false);
        else SM2(2, false);
    } else {
        if (frn == 0) SM2(0, true);
        else if (frn == 1) SM2(1, true);
        else SM2(2, true);
    }
#undef SM2
#undef SM
}

int launch_fr_residual64(int frn, int level, const double* u, const double* rho, const double* f, double* res,
                         double* partials, int n, const FrCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
#define RS(FRN, CO) both_colours(FrResOp<FRN, CO>{u, rho, f, res, c}, partials, g, s)
    if (level == 0) {
        if (frn == 0) return RS(0, false);
        if (frn == 1) return RS(1, false);
        return RS(2, false);
    }
    if (frn == 0) return RS(0, true);
    if (frn == 1) return RS(1, true);
    return RS(2, true);
#undef RS
}

void launch_fr_restrict_residual1_64(int frn, const double* u0, const double* delta, double* rres1, int N,
                                     const FrCoef& c0, hipStream_t s)
{
    const Geo gf = make_geo(N), gc = make_geo(N / 2);
    const ZLaunch L = make_zlaunch(gc.n, gc.n);
    FRN_DISPATCH(frn, k_fr_restrict_residual1_64, u0, delta, rres1, gf, gc, c0, ZARGS);
}

void launch_restrict8_64(const double* fine, double* out1, double* out2, int nc, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const Launch L = make_launch(nc, (long)nc * nc);
    hipLaunchKernelGGL((k_restrict8<double>), L.grid, L.block, 0, s, fine, out1, out2, gf, gc, LAUNCH_ARGS);
}

void launch_restrict8_32(const float* fine, float* out1, float* out2, int nc, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const Launch L = make_launch(nc, (long)nc * nc);
    hipLaunchKernelGGL((k_restrict8<float>), L.grid, L.block, 0, s, fine, out1, out2, gf, gc, LAUNCH_ARGS);
}

void launch_fr_rhs64(int frn, const double* ru, const double* rho, const double* rres, double* f, int n,
                     const FrCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    if (frn == 0) both_colours(FrResOp<0, true>{ru, rho, rres, f, c}, nullptr, g, s);
    else if (frn == 1) both_colours(FrResOp<1, true>{ru, rho, rres, f, c}, nullptr, g, s);
    else both_colours(FrResOp<2, true>{ru, rho, rres, f, c}, nullptr, g, s);
}

void launch_correct64(const double* uH, const double* ru, double* ufine, int nc, int guard, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const ZLaunch L = make_zlaunch(nc, nc);
    hipLaunchKernelGGL(k_correct64, L.grid, L.block, 0, s, uH, ru, ufine, gf, gc, guard, ZARGS);
}

// partials[0..count) -> out[0]; stage 1 writes its chunk sums after the
// partials (the buffer holds count + 1024 doubles)
void launch_sum_partials(const double* partials, int count, double* out, hipStream_t s)
{
    if (count <= 8192) {
        hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(1024), 0, s, partials, count, count, out);
        return;
    }
    const int nchunk = 1024;
    const int chunk = (count + nchunk - 1) / nchunk;
    double* mid = const_cast<double*>(partials) + count;
    hipLaunchKernelGGL(k_sum_chunks, dim3(nchunk), dim3(1024), 0, s, partials, count, chunk, mid);
    hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(1024), 0, s, (const double*)mid, nchunk, nchunk, out);
}

int launch_outer_residual(int frn, const double* u, const double* delta, float* g32, float* xb, float* d,
                          double* partials, int n, const FrCoef& c, const DfCoef& dc, hipStream_t s)
{
    const Geo g = make_geo(n);
    const unsigned half = (unsigned)g.half;
    if (frn == 0) return both_colours(FrOuterOp<0>{u, delta, g32, xb, d, c, dc, half}, partials, g, s);
    if (frn == 1) return both_colours(FrOuterOp<1>{u, delta, g32, xb, d, c, dc, half}, partials, g, s);
    return both_colours(FrOuterOp<2>{u, delta, g32, xb, d, c, dc, half}, partials, g, s);
}

void launch_df_smooth_update(int frn, const float* d, const float* xb, const float* rhs, double* u, int n,
                             const DfCoef& c, double ubg, hipStream_t s)
{
    const Geo g = make_geo(n);
    if (frn == 0) colour_pass(DfSmoothUpdateOp<0>{d, xb, rhs, u, c, ubg}, g, 1, s);
    else if (frn == 1) colour_pass(DfSmoothUpdateOp<1>{d, xb, rhs, u, c, ubg}, g, 1, s);
    else colour_pass(DfSmoothUpdateOp<2>{d, xb, rhs, u, c, ubg}, g, 1, s);
}

void launch_restrict_base(int frn, const float* xb_f, float* xb_c, int nc, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const Launch L = make_launch(nc, (long)nc * nc);
    FRN_DISPATCH(frn, k_restrict_base, xb_f, xb_c, gf, gc, LAUNCH_ARGS);
}

void launch_df_smooth(int frn, float* d, const float* xb, const float* rhs, int n, int colour, const DfCoef& c,
                      hipStream_t s)
{
    const Geo g = make_geo(n);
    if (frn == 0) colour_pass(DfSmoothOp<0>{d, xb, rhs, c}, g, colour, s);
    else if (frn == 1) colour_pass(DfSmoothOp<1>{d, xb, rhs, c}, g, colour, s);
    else colour_pass(DfSmoothOp<2>{d, xb, rhs, c}, g, colour, s);
}

void launch_df_residual(int frn, const float* d, const float* xb, const float* rhs, float* res, int n,
                        const DfCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    if (frn == 0) both_colours(DfResOp<0>{d, xb, rhs, res, c}, nullptr, g, s);
    else if (frn == 1) both_colours(DfResOp<1>{d, xb, rhs, res, c}, nullptr, g, s);
    else both_colours(DfResOp<2>{d, xb, rhs, res, c}, nullptr, g, s);
}

void launch_df_restrict_residual(int frn, const float* d0, const float* xb0, const float* g0, float* rres1, int N,
                                 const DfCoef& c0, hipStream_t s)
{
    const Geo gf = make_geo(N), gc = make_geo(N / 2);
    const ZLaunch L = make_zlaunch(gc.n, gc.n);
    FRN_DISPATCH(frn, k_df_restrict_residual, d0, xb0, g0, rres1, gf, gc, c0, ZARGS);
}

void launch_df_rhs(int frn, const float* rd, const float* xb, const float* rres, float* f, int n, const DfCoef& c,
                   hipStream_t s)
{
    launch_df_residual(frn, rd, xb, rres, f, n, c, s);
}

void launch_df_correct(int frn, const float* dH, const float* rd, float* dfine, const float* xbfine, int nc,
                       hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const ZLaunch L = make_zlaunch(nc, nc);
    FRN_DISPATCH(frn, k_df_correct, dH, rd, dfine, xbfine, gf, gc, ZARGS);
}

void launch_df_coarsest(int frn, float* d, const float* xb, const float* rhs, int n, int sweeps, const DfCoef& c,
                        hipStream_t s)
{
    const Geo g = make_geo(n);
    int threads = (int)g.half;
    if (threads > 1024) threads = 1024;
    threads = ((threads + 63) / 64) * 64;
    if (frn == 0) hipLaunchKernelGGL((k_df_coarsest<0>), dim3(1), dim3(threads), 0, s, d, xb, rhs, g, sweeps, c);
    else if (frn == 1) hipLaunchKernelGGL((k_df_coarsest<1>), dim3(1), dim3(threads), 0, s, d, xb, rhs, g, sweeps, c);
    else hipLaunchKernelGGL((k_df_coarsest<2>), dim3(1), dim3(threads), 0, s, d, xb, rhs, g, sweeps, c);
}

}  // namespace mgs
