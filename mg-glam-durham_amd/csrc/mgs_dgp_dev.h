// mgs_dgp_dev.h -- device-side nDGP operators shared by the standalone
// kernels (mgs_dgp.hip) and the batched serving kernels (mgs_serve.hip):
// the FP64 19-point operator L(phi) of PMP2MGsolver_DGP.f90 and the FP32
// delta-form operator D(e) of the mixed solve, and per-cell bodies of the
// mixed solve's kernels.  The same bodies run in both places, so a request
// served in a batch is bit-identical to mgs_solve.
#pragma once

#include <cmath>

#include "mgs_device.h"
#include "mgs_internal.h"

namespace mgs {

// The 19 points of the stencil at a cell: centre, 6 faces and 12 edges
// (the edges named by their offsets, e.g. pp_xy = (i+1, j+1, k)).  Both the
// per-cell gathers and the z-marching register windows fill one, and the
// operators below read it, so every path adds the same values in the same
// order.
template <typename T>
struct S19 {
    T c0, xp, xm, yp, ym, zp, zm;
    T xy_pp, xy_mm, xy_pm, xy_mp;  // (i+-1, j+-1, k)
    T xz_pp, xz_mm, xz_pm, xz_mp;  // (i+-1, j, k+-1)
    T yz_pp, yz_mm, yz_pm, yz_mp;  // (i, j+-1, k+-1)
};

template <typename T>
__device__ __forceinline__ S19<T> s19_at(const T* __restrict__ u, const Geo& g, int i, int j, int k)
{
    const int n = g.n;
    const int im = wrap_dec(i, n), ip = wrap_inc(i, n);
    const int jm = wrap_dec(j, n), jp = wrap_inc(j, n);
    const int km = wrap_dec(k, n), kp = wrap_inc(k, n);
    S19<T> s;
    s.c0 = u[split_index(g, i, j, k)];
    s.xp = u[split_index(g, ip, j, k)];
    s.xm = u[split_index(g, im, j, k)];
    s.yp = u[split_index(g, i, jp, k)];
    s.ym = u[split_index(g, i, jm, k)];
    s.zp = u[split_index(g, i, j, kp)];
    s.zm = u[split_index(g, i, j, km)];
    s.xy_pp = u[split_index(g, ip, jp, k)];
    s.xy_mm = u[split_index(g, im, jm, k)];
    s.xy_pm = u[split_index(g, ip, jm, k)];
    s.xy_mp = u[split_index(g, im, jp, k)];
    s.xz_pp = u[split_index(g, ip, j, kp)];
    s.xz_mm = u[split_index(g, im, j, km)];
    s.xz_pm = u[split_index(g, ip, j, km)];
    s.xz_mp = u[split_index(g, im, j, kp)];
    s.yz_pp = u[split_index(g, i, jp, kp)];
    s.yz_mm = u[split_index(g, i, jm, km)];
    s.yz_pm = u[split_index(g, i, jp, km)];
    s.yz_mp = u[split_index(g, i, jm, kp)];
    return s;
}

// Sigma's stencil sums and the Laplacian numerator
__device__ __forceinline__ void dgp_sums(const S19<double>& u, double& se0, double& se1, double& se2, double& lap)
{
    const double dxx = u.xp + u.xm - 2.0 * u.c0;
    const double dyy = u.yp + u.ym - 2.0 * u.c0;
    const double dzz = u.zp + u.zm - 2.0 * u.c0;
    const double cxy = u.xy_pp + u.xy_mm - u.xy_pm - u.xy_mp;
    const double cxz = u.xz_pp + u.xz_mm - u.xz_pm - u.xz_mp;
    const double cyz = u.yz_pp + u.yz_mm - u.yz_pm - u.yz_mp;
    se0 = dxx * dxx + dyy * dyy + dzz * dzz;
    se1 = dxx * dyy + dxx * dzz + dyy * dzz;
    se2 = cxy * cxy + cxz * cxz + cyz * cyz;
    const double s6 = u.xp + u.xm + u.yp + u.ym + u.zp + u.zm;
    lap = s6 - u.c0 * 6.0;
}

// L(phi) at a cell from its stencil values (without the FAS right-hand side)
__device__ __forceinline__ double dgp_L19(const S19<double>& u, double rho, const DgpCoef& c)
{
    const double TWOOVERTHREE = 2.0 / 3.0, ONEOVEREIGHT = 0.125, THREEOVERFOUR = 0.75, EIGHTOVERTHREE = 8.0 / 3.0;
    double se0, se1, se2, lap;
    dgp_sums(u, se0, se1, se2, lap);
    double Sigma = c.fct1 * rho;
    Sigma = Sigma + se0 * TWOOVERTHREE / c.dx4 - se1 * TWOOVERTHREE / c.dx4 + se2 * ONEOVEREIGHT / c.dx4;
    return lap / c.dx2 - (copysign(1.0, c.alpha) * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * Sigma) - c.alpha) *
                             THREEOVERFOUR;
}

// L(phi) at absolute cell (i, j, k); c0: the cell's value
__device__ __forceinline__ double dgp_L(const double* __restrict__ u, const Geo& g, int i, int j, int k, double rho,
                                        const DgpCoef& c, double& c0)
{
    const S19<double> s = s19_at(u, g, i, j, k);
    c0 = s.c0;
    return dgp_L19(s, rho, c);
}

// second differences, face sum and centre of a field at a cell
struct D2f {
    float xx, yy, zz, xy, xz, yz, s6, c0;
};

__device__ __forceinline__ D2f d2_of(const S19<float>& u)
{
    D2f d;
    d.c0 = u.c0;
    d.xx = u.xp + u.xm - 2.0f * d.c0;
    d.yy = u.yp + u.ym - 2.0f * d.c0;
    d.zz = u.zp + u.zm - 2.0f * d.c0;
    d.xy = u.xy_pp + u.xy_mm - u.xy_pm - u.xy_mp;
    d.xz = u.xz_pp + u.xz_mm - u.xz_pm - u.xz_mp;
    d.yz = u.yz_pp + u.yz_mm - u.yz_pm - u.yz_mp;
    d.s6 = u.xp + u.xm + u.yp + u.ym + u.zp + u.zm;
    return d;
}

__device__ __forceinline__ D2f d2_at(const float* __restrict__ u, const Geo& g, int i, int j, int k)
{
    return d2_of(s19_at(u, g, i, j, k));
}

struct DgpF {
    float alpha, fct1, inv_dx2, inv_dx4, sgn;
};

inline DgpF dgpf(const DgpCoef& c)
{
    return DgpF{(float)c.alpha, (float)c.fct1, (float)(1.0 / c.dx2), (float)(1.0 / c.dx4),
                (float)std::copysign(1.0, c.alpha)};
}

// D(e) at a cell from the base's and the increment's second differences
__device__ __forceinline__ float dgp_dop(const D2f& b, const D2f& e, float rho, const DgpF& c)
{
    const float Qb = b.xx * b.xx + b.yy * b.yy + b.zz * b.zz - b.xx * b.yy - b.xx * b.zz - b.yy * b.zz;
    const float Cb = b.xy * b.xy + b.xz * b.xz + b.yz * b.yz;
    const float Sig = c.fct1 * rho + ((2.0f / 3.0f) * Qb + 0.125f * Cb) * c.inv_dx4;
    const float Bx = 2.0f * b.xx - b.yy - b.zz, By = 2.0f * b.yy - b.xx - b.zz, Bz = 2.0f * b.zz - b.xx - b.yy;
    const float Qe = e.xx * e.xx + e.yy * e.yy + e.zz * e.zz - e.xx * e.yy - e.xx * e.zz - e.yy * e.zz;
    const float lin = Bx * e.xx + By * e.yy + Bz * e.zz;
    const float cross = 2.0f * (b.xy * e.xy + b.xz * e.xz + b.yz * e.yz) + e.xy * e.xy + e.xz * e.xz + e.yz * e.yz;
    const float dSig = ((2.0f / 3.0f) * (lin + Qe) + 0.125f * cross) * c.inv_dx4;
    const float a2 = c.alpha * c.alpha;
    const float So = sqrtf(fmaxf(a2 + (8.0f / 3.0f) * Sig, 0.0f));
    const float Sn = sqrtf(fmaxf(a2 + (8.0f / 3.0f) * (Sig + dSig), 0.0f));
    const float den = So + Sn;
    const float nl = den > 0.0f ? 2.0f * c.sgn * dSig / den : 0.0f;
    return (e.s6 - 6.0f * e.c0) * c.inv_dx2 - nl;
}

// ---------------------------------------------------------------------------
// per-cell bodies of the mixed nDGP kernels
// ---------------------------------------------------------------------------

// one colour pass of the inner smoother at the cell (m, row = j + n k) of
// colour `colour`, out of place into tmp (one colour half)
__device__ __forceinline__ void dgpm_smooth_cell(const float* __restrict__ e, const float* __restrict__ pb,
                                                 const double* __restrict__ rho, const float* __restrict__ rhs,
                                                 float* __restrict__ tmp, const Geo& g, int colour, const DgpF& c,
                                                 int m, unsigned row)
{
    const int j = (int)(row % (unsigned)g.n), k = (int)(row / (unsigned)g.n);
    const int i = 2 * m + ((colour + j + k) & 1);
    const size_t me = (size_t)colour * g.half + row_base(g, j, k) + m;
    const D2f de = d2_at(e, g, i, j, k), db = d2_at(pb, g, i, j, k);
    const float D = dgp_dop(db, de, (float)rho[me], c) - rhs[me];
    tmp[row_base(g, j, k) + m] = de.c0 + D / (6.0f * c.inv_dx2);
}

// out = D(e) - rhs at the cell (m, row2) of both colours (row2 < n^2:
// colour 0, else colour 1)
__device__ __forceinline__ void dgpm_residual_cell(const float* __restrict__ e, const float* __restrict__ pb,
                                                   const double* __restrict__ rho, const float* __restrict__ rhs,
                                                   float* __restrict__ out, const Geo& g, const DgpF& c, int m,
                                                   unsigned row2)
{
    const unsigned nn = (unsigned)g.n * (unsigned)g.n;
    const int col = row2 >= nn ? 1 : 0;
    const unsigned r2 = row2 - (unsigned)col * nn;
    const int j = (int)(r2 % (unsigned)g.n), k = (int)(r2 / (unsigned)g.n);
    const int i = 2 * m + ((col + j + k) & 1);
    const size_t me = (size_t)col * g.half + row_base(g, j, k) + m;
    out[me] = dgp_dop(d2_at(pb, g, i, j, k), d2_at(e, g, i, j, k), (float)rho[me], c) - rhs[me];
}

// FP64 outer residual r = L(phi) at the cell (m, row2) of both colours, and
// the inner problem's data there (g = -r, phi_b = phi in FP32, e = 0); r^2
__device__ __forceinline__ double dgpm_outer_cell(const double* __restrict__ u, const double* __restrict__ rho,
                                                  float* __restrict__ g32, float* __restrict__ pb,
                                                  float* __restrict__ e, const Geo& g, const DgpCoef& c, int m,
                                                  unsigned row2)
{
    const unsigned nn = (unsigned)g.n * (unsigned)g.n;
    const int col = row2 >= nn ? 1 : 0;
    const unsigned r2 = row2 - (unsigned)col * nn;
    const int j = (int)(r2 % (unsigned)g.n), k = (int)(r2 / (unsigned)g.n);
    const int i = 2 * m + ((col + j + k) & 1);
    const size_t me = (size_t)col * g.half + row_base(g, j, k) + m;
    double c0;
    const double OP = dgp_L(u, g, i, j, k, rho[me], c, c0);
    g32[me] = (float)(-OP);
    pb[me] = (float)c0;
    e[me] = 0.0f;
    return OP * OP;
}

// e_h += P(e_H - R e_h) for the 8 children of coarse cell (I, J, K) (no
// guard for DGP)
__device__ __forceinline__ void dgpm_correct_cell(const float* __restrict__ eH, const float* __restrict__ rd,
                                                  float* __restrict__ ef, const Geo& gf, const Geo& gc, int I, int J,
                                                  int K)
{
    const int nc = gc.n;
    const float a = 1.0f / 64.0f, b = 3.0f * a, c9 = 9.0f * a, dd = 27.0f * a;
    const float wt[8] = {a, b, b, c9, b, c9, c9, dd};
    const int Xs[3] = {wrap_dec(I, nc), I, wrap_inc(I, nc)};
    const int Ys[3] = {wrap_dec(J, nc), J, wrap_inc(J, nc)};
    const int Zs[3] = {wrap_dec(K, nc), K, wrap_inc(K, nc)};
    float ev[27];
#pragma unroll
    for (int z = 0; z < 3; z++)
#pragma unroll
        for (int y = 0; y < 3; y++)
#pragma unroll
            for (int x = 0; x < 3; x++) {
                const size_t q = split_index(gc, Xs[x], Ys[y], Zs[z]);
                ev[x + 3 * y + 9 * z] = eH[q] - rd[q];
            }
    float old[8];
#pragma unroll
    for (int sind = 0; sind < 8; sind++)
        old[sind] = ef[split_index(gf, 2 * I + (sind & 1), 2 * J + ((sind >> 1) & 1), 2 * K + (sind >> 2))];
#pragma unroll
    for (int sind = 0; sind < 8; sind++) {
        const int s1 = sind & 1, s2 = (sind >> 1) & 1, s3 = sind >> 2;
        float P = 0.0f;
#pragma unroll
        for (int ind = 0; ind < 8; ind++) {
            const int x = (ind & 1) ? 1 : (s1 ? 2 : 0);
            const int y = (ind & 2) ? 1 : (s2 ? 2 : 0);
            const int z = (ind & 4) ? 1 : (s3 ? 2 : 0);
            P += wt[ind] * ev[x + 3 * y + 9 * z];
        }
        ef[split_index(gf, 2 * I + s1, 2 * J + s2, 2 * K + s3)] = old[sind] + P;
    }
}

// coarsest level: sweeps of the inner smoother by one workgroup, the update
// evaluated in FP64 (the level holds at most a few thousand cells)
__device__ __forceinline__ void dgpm_coarsest_block(float* __restrict__ e, const float* __restrict__ pb,
                                                    const double* __restrict__ rho, const float* __restrict__ rhs,
                                                    float* __restrict__ tmp, const Geo& g, int sweeps, const DgpF& c)
{
    for (int sw = 0; sw < sweeps; sw++) {
        for (int colour = 0; colour < 2; colour++) {
            for (int t = threadIdx.x; t < (int)g.half; t += blockDim.x) {
                const int m = t % g.h, row = t / g.h;
                const int j = row % g.n, k = row / g.n;
                const int i = 2 * m + ((colour + j + k) & 1);
                const size_t me = (size_t)colour * g.half + t;
                const D2f de = d2_at(e, g, i, j, k);
                const double D = (double)dgp_dop(d2_at(pb, g, i, j, k), de, (float)rho[me], c) - (double)rhs[me];
                tmp[t] = (float)((double)de.c0 + D / (6.0 * (double)c.inv_dx2));
            }
            __syncthreads();
            for (int t = threadIdx.x; t < (int)g.half; t += blockDim.x) e[(size_t)colour * g.half + t] = tmp[t];
            __syncthreads();
        }
    }
}

}  // namespace mgs
