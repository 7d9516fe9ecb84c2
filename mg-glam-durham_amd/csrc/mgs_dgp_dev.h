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
#include "mgs_stencil.h"

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
    // the reference divides by dx^4 and dx^2; dx is a power of two, so the
    // products with the exact reciprocals are the same numbers
    Sigma = Sigma + se0 * TWOOVERTHREE * c.inv_dx4 - se1 * TWOOVERTHREE * c.inv_dx4 + se2 * ONEOVEREIGHT * c.inv_dx4;
    return lap * c.inv_dx2 - (copysign(1.0, c.alpha) * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * Sigma) - c.alpha) *
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

// ---------------------------------------------------------------------------
// z-marching with 19-point register windows (levels with n/2 >= 64, so that
// a wave is 64 consecutive m of one row)
// ---------------------------------------------------------------------------
// A thread owns column (m, j) and marches over a run of DGP_KC planes.  Its
// window (DWin2) holds, for planes k-1, k, k+1 and rows j-1, j, j+1, the
// pair at x = 2m, 2m+1; with x = 2m-1 and 2m+2 at the five slots the
// stencil reaches off the column (taken from the neighbouring lanes by a
// shuffle, or loaded by the lane at the wave's edge), that is every point of
// the 19-point stencil of the cell of either colour at (m, j, k).  A step
// loads the pair at m of plane k+2 in the three rows (both colour halves at
// one offset); plane k+2, the edge lanes' side values and the next cell's
// own data are loaded before plane k is computed.  The stencil values go
// through S19 like the per-cell kernels', so every result is bit-identical
// to theirs.
//
// The field is read through its two colour halves (h0, h1) and the pass
// writes the new values of its colour to `out`, a half of a second array: a
// pair of passes (colour 0, then colour 1 reading the new colour 0) moves
// the field into the second array without the per-cell kernels' copy.
struct DCol {
    unsigned plane, m, mm, mp;
    unsigned rowoff[3];
    int jr[3];
    bool lload, rload;
};

__device__ __forceinline__ DCol dcol_of(const Geo& g, int m, int j)
{
    DCol d;
    d.plane = (unsigned)g.n * (unsigned)g.h;
    d.m = (unsigned)m;
    d.mm = (unsigned)(m == 0 ? g.h - 1 : m - 1);
    d.mp = (unsigned)(m + 1 == g.h ? 0 : m + 1);
    d.jr[0] = wrap_dec(j, g.n);
    d.jr[1] = j;
    d.jr[2] = wrap_inc(j, g.n);
#pragma unroll
    for (int R = 0; R < 3; R++) d.rowoff[R] = (unsigned)d.jr[R] * (unsigned)g.h;
    const unsigned lane = (threadIdx.x + threadIdx.y * blockDim.x) & 63u;
    d.lload = lane == 0;
    d.rload = lane == 63;
    return d;
}

// Single-colour passes hold only x = 2m and 2m+1 of the window (DWin2).
// The cell of parity s needs the third x at five (plane, row) slots, q =
// 0..4: (1,1), (1,0), (1,2), (0,1), (2,1): x = 2m-1 when s = 0 (the o value
// of lane m-1) and x = 2m+2 when s = 1 (the e value of lane m+1), taken by a
// shuffle at use.  A wave is 64 consecutive m of one row and plane, so s is
// the same for all its lanes; only the edge lane on that side loads the five
// values itself (DSide, one step ahead).  The stencil values are the same
// numbers as DWin's, so results are bit-identical to its passes.
template <typename T>
struct DWin2 {
    T e[3][3], o[3][3];  // [plane k-1, k, k+1][row j-1, j, j+1]: x = 2m, 2m+1
};

template <typename T>
struct DRaw2 {
    T ev[3], ov[3];
};

template <typename T>
struct DSide {
    T v[5];
};

__device__ __forceinline__ int side_plane(int q) { return q == 3 ? 0 : (q == 4 ? 2 : 1); }
__device__ __forceinline__ int side_row(int q) { return q == 1 ? 0 : (q == 2 ? 2 : 1); }

// The two colour halves of a field seen from the columns of one wave: the
// value at x = 2m of row jr, plane K lies in p[0] when jr + K has the parity
// par0 of j + k0 (the wave's row and first plane), else in p[1]; the value at
// x = 2m+1 in the other.  A wave is 64 consecutive m of one row and plane
// run, so par0 and both pointers are wave-uniform (scalar registers), and
// with the plane offset dk = K - k0 known when the z-march is unrolled, the
// half of every load is chosen at compile time: (jr + K) & 1 =
// par0 ^ ((R + 1 + dk) & 1) for row jr = j + R - 1 (n even, so the wrap
// keeps the parity).
template <typename T>
struct DHalves {
    const T* p[2];
};

template <typename T>
__device__ __forceinline__ DHalves<T> dhalves(const T* h0, const T* h1, int par0)
{
    DHalves<T> H;
    H.p[0] = par0 ? h1 : h0;
    H.p[1] = par0 ? h0 : h1;
    return H;
}

// the pair at (m, rows j-1..j+1) of plane K (dk = K - k0 modulo 2)
template <typename T>
__device__ __forceinline__ void draw2_load(const DHalves<T>& H, const DCol& d, int K, int dk, DRaw2<T>& r)
{
#pragma unroll
    for (int R = 0; R < 3; R++) {
        const unsigned idx = (unsigned)K * d.plane + d.rowoff[R] + d.m;
        const int f = (R + 1 + dk) & 1;
        r.ev[R] = elem(H.p[f], idx);
        r.ov[R] = elem(H.p[f ^ 1], idx);
    }
}

// the five side values of the cell at plane k (dk = k - k0) with parity s
// (edge lane only)
template <typename T>
__device__ __forceinline__ void dside_load(const DHalves<T>& H, const DCol& d, int k, int dk, int n, bool s,
                                           DSide<T>& sd)
{
#pragma unroll
    for (int q = 0; q < 5; q++) sd.v[q] = T(0);
    if (s ? d.rload : d.lload) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const int P = side_plane(q), R = side_row(q);
            const int K = P == 0 ? wrap_dec(k, n) : (P == 2 ? wrap_inc(k, n) : k);
            const unsigned base = (unsigned)K * d.plane + d.rowoff[R];
            const int f = (R + P + dk) & 1;
            // x = 2m+2 has the colour of x = 2m, x = 2m-1 that of 2m+1
            sd.v[q] = s ? elem(H.p[f], base + d.mp) : elem(H.p[f ^ 1], base + d.mm);
        }
    }
}

// both sides at once, for passes that evaluate both cells of the pair: the
// left edge lane loads x = 2m-1 (the side of the cell at 2m), the right edge
// lane x = 2m+2 (the side of the cell at 2m+1), into the same five slots.
// Measured at 512^3: FP64 windows run faster with one branch per edge
// (scalar half bases); the two-field FP32 window of the mixed residual needs
// 129 registers that way (three waves per SIMD, 10% slower) and 118 with
// one branch choosing the half per lane.
template <typename T>
__device__ __forceinline__ void dside_load_edges(const DHalves<T>& H, const DCol& d, int k, int dk, int n,
                                                 DSide<T>& sd)
{
#pragma unroll
    for (int q = 0; q < 5; q++) sd.v[q] = T(0);
    if constexpr (sizeof(T) == 8) {
        if (d.lload) {
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const int P = side_plane(q), R = side_row(q);
                const int K = P == 0 ? wrap_dec(k, n) : (P == 2 ? wrap_inc(k, n) : k);
                sd.v[q] = elem(H.p[((R + P + dk) & 1) ^ 1], (unsigned)K * d.plane + d.rowoff[R] + d.mm);
            }
        } else if (d.rload) {
#pragma unroll
            for (int q = 0; q < 5; q++) {
                const int P = side_plane(q), R = side_row(q);
                const int K = P == 0 ? wrap_dec(k, n) : (P == 2 ? wrap_inc(k, n) : k);
                sd.v[q] = elem(H.p[(R + P + dk) & 1], (unsigned)K * d.plane + d.rowoff[R] + d.mp);
            }
        }
    } else if (d.lload || d.rload) {
        const unsigned mo = d.rload ? d.mp : d.mm;
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const int P = side_plane(q), R = side_row(q);
            const int K = P == 0 ? wrap_dec(k, n) : (P == 2 ? wrap_inc(k, n) : k);
            const int f = (R + P + dk) & 1;
            sd.v[q] = elem(d.rload ? H.p[f] : H.p[f ^ 1], (unsigned)K * d.plane + d.rowoff[R] + mo);
        }
    }
}

template <typename T>
__device__ __forceinline__ void dwin2_set(DWin2<T>& w, int P, const DRaw2<T>& r)
{
#pragma unroll
    for (int R = 0; R < 3; R++) {
        w.e[P][R] = r.ev[R];
        w.o[P][R] = r.ov[R];
    }
}

template <typename T>
__device__ __forceinline__ void dwin2_shift(DWin2<T>& w)
{
#pragma unroll
    for (int R = 0; R < 3; R++) {
        w.e[0][R] = w.e[1][R];
        w.o[0][R] = w.o[1][R];
        w.e[1][R] = w.e[2][R];
        w.o[1][R] = w.o[2][R];
    }
}

template <typename T>
__device__ __forceinline__ void dwin2_init(DWin2<T>& w, const DHalves<T>& H, const DCol& d, int k0, int n)
{
    DRaw2<T> r0, r1, r2;
    draw2_load(H, d, wrap_dec(k0, n), -1, r0);
    draw2_load(H, d, k0, 0, r1);
    draw2_load(H, d, wrap_inc(k0, n), 1, r2);
    dwin2_set(w, 0, r0);
    dwin2_set(w, 1, r1);
    dwin2_set(w, 2, r2);
}

// the value of lane + 1 (FROM_RIGHT) or lane - 1 of the wave, by a DPP
// whole-wave shift (wave_shl:1 / wave_shr:1); the last / first lane, which
// has no such neighbour, keeps its own value
template <bool FROM_RIGHT>
__device__ __forceinline__ int wave_shift1(int v)
{
    if constexpr (FROM_RIGHT)
        return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false);
    else
        return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);
}

template <bool FROM_RIGHT>
__device__ __forceinline__ float wave_shift1(float v)
{
    return __int_as_float(wave_shift1<FROM_RIGHT>(__float_as_int(v)));
}

template <bool FROM_RIGHT>
__device__ __forceinline__ double wave_shift1(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)wave_shift1<FROM_RIGHT>((int)(unsigned)u);
    const unsigned hi = (unsigned)wave_shift1<FROM_RIGHT>((int)(unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// the 19 stencil values of the cell of parity s (all lanes of the wave
// execute the shifts)
template <typename T>
__device__ __forceinline__ S19<T> s19_win2(const DWin2<T>& w, bool s, const DSide<T>& sd, const DCol& d)
{
    const bool edge = s ? d.rload : d.lload;  // the lane whose neighbour is outside the wave keeps sd
    T v[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const int P = side_plane(q), R = side_row(q);
        const T sh = s ? wave_shift1<true>(w.e[P][R]) : wave_shift1<false>(w.o[P][R]);
        v[q] = edge ? sd.v[q] : sh;
    }
    // x = i + dx at slot (P, R), i = 2m + s; q: the side slot of (P, R)
    auto X = [&](int P, int R, int dx, int q) -> T {
        if (dx == 0) return s ? w.o[P][R] : w.e[P][R];
        if (dx < 0) return s ? w.e[P][R] : v[q];
        return s ? v[q] : w.o[P][R];
    };
    S19<T> u;
    u.c0 = X(1, 1, 0, 0);
    u.xp = X(1, 1, 1, 0);
    u.xm = X(1, 1, -1, 0);
    u.yp = X(1, 2, 0, 0);
    u.ym = X(1, 0, 0, 0);
    u.zp = X(2, 1, 0, 0);
    u.zm = X(0, 1, 0, 0);
    u.xy_pp = X(1, 2, 1, 2);
    u.xy_mm = X(1, 0, -1, 1);
    u.xy_pm = X(1, 0, 1, 1);
    u.xy_mp = X(1, 2, -1, 2);
    u.xz_pp = X(2, 1, 1, 4);
    u.xz_mm = X(0, 1, -1, 3);
    u.xz_pm = X(0, 1, 1, 3);
    u.xz_mp = X(2, 1, -1, 4);
    u.yz_pp = X(2, 2, 0, 0);
    u.yz_mm = X(0, 0, 0, 0);
    u.yz_pm = X(0, 2, 0, 0);
    u.yz_mp = X(2, 0, 0, 0);
    return u;
}

// Both cells of the pair at (m, j) over the run [k0, k0 + DGP_KC), from one
// window per field (OP::NF fields of type OP::T, colour halves
// op.half(f, 0/1)).  Per plane the cell at x = 2m comes first, then x = 2m+1;
// OP::load(index) loads a cell's own data, OP::cell(index, own, stencils)
// returns a value whose square is accumulated (the sum is returned).  The
// next plane's window values and own data are loaded before a plane is
// evaluated.
template <class OP>
__device__ __forceinline__ double dgp_both_zcol(const OP& op, const Geo& g, int m, int j, int k0)
{
    using T = typename OP::T;
    using Own = typename OP::Own;
    constexpr int NF = OP::NF;
    const DCol d = dcol_of(g, m, j);
    const unsigned half = (unsigned)g.half;
    const int par0 = __builtin_amdgcn_readfirstlane((j + k0) & 1);  // wave-uniform
    DHalves<T> H[NF];
    DWin2<T> w[NF];
    DSide<T> sd[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        H[f] = dhalves(op.half(f, 0), op.half(f, 1), par0);
        dwin2_init(w[f], H[f], d, k0, g.n);
        dside_load_edges(H[f], d, k0, 0, g.n, sd[f]);
    }
    unsigned off = (unsigned)k0 * d.plane + d.rowoff[1] + d.m;
    unsigned par = (unsigned)((j + k0) & 1);  // colour of x = 2m
    Own o0 = op.load(par * half + off), o1 = op.load((1u - par) * half + off);
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < DGP_KC; t++) {
        const int k = k0 + t;
        DRaw2<T> r[NF];
        DSide<T> sdn[NF];
        Own n0 = o0, n1 = o1;
        const unsigned offn = off + d.plane, parn = par ^ 1u;
        if (t + 1 < DGP_KC) {
#pragma unroll
            for (int f = 0; f < NF; f++) {
                draw2_load(H[f], d, wrap_inc(k + 1, g.n), t + 2, r[f]);
                dside_load_edges(H[f], d, wrap_inc(k, g.n), t + 1, g.n, sdn[f]);
            }
            n0 = op.load(parn * half + offn);
            n1 = op.load((1u - parn) * half + offn);
        }
        // one cell's stencils at a time (fewer live values)
        double r0, r1;
        {
            S19<T> s0[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) s0[f] = s19_win2(w[f], false, sd[f], d);
            r0 = op.cell(par * half + off, o0, s0);
        }
        {
            S19<T> s1[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) s1[f] = s19_win2(w[f], true, sd[f], d);
            r1 = op.cell((1u - par) * half + off, o1, s1);
        }
        acc += r0 * r0 + r1 * r1;
        if (t + 1 < DGP_KC) {
#pragma unroll
            for (int f = 0; f < NF; f++) {
                dwin2_shift(w[f]);
                dwin2_set(w[f], 2, r[f]);
                sd[f] = sdn[f];
            }
            off = offn;
            par = parn;
            o0 = n0;
            o1 = n1;
        }
    }
    return acc;
}

// FP64 residual L(phi) (- f on coarse levels), stored in res when non-null
template <bool COARSE, bool NT = false>
struct DgpRes64Op {
    using T = double;
    static constexpr int NF = 1;
    struct Own {
        double rho, f;
    };
    const double* u;
    const double* rho;
    const double* f;
    double* res;
    DgpCoef c;
    unsigned hf;
    __device__ const double* half(int, int colour) const { return u + (colour ? hf : 0u); }
    __device__ __forceinline__ Own load(unsigned i) const
    {
        return Own{ld_once<NT>(rho[i]), COARSE ? ld_once<NT>(f[i]) : 0.0};
    }
    __device__ __forceinline__ double cell(unsigned i, const Own& o, const S19<double>* s) const
    {
        double OP = dgp_L19(s[0], o.rho, c);
        if constexpr (COARSE) OP = OP - o.f;
        if (res) st_once<NT>(res[i], OP);
        return OP;
    }
};

// the mixed solve's FP64 outer residual r = L(phi), with the inner problem's
// data g = -r, phi_b = phi (FP32) and e = 0 (as dgpm_outer_cell)
template <bool NT = false>
struct DgpmOuterOp {
    using T = double;
    static constexpr int NF = 1;
    struct Own {
        double rho;
    };
    const double* u;
    const double* rho;
    float* g32;
    float* pb;
    float* e;
    DgpCoef c;
    unsigned hf;
    __device__ const double* half(int, int colour) const { return u + (colour ? hf : 0u); }
    __device__ __forceinline__ Own load(unsigned i) const { return Own{ld_once<NT>(rho[i])}; }
    __device__ __forceinline__ double cell(unsigned i, const Own& o, const S19<double>* s) const
    {
        const double OP = dgp_L19(s[0], o.rho, c);
        st_once<NT>(g32[i], (float)(-OP));
        st_once<NT>(pb[i], (float)s[0].c0);
        st_once<NT>(e[i], 0.0f);
        return OP;
    }
};

// the mixed inner residual out = D(e) - rhs (as dgpm_residual_cell, which
// rounds the FP64 density to FP32: rho32 holds those values); no norm
template <bool NT = false>
struct DgpmResOp {
    using T = float;
    static constexpr int NF = 2;  // e, phi_b
    struct Own {
        float rho, rhs;
    };
    const float* e;
    const float* pb;
    const float* rho32;
    const float* rhs;
    float* out;
    DgpF c;
    unsigned hf;
    __device__ const float* half(int f, int colour) const { return (f ? pb : e) + (colour ? hf : 0u); }
    __device__ __forceinline__ Own load(unsigned i) const { return Own{ld_once<NT>(rho32[i]), ld_once<NT>(rhs[i])}; }
    __device__ __forceinline__ double cell(unsigned i, const Own& o, const S19<float>* s) const
    {
        st_once<NT>(out[i], dgp_dop(d2_of(s[1]), d2_of(s[0]), o.rho, c) - o.rhs);
        return 0.0;
    }
};

}  // namespace mgs
