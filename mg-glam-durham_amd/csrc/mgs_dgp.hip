// mgs_dgp.hip -- nDGP brane-bending multigrid kernels for gfx950.
//
// The reference's DGP operators (PMP2MGsolver_DGP.f90) in FP64 on the
// colour-split layout:
//   L(phi) = lap(phi)/dx^2 - 3/4 [sgn(alpha) sqrt(alpha^2 + 8/3 Sigma) - alpha] (- f)
//   Sigma  = fct1 rho + 2/3 (Dxx^2+Dyy^2+Dzz^2) - 2/3 (DxxDyy+DxxDzz+DyyDzz)
//            + 1/8 (Cxy^2+Cxz^2+Cyz^2),   all / dx^4
// on the 19-point stencil (faces + 12 edges).  The smoother updates
// phi <- phi - L / (-6/dx^2) (:226-230).
//
// Red-black on a 19-point stencil: a cell's 12 edge neighbours have the
// cell's own colour.  The reference sweeps a colour sequentially, so a cell
// sees the already-updated values of the edge neighbours that precede it in
// memory order.  A parallel colour pass cannot reproduce that order; here
// every cell of the colour is updated from the values the colour had before
// the pass (the pass writes into a scratch half-array that is then copied
// back).  The CPU oracle implements both orders (orc_dgp_smooth, lexi=0/1)
// and both converge to the same discrete solution.

#include <math.h>

#include "mgs_device.h"
#include "mgs_internal.h"
#include "mgs_kernels.h"

namespace mgs {

DgpCoef dgp_coefs(const mgs_params& p, int level)
{
    DgpCoef c;
    const double a = p.aexpn;
    const double dx = (double)(1 << level);
    c.dx2 = dx * dx;
    c.dx4 = c.dx2 * c.dx2;
    const double Orc = 1.0 / (4.0 * p.h0rc * p.h0rc);
    const double a3 = a * a * a;
    if (p.branch >= 0)
        c.beta = 1.0 + (0.5 * p.omega_m / a3 + p.omega_l) / std::sqrt(Orc * (p.omega_m / a3 + p.omega_l));
    else
        c.beta = -(0.5 * p.omega_m / a3 + Orc) / std::sqrt(Orc * (p.omega_m / a3 + Orc));
    const double Rc_sq = 1.0 / (4.0 * Orc);
    c.alpha = 3.0 * c.beta * (a * a) / Rc_sq;
    c.fct1 = c.alpha * p.omega_m / c.beta / a;
    c.fct2_1 = -0.75 * c.alpha;
    c.fct3_1 = std::copysign(0.75, c.alpha);
    return c;
}

// Sigma's stencil sums and the Laplacian numerator at absolute cell (i,j,k)
__device__ __forceinline__ void dgp_stencil(const double* __restrict__ u, const Geo& g, int i, int j, int k,
                                            double& se0, double& se1, double& se2, double& lap, double& c0)
{
    const int n = g.n;
    const int im = wrap_dec(i, n), ip = wrap_inc(i, n);
    const int jm = wrap_dec(j, n), jp = wrap_inc(j, n);
    const int km = wrap_dec(k, n), kp = wrap_inc(k, n);
    c0 = u[split_index(g, i, j, k)];
    const double xp = u[split_index(g, ip, j, k)], xm = u[split_index(g, im, j, k)];
    const double yp = u[split_index(g, i, jp, k)], ym = u[split_index(g, i, jm, k)];
    const double zp = u[split_index(g, i, j, kp)], zm = u[split_index(g, i, j, km)];
    const double dxx = xp + xm - 2.0 * c0;
    const double dyy = yp + ym - 2.0 * c0;
    const double dzz = zp + zm - 2.0 * c0;
    const double cxy = u[split_index(g, ip, jp, k)] + u[split_index(g, im, jm, k)] - u[split_index(g, ip, jm, k)] -
                       u[split_index(g, im, jp, k)];
    const double cxz = u[split_index(g, ip, j, kp)] + u[split_index(g, im, j, km)] - u[split_index(g, ip, j, km)] -
                       u[split_index(g, im, j, kp)];
    const double cyz = u[split_index(g, i, jp, kp)] + u[split_index(g, i, jm, km)] - u[split_index(g, i, jp, km)] -
                       u[split_index(g, i, jm, kp)];
    se0 = dxx * dxx + dyy * dyy + dzz * dzz;
    se1 = dxx * dyy + dxx * dzz + dyy * dzz;
    se2 = cxy * cxy + cxz * cxz + cyz * cyz;
    const double s6 = xp + xm + yp + ym + zp + zm;
    lap = s6 - c0 * 6.0;
}

// L(phi) at a cell (without the FAS right-hand side)
__device__ __forceinline__ double dgp_L(const double* __restrict__ u, const Geo& g, int i, int j, int k, double rho,
                                        const DgpCoef& c, double& c0)
{
    const double TWOOVERTHREE = 2.0 / 3.0, ONEOVEREIGHT = 0.125, THREEOVERFOUR = 0.75, EIGHTOVERTHREE = 8.0 / 3.0;
    double se0, se1, se2, lap;
    dgp_stencil(u, g, i, j, k, se0, se1, se2, lap, c0);
    double Sigma = c.fct1 * rho;
    Sigma = Sigma + se0 * TWOOVERTHREE / c.dx4 - se1 * TWOOVERTHREE / c.dx4 + se2 * ONEOVEREIGHT / c.dx4;
    return lap / c.dx2 - (copysign(1.0, c.alpha) * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * Sigma) - c.alpha) *
                             THREEOVERFOUR;
}

// relaxation_iterations_DGP, one colour, out of place into tmp (one colour half)
template <bool COARSE>
__global__ __launch_bounds__(256) void k_dgp_smooth64(const double* __restrict__ u, const double* __restrict__ rho,
                                                      const double* __restrict__ f, double* __restrict__ tmp, Geo g,
                                                      int colour, DgpCoef c, int TX, int TY, int XB, unsigned NB)
{
    int m;
    unsigned row;
    if (!map_xy(g.h, (long)g.n * g.n, TX, TY, XB, NB, m, row)) return;
    const int j = (int)(row % (unsigned)g.n), k = (int)(row / (unsigned)g.n);
    const int i = 2 * m + ((colour + j + k) & 1);
    const size_t me = (size_t)colour * g.half + row_base(g, j, k) + m;
    double c0;
    double L = dgp_L(u, g, i, j, k, rho[me], c, c0);
    if constexpr (COARSE) L = L - f[me];
    const double dLphi = -6.0 / c.dx2;
    tmp[row_base(g, j, k) + m] = c0 - L / dLphi;
}

template <bool COARSE>
__global__ __launch_bounds__(256) void k_dgp_residual64(const double* __restrict__ u, const double* __restrict__ rho,
                                                        const double* __restrict__ f, double* __restrict__ res,
                                                        double* __restrict__ partials, Geo g, DgpCoef c, int TX,
                                                        int TY, int XB, unsigned NB)
{
    int m;
    unsigned row;
    double acc = 0.0;
    if (map_xy(g.h, 2L * g.n * g.n, TX, TY, XB, NB, m, row)) {
        const unsigned nn = (unsigned)g.n * (unsigned)g.n;
        const int col = row >= nn ? 1 : 0;
        const unsigned r2 = row - (unsigned)col * nn;
        const int j = (int)(r2 % (unsigned)g.n), k = (int)(r2 / (unsigned)g.n);
        const int i = 2 * m + ((col + j + k) & 1);
        const size_t me = (size_t)col * g.half + row_base(g, j, k) + m;
        double c0;
        double OP = dgp_L(u, g, i, j, k, rho[me], c, c0);
        if constexpr (COARSE) OP = OP - f[me];
        if (res) res[me] = OP;
        acc = OP * OP;
    }
    store_partial(partials, NB, block_sum(acc));
}

// restrict_residual_DGP(1), PMP2MGsolver_DGP.f90:459-763 (rearranged Sigma)
__global__ __launch_bounds__(256) void k_dgp_restrict_residual1_64(const double* __restrict__ u0,
                                                                   const double* __restrict__ delta,
                                                                   double* __restrict__ rres1, Geo gf, Geo gc,
                                                                   DgpCoef c, int TX, int TY, int XB, unsigned NB)
{
    int I;
    unsigned row;
    if (!map_xy(gc.n, (long)gc.n * gc.n, TX, TY, XB, NB, I, row)) return;
    const int J = (int)(row % (unsigned)gc.n), K = (int)(row / (unsigned)gc.n);
    const int N = gf.n;
    const double ONEOVEREIGHT = 1.0 / 8.0, ONEOVERTHREE = 1.0 / 3.0, EIGHTOVERTHREE = 8.0 / 3.0;
    const double dx2 = 1.0, dx4 = 1.0;
    double Q = -8.0 * c.fct2_1;
    double sig[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + ((q >> 2) & 1);
        const int im = wrap_dec(i, N), ip = wrap_inc(i, N);
        const int jm = wrap_dec(j, N), jp = wrap_inc(j, N);
        const int km = wrap_dec(k, N), kp = wrap_inc(k, N);
#define U_(x, y, z) u0[split_index(gf, x, y, z)]
        const double Sx = U_(ip, j, k) + U_(im, j, k);
        const double Sy = U_(i, jp, k) + U_(i, jm, k);
        const double Sz = U_(i, j, kp) + U_(i, j, km);
        const double Dxy = (U_(ip, jp, k) + U_(im, jm, k)) - (U_(ip, jm, k) + U_(im, jp, k));
        const double Dxz = (U_(ip, j, kp) + U_(im, j, km)) - (U_(ip, j, km) + U_(im, j, kp));
        const double Dyz = (U_(i, jp, kp) + U_(i, jm, km)) - (U_(i, jp, km) + U_(i, jm, kp));
#undef U_
        double s = c.fct1 * delta[split_index(gf, i, j, k)];
        s = s + Sx * (2.0 * Sx - Sy - Sz) * ONEOVERTHREE / dx4 + Sy * (2.0 * Sy - Sx - Sz) * ONEOVERTHREE / dx4 +
            Sz * (2.0 * Sz - Sx - Sy) * ONEOVERTHREE / dx4 + (Dxy * Dxy) * ONEOVEREIGHT / dx4 +
            (Dxz * Dxz) * ONEOVEREIGHT / dx4 + (Dyz * Dyz) * ONEOVEREIGHT / dx4;
        sig[q] = s;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) Q = Q - c.fct3_1 * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * sig[q]);
    const int x0 = 2 * I, x1 = 2 * I + 1, y0 = 2 * J, y1 = 2 * J + 1, z0 = 2 * K, z1 = 2 * K + 1;
    const int xl = (I == 0) ? N - 1 : 2 * I - 1, xu = (2 * I + 2 == N) ? 0 : 2 * I + 2;
    const int yl = (J == 0) ? N - 1 : 2 * J - 1, yu = (2 * J + 2 == N) ? 0 : 2 * J + 2;
    const int zl = (K == 0) ? N - 1 : 2 * K - 1, zu = (2 * K + 2 == N) ? 0 : 2 * K + 2;
    double P = 0.0;
#define V_(x, y, z) u0[split_index(gf, x, y, z)]
    P = P - 3.0 * V_(x0, y0, z0);
    P = P - 3.0 * V_(x1, y0, z0);
    P = P - 3.0 * V_(x0, y1, z0);
    P = P - 3.0 * V_(x1, y1, z0);
    P = P - 3.0 * V_(x0, y0, z1);
    P = P - 3.0 * V_(x1, y0, z1);
    P = P - 3.0 * V_(x0, y1, z1);
    P = P - 3.0 * V_(x1, y1, z1);
    P = P + V_(xl, y0, z0); P = P + V_(xl, y1, z0); P = P + V_(xl, y0, z1); P = P + V_(xl, y1, z1);
    P = P + V_(xu, y0, z0); P = P + V_(xu, y1, z0); P = P + V_(xu, y0, z1); P = P + V_(xu, y1, z1);
    P = P + V_(x0, yl, z0); P = P + V_(x1, yl, z0); P = P + V_(x0, yl, z1); P = P + V_(x1, yl, z1);
    P = P + V_(x0, yu, z0); P = P + V_(x1, yu, z0); P = P + V_(x0, yu, z1); P = P + V_(x1, yu, z1);
    P = P + V_(x0, y0, zl); P = P + V_(x1, y0, zl); P = P + V_(x0, y1, zl); P = P + V_(x1, y1, zl);
    P = P + V_(x0, y0, zu); P = P + V_(x1, y0, zu); P = P + V_(x0, y1, zu); P = P + V_(x1, y1, zu);
#undef V_
    P = P / dx2 + Q;
    rres1[split_index(gc, I, J, K)] = P / 8.0;
}

// calculate_physical_right_hand_side_DGP, PMP2MGsolver_DGP.f90:793-885
__global__ __launch_bounds__(256) void k_dgp_rhs64(const double* __restrict__ ru, const double* __restrict__ rho,
                                                   const double* __restrict__ rres, double* __restrict__ f, Geo g,
                                                   DgpCoef c, int TX, int TY, int XB, unsigned NB)
{
    int m;
    unsigned row;
    if (!map_xy(g.h, 2L * g.n * g.n, TX, TY, XB, NB, m, row)) return;
    const unsigned nn = (unsigned)g.n * (unsigned)g.n;
    const int col = row >= nn ? 1 : 0;
    const unsigned r2 = row - (unsigned)col * nn;
    const int j = (int)(r2 % (unsigned)g.n), k = (int)(r2 / (unsigned)g.n);
    const int i = 2 * m + ((col + j + k) & 1);
    const size_t me = (size_t)col * g.half + row_base(g, j, k) + m;
    double c0;
    double OP = dgp_L(ru, g, i, j, k, rho[me], c, c0);
    OP = OP - rres[me];
    f[me] = OP;
}

void launch_dgp_smooth64(int level, double* u, const double* rho, const double* f, double* tmp, int n, int colour,
                         const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(g.h, (long)n * n);
    if (level == 0)
        hipLaunchKernelGGL((k_dgp_smooth64<false>), L.grid, L.block, 0, s, u, rho, f, tmp, g, colour, c, L.TX, L.TY,
                           L.XB, L.NB);
    else
        hipLaunchKernelGGL((k_dgp_smooth64<true>), L.grid, L.block, 0, s, u, rho, f, tmp, g, colour, c, L.TX, L.TY,
                           L.XB, L.NB);
    launch_copy_f64(tmp, u + (size_t)colour * g.half, g.half, s);
}

int launch_dgp_residual64(int level, const double* u, const double* rho, const double* f, double* res,
                          double* partials, int n, const DgpCoef& c, double* scratch, hipStream_t s)
{
    (void)scratch;
    const Geo g = make_geo(n);
    const Launch L = make_launch(g.h, 2L * n * n);
    if (level == 0)
        hipLaunchKernelGGL((k_dgp_residual64<false>), L.grid, L.block, 0, s, u, rho, f, res, partials, g, c, L.TX,
                           L.TY, L.XB, L.NB);
    else
        hipLaunchKernelGGL((k_dgp_residual64<true>), L.grid, L.block, 0, s, u, rho, f, res, partials, g, c, L.TX,
                           L.TY, L.XB, L.NB);
    return (int)L.NB;
}

void launch_dgp_restrict_residual1_64(const double* u0, const double* delta, double* rres1, int N, const DgpCoef& c0,
                                      hipStream_t s)
{
    const Geo gf = make_geo(N), gc = make_geo(N / 2);
    const Launch L = make_launch(gc.n, (long)gc.n * gc.n);
    hipLaunchKernelGGL(k_dgp_restrict_residual1_64, L.grid, L.block, 0, s, u0, delta, rres1, gf, gc, c0, L.TX, L.TY,
                       L.XB, L.NB);
}

void launch_dgp_rhs64(const double* ru, const double* rho, const double* rres, double* f, int n, const DgpCoef& c,
                      hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(g.h, 2L * n * n);
    hipLaunchKernelGGL(k_dgp_rhs64, L.grid, L.block, 0, s, ru, rho, rres, f, g, c, L.TX, L.TY, L.XB, L.NB);
}

int dgp_alloc(mgs_handle* h)
{
    const size_t half = (size_t)h->p.ngrid * h->p.ngrid * (h->p.ngrid / 2);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, half * sizeof(double) + 256);
    if (e != hipSuccess) return set_err(h, MGS_E_NOMEM, "hipMalloc DGP scratch");
    h->allocs.push_back(q);
    h->dgp_tmp = (double*)q;
    return 0;
}

// ===========================================================================
// Mixed precision: FP32 delta-form FAS inside the FP64 defect correction
// ===========================================================================
// The inner unknown is the increment e of the field over an FP32 base
// phi_b (the current FP64 iterate rounded): D(e) = L(phi_b + e) - L(phi_b),
// written so that every term is proportional to e,
//   D(e) = lap(e)/dx^2 - 2 sgn(alpha) dSigma / (S(Sigma_b + dSigma) + S(Sigma_b)),
//   S(x) = sqrt(alpha^2 + 8/3 x),
//   dSigma = [2/3 (B.eps + Q(eps)) + 1/8 (2 C_b.eta + |eta|^2)] / dx^4,
// where eps = (e_xx, e_yy, e_zz) and eta = (e_xy, e_xz, e_yz) are e's second
// differences, Q(a) = a_x^2 + a_y^2 + a_z^2 - a_x a_y - a_x a_z - a_y a_z,
// B = grad Q at the base's (a_xx, a_yy, a_zz) and C_b the base's cross
// differences: Sigma(phi_b + e) - Sigma(phi_b) exactly, Sigma being quadratic.
// FP32 rounding then scales with the increment and the FP64 outer residual
// keeps converging.  The smoother is the reference's update
// e <- e - (D(e) - rhs) / (-6/dx^2), colour by colour, out of place.

// second differences, face sum and centre of a field at (i, j, k)
struct D2f {
    float xx, yy, zz, xy, xz, yz, s6, c0;
};

__device__ __forceinline__ D2f d2_at(const float* __restrict__ u, const Geo& g, int i, int j, int k)
{
    const int n = g.n;
    const int im = wrap_dec(i, n), ip = wrap_inc(i, n);
    const int jm = wrap_dec(j, n), jp = wrap_inc(j, n);
    const int km = wrap_dec(k, n), kp = wrap_inc(k, n);
    D2f d;
    d.c0 = u[split_index(g, i, j, k)];
    const float xp = u[split_index(g, ip, j, k)], xm = u[split_index(g, im, j, k)];
    const float yp = u[split_index(g, i, jp, k)], ym = u[split_index(g, i, jm, k)];
    const float zp = u[split_index(g, i, j, kp)], zm = u[split_index(g, i, j, km)];
    d.xx = xp + xm - 2.0f * d.c0;
    d.yy = yp + ym - 2.0f * d.c0;
    d.zz = zp + zm - 2.0f * d.c0;
    d.xy = u[split_index(g, ip, jp, k)] + u[split_index(g, im, jm, k)] - u[split_index(g, ip, jm, k)] -
           u[split_index(g, im, jp, k)];
    d.xz = u[split_index(g, ip, j, kp)] + u[split_index(g, im, j, km)] - u[split_index(g, ip, j, km)] -
           u[split_index(g, im, j, kp)];
    d.yz = u[split_index(g, i, jp, kp)] + u[split_index(g, i, jm, km)] - u[split_index(g, i, jp, km)] -
           u[split_index(g, i, jm, kp)];
    d.s6 = xp + xm + yp + ym + zp + zm;
    return d;
}

struct DgpF {
    float alpha, fct1, inv_dx2, inv_dx4, sgn;
};

static DgpF dgpf(const DgpCoef& c)
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

// one colour of the inner smoother, out of place into tmp (one colour half)
__global__ __launch_bounds__(256) void k_dgpm_smooth(const float* __restrict__ e, const float* __restrict__ pb,
                                                     const double* __restrict__ rho, const float* __restrict__ rhs,
                                                     float* __restrict__ tmp, Geo g, int colour, DgpF c, int TX,
                                                     int TY, int XB, unsigned NB)
{
    int m;
    unsigned row;
    if (!map_xy(g.h, (long)g.n * g.n, TX, TY, XB, NB, m, row)) return;
    const int j = (int)(row % (unsigned)g.n), k = (int)(row / (unsigned)g.n);
    const int i = 2 * m + ((colour + j + k) & 1);
    const size_t me = (size_t)colour * g.half + row_base(g, j, k) + m;
    const D2f de = d2_at(e, g, i, j, k), db = d2_at(pb, g, i, j, k);
    const float D = dgp_dop(db, de, (float)rho[me], c) - rhs[me];
    tmp[row_base(g, j, k) + m] = de.c0 + D / (6.0f * c.inv_dx2);
}

// out = D(e) - rhs on both colours (inner residual; with (e, rhs, out) =
// (R e, R r, f) the coarse FAS right-hand side)
__global__ __launch_bounds__(256) void k_dgpm_residual(const float* __restrict__ e, const float* __restrict__ pb,
                                                       const double* __restrict__ rho,
                                                       const float* __restrict__ rhs, float* __restrict__ out, Geo g,
                                                       DgpF c, int TX, int TY, int XB, unsigned NB)
{
    int m;
    unsigned row;
    if (!map_xy(g.h, 2L * g.n * g.n, TX, TY, XB, NB, m, row)) return;
    const unsigned nn = (unsigned)g.n * (unsigned)g.n;
    const int col = row >= nn ? 1 : 0;
    const unsigned r2 = row - (unsigned)col * nn;
    const int j = (int)(r2 % (unsigned)g.n), k = (int)(r2 / (unsigned)g.n);
    const int i = 2 * m + ((col + j + k) & 1);
    const size_t me = (size_t)col * g.half + row_base(g, j, k) + m;
    out[me] = dgp_dop(d2_at(pb, g, i, j, k), d2_at(e, g, i, j, k), (float)rho[me], c) - rhs[me];
}

// FP64 outer residual r = L(phi) (partials of r^2), and the inner problem's
// data: g = -r and phi_b = phi in FP32, e = 0
__global__ __launch_bounds__(256) void k_dgpm_outer(const double* __restrict__ u, const double* __restrict__ rho,
                                                    float* __restrict__ g32, float* __restrict__ pb,
                                                    float* __restrict__ e, double* __restrict__ partials, Geo g,
                                                    DgpCoef c, int TX, int TY, int XB, unsigned NB)
{
    int m;
    unsigned row;
    double acc = 0.0;
    if (map_xy(g.h, 2L * g.n * g.n, TX, TY, XB, NB, m, row)) {
        const unsigned nn = (unsigned)g.n * (unsigned)g.n;
        const int col = row >= nn ? 1 : 0;
        const unsigned r2 = row - (unsigned)col * nn;
        const int j = (int)(r2 % (unsigned)g.n), k = (int)(r2 / (unsigned)g.n);
        const int i = 2 * m + ((col + j + k) & 1);
        const size_t me = (size_t)col * g.half + row_base(g, j, k) + m;
        double c0;
        const double OP = dgp_L(u, g, i, j, k, rho[me], c, c0);
        g32[me] = (float)(-OP);
        pb[me] = (float)c0;
        e[me] = 0.0f;
        acc = OP * OP;
    }
    store_partial(partials, NB, block_sum(acc));
}

// coarsest level: sweeps of the inner smoother by one workgroup, the update
// evaluated in FP64 (the level holds at most a few thousand cells)
__global__ __launch_bounds__(1024) void k_dgpm_coarsest(float* __restrict__ e, const float* __restrict__ pb,
                                                        const double* __restrict__ rho,
                                                        const float* __restrict__ rhs, float* __restrict__ tmp, Geo g,
                                                        int sweeps, DgpF c)
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

// e_h += P(e_H - R e_h) (no guard for DGP)
__global__ __launch_bounds__(256) void k_dgpm_correct(const float* __restrict__ eH, const float* __restrict__ rd,
                                                      float* __restrict__ ef, Geo gf, Geo gc, int TX, int TY, int XB,
                                                      unsigned NB)
{
    int I;
    unsigned row;
    if (!map_xy(gc.n, (long)gc.n * gc.n, TX, TY, XB, NB, I, row)) return;
    const int J = (int)(row % (unsigned)gc.n), K = (int)(row / (unsigned)gc.n);
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

// phi += e
__global__ __launch_bounds__(256) void k_dgpm_update(double* __restrict__ u, const float* __restrict__ e, size_t count)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) u[i] = u[i] + (double)e[i];
}

int launch_dgpm_outer(const double* u, const double* rho, float* g32, float* pb, float* e, double* partials, int n,
                      const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(g.h, 2L * n * n);
    hipLaunchKernelGGL(k_dgpm_outer, L.grid, L.block, 0, s, u, rho, g32, pb, e, partials, g, c, L.TX, L.TY, L.XB,
                       L.NB);
    return (int)L.NB;
}

void launch_dgpm_smooth(float* e, const float* pb, const double* rho, const float* rhs, float* tmp, int n, int colour,
                        const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(g.h, (long)n * n);
    hipLaunchKernelGGL(k_dgpm_smooth, L.grid, L.block, 0, s, e, pb, rho, rhs, tmp, g, colour, dgpf(c), L.TX, L.TY,
                       L.XB, L.NB);
    launch_copy_f32(tmp, e + (size_t)colour * g.half, g.half, s);
}

void launch_dgpm_residual(const float* e, const float* pb, const double* rho, const float* rhs, float* out, int n,
                          const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(g.h, 2L * n * n);
    hipLaunchKernelGGL(k_dgpm_residual, L.grid, L.block, 0, s, e, pb, rho, rhs, out, g, dgpf(c), L.TX, L.TY, L.XB,
                       L.NB);
}

void launch_dgpm_coarsest(float* e, const float* pb, const double* rho, const float* rhs, float* tmp, int n,
                          int sweeps, const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    int threads = (int)g.half;
    if (threads > 1024) threads = 1024;
    threads = ((threads + 63) / 64) * 64;
    hipLaunchKernelGGL(k_dgpm_coarsest, dim3(1), dim3(threads), 0, s, e, pb, rho, rhs, tmp, g, sweeps, dgpf(c));
}

void launch_dgpm_correct(const float* eH, const float* rd, float* ef, int nc, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const Launch L = make_launch(nc, (long)nc * nc);
    hipLaunchKernelGGL(k_dgpm_correct, L.grid, L.block, 0, s, eH, rd, ef, gf, gc, L.TX, L.TY, L.XB, L.NB);
}

void launch_dgpm_update(double* u, const float* e, int n, hipStream_t s)
{
    const size_t count = (size_t)n * n * n;
    size_t b = (count + 255) / 256;
    if (b > 65536) b = 65536;
    hipLaunchKernelGGL(k_dgpm_update, dim3((unsigned)b), dim3(256), 0, s, u, e, count);
}

}  // namespace mgs
