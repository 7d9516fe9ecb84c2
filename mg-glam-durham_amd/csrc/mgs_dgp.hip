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
#include "mgs_dgp_dev.h"
#include "mgs_fr_dev.h"
#include "mgs_stencil.h"

namespace mgs {

DgpCoef dgp_coefs(const mgs_params& p, int level)
{
    DgpCoef c;
    const double a = p.aexpn;
    const double dx = (double)(1 << level);
    c.dx2 = dx * dx;
    c.dx4 = c.dx2 * c.dx2;
    c.inv_dx2 = 1.0 / c.dx2;
    c.inv_dx4 = 1.0 / c.dx4;
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

// z-marching launches (the windows and per-column bodies are in mgs_dgp_dev.h)
static ZLaunch dgp_zlaunch(int n) { return make_zlaunch(n / 2, n, DGP_KC); }

// both colours of a level (residual-type OP), one block partial of the sum of
// squares per block when partials is non-null
template <class OP>
__global__ __launch_bounds__(256) void k_dgp_both_z(OP op, double* __restrict__ partials, Geo g, int TX, int TY,
                                                    int XB, int JB, unsigned NB)
{
    ZCol cc;
    double acc = 0.0;
    if (zcol(g, TX, TY, XB, JB, DGP_KC, NB, cc)) acc = dgp_both_zcol(op, g, cc.m, cc.j, cc.k0);
    if (partials) store_partial(partials, NB, block_sum(acc));
}

template <class OP>
static unsigned launch_both_z(const OP& op, double* partials, int n, hipStream_t s)
{
    const Geo g = make_geo(n);
    const ZLaunch L = dgp_zlaunch(n);
    hipLaunchKernelGGL((k_dgp_both_z<OP>), L.grid, L.block, 0, s, op, partials, g, L.TX, L.TY, L.XB, L.JB, L.NB);
    return L.NB;
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

// restrict_residual_DGP(1), PMP2MGsolver_DGP.f90:459-763 (rearranged Sigma).
// A coarse cell (I, J, K) reads the 4x4x4 block of fine u with corner
// (2I-1, 2J-1, 2K-1): Sigma of each of its 8 children takes the 18
// neighbours around it, and the 8 children and their 24 outer face
// neighbours make P.  A thread owns coarse column (I, J) and marches over a
// run of K, holding one 4x4 in-plane patch per fine plane: the patches of
// fine planes 2K+1 and 2K+2 are planes 2K'-1 and 2K' of K' = K+1, so each
// step loads two fine planes (32 values; the x-neighbour columns hit the
// cache lines of the adjacent lanes) and the children's densities.  Every
// sum keeps the reference's operation order.
struct DPatch {
    double v[4][4];  // [y][x]: fine x = 2I-1+xx, y = 2J-1+yy (periodic)
};

__device__ __forceinline__ double dgp_sigma1(const DPatch& A, const DPatch& B, const DPatch& C, int yy, int xx,
                                             double dens, const DgpCoef& c)
{
    const double ONEOVEREIGHT = 1.0 / 8.0, ONEOVERTHREE = 1.0 / 3.0;
    const double dx4 = 1.0;
    const double Sx = B.v[yy][xx + 1] + B.v[yy][xx - 1];
    const double Sy = B.v[yy + 1][xx] + B.v[yy - 1][xx];
    const double Sz = C.v[yy][xx] + A.v[yy][xx];
    const double Dxy = (B.v[yy + 1][xx + 1] + B.v[yy - 1][xx - 1]) - (B.v[yy - 1][xx + 1] + B.v[yy + 1][xx - 1]);
    const double Dxz = (C.v[yy][xx + 1] + A.v[yy][xx - 1]) - (A.v[yy][xx + 1] + C.v[yy][xx - 1]);
    const double Dyz = (C.v[yy + 1][xx] + A.v[yy - 1][xx]) - (A.v[yy + 1][xx] + C.v[yy - 1][xx]);
    double s = c.fct1 * dens;
    s = s + Sx * (2.0 * Sx - Sy - Sz) * ONEOVERTHREE / dx4 + Sy * (2.0 * Sy - Sx - Sz) * ONEOVERTHREE / dx4 +
        Sz * (2.0 * Sz - Sx - Sy) * ONEOVERTHREE / dx4 + (Dxy * Dxy) * ONEOVEREIGHT / dx4 +
        (Dxz * Dxz) * ONEOVEREIGHT / dx4 + (Dyz * Dyz) * ONEOVEREIGHT / dx4;
    return s;
}

__global__ __launch_bounds__(256) void k_dgp_restrict_residual1_64(const double* __restrict__ u0,
                                                                   const double* __restrict__ delta,
                                                                   double* __restrict__ rres1, Geo gf, Geo gc,
                                                                   DgpCoef c, int TX, int TY, int XB, int JB, int KC,
                                                                   unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    const int I = cc.m, J = cc.j, N = gf.n;
    const unsigned plane = (unsigned)gf.n * (unsigned)gf.h, half = (unsigned)gf.half;
    // fine columns and rows of the patch: m = x/2 with its parity, row j*h
    const int xs[4] = {I == 0 ? N - 1 : 2 * I - 1, 2 * I, 2 * I + 1, 2 * I + 2 == N ? 0 : 2 * I + 2};
    const int ys[4] = {J == 0 ? N - 1 : 2 * J - 1, 2 * J, 2 * J + 1, 2 * J + 2 == N ? 0 : 2 * J + 2};
    auto at = [&](int yy, int xx, int z) -> unsigned {
        const unsigned colour = (unsigned)((xs[xx] + ys[yy] + z) & 1);
        return colour * half + (unsigned)z * plane + (unsigned)ys[yy] * (unsigned)gf.h + (unsigned)(xs[xx] >> 1);
    };
    auto load = [&](int z, DPatch& P) {
#pragma unroll
        for (int yy = 0; yy < 4; yy++)
#pragma unroll
            for (int xx = 0; xx < 4; xx++) P.v[yy][xx] = u0[at(yy, xx, z)];
    };
    const double EIGHTOVERTHREE = 8.0 / 3.0;
    const double dx2 = 1.0;
    // B: fine plane 2K, C: 2K+1; Sigma of B's children before plane 2K+2 is
    // loaded, so of plane 2K-1 only the 4 centre values (z-1 faces) stay live
    DPatch A, B, C;
    load(cc.k0 == 0 ? N - 1 : 2 * cc.k0 - 1, A);
    load(2 * cc.k0, B);
    for (int K = cc.k0; K < cc.k1; K++) {
        const int z0 = 2 * K, z1 = 2 * K + 1, zu = 2 * K + 2 == N ? 0 : 2 * K + 2;
        load(z1, C);
        double sig[8];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int yy = 1 + ((q >> 1) & 1), xx = 1 + (q & 1);
            sig[q] = dgp_sigma1(A, B, C, yy, xx, delta[at(yy, xx, z0)], c);
        }
        const double zl0 = A.v[1][1], zl1 = A.v[1][2], zl2 = A.v[2][1], zl3 = A.v[2][2];
        load(zu, A);  // A now holds fine plane 2K+2
#pragma unroll
        for (int q = 4; q < 8; q++) {
            const int yy = 1 + ((q >> 1) & 1), xx = 1 + (q & 1);
            sig[q] = dgp_sigma1(B, C, A, yy, xx, delta[at(yy, xx, z1)], c);
        }
        double Q = -8.0 * c.fct2_1;
#pragma unroll
        for (int q = 0; q < 8; q++) Q = Q - c.fct3_1 * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * sig[q]);
        double P = 0.0;
        P = P - 3.0 * B.v[1][1];
        P = P - 3.0 * B.v[1][2];
        P = P - 3.0 * B.v[2][1];
        P = P - 3.0 * B.v[2][2];
        P = P - 3.0 * C.v[1][1];
        P = P - 3.0 * C.v[1][2];
        P = P - 3.0 * C.v[2][1];
        P = P - 3.0 * C.v[2][2];
        P = P + B.v[1][0]; P = P + B.v[2][0]; P = P + C.v[1][0]; P = P + C.v[2][0];  // x-1 faces
        P = P + B.v[1][3]; P = P + B.v[2][3]; P = P + C.v[1][3]; P = P + C.v[2][3];  // x+2 faces
        P = P + B.v[0][1]; P = P + B.v[0][2]; P = P + C.v[0][1]; P = P + C.v[0][2];  // y-1 faces
        P = P + B.v[3][1]; P = P + B.v[3][2]; P = P + C.v[3][1]; P = P + C.v[3][2];  // y+2 faces
        P = P + zl0; P = P + zl1; P = P + zl2; P = P + zl3;                          // z-1 faces
        P = P + A.v[1][1]; P = P + A.v[1][2]; P = P + A.v[2][1]; P = P + A.v[2][2];  // z+2 faces
        P = P / dx2 + Q;
        rres1[split_index(gc, I, J, K)] = P / 8.0;
        // planes 2K+1 and 2K+2 are 2K'-1 and 2K' of the next coarse plane
        const DPatch t = A;
        A = C;
        B = t;
    }
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
    if (dgp_zmarch(n)) {
        const unsigned hf = (unsigned)g.half;
        if (level == 0) {
            if (n >= NT_MIN_N) return (int)launch_both_z(DgpRes64Op<false, true>{u, rho, f, res, c, hf}, partials, n, s);
            return (int)launch_both_z(DgpRes64Op<false>{u, rho, f, res, c, hf}, partials, n, s);
        }
        if (n >= NT_MIN_N) return (int)launch_both_z(DgpRes64Op<true, true>{u, rho, f, res, c, hf}, partials, n, s);
        return (int)launch_both_z(DgpRes64Op<true>{u, rho, f, res, c, hf}, partials, n, s);
    }
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
    const ZLaunch L = make_zlaunch(gc.n, gc.n);
    hipLaunchKernelGGL(k_dgp_restrict_residual1_64, L.grid, L.block, 0, s, u0, delta, rres1, gf, gc, c0, L.TX, L.TY,
                       L.XB, L.JB, L.KC, L.NB);
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
    int rc = dev_alloc(h, half * sizeof(double) + 256, &q, "DGP scratch");
    if (rc) return rc;
    h->dgp_tmp = (double*)q;
    // FP64 z-marching passes: the second array of the field's ping-pong
    if (h->p.precision == MGS_PREC_FP64 && dgp_zmarch(h->p.ngrid)) {
        rc = dev_alloc(h, 2 * half * sizeof(double) + 256, &q, "DGP second array");
        if (rc) return rc;
        h->dgp_spare64 = (double*)q;
    }
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

// one colour of the inner smoother, out of place into tmp (one colour half)
__global__ __launch_bounds__(256) void k_dgpm_smooth(const float* __restrict__ e, const float* __restrict__ pb,
                                                     const double* __restrict__ rho, const float* __restrict__ rhs,
                                                     float* __restrict__ tmp, Geo g, int colour, DgpF c, int TX,
                                                     int TY, int XB, unsigned NB)
{
    int m;
    unsigned row;
    if (!map_xy(g.h, (long)g.n * g.n, TX, TY, XB, NB, m, row)) return;
    dgpm_smooth_cell(e, pb, rho, rhs, tmp, g, colour, c, m, row);
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
    dgpm_residual_cell(e, pb, rho, rhs, out, g, c, m, row);
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
    if (map_xy(g.h, 2L * g.n * g.n, TX, TY, XB, NB, m, row)) acc = dgpm_outer_cell(u, rho, g32, pb, e, g, c, m, row);
    store_partial(partials, NB, block_sum(acc));
}

__global__ __launch_bounds__(1024) void k_dgpm_coarsest(float* __restrict__ e, const float* __restrict__ pb,
                                                        const double* __restrict__ rho,
                                                        const float* __restrict__ rhs, float* __restrict__ tmp, Geo g,
                                                        int sweeps, DgpF c)
{
    dgpm_coarsest_block(e, pb, rho, rhs, tmp, g, sweeps, c);
}

// The same correction z-marching over coarse columns, as k_df_correct: planes
// of e = e_H - R e_h in registers (x-neighbours from the adjacent lanes), all
// 8 sons of a coarse cell, software-pipelined (plane K's sons computed, the
// loads of plane K+1 issued, then plane K's sons stored).  Sums in
// dgpm_correct_cell's order.
template <int S3, bool NT>
__device__ __forceinline__ void dgpm_sons_load(const float* __restrict__ ef, const Geo& gf, int I, int J, int K,
                                               float* old)
{
#pragma unroll
    for (int t = 0; t < 4; t++) old[t] = ld_once<NT>(elem(ef, son_index(gf, I, J, K, t & 1, t >> 1, S3)));
}

template <int S3>
__device__ __forceinline__ void dgpm_sons_new(const float* ez, const float* ek, const float* old, float* nv)
{
    const float a = 1.0f / 64.0f, b = 3.0f * a, c9 = 9.0f * a, dd = 27.0f * a;
    const float wt[8] = {a, b, b, c9, b, c9, c9, dd};
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int s1 = t & 1, s2 = t >> 1;
        float P = 0.0f;
#pragma unroll
        for (int ind = 0; ind < 8; ind++) {
            const int x = (ind & 1) ? 1 : (s1 ? 2 : 0);
            const int y = (ind & 2) ? 1 : (s2 ? 2 : 0);
            P += wt[ind] * ((ind & 4) ? ek[x + 3 * y] : ez[x + 3 * y]);
        }
        nv[t] = old[t] + P;
    }
}

template <int S3, bool NT>
__device__ __forceinline__ void dgpm_sons_store(float* __restrict__ ef, const Geo& gf, int I, int J, int K,
                                                const float* nv)
{
#pragma unroll
    for (int t = 0; t < 4; t++) st_once<NT>(elem(ef, son_index(gf, I, J, K, t & 1, t >> 1, S3)), nv[t]);
}

template <bool NT>
__global__ __launch_bounds__(256) void k_dgpm_zcorrect(const float* __restrict__ eH, const float* __restrict__ rd,
                                                       float* __restrict__ ef, Geo gf, Geo gc, int TX, int TY, int XB,
                                                       int JB, int KC, unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    // lanes t-1 / t+1 hold columns I-1 / I+1 of the same row and plane run
    // when they are in the same block row, in this wave and inside the level
    const unsigned t = threadIdx.y * (unsigned)TX + threadIdx.x;
    const bool lok = threadIdx.x > 0 && (t & 63u) != 0;
    const bool rok = (int)threadIdx.x + 1 < TX && cc.m + 1 < gc.n && ((t + 1) & 63u) != 0;
    const int I = cc.m, J = cc.j;
    Block9 B;
    B.init(gc, I, J);
    ERaw wm, wk, wp;  // coarse planes K-1, K, K+1
    float o0[4], o1[4];
    e_load<true>(eH, rd, B, wrap_dec(cc.k0, gc.n), lok, rok, wm);
    e_load<true>(eH, rd, B, cc.k0, lok, rok, wk);
    e_load<true>(eH, rd, B, wrap_inc(cc.k0, gc.n), lok, rok, wp);
    dgpm_sons_load<0, NT>(ef, gf, I, J, cc.k0, o0);
    dgpm_sons_load<1, NT>(ef, gf, I, J, cc.k0, o1);
    float em[9], ek[9], ep[9];
    e_form<true>(wm, lok, rok, em);
    e_form<true>(wk, lok, rok, ek);
    for (int K = cc.k0; K < cc.k1; K++) {
        e_form<true>(wp, lok, rok, ep);
        float n0[4], n1[4];
        dgpm_sons_new<0>(em, ek, o0, n0);
        dgpm_sons_new<1>(ep, ek, o1, n1);
        if (K + 1 < cc.k1) {
            dgpm_sons_load<0, NT>(ef, gf, I, J, K + 1, o0);
            dgpm_sons_load<1, NT>(ef, gf, I, J, K + 1, o1);
            e_load<true>(eH, rd, B, wrap_inc(K + 1, gc.n), lok, rok, wp);
        }
        dgpm_sons_store<0, NT>(ef, gf, I, J, K, n0);
        dgpm_sons_store<1, NT>(ef, gf, I, J, K, n1);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            em[q] = ek[q];
            ek[q] = ep[q];
        }
    }
}

// phi += e
template <bool NT>
__global__ __launch_bounds__(256) void k_dgpm_update(double* __restrict__ u, const float* __restrict__ e, size_t count)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        st_once<NT>(u[i], ld_once<NT>(u[i]) + (double)ld_once<NT>(e[i]));
}

int launch_dgpm_outer(const double* u, const double* rho, float* g32, float* pb, float* e, double* partials, int n,
                      const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    if (dgp_zmarch(n)) {
        if (n >= NT_MIN_N)
            return (int)launch_both_z(DgpmOuterOp<true>{u, rho, g32, pb, e, c, (unsigned)g.half}, partials, n, s);
        return (int)launch_both_z(DgpmOuterOp<>{u, rho, g32, pb, e, c, (unsigned)g.half}, partials, n, s);
    }
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

void launch_dgpm_residual(const float* e, const float* pb, const double* rho, const float* rho32, const float* rhs,
                          float* out, int n, const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    if (dgp_zmarch(n) && rho32) {
        if (n >= NT_MIN_N)
            launch_both_z(DgpmResOp<true>{e, pb, rho32, rhs, out, dgpf(c), (unsigned)g.half}, nullptr, n, s);
        else
            launch_both_z(DgpmResOp<>{e, pb, rho32, rhs, out, dgpf(c), (unsigned)g.half}, nullptr, n, s);
        return;
    }
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
    const ZLaunch L = make_zlaunch(nc, nc);
    if (2 * nc >= NT_MIN_N)
        hipLaunchKernelGGL(k_dgpm_zcorrect<true>, L.grid, L.block, 0, s, eH, rd, ef, gf, gc, L.TX, L.TY, L.XB, L.JB, L.KC,
                           L.NB);
    else
        hipLaunchKernelGGL(k_dgpm_zcorrect<false>, L.grid, L.block, 0, s, eH, rd, ef, gf, gc, L.TX, L.TY, L.XB, L.JB,
                           L.KC, L.NB);
}

void launch_dgpm_update(double* u, const float* e, int n, hipStream_t s)
{
    const size_t count = (size_t)n * n * n;
    size_t b = (count + 255) / 256;
    if (b > 65536) b = 65536;
    if (n >= NT_MIN_N)
        hipLaunchKernelGGL(k_dgpm_update<true>, dim3((unsigned)b), dim3(256), 0, s, u, e, count);
    else
        hipLaunchKernelGGL(k_dgpm_update<false>, dim3((unsigned)b), dim3(256), 0, s, u, e, count);
}

// ===========================================================================
// z-marching colour passes (the windows are in mgs_dgp_dev.h)
// ===========================================================================

// relaxation_iterations_DGP (as k_dgp_smooth64), one colour.  S0: parity of
// the column's first cell (the same for a whole wave); with the run unrolled
// the parity of every step is a constant, so the stencil values are the
// window's registers themselves.
template <bool COARSE, bool S0, bool NT>
__device__ __forceinline__ void dgp_zsmooth64_run(const double* __restrict__ h0, const double* __restrict__ h1,
                                                  double* __restrict__ out, const double* __restrict__ rho,
                                                  const double* __restrict__ f, const Geo& g, int colour,
                                                  const DgpCoef& c, const DCol& d, int k0)
{
    // (j + k0) & 1 = S0 ^ colour: wave-uniform
    const DHalves<double> H = dhalves(h0, h1, (S0 ? 1 : 0) ^ (colour & 1));
    const double* rhoc = rho + (size_t)colour * g.half;
    const double* fc = COARSE ? f + (size_t)colour * g.half : nullptr;
    DWin2<double> w;
    dwin2_init(w, H, d, k0, g.n);
    DSide<double> sd;
    dside_load(H, d, k0, 0, g.n, S0, sd);
    const double dLphi = -6.0 / c.dx2;
    unsigned off = (unsigned)k0 * d.plane + d.rowoff[1] + d.m;
    double rh = ld_once<NT>(elem(rhoc, off));
    double fv = COARSE ? ld_once<NT>(elem(fc, off)) : 0.0;
#pragma unroll
    for (int t = 0; t < DGP_KC; t++) {
        const int k = k0 + t;
        const bool s = S0 ^ ((t & 1) != 0);
        DRaw2<double> r;
        DSide<double> sdn;
        double rhn = 0.0, fvn = 0.0;
        const unsigned offn = off + d.plane;
        if (t + 1 < DGP_KC) {
            draw2_load(H, d, wrap_inc(k + 1, g.n), t + 2, r);
            dside_load(H, d, wrap_inc(k, g.n), t + 1, g.n, !s, sdn);
            rhn = ld_once<NT>(elem(rhoc, offn));
            if constexpr (COARSE) fvn = ld_once<NT>(elem(fc, offn));
        }
        const S19<double> u = s19_win2(w, s, sd, d);
        double L = dgp_L19(u, rh, c);
        if constexpr (COARSE) L = L - fv;
        st_once<NT>(elem(out, off), u.c0 - L / dLphi);
        if (t + 1 < DGP_KC) {
            dwin2_shift(w);
            dwin2_set(w, 2, r);
            sd = sdn;
            off = offn;
            rh = rhn;
            fv = fvn;
        }
    }
}

template <bool COARSE, bool NT>
__global__ __launch_bounds__(256) void k_dgp_zsmooth64(const double* __restrict__ h0, const double* __restrict__ h1,
                                                       double* __restrict__ out, const double* __restrict__ rho,
                                                       const double* __restrict__ f, Geo g, int colour, DgpCoef c,
                                                       int TX, int TY, int XB, int JB, unsigned NB)
{
    ZCol cc;
    if (!zcol(g, TX, TY, XB, JB, DGP_KC, NB, cc)) return;
    const DCol d = dcol_of(g, cc.m, cc.j);
    if ((colour + cc.j + cc.k0) & 1)
        dgp_zsmooth64_run<COARSE, true, NT>(h0, h1, out, rho, f, g, colour, c, d, cc.k0);
    else
        dgp_zsmooth64_run<COARSE, false, NT>(h0, h1, out, rho, f, g, colour, c, d, cc.k0);
}

// one colour of the mixed inner smoother (as k_dgpm_smooth); S0 as above
template <bool S0, bool NT>
__device__ __forceinline__ void dgpm_zsmooth_run(const float* __restrict__ e0, const float* __restrict__ e1,
                                                 float* __restrict__ out, const float* __restrict__ pb,
                                                 const float* __restrict__ rho32, const float* __restrict__ rhs,
                                                 const Geo& g, int colour, const DgpF& c, const DCol& d, int k0)
{
    // (j + k0) & 1 = S0 ^ colour: wave-uniform
    const int par0 = (S0 ? 1 : 0) ^ (colour & 1);
    const DHalves<float> He = dhalves(e0, e1, par0), Hb = dhalves(pb, pb + g.half, par0);
    const float* rhoc = rho32 + (size_t)colour * g.half;
    const float* rhsc = rhs + (size_t)colour * g.half;
    DWin2<float> we, wb;
    dwin2_init(we, He, d, k0, g.n);
    dwin2_init(wb, Hb, d, k0, g.n);
    DSide<float> se, sb;
    dside_load(He, d, k0, 0, g.n, S0, se);
    dside_load(Hb, d, k0, 0, g.n, S0, sb);
    unsigned off = (unsigned)k0 * d.plane + d.rowoff[1] + d.m;
    float rh = ld_once<NT>(elem(rhoc, off));
    float fv = ld_once<NT>(elem(rhsc, off));
#pragma unroll
    for (int t = 0; t < DGP_KC; t++) {
        const int k = k0 + t;
        const bool s = S0 ^ ((t & 1) != 0);
        DRaw2<float> re, rb;
        DSide<float> sen, sbn;
        float rhn = 0.0f, fvn = 0.0f;
        const unsigned offn = off + d.plane;
        if (t + 1 < DGP_KC) {
            const int kn = wrap_inc(k + 1, g.n);
            draw2_load(He, d, kn, t + 2, re);
            draw2_load(Hb, d, kn, t + 2, rb);
            dside_load(He, d, wrap_inc(k, g.n), t + 1, g.n, !s, sen);
            dside_load(Hb, d, wrap_inc(k, g.n), t + 1, g.n, !s, sbn);
            rhn = ld_once<NT>(elem(rhoc, offn));
            fvn = ld_once<NT>(elem(rhsc, offn));
        }
        const D2f de = d2_of(s19_win2(we, s, se, d)), db = d2_of(s19_win2(wb, s, sb, d));
        const float D = dgp_dop(db, de, rh, c) - fv;
        st_once<NT>(elem(out, off), de.c0 + D / (6.0f * c.inv_dx2));
        if (t + 1 < DGP_KC) {
            dwin2_shift(we);
            dwin2_shift(wb);
            dwin2_set(we, 2, re);
            dwin2_set(wb, 2, rb);
            se = sen;
            sb = sbn;
            off = offn;
            rh = rhn;
            fv = fvn;
        }
    }
}

template <bool NT>
__global__ __launch_bounds__(256) void k_dgpm_zsmooth(const float* __restrict__ e0, const float* __restrict__ e1,
                                                      float* __restrict__ out, const float* __restrict__ pb,
                                                      const float* __restrict__ rho32, const float* __restrict__ rhs,
                                                      Geo g, int colour, DgpF c, int TX, int TY, int XB, int JB,
                                                      unsigned NB)
{
    ZCol cc;
    if (!zcol(g, TX, TY, XB, JB, DGP_KC, NB, cc)) return;
    const DCol d = dcol_of(g, cc.m, cc.j);
    if ((colour + cc.j + cc.k0) & 1)
        dgpm_zsmooth_run<true, NT>(e0, e1, out, pb, rho32, rhs, g, colour, c, d, cc.k0);
    else
        dgpm_zsmooth_run<false, NT>(e0, e1, out, pb, rho32, rhs, g, colour, c, d, cc.k0);
}

void launch_dgp_zpass64(int level, const double* h0, const double* h1, double* out, const double* rho,
                        const double* f, int n, int colour, const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const ZLaunch L = dgp_zlaunch(n);
#define ZS64(CO, NT)                                                                                    \
    hipLaunchKernelGGL((k_dgp_zsmooth64<CO, NT>), L.grid, L.block, 0, s, h0, h1, out, rho, f, g, colour, c, L.TX, \
                       L.TY, L.XB, L.JB, L.NB)
    const bool nt = n >= NT_MIN_N;
    if (level == 0) {
        if (nt) ZS64(false, true);
        else ZS64(false, false);
    } else {
        if (nt) ZS64(true, true);
        else ZS64(true, false);
    }
#undef ZS64
}

void launch_dgpm_zpass(const float* e0, const float* e1, float* out, const float* pb, const float* rho32,
                       const float* rhs, int n, int colour, const DgpCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const ZLaunch L = dgp_zlaunch(n);
    if (n >= NT_MIN_N)
        hipLaunchKernelGGL(k_dgpm_zsmooth<true>, L.grid, L.block, 0, s, e0, e1, out, pb, rho32, rhs, g, colour, dgpf(c),
                           L.TX, L.TY, L.XB, L.JB, L.NB);
    else
        hipLaunchKernelGGL(k_dgpm_zsmooth<false>, L.grid, L.block, 0, s, e0, e1, out, pb, rho32, rhs, g, colour,
                           dgpf(c), L.TX, L.TY, L.XB, L.JB, L.NB);
}

__global__ __launch_bounds__(256) void k_round_f32(const double* __restrict__ a, float* __restrict__ b, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = (float)a[i];
}

void launch_round_f32(const double* a, float* b, size_t n, hipStream_t s)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_round_f32, dim3((unsigned)blocks), dim3(256), 0, s, a, b, n);
}

}  // namespace mgs
