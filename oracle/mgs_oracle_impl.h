/*
 * Oracle kernels, generated twice by mgs_oracle.c with
 *   STOR = float  / SFX = f32   (reference-faithful FP32 storage)
 *   STOR = double / SFX = f64   (FP64 oracle)
 *
 * TEST INFRASTRUCTURE ONLY (see mgs_oracle.h).
 *
 * Expressions keep the reference's operand order and its typing: a
 * sub-expression whose operands are all STOR values is evaluated in STOR (as
 * Fortran evaluates real*4 op real*4 in real*4), and is promoted to double
 * where the reference promotes it.  With STOR = double everything is FP64.
 */

#define CAT2_(a, b) a##_##b
#define CAT2(a, b) CAT2_(a, b)
#define FN(name) CAT2(name, SFX)

/* w = u^(n+1) of one stored value, evaluated in STOR like the reference's
 * FI2*FI2 (n=1) and FI2*FI2*FI2 (n=2). */
static inline STOR FN(wpow)(STOR x, int n)
{
    if (n == 0) return x;
    if (n == 1) return x * x;
    return x * x * x;
}

/* Newton's method on f(u) = u^{n+2} + P u + Q from the closed-form root r,
 * to convergence (|step| <= 1e-15 u, at most 8 steps).  The closed forms
 * cancel where u << |P|^{1/(n+1)} (dense, screened cells: Cardano's
 * cbrt(t1) - cbrt(t2) with t1 ~ t2, the quadratic's -P + sqrt(...) with
 * P > 0, Ferrari's -S + sqrt(...)/2); f is evaluated with an absolute error
 * ~ eps |Q| there, so the iteration restores full relative precision. */
static inline double FN(polish_root)(double r, double P, double Q, int n)
{
    for (int it = 0; it < 8; it++) {
        const double rn1 = n == 0 ? r : (n == 1 ? r * r : r * r * r);
        const double f = rn1 * r + P * r + Q;
        const double fp = (double)(n + 2) * rn1 + P;
        if (!(fp > 0.0)) return r;
        const double d = f / fp;
        const double r1 = r - d;
        if (!(r1 > 0.0)) return r;
        r = r1;
        if (fabs(d) <= 1.0e-15 * r) break;
    }
    return r;
}

/* ---------------------------------------------------------------------- */
/* f(R) red-black smoother, relaxation_iterations_fR                       */
/* PMP2MGsolver_fR.f90:20-301                                              */
/* ---------------------------------------------------------------------- */
void FN(orc_fr_smooth)(const orc_params* p, int level, int red, STOR* u, const STOR* rho,
                       const STOR* f)
{
    const int n = p->N >> level;
    const int frn = p->fr_n;
    orc_fr_coef c;
    orc_fr_coefs(p, level, &c);
    const double ZERO = 1.0e-20;
    const double ONEOVERTHREE = 1.0 / 3.0;
    const double TWOOVERTHREE = 2.0 / 3.0;
    const double TWOPIOVERTHREE = 2.0 / 3.0 * acos(-1.0);
    /* smoother constants, PMP2MGsolver_fR.f90:59-71 */
    const double fct2 = c.dx2 / (c.ctilde * c.ctilde * c.sgn) / 6.0;
    const double fct1 = -(c.om_a) * fct2;
    const double Q2 = c.R_bg * c.a2 * fct2 / 3.0;
    const double Q = Q2 * pow(fabs(c.fR_bg), 1.0 / ((double)frn + 1.0));
    const double Q2onethird = pow(-Q, ONEOVERTHREE);
    const double Q4 = 4.0 * Q;
    const double Delta11 = 27.0 * Q;
    const double Delta20 = 12.0 * Q;

#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < n; k++) {
        for (int j = 0; j < n; j++) {
            const int km = (k == 0) ? n - 1 : k - 1, kp = (k == n - 1) ? 0 : k + 1;
            const int jm = (j == 0) ? n - 1 : j - 1, jp = (j == n - 1) ? 0 : j + 1;
            /* red = reference M1+M2+M3 odd = i+j+k even */
            const int i0 = red ? ((j + k) & 1) : (1 - ((j + k) & 1));
            for (int i = i0; i < n; i += 2) {
                const int im = (i == 0) ? n - 1 : i - 1, ip = (i == n - 1) ? 0 : i + 1;
                const size_t cc = ORC_IDX(i, j, k, n);
                /* neighbour order M1u, M1l, M2u, M2l, M3u, M3l */
                const STOR s = FN(wpow)(u[ORC_IDX(ip, j, k, n)], frn) + FN(wpow)(u[ORC_IDX(im, j, k, n)], frn) +
                               FN(wpow)(u[ORC_IDX(i, jp, k, n)], frn) + FN(wpow)(u[ORC_IDX(i, jm, k, n)], frn) +
                               FN(wpow)(u[ORC_IDX(i, j, kp, n)], frn) + FN(wpow)(u[ORC_IDX(i, j, km, n)], frn);
                double P;
                if (level == 0)
                    P = fct1 * (double)rho[cc] - Q2 - (double)s / 6.0;
                else
                    P = fct1 * (double)rho[cc] - Q2 + fct2 * (double)f[cc] - (double)s / 6.0;
                double r;
                if (frn == 0) {
                    r = 0.5 * (-P + sqrt(P * P - Q4));
                } else if (frn == 1) {
                    const double Delta10 = -3.0 * P;
                    const double CC = Delta11 * Delta11 - 4.0 * Delta10 * Delta10 * Delta10;
                    if (CC >= 0.0) {
                        const double tmp1 = 0.5 * (sqrt(CC) + Delta11);
                        const double tmp2 = tmp1 - Delta11;
                        if (fabs(P) < ZERO)
                            r = Q2onethird;
                        else
                            r = -(copysign(pow(fabs(tmp1), ONEOVERTHREE), tmp1) -
                                  copysign(pow(fabs(tmp2), ONEOVERTHREE), tmp2)) *
                                ONEOVERTHREE;
                    } else {
                        const double sd = sqrt(Delta10);
                        const double S = acos(0.5 * Delta11 / (sd * sd * sd));
                        r = -TWOOVERTHREE * sd * cos(S * ONEOVERTHREE + TWOPIOVERTHREE);
                    }
                } else {
                    const double Delta21 = 27.0 * P * P;
                    const double tmp1 = 0.5 * (sqrt(Delta21 * Delta21 - 4.0 * Delta20 * Delta20 * Delta20) + Delta21);
                    const double cb = copysign(pow(fabs(tmp1), ONEOVERTHREE), tmp1);
                    const double S = sqrt(cb + Delta20 / cb) / sqrt(12.0);
                    if (P > ZERO)
                        r = -S + 0.5 * sqrt(-4.0 * S * S + P / S);
                    else if (P < -ZERO)
                        r = S + 0.5 * sqrt(-4.0 * S * S - P / S);
                    else
                        r = sqrt(sqrt(-Q));
                }
                if (p->polish) r = FN(polish_root)(r, P, Q, frn);
                u[cc] = (STOR)r;
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* f(R) residual, calculate_residual_fR, PMP2MGsolver_fR.f90:308-561        */
/* level 0: r = N(u) with delta; level>0: r = N_l(u) - f.  res may be NULL  */
/* (norm only).  Returns sqrt(sum r^2 / n^3) (the reference reports it at   */
/* level 0 only).                                                           */
/* ---------------------------------------------------------------------- */
double FN(orc_fr_residual)(const orc_params* p, int level, const STOR* u, const STOR* rho,
                           const STOR* f, STOR* res)
{
    const int n = p->N >> level;
    const int frn = p->fr_n;
    orc_fr_coef c;
    orc_fr_coefs(p, level, &c);
    const double fct1 = c.om_a;
    const double fct2 = -c.R_bg * c.a2 / 3.0;
    const double fct3 = c.ctilde * c.ctilde * c.sgn / c.dx2;
    const double a_fR_rt = pow(fabs(c.fR_bg), 1.0 / (double)(frn + 1));
    double* part = (double*)calloc((size_t)n * n, sizeof(double));

#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < n; k++) {
        for (int j = 0; j < n; j++) {
            const int km = (k == 0) ? n - 1 : k - 1, kp = (k == n - 1) ? 0 : k + 1;
            const int jm = (j == 0) ? n - 1 : j - 1, jp = (j == n - 1) ? 0 : j + 1;
            double acc = 0.0;
            for (int i = 0; i < n; i++) {
                const int im = (i == 0) ? n - 1 : i - 1, ip = (i == n - 1) ? 0 : i + 1;
                const size_t cc = ORC_IDX(i, j, k, n);
                const STOR s = FN(wpow)(u[ORC_IDX(ip, j, k, n)], frn) + FN(wpow)(u[ORC_IDX(im, j, k, n)], frn) +
                               FN(wpow)(u[ORC_IDX(i, jp, k, n)], frn) + FN(wpow)(u[ORC_IDX(i, jm, k, n)], frn) +
                               FN(wpow)(u[ORC_IDX(i, j, kp, n)], frn) + FN(wpow)(u[ORC_IDX(i, j, km, n)], frn);
                double OP = (double)s - (double)FN(wpow)(u[cc], frn) * 6.0;
                OP = OP * fct3;
                OP = OP + fct1 * (double)rho[cc];
                OP = OP + fct2 * (a_fR_rt / (double)u[cc] - 1.0);
                if (level > 0) OP = OP - (double)f[cc];
                if (res) res[cc] = (STOR)OP;
                acc += OP * OP;
            }
            part[(size_t)k * n + j] = acc;
        }
    }
    double RES = 0.0;
    for (size_t q = 0; q < (size_t)n * n; q++) RES += part[q];
    free(part);
    return sqrt(RES / ((double)n * (double)n * (double)n));
}

/* ---------------------------------------------------------------------- */
/* f(R) fused level-0 residual restriction to level 1,                      */
/* restrict_residual_fR(1), PMP2MGsolver_fR.f90:607-749                     */
/* ---------------------------------------------------------------------- */
void FN(orc_fr_restrict_residual1)(const orc_params* p, const STOR* u0, const STOR* delta,
                                   STOR* rres1)
{
    const int N = p->N, nc = N / 2;
    const int frn = p->fr_n;
    orc_fr_coef c;
    orc_fr_coefs(p, 0, &c);
    const double fct1 = c.om_a;
    const double fct2 = -c.R_bg * c.a2 / 3.0;
    const double fct3 = c.ctilde * c.ctilde * c.sgn / 1.0;
    const double fct4 = fct2 * pow(fabs(c.fR_bg), 1.0 / (1.0 + (double)frn));

#pragma omp parallel for collapse(2) schedule(static)
    for (int K = 0; K < nc; K++) {
        for (int J = 0; J < nc; J++) {
            for (int I = 0; I < nc; I++) {
                const int x0 = 2 * I, x1 = 2 * I + 1, y0 = 2 * J, y1 = 2 * J + 1, z0 = 2 * K, z1 = 2 * K + 1;
                const int xl = (I == 0) ? N - 1 : 2 * I - 1, xu = (2 * I + 2 == N) ? 0 : 2 * I + 2;
                const int yl = (J == 0) ? N - 1 : 2 * J - 1, yu = (2 * J + 2 == N) ? 0 : 2 * J + 2;
                const int zl = (K == 0) ? N - 1 : 2 * K - 1, zu = (2 * K + 2 == N) ? 0 : 2 * K + 2;
                const size_t ch[8] = {ORC_IDX(x0, y0, z0, N), ORC_IDX(x1, y0, z0, N), ORC_IDX(x0, y1, z0, N),
                                      ORC_IDX(x1, y1, z0, N), ORC_IDX(x0, y0, z1, N), ORC_IDX(x1, y0, z1, N),
                                      ORC_IDX(x0, y1, z1, N), ORC_IDX(x1, y1, z1, N)};
                double Q = -8.0 * fct2;
                for (int q = 0; q < 8; q++) Q = Q + fct1 * (double)delta[ch[q]] + fct4 / (double)u0[ch[q]];
                STOR s8 = FN(wpow)(u0[ch[0]], frn);
                for (int q = 1; q < 8; q++) s8 = s8 + FN(wpow)(u0[ch[q]], frn);
                double P = -3.0 * (double)s8;
#define W_(x, y, z) ((double)FN(wpow)(u0[ORC_IDX(x, y, z, N)], frn))
                P = P + W_(xl, y0, z0) + W_(xl, y1, z0) + W_(xl, y0, z1) + W_(xl, y1, z1);
                P = P + W_(xu, y0, z0) + W_(xu, y1, z0) + W_(xu, y0, z1) + W_(xu, y1, z1);
                P = P + W_(x0, yl, z0) + W_(x1, yl, z0) + W_(x0, yl, z1) + W_(x1, yl, z1);
                P = P + W_(x0, yu, z0) + W_(x1, yu, z0) + W_(x0, yu, z1) + W_(x1, yu, z1);
                P = P + W_(x0, y0, zl) + W_(x1, y0, zl) + W_(x0, y1, zl) + W_(x1, y1, zl);
                P = P + W_(x0, y0, zu) + W_(x1, y0, zu) + W_(x0, y1, zu) + W_(x1, y1, zu);
#undef W_
                rres1[ORC_IDX(I, J, K, nc)] = (STOR)((P * fct3 + Q) / 8.0);
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* 8-child average: restrict_density / restrict_scalar_field               */
/* (PMP2extradof.f90:1198-1247, 1293-1345) and restrict_residual_fR/_DGP   */
/* at levels > 1 (PMP2MGsolver_fR.f90:750-779).                             */
/* ---------------------------------------------------------------------- */
void FN(orc_restrict8)(int nc, const STOR* fine, STOR* coarse)
{
    const int nf = 2 * nc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int K = 0; K < nc; K++) {
        for (int J = 0; J < nc; J++) {
            for (int I = 0; I < nc; I++) {
                double P = (double)fine[ORC_IDX(2 * I, 2 * J, 2 * K, nf)];
                P = P + (double)fine[ORC_IDX(2 * I + 1, 2 * J, 2 * K, nf)];
                P = P + (double)fine[ORC_IDX(2 * I, 2 * J + 1, 2 * K, nf)];
                P = P + (double)fine[ORC_IDX(2 * I + 1, 2 * J + 1, 2 * K, nf)];
                P = P + (double)fine[ORC_IDX(2 * I, 2 * J, 2 * K + 1, nf)];
                P = P + (double)fine[ORC_IDX(2 * I + 1, 2 * J, 2 * K + 1, nf)];
                P = P + (double)fine[ORC_IDX(2 * I, 2 * J + 1, 2 * K + 1, nf)];
                P = P + (double)fine[ORC_IDX(2 * I + 1, 2 * J + 1, 2 * K + 1, nf)];
                coarse[ORC_IDX(I, J, K, nc)] = (STOR)(P / 8.0);
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* f(R) coarse FAS right-hand side f_l = N_l(R u) - R r                     */
/* calculate_physical_right_hand_side_fR, PMP2MGsolver_fR.f90:789-927      */
/* ---------------------------------------------------------------------- */
void FN(orc_fr_rhs)(const orc_params* p, int level, const STOR* ru, const STOR* rho, const STOR* rres,
                    STOR* f)
{
    const int n = p->N >> level;
    const int frn = p->fr_n;
    orc_fr_coef c;
    orc_fr_coefs(p, level, &c);
    const double fct1 = c.om_a;
    const double fct2 = -c.R_bg * c.a2 / 3.0;
    const double fct3 = c.ctilde * c.ctilde * c.sgn / c.dx2;
    const double a_fR_rt = pow(fabs(c.fR_bg), 1.0 / (double)(frn + 1));
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < n; k++) {
        for (int j = 0; j < n; j++) {
            const int km = (k == 0) ? n - 1 : k - 1, kp = (k == n - 1) ? 0 : k + 1;
            const int jm = (j == 0) ? n - 1 : j - 1, jp = (j == n - 1) ? 0 : j + 1;
            for (int i = 0; i < n; i++) {
                const int im = (i == 0) ? n - 1 : i - 1, ip = (i == n - 1) ? 0 : i + 1;
                const size_t cc = ORC_IDX(i, j, k, n);
                const STOR s = FN(wpow)(ru[ORC_IDX(ip, j, k, n)], frn) + FN(wpow)(ru[ORC_IDX(im, j, k, n)], frn) +
                               FN(wpow)(ru[ORC_IDX(i, jp, k, n)], frn) + FN(wpow)(ru[ORC_IDX(i, jm, k, n)], frn) +
                               FN(wpow)(ru[ORC_IDX(i, j, kp, n)], frn) + FN(wpow)(ru[ORC_IDX(i, j, km, n)], frn);
                double OP = (double)s - (double)FN(wpow)(ru[cc], frn) * 6.0;
                OP = OP * fct3;
                OP = OP + fct1 * (double)rho[cc];
                OP = OP + fct2 * (a_fR_rt / (double)ru[cc] - 1.0);
                OP = OP - (double)rres[cc];
                f[cc] = (STOR)OP;
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Prolongation + correction, correct_solution, PMP2extradof.f90:1361-1578 */
/* u_h += P(u_H - R u_h) with cell-centred trilinear weights 27/9/3/1 /64.  */
/* guard != 0: keep the old fine value unless the corrected one is > 0     */
/* (f(R), PMP2extradof.f90:1496-1499).                                      */
/* ---------------------------------------------------------------------- */
void FN(orc_correct)(int nc, const STOR* uH, const STOR* ru, STOR* ufine, int guard)
{
    const int nf = 2 * nc;
    const double a = 1.0 / 64.0, b = 3.0 * a, c9 = 9.0 * a, d = 27.0 * a;
    /* weight of contributor ind (bit set = parent coordinate, clear = neighbour) */
    const double wt[8] = {a, b, b, c9, b, c9, c9, d};
#pragma omp parallel for collapse(2) schedule(static)
    for (int K = 0; K < nc; K++) {
        for (int J = 0; J < nc; J++) {
            for (int I = 0; I < nc; I++) {
                for (int sind = 0; sind < 8; sind++) {
                    const int s1 = sind & 1, s2 = (sind >> 1) & 1, s3 = (sind >> 2) & 1;
                    /* neighbour coarse coordinate on the son's side, periodic */
                    const int nx = s1 ? ((I == nc - 1) ? 0 : I + 1) : ((I == 0) ? nc - 1 : I - 1);
                    const int ny = s2 ? ((J == nc - 1) ? 0 : J + 1) : ((J == 0) ? nc - 1 : J - 1);
                    const int nz = s3 ? ((K == nc - 1) ? 0 : K + 1) : ((K == 0) ? nc - 1 : K - 1);
                    double P = 0.0;
                    for (int ind = 0; ind < 8; ind++) {
                        const int X = (ind & 1) ? I : nx;
                        const int Y = (ind & 2) ? J : ny;
                        const int Z = (ind & 4) ? K : nz;
                        const size_t q = ORC_IDX(X, Y, Z, nc);
                        P = P + wt[ind] * (double)uH[q];
                        P = P - wt[ind] * (double)ru[q];
                    }
                    const size_t fi = ORC_IDX(2 * I + s1, 2 * J + s2, 2 * K + s3, nf);
                    const double tmpt = (double)ufine[fi] + P;
                    if (!guard || tmpt > 0.0) ufine[fi] = (STOR)tmpt;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* DGP: Sigma and Laplacian of one cell on the 19-point stencil           */
/* (PMP2MGsolver_DGP.f90:166-205 smoother form, also residual and RHS)     */
/* ---------------------------------------------------------------------- */
static inline void FN(dgp_cell)(const STOR* u, int n, int i, int j, int k, double* sig_extra, double* lap_num)
{
    const int km = (k == 0) ? n - 1 : k - 1, kp = (k == n - 1) ? 0 : k + 1;
    const int jm = (j == 0) ? n - 1 : j - 1, jp = (j == n - 1) ? 0 : j + 1;
    const int im = (i == 0) ? n - 1 : i - 1, ip = (i == n - 1) ? 0 : i + 1;
    const STOR c0 = u[ORC_IDX(i, j, k, n)];
    const STOR xp = u[ORC_IDX(ip, j, k, n)], xm = u[ORC_IDX(im, j, k, n)];
    const STOR yp = u[ORC_IDX(i, jp, k, n)], ym = u[ORC_IDX(i, jm, k, n)];
    const STOR zp = u[ORC_IDX(i, j, kp, n)], zm = u[ORC_IDX(i, j, km, n)];
    const double dxx = xp + xm - 2.0 * c0;
    const double dyy = yp + ym - 2.0 * c0;
    const double dzz = zp + zm - 2.0 * c0;
    const STOR cxy = u[ORC_IDX(ip, jp, k, n)] + u[ORC_IDX(im, jm, k, n)] - u[ORC_IDX(ip, jm, k, n)] - u[ORC_IDX(im, jp, k, n)];
    const STOR cxz = u[ORC_IDX(ip, j, kp, n)] + u[ORC_IDX(im, j, km, n)] - u[ORC_IDX(ip, j, km, n)] - u[ORC_IDX(im, j, kp, n)];
    const STOR cyz = u[ORC_IDX(i, jp, kp, n)] + u[ORC_IDX(i, jm, km, n)] - u[ORC_IDX(i, jp, km, n)] - u[ORC_IDX(i, jm, kp, n)];
    const STOR diag = cxy * cxy + cxz * cxz + cyz * cyz;
    sig_extra[0] = (dxx * dxx + dyy * dyy + dzz * dzz);
    sig_extra[1] = (dxx * dyy + dxx * dzz + dyy * dzz);
    sig_extra[2] = (double)diag;
    const STOR s6 = xp + xm + yp + ym + zp + zm;
    *lap_num = (double)s6 - (double)c0 * 6.0;
}

/* DGP smoother, relaxation_iterations_DGP, PMP2MGsolver_DGP.f90:8-237.
 * lexi = 1: the reference's sequential lexicographic sweep over the colour
 * (same-colour edge neighbours of the 19-point stencil already updated when
 * they precede the cell).  lexi = 0: every cell of the colour is updated
 * from the values the colour had before the pass (the order-independent
 * form a parallel sweep computes). */
void FN(orc_dgp_smooth)(const orc_params* p, int level, int red, int lexi, STOR* u, const STOR* rho,
                        const STOR* f)
{
    const int n = p->N >> level;
    orc_dgp_coef c;
    orc_dgp_coefs(p, level, &c);
    const double TWOOVERTHREE = 2.0 / 3.0, ONEOVEREIGHT = 0.125, THREEOVERFOUR = 0.75, EIGHTOVERTHREE = 8.0 / 3.0;
    const double dLphi = -6.0 / c.dx2;
    const size_t n3 = (size_t)n * n * n;
    STOR* src = u;
    STOR* snap = NULL;
    if (!lexi) {
        snap = (STOR*)malloc(n3 * sizeof(STOR));
        memcpy(snap, u, n3 * sizeof(STOR));
        src = snap;
    }
#define DGP_BODY                                                                                       \
    {                                                                                                  \
        const size_t cc = ORC_IDX(i, j, k, n);                                                         \
        double se[3], lap;                                                                             \
        FN(dgp_cell)(src, n, i, j, k, se, &lap);                                                       \
        double Sigma = c.fct1 * (double)rho[cc];                                                       \
        Sigma = Sigma + se[0] * TWOOVERTHREE / c.dx4 - se[1] * TWOOVERTHREE / c.dx4 +                  \
                se[2] * ONEOVEREIGHT / c.dx4;                                                          \
        double L = lap / c.dx2 -                                                                       \
                   (copysign(1.0, c.alpha) * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * Sigma) - c.alpha) * \
                       THREEOVERFOUR;                                                                  \
        if (level > 0) L = L - (double)f[cc];                                                          \
        u[cc] = (STOR)((double)src[cc] - L / dLphi);                                                   \
    }
    if (lexi) {
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    const int odd = (i + j + k) & 1;
                    if (red ? odd : !odd) continue;
                    DGP_BODY
                }
    } else {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++) {
                const int i0 = red ? ((j + k) & 1) : (1 - ((j + k) & 1));
                for (int i = i0; i < n; i += 2) DGP_BODY
            }
    }
#undef DGP_BODY
    free(snap);
}

/* DGP residual, calculate_residual_DGP, PMP2MGsolver_DGP.f90:240-388.
 * The RMS is taken over the stored residual squared in storage precision,
 * as the reference's second pass does (:368-383). */
double FN(orc_dgp_residual)(const orc_params* p, int level, const STOR* u, const STOR* rho, const STOR* f,
                            STOR* res)
{
    const int n = p->N >> level;
    orc_dgp_coef c;
    orc_dgp_coefs(p, level, &c);
    const double TWOOVERTHREE = 2.0 / 3.0, ONEOVEREIGHT = 1.0 / 8.0, THREEOVERFOUR = 3.0 / 4.0, EIGHTOVERTHREE = 8.0 / 3.0;
    double* part = (double*)calloc((size_t)n * n, sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
            double acc = 0.0;
            for (int i = 0; i < n; i++) {
                const size_t cc = ORC_IDX(i, j, k, n);
                double se[3], lap;
                FN(dgp_cell)(u, n, i, j, k, se, &lap);
                double Sigma = c.fct1 * (double)rho[cc];
                Sigma = Sigma + se[0] * TWOOVERTHREE / c.dx4 - se[1] * TWOOVERTHREE / c.dx4 + se[2] * ONEOVEREIGHT / c.dx4;
                double OP = lap / c.dx2;
                OP = OP - (copysign(1.0, c.alpha) * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * Sigma) - c.alpha) * THREEOVERFOUR;
                if (level > 0) OP = OP - (double)f[cc];
                const STOR st = (STOR)OP;
                if (res) res[cc] = st;
                acc += (double)(st * st);
            }
            part[(size_t)k * n + j] = acc;
        }
    double RES = 0.0;
    for (size_t q = 0; q < (size_t)n * n; q++) RES += part[q];
    free(part);
    return sqrt(RES / ((double)n * (double)n * (double)n));
}

/* DGP fused level-0 residual restriction, restrict_residual_DGP(1),
 * PMP2MGsolver_DGP.f90:459-763: each son's Sigma in the rearranged form
 * S_x (2 S_x - S_y - S_z)/3 + ... with S_x = u(x+1) + u(x-1). */
void FN(orc_dgp_restrict_residual1)(const orc_params* p, const STOR* u0, const STOR* delta, STOR* rres1)
{
    const int N = p->N, nc = N / 2;
    orc_dgp_coef c;
    orc_dgp_coefs(p, 0, &c);
    const double ONEOVEREIGHT = 1.0 / 8.0, ONEOVERTHREE = 1.0 / 3.0, EIGHTOVERTHREE = 8.0 / 3.0;
    const double dx2 = 1.0, dx4 = 1.0;
    const double fct2 = -0.75 * c.alpha;
    const double fct3 = copysign(0.75, c.alpha);
#pragma omp parallel for collapse(2) schedule(static)
    for (int K = 0; K < nc; K++)
        for (int J = 0; J < nc; J++)
            for (int I = 0; I < nc; I++) {
                double Q = -8.0 * fct2;
                double sig[8];
                for (int q = 0; q < 8; q++) {
                    const int i = 2 * I + (q & 1), j = 2 * J + ((q >> 1) & 1), k = 2 * K + ((q >> 2) & 1);
                    const int im = (i == 0) ? N - 1 : i - 1, ip = (i == N - 1) ? 0 : i + 1;
                    const int jm = (j == 0) ? N - 1 : j - 1, jp = (j == N - 1) ? 0 : j + 1;
                    const int km = (k == 0) ? N - 1 : k - 1, kp = (k == N - 1) ? 0 : k + 1;
#define U_(x, y, z) u0[ORC_IDX(x, y, z, N)]
                    const STOR Sx = U_(ip, j, k) + U_(im, j, k);
                    const STOR Sy = U_(i, jp, k) + U_(i, jm, k);
                    const STOR Sz = U_(i, j, kp) + U_(i, j, km);
                    const STOR Dxy = (U_(ip, jp, k) + U_(im, jm, k)) - (U_(ip, jm, k) + U_(im, jp, k));
                    const STOR Dxz = (U_(ip, j, kp) + U_(im, j, km)) - (U_(ip, j, km) + U_(im, j, kp));
                    const STOR Dyz = (U_(i, jp, kp) + U_(i, jm, km)) - (U_(i, jp, km) + U_(i, jm, kp));
#undef U_
                    double s = c.fct1 * (double)delta[ORC_IDX(i, j, k, N)];
                    s = s + (double)Sx * (2.0 * Sx - (double)Sy - (double)Sz) * ONEOVERTHREE / dx4 +
                        (double)Sy * (2.0 * Sy - (double)Sx - (double)Sz) * ONEOVERTHREE / dx4 +
                        (double)Sz * (2.0 * Sz - (double)Sx - (double)Sy) * ONEOVERTHREE / dx4 +
                        (double)(Dxy * Dxy) * ONEOVEREIGHT / dx4 + (double)(Dxz * Dxz) * ONEOVEREIGHT / dx4 +
                        (double)(Dyz * Dyz) * ONEOVEREIGHT / dx4;
                    sig[q] = s;
                }
                for (int q = 0; q < 8; q++) Q = Q - fct3 * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * sig[q]);
                const int x0 = 2 * I, x1 = 2 * I + 1, y0 = 2 * J, y1 = 2 * J + 1, z0 = 2 * K, z1 = 2 * K + 1;
                const int xl = (I == 0) ? N - 1 : 2 * I - 1, xu = (2 * I + 2 == N) ? 0 : 2 * I + 2;
                const int yl = (J == 0) ? N - 1 : 2 * J - 1, yu = (2 * J + 2 == N) ? 0 : 2 * J + 2;
                const int zl = (K == 0) ? N - 1 : 2 * K - 1, zu = (2 * K + 2 == N) ? 0 : 2 * K + 2;
                double P = 0.0;
#define V_(x, y, z) ((double)u0[ORC_IDX(x, y, z, N)])
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
                rres1[ORC_IDX(I, J, K, nc)] = (STOR)(P / 8.0);
            }
}

/* DGP coarse FAS right-hand side, calculate_physical_right_hand_side_DGP,
 * PMP2MGsolver_DGP.f90:793-885 */
void FN(orc_dgp_rhs)(const orc_params* p, int level, const STOR* ru, const STOR* rho, const STOR* rres, STOR* f)
{
    const int n = p->N >> level;
    orc_dgp_coef c;
    orc_dgp_coefs(p, level, &c);
    const double TWOOVERTHREE = 2.0 / 3.0, ONEOVEREIGHT = 1.0 / 8.0, THREEOVERFOUR = 3.0 / 4.0, EIGHTOVERTHREE = 8.0 / 3.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const size_t cc = ORC_IDX(i, j, k, n);
                double se[3], lap;
                FN(dgp_cell)(ru, n, i, j, k, se, &lap);
                double Sigma = c.fct1 * (double)rho[cc];
                Sigma = Sigma + se[0] * TWOOVERTHREE / c.dx4 - se[1] * TWOOVERTHREE / c.dx4 + se[2] * ONEOVEREIGHT / c.dx4;
                double OP = lap / c.dx2;
                OP = OP - (copysign(1.0, c.alpha) * sqrt(c.alpha * c.alpha + EIGHTOVERTHREE * Sigma) - c.alpha) * THREEOVERFOUR;
                OP = OP - (double)rres[cc];
                f[cc] = (STOR)OP;
            }
}

/* ---------------------------------------------------------------------- */
/* V-cycle and outer loop: V_cycle (PMP2extradof.f90:352-555) and          */
/* scalar_field (PMP2extradof.f90:12-178)                                  */
/* ---------------------------------------------------------------------- */
typedef struct {
    int nlev;     /* coarsest level index + 1 */
    int lc;       /* coarsest level index */
    STOR *u[32], *rho[32], *f[32], *res[32], *rres[32], *ru[32];
} FN(orc_hier);

static void FN(hier_alloc)(const orc_params* p, FN(orc_hier) * h, STOR* u0, const STOR* delta)
{
    const int lmax = orc_levelmax(p->N), lmin = orc_levelmin(p->N);
    h->lc = lmax - lmin;
    if (h->lc < 0) h->lc = 0;
    h->nlev = h->lc + 1;
    memset(h->u, 0, sizeof(h->u));
    memset(h->rho, 0, sizeof(h->rho));
    memset(h->f, 0, sizeof(h->f));
    memset(h->res, 0, sizeof(h->res));
    memset(h->rres, 0, sizeof(h->rres));
    memset(h->ru, 0, sizeof(h->ru));
    h->u[0] = u0;
    h->rho[0] = (STOR*)delta;
    for (int l = 1; l <= h->lc; l++) {
        const size_t n = (size_t)(p->N >> l);
        const size_t n3 = n * n * n;
        h->u[l] = (STOR*)calloc(n3, sizeof(STOR));
        h->rho[l] = (STOR*)calloc(n3, sizeof(STOR));
        h->f[l] = (STOR*)calloc(n3, sizeof(STOR));
        h->res[l] = (STOR*)calloc(n3, sizeof(STOR));
        h->rres[l] = (STOR*)calloc(n3, sizeof(STOR));
        h->ru[l] = (STOR*)calloc(n3, sizeof(STOR));
    }
}

static void FN(hier_free)(FN(orc_hier) * h)
{
    for (int l = 1; l <= h->lc; l++) {
        free(h->u[l]);
        free(h->rho[l]);
        free(h->f[l]);
        free(h->res[l]);
        free(h->rres[l]);
        free(h->ru[l]);
    }
}

static void FN(smooth_any)(const orc_params* p, int l, int red, int dgp_lexi, FN(orc_hier) * h)
{
    if (p->model == 1)
        FN(orc_fr_smooth)(p, l, red, h->u[l], h->rho[l], h->f[l]);
    else
        FN(orc_dgp_smooth)(p, l, red, dgp_lexi, h->u[l], h->rho[l], h->f[l]);
}

static void FN(cycle)(const orc_params* p, int l, int kind, int dgp_lexi, FN(orc_hier) * h);

/* 2 x (red, black) sweeps, the Nsmth = 2 loop of every cycle type */
static void FN(smooth2)(const orc_params* p, int l, int dgp_lexi, FN(orc_hier) * h)
{
    const int Nsmth = 2;
    for (int s = 0; s < Nsmth; s++) {
        FN(smooth_any)(p, l, 1, dgp_lexi, h);
        FN(smooth_any)(p, l, 0, dgp_lexi, h);
    }
}

/* coarse-grid correction from level l: residual, restrictions, FAS RHS, a
 * `kind` cycle on level l+1 and the prolongation (PMP2extradof.f90:418-545
 * for V_cycle; the same sequence in F_cycle :650-840 and W_cycle :935-1140) */
static void FN(cgc)(const orc_params* p, int l, int kind, int dgp_lexi, FN(orc_hier) * h)
{
    const int nc = p->N >> (l + 1);
    if (l == 0) {
        /* the level-0 residual the reference computes here is never read:
         * the level-1 restriction recomputes it from u and delta */
        if (p->model == 1)
            FN(orc_fr_restrict_residual1)(p, h->u[0], h->rho[0], h->rres[1]);
        else
            FN(orc_dgp_restrict_residual1)(p, h->u[0], h->rho[0], h->rres[1]);
    } else {
        if (p->model == 1)
            FN(orc_fr_residual)(p, l, h->u[l], h->rho[l], h->f[l], h->res[l]);
        else
            FN(orc_dgp_residual)(p, l, h->u[l], h->rho[l], h->f[l], h->res[l]);
        FN(orc_restrict8)(nc, h->res[l], h->rres[l + 1]);
    }
    /* restrict_scalar_field: R u to both the frozen copy and the unknown */
    FN(orc_restrict8)(nc, h->u[l], h->ru[l + 1]);
    memcpy(h->u[l + 1], h->ru[l + 1], (size_t)nc * nc * nc * sizeof(STOR));
    FN(orc_restrict8)(nc, h->rho[l], h->rho[l + 1]);
    if (p->model == 1)
        FN(orc_fr_rhs)(p, l + 1, h->ru[l + 1], h->rho[l + 1], h->rres[l + 1], h->f[l + 1]);
    else
        FN(orc_dgp_rhs)(p, l + 1, h->ru[l + 1], h->rho[l + 1], h->rres[l + 1], h->f[l + 1]);
    FN(cycle)(p, l + 1, kind, dgp_lexi, h);
    FN(orc_correct)(nc, h->u[l + 1], h->ru[l + 1], h->u[l], p->model == 1);
}

/* V_cycle (PMP2extradof.f90:352-555): S2, CGC(V), S2.
 * F_cycle (:563-858): S2, CGC(F), S2, CGC(V), S2.
 * W_cycle (:867-1157): S2, CGC(W), S2, CGC(W), S2.
 * On the coarsest level each is S2 only. */
static void FN(cycle)(const orc_params* p, int l, int kind, int dgp_lexi, FN(orc_hier) * h)
{
    FN(smooth2)(p, l, dgp_lexi, h);
    if (l >= h->lc) return;
    FN(cgc)(p, l, kind, dgp_lexi, h);
    FN(smooth2)(p, l, dgp_lexi, h);
    if (kind == 0) return;
    FN(cgc)(p, l, kind == 1 ? 0 : 2, dgp_lexi, h);
    FN(smooth2)(p, l, dgp_lexi, h);
}

static void FN(vcycle)(const orc_params* p, int l, int dgp_lexi, FN(orc_hier) * h)
{
    FN(cycle)(p, l, 0, dgp_lexi, h);
}

static double FN(res0)(const orc_params* p, FN(orc_hier) * h)
{
    if (p->model == 1) return FN(orc_fr_residual)(p, 0, h->u[0], h->rho[0], NULL, NULL);
    return FN(orc_dgp_residual)(p, 0, h->u[0], h->rho[0], NULL, NULL);
}

int FN(orc_vcycles)(const orc_params* p, STOR* u, const STOR* delta, int ncycles, int dgp_lexi, double* res_after)
{
    FN(orc_hier) h;
    FN(hier_alloc)(p, &h, u, delta);
    for (int c = 0; c < ncycles; c++) {
        FN(vcycle)(p, 0, dgp_lexi, &h);
        if (res_after) res_after[c] = FN(res0)(p, &h);
    }
    FN(hier_free)(&h);
    return 0;
}

int FN(orc_cycles)(const orc_params* p, STOR* u, const STOR* delta, int ncycles, int cycle, int dgp_lexi,
                   double* res_after)
{
    FN(orc_hier) h;
    FN(hier_alloc)(p, &h, u, delta);
    for (int c = 0; c < ncycles; c++) {
        FN(cycle)(p, 0, cycle, dgp_lexi, &h);
        if (res_after) res_after[c] = FN(res0)(p, &h);
    }
    FN(hier_free)(&h);
    return 0;
}

int FN(orc_solve)(const orc_params* p, STOR* u, const STOR* delta, const orc_solve_opts* o, orc_solve_report* rep)
{
    const size_t N3 = (size_t)p->N * p->N * p->N;
    memset(rep, 0, sizeof(*rep));
    if (o->init_guess) {
        const STOR v = (p->model == 1) ? (STOR)orc_fr_ubg(p) : (STOR)0.0;
        for (size_t q = 0; q < N3; q++) u[q] = v;
    }
    FN(orc_hier) h;
    FN(hier_alloc)(p, &h, u, delta);
    rep->res0 = FN(res0)(p, &h);
    double res_old_1 = 1.0e5, res_old_2 = 1.0e10;
    int iter = 1;
    for (;;) {
        FN(cycle)(p, 0, o->cycle, o->dgp_lexi, &h);
        const double res = FN(res0)(p, &h);
        if (iter <= 256) rep->history[iter - 1] = res;
        rep->cycles = iter;
        rep->res_final = res;
        int can_exit = 0;
        if (o->mode == 0) {
            if (res <= o->res_conv) { can_exit = 1; rep->converged = 1; }
            if (iter > 2) {
                if (fabs(res / res_old_1 - 1.0) <= 5.0e-3) can_exit = 1;
                if (fabs(res / res_old_1 - 1.0) <= 1.0e-2 && fabs(res_old_1 / res_old_2 - 1.0) <= 1.0e-2) can_exit = 1;
            }
            if (iter >= o->max_cycles) can_exit = 1;
        } else {
            if (res < o->rel_tol * rep->res0) { can_exit = 1; rep->converged = 1; }
            if (iter >= o->max_cycles) can_exit = 1;
        }
        if (can_exit) break;
        iter++;
        res_old_2 = res_old_1;
        res_old_1 = res;
    }
    if (p->model == 1 && o->normalize) {
        const double inv = pow(fabs(orc_fr_fbg(p)), -1.0 / (1.0 + (double)p->fr_n));
        for (size_t q = 0; q < N3; q++) u[q] = (STOR)((double)u[q] * inv);
    }
    FN(hier_free)(&h);
    return 0;
}

#undef FN
#undef CAT2
#undef CAT2_
