/*
 * CPU oracle for the f(R) / nDGP multigrid solve -- TEST INFRASTRUCTURE ONLY.
 * See mgs_oracle.h.  Built by oracle/Makefile into oracle/liboracle.so.
 */
#include "mgs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_IDX(i, j, k, n) ((size_t)(i) + (size_t)(n) * ((size_t)(j) + (size_t)(n) * (size_t)(k)))

/* level bookkeeping, initialise_extradof, PMP2extradof.f90:240-253 */
int orc_levelmax(int N)
{
    int l = 0;
    while ((1 << l) < N) l++;
    return l;
}

int orc_levelmin(int N)
{
    int odd = N;
    while (odd % 2 == 0) odd /= 2;
    int l = 0;
    while ((1 << l) < odd) l++;
    return l + 1;
}

/* integer power with the repeated-squaring order of the compiler runtime
 * routine the reference's x**(fr_n+1) lowers to */
static double powi_d(double a, int b)
{
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return r;
}

/* f(R) background quantities shared by all f(R) kernels,
 * PMP2MGsolver_fR.f90:54-61 (and :333-346, :590-605, :809-823) */
typedef struct {
    double ctilde, dx2, R_bg, R0_bg, fR_bg, sgn, om_a, a2;
} orc_fr_coef;

static void orc_fr_coefs(const orc_params* p, int level, orc_fr_coef* c)
{
    const double a = p->a;
    c->ctilde = 2.99792458e3 * (double)p->N / p->box;
    const double dx = (double)(1 << level);
    c->dx2 = dx * dx;
    c->R_bg = 3.0 * (p->om / (a * a * a) + 4.0 * p->oml);
    c->R0_bg = 3.0 * (p->om + 4.0 * p->oml);
    c->fR_bg = p->fr0 * powi_d(c->R0_bg / c->R_bg, p->fr_n + 1);
    c->sgn = copysign(1.0, c->fR_bg);
    c->om_a = p->om / a;
    c->a2 = a * a;
}

double orc_fr_fbg(const orc_params* p)
{
    orc_fr_coef c;
    orc_fr_coefs(p, 0, &c);
    return c.fR_bg;
}

double orc_fr_ubg(const orc_params* p)
{
    return pow(fabs(orc_fr_fbg(p)), 1.0 / (1.0 + (double)p->fr_n));
}

/* DGP constants, PMP2MGsolver_DGP.f90:36-48 */
typedef struct {
    double dx2, dx4, alpha, beta, fct1;
} orc_dgp_coef;

static void orc_dgp_coefs(const orc_params* p, int level, orc_dgp_coef* c)
{
    const double a = p->a;
    const double dx = (double)(1 << level);
    c->dx2 = dx * dx;
    c->dx4 = c->dx2 * c->dx2;
    const double Orc = 1.0 / (4.0 * p->h0rc * p->h0rc);
    const double a3 = a * a * a;
    if (p->branch >= 0)
        c->beta = 1.0 + (0.5 * p->om / a3 + p->oml) / sqrt(Orc * (p->om / a3 + p->oml));
    else
        c->beta = -(0.5 * p->om / a3 + Orc) / sqrt(Orc * (p->om / a3 + Orc));
    const double Rc_sq = 1.0 / (4.0 * Orc);
    c->alpha = 3.0 * c->beta * (a * a) / Rc_sq;
    c->fct1 = c->alpha * p->om / c->beta / a;
}

#define STOR float
#define SFX f32
#include "mgs_oracle_impl.h"
#undef STOR
#undef SFX

#define STOR double
#define SFX f64
#include "mgs_oracle_impl.h"
#undef STOR
#undef SFX
