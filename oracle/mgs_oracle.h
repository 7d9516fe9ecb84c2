/*
 * CPU oracle for the f(R) / nDGP multigrid scalar-field solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libmgsolve) never links, loads or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's multigrid path:
 *   PMP2MGsolver_fR.f90   (f(R) smoother, residual, restriction, FAS RHS)
 *   PMP2MGsolver_DGP.f90  (nDGP smoother, residual, restriction, FAS RHS)
 *   PMP2extradof.f90      (scalar_field, V_cycle, restrict_density,
 *                          restrict_scalar_field, correct_solution)
 *
 * Every kernel exists twice, generated from one source text
 * (mgs_oracle_impl.h) with a storage type parameter:
 *   *_f32  storage float, arithmetic following Fortran's typing rules exactly
 *          as the reference writes its expressions (real*4 sub-expressions
 *          stay real*4, promotion to real*8 where a real*8 operand appears).
 *          This variant is pinned bit-for-bit against the reference compiled
 *          with flang (oracle/_ref).
 *   *_f64  storage double, the same expressions in FP64: the true oracle the
 *          GPU FP64 path and the mixed-precision path are measured against.
 *
 * Layout: every level is one array of n^3 values, 0-based, x fastest
 * (index i + n*(j + n*k)), the same memory order as the reference's
 * column-major FI/FI2 arrays.  Cell (i,j,k) here is the reference's
 * (M1,M2,M3) = (i+1,j+1,k+1).  The reference's "red" pass (redstep=.TRUE.)
 * updates cells with M1+M2+M3 odd, i.e. i+j+k even here.
 */
#ifndef MGS_ORACLE_H
#define MGS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Model parameters.  Field meaning follows the reference's Tools module
 * (PMP2mod_tools.f90:9-16, 52-104). */
typedef struct orc_params {
    int32_t N;        /* NGRID: grid points per side (level 0) */
    int32_t model;    /* 1 = f(R) Hu-Sawicki, 2 = DGP (MG_model) */
    int32_t fr_n;     /* Hu-Sawicki n in {0,1,2} */
    int32_t branch;   /* DGP: +1 normal branch (N_branch), -1 self-accelerating (S_branch) */
    double box;       /* Box, Mpc/h */
    double om;        /* Om */
    double oml;       /* OmL */
    double a;         /* AEXPN */
    double fr0;       /* fR0 (negative for the usual Hu-Sawicki sign) */
    double h0rc;      /* H0rc */
    int32_t polish;   /* f(R) smoother: 0 = the reference's closed form as written;
                         1 = closed form followed by one Newton step on the local
                         polynomial u^{n+2} + P u + Q (same root, full relative
                         precision where the closed form cancels) */
    int32_t pad_;
} orc_params;

/* Solve control for orc_solve_*.  mode 0 = the reference's own exits
 * (absolute RMS <= res_conv, 0.5 % stagnation tests, iteration cap;
 * PMP2extradof.f90:107-122); mode 1 = relative residual only:
 * stop when RMS(r_k)/RMS(r_0) < rel_tol, r_0 = residual of the initial
 * guess, or at max_cycles. */
typedef struct orc_solve_opts {
    int32_t mode;
    int32_t max_cycles;
    double res_conv;
    double rel_tol;
    int32_t init_guess;   /* 1: start from u = u_bg (f(R)) / phi = 0 (DGP); 0: use given field */
    int32_t normalize;    /* f(R): 1 = divide the result by u_bg on exit (PMP2extradof.f90:131-146) */
    int32_t dgp_lexi;     /* DGP smoother order: 1 = reference sequential, 0 = colour-parallel */
    int32_t cycle;        /* 0 V_cycle, 1 F_cycle, 2 W_cycle (the reference's Vcycle/Fcycle/Wcycle) */
} orc_solve_opts;

typedef struct orc_solve_report {
    int32_t cycles;
    int32_t converged;
    double res0;          /* RMS residual of the initial field */
    double res_final;
    double history[256];  /* RMS residual after each V-cycle */
} orc_solve_report;

/* level bookkeeping, PMP2extradof.f90:240-253 */
int orc_levelmax(int N);
int orc_levelmin(int N);

/* f(R) background helpers (double): u_bg = |fR_bg|^(1/(n+1)) */
double orc_fr_ubg(const orc_params* p);
double orc_fr_fbg(const orc_params* p);

#define ORC_DECLARE(SFX, STOR)                                                              \
    void orc_fr_smooth_##SFX(const orc_params* p, int level, int red, STOR* u,               \
                             const STOR* rho, const STOR* f);                               \
    double orc_fr_residual_##SFX(const orc_params* p, int level, const STOR* u,              \
                                 const STOR* rho, const STOR* f, STOR* res);                \
    void orc_fr_restrict_residual1_##SFX(const orc_params* p, const STOR* u0,                \
                                         const STOR* delta, STOR* rres1);                   \
    void orc_restrict8_##SFX(int nc, const STOR* fine, STOR* coarse);                        \
    void orc_fr_rhs_##SFX(const orc_params* p, int level, const STOR* ru, const STOR* rho,   \
                          const STOR* rres, STOR* f);                                       \
    void orc_correct_##SFX(int nc, const STOR* uH, const STOR* ru, STOR* ufine, int guard);  \
    void orc_dgp_smooth_##SFX(const orc_params* p, int level, int red, int lexi, STOR* u,     \
                              const STOR* rho, const STOR* f);                              \
    double orc_dgp_residual_##SFX(const orc_params* p, int level, const STOR* u,             \
                                  const STOR* rho, const STOR* f, STOR* res);               \
    void orc_dgp_restrict_residual1_##SFX(const orc_params* p, const STOR* u0,               \
                                          const STOR* delta, STOR* rres1);                  \
    void orc_dgp_rhs_##SFX(const orc_params* p, int level, const STOR* ru, const STOR* rho,  \
                           const STOR* rres, STOR* f);                                      \
    int orc_solve_##SFX(const orc_params* p, STOR* u, const STOR* delta,                     \
                        const orc_solve_opts* o, orc_solve_report* rep);                    \
    int orc_vcycles_##SFX(const orc_params* p, STOR* u, const STOR* delta, int ncycles,      \
                          int dgp_lexi, double* res_after);                                 \
    int orc_cycles_##SFX(const orc_params* p, STOR* u, const STOR* delta, int ncycles,       \
                         int cycle, int dgp_lexi, double* res_after);

ORC_DECLARE(f32, float)
ORC_DECLARE(f64, double)

#ifdef __cplusplus
}
#endif
#endif
