/*
 * mgsolve.h -- C-ABI of libmgsolve, the MI355X (gfx950) multigrid solver for
 * the modified-gravity scalar field of MG-GLAM-type particle-mesh codes.
 *
 * Plain pointers and sizes only.  All functions return 0 on success and a
 * negative MGS_E* code on failure; mgs_last_error() then describes it.
 * A handle owns a device-resident multigrid hierarchy and one HIP stream;
 * it is single-threaded (like the reference's module state) and persists
 * across calls (warm start between time steps).
 *
 * Field arrays crossing this boundary are 3-D grids of N^3 values in the
 * reference's memory order: x fastest, 0-based, i.e. FI(M1,M2,M3) at
 * offset (M1-1) + N*((M2-1) + N*(M3-1)) (PMP2mod_tools.f90:40-42).
 *
 * Reference interfaces replaced (file:line in chzruan/MG-GLAM-Durham):
 *   scalar_field(i_step)                        PMP2extradof.f90:12
 *   initialise_extradof                         PMP2extradof.f90:181
 *   V_cycle(ilevel)                             PMP2extradof.f90:352
 *   F_cycle(ilevel), W_cycle(ilevel)            PMP2extradof.f90:563, 867
 *   restrict_density(ilevel)                    PMP2extradof.f90:1168
 *   restrict_scalar_field(ilevel)               PMP2extradof.f90:1262
 *   correct_solution(ilevel)                    PMP2extradof.f90:1361
 *   relaxation_iterations_fR(ilevel,redstep)    PMP2MGsolver_fR.f90:20
 *   calculate_residual_fR(ilevel,res)           PMP2MGsolver_fR.f90:308
 *   restrict_residual_fR(ilevel)                PMP2MGsolver_fR.f90:569
 *   calculate_physical_right_hand_side_fR(l)    PMP2MGsolver_fR.f90:789
 *   relaxation_iterations_DGP(ilevel,redstep)   PMP2MGsolver_DGP.f90:8
 *   calculate_residual_DGP(ilevel,res)          PMP2MGsolver_DGP.f90:240
 *   restrict_residual_DGP(ilevel)               PMP2MGsolver_DGP.f90:418
 *   calculate_physical_right_hand_side_DGP(l)   PMP2MGsolver_DGP.f90:793
 *   POTENT_FOR_MG (MG potential)                PMP2main.f90:869
 *   MOVE gradient stencil (2x grad by central difference) PMP2main.f90:565-594
 *   POWERfft5 (NGP-binned power spectrum)       PMP2mod_power.f90:209
 *   DENSIT, CIC deposit branch                  PMP2mod_density.f90:235-281
 *   POTENTfft5 (FFT Poisson solve)              PMP2main.f90:705
 *   MOVE (f(R)/nDGP particle push)              PMP2main.f90:304
 */
#ifndef MGSOLVE_H
#define MGSOLVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define MGS_OK 0
#define MGS_E_ARG -1         /* invalid argument (size, level, colour, model, n) */
#define MGS_E_HIP -2         /* HIP runtime error */
#define MGS_E_NOMEM -3       /* device allocation failed */
#define MGS_E_STATE -4       /* call order (e.g. solve before density) */
#define MGS_E_DOMAIN -5      /* input outside the model's domain (NaN, delta <= -1) */
#define MGS_E_BUSY -6        /* serving: queue full / admission refused */

/* models (MG_model, PMP2mod_tools.f90:58-62) */
#define MGS_MODEL_FR 1
#define MGS_MODEL_DGP 2

/* precision of a solve */
#define MGS_PREC_FP64 0      /* plain: the reference's method in FP64 throughout */
#define MGS_PREC_MIXED 1     /* FP32 delta-form FAS inside an FP64 defect-correction loop */

/* element type of host/device arrays crossing the boundary */
#define MGS_F32 0            /* real(4), the reference's FI/FI2 storage */
#define MGS_F64 1            /* real(8) */

/* where a pointer lives */
#define MGS_HOST 0
#define MGS_DEVICE 1

/* per-level arrays of the FP64 hierarchy (the reference's FI3 sections,
 * PMP2MGsolver_fR.f90:76-83); level 0 u = FI2, level 0 rho = FI */
#define MGS_ARR_U 0          /* scalar field u (f(R)) / phi (DGP)      FI3 S4 */
#define MGS_ARR_RHO 1        /* density (level 0: delta)               FI3 S2 */
#define MGS_ARR_F 2          /* FAS right-hand side f_l                FI3 S5 */
#define MGS_ARR_RES 3        /* level residual                        FI3 S6 */
#define MGS_ARR_RRES 4       /* restricted residual R r               FI3 S1 */
#define MGS_ARR_RU 5         /* frozen restricted field R u           FI3 S3 */

typedef struct mgs_params {
    int32_t ngrid;           /* NGRID, cells per side at level 0 (even, <= 1024) */
    int32_t model;           /* MGS_MODEL_FR or MGS_MODEL_DGP */
    int32_t fr_n;            /* Hu-Sawicki n in {0,1,2} */
    int32_t branch;          /* DGP: +1 normal (N_branch), -1 self-accelerating (S_branch) */
    double box;              /* Box [Mpc/h] */
    double omega_m;          /* Om */
    double omega_l;          /* OmL */
    double aexpn;            /* AEXPN */
    double fr0;              /* fR0 (< 0 for Hu-Sawicki) */
    double h0rc;             /* H0rc */
    int32_t precision;       /* MGS_PREC_FP64 / MGS_PREC_MIXED */
    int32_t flags;           /* MGS_FLAG_* */
} mgs_params;

/* f(R) FP64 smoother: evaluate the reference's closed-form root literally
 * (Cardano / trigonometric / Ferrari forms, PMP2MGsolver_fR.f90:105-256).
 * Default (flag clear): the same root computed from a cancellation-free FP32
 * closed-form seed and FP64 Newton steps on u^{n+2} + P u + Q = 0, which
 * keeps full relative precision in dense cells where the literal forms
 * cancel (there the literal forms floor the solve's relative residual). */
#define MGS_FLAG_REFERENCE_ROOTS 1

typedef struct mgs_solve_opts {
    int32_t mode;            /* 0: the reference's exits (absolute RMS <= res_conv, 0.5 % stagnation,
                                iteration cap; PMP2extradof.f90:107-122); 1: relative residual */
    int32_t max_cycles;      /* iter_count_max (mode 0) / cap on V-cycles or outer iterations (mode 1) */
    double res_conv;         /* mode 0 threshold on the level-0 RMS residual */
    double rel_tol;          /* mode 1: stop when RMS(r)/RMS(r0) < rel_tol */
    int32_t init_guess;      /* 1: start from u = u_bg (f(R)) / phi = 0 (DGP), as initialise_extradof */
    int32_t inner_cycles;    /* mixed: inner FP32 V-cycles per outer FP64 iteration; 0: mode 1 picks each
                                outer iteration's count (1 to 4) from the contraction measured so far,
                                mode 0 runs 1 */
    int32_t cycle;           /* 0 V_cycle, 1 F_cycle, 2 W_cycle (the reference's Vcycle/Fcycle/Wcycle
                                switches, PMP2extradof.f90:86-88); mixed: the inner cycle's kind */
    int32_t reserved;
} mgs_solve_opts;

typedef struct mgs_solve_report {
    int32_t cycles;          /* V-cycles (plain) or inner V-cycles (mixed) performed */
    int32_t outer;           /* outer FP64 iterations (mixed), = cycles for plain */
    int32_t converged;
    int32_t reserved;
    double res0;             /* RMS residual of the initial field (FP64) */
    double res_final;        /* RMS residual of the returned field (FP64) */
    double seconds;          /* device time of the solve (hipEvent), excl. host<->device copies */
    double history[128];     /* RMS residual after each V-cycle / outer iteration */
} mgs_solve_report;

typedef struct mgs_handle mgs_handle;

/* ---- lifecycle ------------------------------------------------------ */
int mgs_create(const mgs_params* params, mgs_handle** out);
int mgs_destroy(mgs_handle* h);
const char* mgs_last_error(const mgs_handle* h);   /* also valid for h == NULL (creation errors) */
int mgs_levels(const mgs_handle* h, int32_t* levelmax, int32_t* levelmin, int32_t* coarsest);
/* change the cosmology/model scalars of an existing hierarchy (new time step) */
int mgs_set_params(mgs_handle* h, const mgs_params* params);

/* ---- data in / out -------------------------------------------------- */
/* density contrast delta (FI) in the reference's layout */
int mgs_set_density(mgs_handle* h, const void* delta, int32_t dtype, int32_t where);
/* scalar field u (FI2), un-normalised (the value the kernels work on) */
int mgs_set_field(mgs_handle* h, const void* u, int32_t dtype, int32_t where);
/* field out, multiplied by `scale` (1 for the raw field, 1/u_bg for the
 * normalised field scalar_field leaves in FI2, PMP2extradof.f90:131-146) */
int mgs_get_field(mgs_handle* h, void* u, int32_t dtype, int32_t where, double scale);
/* one per-level array (MGS_ARR_*) of the FP64 hierarchy, n_l^3 values */
int mgs_set_level_array(mgs_handle* h, int32_t level, int32_t which, const void* src, int32_t dtype, int32_t where);
int mgs_get_level_array(mgs_handle* h, int32_t level, int32_t which, void* dst, int32_t dtype, int32_t where);
/* density back out (e.g. a generated one), in the reference's layout */
int mgs_get_density(mgs_handle* h, void* delta, int32_t dtype, int32_t where);
/* u_bg = |fR_bg|^{1/(n+1)} (f(R)) */
int mgs_background(const mgs_handle* h, double* fr_bg, double* u_bg);

/* ---- single operators on the FP64 hierarchy (drop-in for the Fortran
 *      routines above; each call is one or two kernel launches) --------- */
int mgs_initialise(mgs_handle* h);                                   /* initialise_extradof */
int mgs_smooth(mgs_handle* h, int32_t level, int32_t red);          /* relaxation_iterations_* */
int mgs_residual(mgs_handle* h, int32_t level, double* rms);        /* calculate_residual_* (stores r) */
int mgs_restrict_residual(mgs_handle* h, int32_t level);            /* restrict_residual_* */
int mgs_restrict_field(mgs_handle* h, int32_t level);               /* restrict_scalar_field */
int mgs_restrict_density(mgs_handle* h, int32_t level);             /* restrict_density */
int mgs_rhs(mgs_handle* h, int32_t level);                          /* calculate_physical_right_hand_side_* */
int mgs_correct(mgs_handle* h, int32_t level);                      /* correct_solution */
int mgs_vcycle(mgs_handle* h, int32_t level);                       /* V_cycle */
int mgs_fcycle(mgs_handle* h, int32_t level);                       /* F_cycle, PMP2extradof.f90:563 */
int mgs_wcycle(mgs_handle* h, int32_t level);                       /* W_cycle, PMP2extradof.f90:867 */

/* ---- whole solve (scalar_field) -------------------------------------- */
int mgs_solve(mgs_handle* h, const mgs_solve_opts* opts, mgs_solve_report* report);

/* ---- fifth force ------------------------------------------------------ */
/* psi = -1/2 c~^2 fR_bg (u/u_bg)^{n+1} (f(R)) or phi/2 (DGP), the term
 * POTENT_FOR_MG adds to the potential, and its central-difference gradient
 * (psi[i+1]-psi[i-1])/2 per axis in grid units (the on-grid form of MOVE's
 * stencil).  Any output pointer may be NULL. */
int mgs_fifth_force(mgs_handle* h, void* psi, void* gx, void* gy, void* gz, int32_t dtype, int32_t where);
/* NGP-binned power spectrum of psi (POWERfft5 binning): nbins = N/2 bins,
 * k [h/Mpc], P [(Mpc/h)^3], count of modes */
int mgs_power_spectrum(mgs_handle* h, int32_t which, double* k, double* pk, double* count);
/* POWERfft5(iPol, .) with its k-space binning iCIC (PMP2mod_power.f90:
 * 209-463; the reference's module Power sets iCIC = 0 and dLog = 0.02):
 * which as above; ipol 0 monopole, 1 quadrupole (weights P2(mu) = 1.5 mu^2 -
 * 0.5, mu = kz/|k|, z the reference's line of sight after its transposition,
 * times 5/2); icic 0 nearest bin int(|m| + 0.5) plus the log bins
 * int(log10|m| / 0.02) (their power monopole only: zero for ipol 1), icic 1
 * linear weights over bins int(|m| + 1e-4) and the next one (no log bins:
 * zeros).  |m| and the bin rules in the reference's REAL*4 arithmetic.
 * N/2 values each: k [h/Mpc], P [(Mpc/h)^3], count of modes (summed
 * weights for icic 1); a bin without a positive count gives zeros.  klog,
 * pklog, countlog and count may be NULL. */
int mgs_power_spectrum_ex(mgs_handle* h, int32_t which, int32_t ipol, int32_t icic, double* k, double* pk,
                          double* count, double* klog, double* pklog, double* countlog);

/* ---- PM step around the solve: particles, CIC deposit, Poisson FFT, push -- */
/* Particles are the reference's REAL*4 XPAR, YPAR, ZPAR, VX, VY, VZ
 * (PMP2mod_tools.f90:46): positions in grid units in [1, N+1), cell
 * I = INT(X).  mgs_particles_set copies them into device arrays owned by the
 * handle (where = MGS_HOST or MGS_DEVICE for the source; NULL velocities are
 * zero); they stay resident across calls.  mgs_particles_get copies them out
 * (NULL outputs are skipped). */
int mgs_particles_set(mgs_handle* h, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                      const float* vz, int64_t npart, int32_t where);
int mgs_particles_get(mgs_handle* h, float* x, float* y, float* z, float* vx, float* vy, float* vz, int32_t where);
/* DENSIT's CIC deposit, PMP2mod_density.f90:235-281 (FI = -1, then each
 * particle adds WPAR = N^3/npart over its 8 nearest cells): writes the
 * handle's density contrast, which mgs_solve then uses.  FP64 accumulation.
 * MGS_E_DOMAIN if any particle lies outside [1, N+1) or is not finite (it is
 * skipped). */
int mgs_cic_deposit(mgs_handle* h);
/* POTENTfft5, PMP2main.f90:705-867: the potential phi with
 * L_d phi = trfi delta (L_d the 7-point discrete Laplacian, grid units; the
 * mean mode is multiplied by trfi), trfi = 1.5 Omega_m / a as the reference
 * sets it (:727-731).  phi is kept on the device (N^3, x fastest) in REAL*4
 * like the reference's FI, rounded once from the FP64 transforms. */
int mgs_potential(mgs_handle* h, double trfi);
/* POTENT_FOR_MG, PMP2main.f90:869-935 (f(R) and nDGP branches): phi += the
 * scalar field's potential (psi of mgs_fifth_force) from the current u,
 * summed in FP64 and rounded to REAL*4 as FI = FI + sf there. */
int mgs_potential_add_fifth_force(mgs_handle* h);
/* phi to dst (N^3 in the reference's FI order) as MGS_F32 or MGS_F64 */
int mgs_get_potential(mgs_handle* h, void* dst, int32_t dtype, int32_t where);
/* phi from src (stored as REAL*4): a potential the host holds in FI */
int mgs_set_potential(mgs_handle* h, const void* src, int32_t dtype, int32_t where);
/* MOVE, PMP2main.f90:304-685, f(R)/nDGP branch: each particle's velocity
 * gets pconst times the CIC interpolation of the doubled central-difference
 * gradients of phi at its 8 nodes, its position xconst times the new
 * velocity, with the periodic wrap and the REAL*4 edge rule of :677-679;
 * *svel = sum of v^2 WPAR (may be NULL).  pconst and xconst are the
 * reference's PCONST and XCONST (:390-393). */
int mgs_move(mgs_handle* h, double pconst, double xconst, double* svel);

/* ---- synthetic input ------------------------------------------------- */
/* seeded lognormal density delta = exp(g - sigma_g^2/2) - 1, g Gaussian with
 * P_g(k) ~ k^slope exp(-k^2/(2 k_c^2)), k_c = kcut_frac k_Nyq (Philox-4x32-10
 * white noise + FFT filter), written as the handle's density */
int mgs_make_lognormal_density(mgs_handle* h, uint64_t seed, double sigma_g, double slope, double kcut_frac);

/* ---- profiling ------------------------------------------------------- */
/* per-kernel-class device timing with hipEvents on the handle's stream */
#define MGS_KC_SMOOTH 0      /* level-0 smoother colour passes (plain FP64 or mixed FP32) */
#define MGS_KC_RESIDUAL 1    /* level-0 residual / outer residual */
#define MGS_KC_RESTRICT 2    /* level-0 -> 1 restrictions */
#define MGS_KC_CORRECT 3     /* level-1 -> 0 prolongation */
#define MGS_KC_COARSE 4      /* everything on levels >= 1 */
#define MGS_KC_UPDATE 5      /* mixed: outer update */
#define MGS_KC_PM 6          /* PM step: deposit, Poisson FFT, fifth-force potential, move */
#define MGS_KC_COUNT 7
int mgs_profile(mgs_handle* h, int32_t enable);
/* totals since enable: launches, milliseconds, algorithmic bytes */
int mgs_profile_get(mgs_handle* h, int32_t kclass, int64_t* launches, double* ms, double* bytes);

/* ---- device plumbing -------------------------------------------------- */
int mgs_synchronize(mgs_handle* h);
/* launch the coarse levels of each V-cycle from a captured hipGraph (default
 * 1) or as individual kernels (0); results are identical either way */
int mgs_use_graphs(mgs_handle* h, int32_t enable);
/* device pointer of the level-0 field in the internal colour-split layout
 * (for zero-copy consumers on the same device) */
int mgs_device_field(mgs_handle* h, void** ptr, size_t* bytes);

/* ---- whole solve with no device allocation ---------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mgs_create + mgs_set_density + mgs_solve + mgs_get_field
 * in one call that calls no hipMalloc/hipFree and no allocating library
 * routine.  delta (input) and u (input when opts->init_guess == 0, output)
 * are DEVICE arrays of ngrid^3 values in the reference's order; u comes back
 * raw (scale 1).  The hierarchy lives in a fixed device array of
 * MGS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends (enough for
 * ngrid <= 64 in every model and precision); a grid that needs more is
 * rejected with MGS_E_ARG.  Runs on the device's null stream and returns when
 * the field is in u.  Calls on one GPU share that GPU's array: calls that
 * overlap on one GPU race.  Bitwise the field of the handle-based path;
 * report->seconds is host wall time.  Errors: mgs_last_error(NULL). */
#define MGS_NOALLOC_SCRATCH_BYTES (16u << 20)
int mgs_solve_noalloc(const mgs_params* params, const mgs_solve_opts* opts, const void* delta, int32_t delta_dtype,
                      void* u, int32_t u_dtype, mgs_solve_report* report);

/* ---- serving: continuous-batching solve server ------------------------
 * No counterpart in the reference (a build-only component).  A server owns a
 * worker thread and one HIP stream on the current device.  Requests (f(R) or
 * nDGP, MGS_PREC_MIXED, ngrid >= 16; nDGP: ngrid a power of two) wait in a
 * FIFO queue and are admitted into
 * one of max_active slots when a slot is free and the request's hierarchy
 * fits in hbm_budget; admitted requests keep their hierarchies in HBM and
 * advance together: every batched step (outer residual, each level's colour
 * passes, transfers, coarsest solves, update) is one persistent launch over
 * the tiles of all active requests, replayed from a captured hipGraph.
 * Requests join at the next step after admission and leave when converged;
 * each step runs one inner V-cycle of every admitted request, and a request
 * takes its FP64 outer residual when its outer iteration's inner cycles are
 * done.  f(R) requests with MGS_EXIT_REL_TOL choose their inner cycles per
 * outer iteration as mgs_solve's mode 1 with inner_cycles 0 does; the others
 * run one.  Each request's result is therefore bit-identical alone, in any
 * mix, and to mgs_solve with the same parameters: mode 1 and inner_cycles 0
 * for f(R) with MGS_EXIT_REL_TOL, mode 1 and inner_cycles 1 for nDGP with
 * MGS_EXIT_REL_TOL, mode 0 (one inner cycle) for MGS_EXIT_REFERENCE. */
typedef struct mgs_server mgs_server;

typedef struct mgs_server_opts {
    uint64_t hbm_budget;     /* bytes for request hierarchies; 0: 60 % of the device's free memory */
    int32_t max_active;      /* requests solved together (slots), 1..64 */
    int32_t queue_capacity;  /* waiting requests; beyond it mgs_server_submit returns MGS_E_BUSY */
    int32_t use_graphs;      /* 1: replay each batch step from a captured hipGraph */
    int32_t reserved;
} mgs_server_opts;

enum { MGS_EXIT_REL_TOL = 0, MGS_EXIT_REFERENCE = 1 };

typedef struct mgs_request {
    mgs_params params;       /* precision MGS_PREC_MIXED, 16 <= ngrid <= 1024; f(R) (fr0 < 0) or nDGP
                                (ngrid a power of two) */
    const void* delta;       /* host density contrast, ngrid^3 in the reference's order; copied by submit */
    int32_t dtype;           /* MGS_F32 / MGS_F64 */
    int32_t max_outer;       /* cap on outer iterations */
    double rel_tol;          /* MGS_EXIT_REL_TOL: stop when RMS(r)/RMS(r0) < rel_tol */
    int32_t exit_rule;       /* MGS_EXIT_REL_TOL (mgs_solve's mode 1) or MGS_EXIT_REFERENCE (mode 0: res_conv
                                and the stagnation exits of PMP2extradof.f90:107-122) */
    int32_t verify_fp64;     /* 1: mgs_server_wait also solves the request with the plain FP64 method
                                (a handle of its own, same exit rule) and reports the served field's
                                difference from that solve in the result; the density stays in host
                                memory until then */
    double res_conv;         /* MGS_EXIT_REFERENCE: converged when RMS(r) <= res_conv */
} mgs_request;

typedef struct mgs_request_result {
    int32_t status;          /* 0 or the MGS_E_* code the request failed with */
    int32_t outer;           /* outer iterations */
    int32_t converged;
    int32_t verified;        /* 1 when the fp64_* fields hold a comparison (verify_fp64) */
    double res0, res_final;  /* FP64 RMS residuals under the FP64 operator: res_final / res0 is the
                                served field's relative residual, which every result reports */
    double queued_ms;        /* submit -> admission */
    double latency_ms;       /* submit -> result ready */
    double fp64_rel_linf;    /* verify_fp64: max |u - u64| / max |u64| against the plain FP64 solve */
    double fp64_rel_l2;      /* verify_fp64: ||u - u64||_2 / ||u64||_2 */
    double fp64_res_final;   /* verify_fp64: the FP64 solve's final RMS residual */
} mgs_request_result;

typedef struct mgs_server_stats {
    int64_t submitted, completed, rejected;
    int64_t steps;           /* batched outer-iteration steps executed */
    double p50_ms, p99_ms, mean_ms;   /* latency submit -> result ready in HBM, completed requests */
    double solves_per_s;     /* completed / (last completion - first submission) */
    int32_t max_batch;       /* most requests advanced by one batched step */
    int32_t reserved;
    double step_ms;          /* worker time in batched steps (launch to results on the host; admission
                                made while a step runs is inside this span) */
    double admit_ms;         /* worker time admitting requests (density upload included; while requests
                                are active it runs beside the step on a second stream) */
    double finish_ms;        /* worker time applying the stopping rule and freeing slots */
    int64_t spilled;         /* unclaimed results copied to the host to admit others */
} mgs_server_stats;

int mgs_server_create(const mgs_server_opts* opts, mgs_server** out);
int mgs_server_destroy(mgs_server* srv);
/* the calling thread's last server error message (srv may be NULL; submit
 * and wait may be called from different threads) */
const char* mgs_server_last_error(const mgs_server* srv);
/* enqueue a request; *ticket identifies it for mgs_server_wait */
int mgs_server_submit(mgs_server* srv, const mgs_request* req, int64_t* ticket);
/* block until the request is done; copy its field u (raw, x scale) to the
 * host array `u` of u_count elements (may be NULL; MGS_E_ARG, with nothing
 * written, when u_count is below the request's ngrid^3) and its report to
 * `res` (may be NULL).  A ticket can be waited for once: a second wait on it,
 * also one started while the first still blocks, returns MGS_E_STATE.  A
 * finished request's field stays in HBM (counted against hbm_budget) until
 * this call copies it out, or until admission needs the memory and copies it
 * to the host (stats.spilled).  mgs_server_destroy wakes blocked waiters,
 * which return MGS_E_STATE, and returns once none is left inside. */
int mgs_server_wait(mgs_server* srv, int64_t ticket, void* u, size_t u_count, int32_t dtype, double scale,
                    mgs_request_result* res);
/* *done = 1 when the request has finished (mgs_server_wait then returns
 * without waiting for a solve), else 0; never blocks.  Lets a client claim
 * results in the order they finish. */
int mgs_server_test(mgs_server* srv, int64_t ticket, int32_t* done);
int mgs_server_get_stats(mgs_server* srv, mgs_server_stats* st);

#ifdef __cplusplus
}
#endif
#endif
