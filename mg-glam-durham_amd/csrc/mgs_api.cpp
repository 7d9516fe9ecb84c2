// mgs_api.cpp -- host side of libmgsolve: the C-ABI of include/mgsolve.h,
// the device-resident multigrid hierarchy, V-cycle control and the plain
// FP64 / mixed-precision solve loops.
//
// Control flow follows the reference's driver:
//   scalar_field     PMP2extradof.f90:12-178   -> mgs_solve / solve_plain
//   V_cycle          PMP2extradof.f90:352-555  -> vcycle_plain
// The reference keeps every coarse level in one packed FI3 array
// (PMP2MGsolver_fR.f90:76-83); here each level owns separate colour-split
// arrays (mgs_common.h).  The level-0 residual the reference computes inside
// V_cycle is never read (restrict_residual_fR(1) recomputes it from u and
// delta), so it is not launched; the restricted densities are constant during
// a solve and are computed once per density instead of once per cycle.
// Neither change alters any value.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mgsolve.h"
#include "mgs_internal.h"
#include "mgs_kernels.h"

using namespace mgs;

static thread_local std::string g_create_error;

#define HIPCHK(h, call)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return set_err((h), MGS_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int set_err(mgs_handle* h, int code, const std::string& msg)
{
    if (h)
        h->err = msg;
    else
        g_create_error = msg;
    return code;
}

namespace mgs {

// ---------------------------------------------------------------------------
// coefficients (double, host): PMP2MGsolver_fR.f90:54-71, 333-346, 590-605
// ---------------------------------------------------------------------------
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

double fr_background(const mgs_params& p, double* ubg)
{
    const double a = p.aexpn;
    const double R_bg = 3.0 * (p.omega_m / (a * a * a) + 4.0 * p.omega_l);
    const double R0_bg = 3.0 * (p.omega_m + 4.0 * p.omega_l);
    const double fR_bg = p.fr0 * powi_d(R0_bg / R_bg, p.fr_n + 1);
    if (ubg) *ubg = std::pow(std::fabs(fR_bg), 1.0 / (1.0 + (double)p.fr_n));
    return fR_bg;
}

FrCoef fr_coefs(const mgs_params& p, int level)
{
    FrCoef c;
    const double a = p.aexpn;
    const double ctilde = 2.99792458e3 * (double)p.ngrid / p.box;
    const double dx = (double)(1 << level);
    const double dx2 = dx * dx;
    const double R_bg = 3.0 * (p.omega_m / (a * a * a) + 4.0 * p.omega_l);
    const double R0_bg = 3.0 * (p.omega_m + 4.0 * p.omega_l);
    const double fR_bg = p.fr0 * powi_d(R0_bg / R_bg, p.fr_n + 1);
    const double sgn = std::copysign(1.0, fR_bg);
    const double om_a = p.omega_m / a;
    const double a2 = a * a;
    const int n = p.fr_n;
    c.fct2s = dx2 / (ctilde * ctilde * sgn) / 6.0;
    c.fct1s = -(om_a)*c.fct2s;
    c.Q2 = R_bg * a2 * c.fct2s / 3.0;
    c.Q = c.Q2 * std::pow(std::fabs(fR_bg), 1.0 / ((double)n + 1.0));
    c.Q2onethird = std::pow(-c.Q, 1.0 / 3.0);
    c.Q4 = 4.0 * c.Q;
    c.Delta11 = 27.0 * c.Q;
    c.Delta20 = 12.0 * c.Q;
    c.fct1r = om_a;
    c.fct2r = -R_bg * a2 / 3.0;
    c.fct3r = ctilde * ctilde * sgn / dx2;
    c.a_fR_rt = std::pow(std::fabs(fR_bg), 1.0 / (double)(n + 1));
    c.fct3_1 = ctilde * ctilde * sgn / 1.0;
    c.fct4_1 = c.fct2r * std::pow(std::fabs(fR_bg), 1.0 / (1.0 + (double)n));
    c.ubg = std::pow(std::fabs(fR_bg), 1.0 / (1.0 + (double)n));
    c.inv_w = 1.0 / powi_d(c.ubg, n + 1);
    c.inv_ubg = 1.0 / c.ubg;
    c.reference_roots = (p.flags & MGS_FLAG_REFERENCE_ROOTS) ? 1 : 0;
    return c;
}

DfCoef df_coefs(const mgs_params& p, int level)
{
    const FrCoef c = fr_coefs(p, level);
    DfCoef d;
    const double A = c.fct3r * powi_d(c.ubg, p.fr_n + 1);
    const double C = -c.fct2r;
    d.A64 = A;
    d.C64 = C;
    d.A = (float)A;
    d.C = (float)C;
    d.inv6A = (float)(1.0 / (6.0 * A));
    return d;
}

// ---------------------------------------------------------------------------
// hierarchy
// ---------------------------------------------------------------------------
int levelmax_of(int N)
{
    int l = 0;
    while ((1 << l) < N) l++;
    return l;
}

int levelmin_of(int N)
{
    int odd = N;
    while (odd % 2 == 0) odd /= 2;
    int l = 0;
    while ((1 << l) < odd) l++;
    return l + 1;
}

int dev_alloc(mgs_handle* h, size_t bytes, void** out, const char* what)
{
    *out = nullptr;
    if (h->use_arena) {
        const size_t off = (h->arena_used + 255) & ~(size_t)255;
        if (h->arena && off + bytes > h->arena_cap)
            return set_err(h, MGS_E_NOMEM, std::string("scratch arena too small for ") + what);
        h->arena_used = off + bytes;
        if (h->arena) *out = h->arena + off;
        return 0;
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return set_err(h, MGS_E_NOMEM, std::string("hipMalloc ") + what + ": " + hipGetErrorString(e));
    *out = q;
    h->allocs.push_back(q);
    return 0;
}

template <typename T>
static int dalloc(mgs_handle* h, T** p, size_t count)
{
    void* q = nullptr;
    int rc = dev_alloc(h, count * sizeof(T) + 256, &q, "hierarchy");
    *p = (T*)q;
    return rc;
}

static int alloc_hierarchy(mgs_handle* h)
{
    const int N = h->p.ngrid;
    h->L.assign(h->lc + 1, LevelD());
    for (int l = 0; l <= h->lc; l++) {
        LevelD& v = h->L[l];
        v.n = N >> l;
        const size_t n3 = (size_t)v.n * v.n * v.n;
        int rc;
        if ((rc = dalloc(h, &v.u, n3))) return rc;
        if ((rc = dalloc(h, &v.rho, n3))) return rc;
        if (l > 0) {
            if ((rc = dalloc(h, &v.f, n3))) return rc;
            if ((rc = dalloc(h, &v.res, n3))) return rc;
            if ((rc = dalloc(h, &v.rres, n3))) return rc;
            if ((rc = dalloc(h, &v.ru, n3))) return rc;
        }
    }
    if (h->p.precision == MGS_PREC_MIXED) {
        h->M.assign(h->lc + 1, LevelF());
        for (int l = 0; l <= h->lc; l++) {
            LevelF& v = h->M[l];
            v.n = N >> l;
            const size_t n3 = (size_t)v.n * v.n * v.n;
            int rc;
            if ((rc = dalloc(h, &v.d, n3))) return rc;
            if ((rc = dalloc(h, &v.xb, n3))) return rc;
            if ((rc = dalloc(h, &v.f, n3))) return rc;
            if (l > 0 || h->p.model == MGS_MODEL_DGP) {
                if ((rc = dalloc(h, &v.res, n3))) return rc;
            }
            if (h->p.model == MGS_MODEL_DGP && dgp_zmarch(v.n)) {
                if ((rc = dalloc(h, &v.rho, n3))) return rc;
            }
            if (l > 0) {
                if ((rc = dalloc(h, &v.rres, n3))) return rc;
                if ((rc = dalloc(h, &v.rd, n3))) return rc;
            }
        }
    }
    // block partials of the largest reduction: 2 n^3 / 2 cells in rows of n/2
    const size_t cells = (size_t)N * N * N;
    h->npart = cells / 64 + 1024;
    int rc;
    if ((rc = dalloc(h, &h->partials, h->npart))) return rc;
    if ((rc = dalloc(h, &h->dsum, 16))) return rc;
    if (h->use_arena) {
        h->hsum = h->hsum_host;
    } else {
        HIPCHK(h, hipHostMalloc((void**)&h->hsum, 16 * sizeof(double), hipHostMallocDefault));
        HIPCHK(h, hipEventCreateWithFlags(&h->ev_norm, hipEventDisableTiming));
    }
    return 0;
}

int ensure_res0(mgs_handle* h)
{
    if (h->L[0].res) return 0;
    const size_t n3 = (size_t)h->p.ngrid * h->p.ngrid * h->p.ngrid;
    return dalloc(h, &h->L[0].res, n3);
}

int ensure_staging(mgs_handle* h, size_t bytes)
{
    if (h->staging_bytes >= bytes) return 0;
    if (h->staging) {
        (void)hipFree(h->staging);
        h->staging = nullptr;
        h->staging_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->staging, bytes);
    if (e != hipSuccess) return set_err(h, MGS_E_NOMEM, std::string("hipMalloc staging: ") + hipGetErrorString(e));
    h->staging_bytes = bytes;
    return 0;
}

// ---------------------------------------------------------------------------
// profiling
// ---------------------------------------------------------------------------
// a launch whose events cannot be created or recorded goes unmeasured
void prof_begin(mgs_handle* h, int kc, double bytes)
{
    if (!h->prof) return;
    ProfRec r;
    r.kc = kc;
    r.bytes = bytes;
    r.a = r.b = nullptr;
    r.ok = hipEventCreate(&r.a) == hipSuccess && hipEventCreate(&r.b) == hipSuccess &&
           hipEventRecord(r.a, h->s) == hipSuccess;
    h->recs.push_back(r);
}

void prof_end(mgs_handle* h)
{
    if (!h->prof || h->recs.empty()) return;
    ProfRec& r = h->recs.back();
    r.ok = r.ok && hipEventRecord(r.b, h->s) == hipSuccess;
}

static void prof_flush(mgs_handle* h)
{
    if (h->recs.empty()) return;
    const bool synced = hipStreamSynchronize(h->s) == hipSuccess;
    for (auto& r : h->recs) {
        float ms = 0.0f;
        if (synced && r.ok && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            h->pc[r.kc].launches += 1;
            h->pc[r.kc].ms += ms;
            h->pc[r.kc].bytes += r.bytes;
        }
        if (r.a) (void)hipEventDestroy(r.a);
        if (r.b) (void)hipEventDestroy(r.b);
    }
    h->recs.clear();
}

// ---------------------------------------------------------------------------
// plain FP64 operators on the hierarchy
// ---------------------------------------------------------------------------
static double n3d(int n) { return (double)n * n * n; }

int op_smooth(mgs_handle* h, int l, int red)
{
    LevelD& v = h->L[l];
    const int colour = red ? 0 : 1;
    const double s = 8.0;
    if (l == 0) prof_begin(h, MGS_KC_SMOOTH, 1.5 * n3d(v.n) * s);
    else prof_begin(h, MGS_KC_COARSE, 2.0 * n3d(v.n) * s);
    if (h->p.model == MGS_MODEL_FR)
        launch_fr_smooth64(h->p.fr_n, l, v.u, v.rho, v.f, v.n, colour, fr_coefs(h->p, l), h->s);
    else
        launch_dgp_smooth64(l, v.u, v.rho, v.f, h->dgp_tmp, v.n, colour, dgp_coefs(h->p, l), h->s);
    prof_end(h);
    return 0;
}

// Nsmth = 2 red-black pairs (colour 0, then 1) on level l.  nDGP levels with
// z-marching passes ping-pong between u and the second array: after two
// pairs the field is back in u.
static void op_smooth2(mgs_handle* h, int l)
{
    LevelD& v = h->L[l];
    if (h->p.model == MGS_MODEL_DGP && h->dgp_spare64 && dgp_zmarch(v.n)) {
        const DgpCoef c = dgp_coefs(h->p, l);
        const size_t half = (size_t)v.n * v.n * (v.n / 2);
        double* A = v.u;
        double* B = h->dgp_spare64;
        for (int s = 0; s < 2; s++) {
            for (int colour = 0; colour < 2; colour++) {
                if (l == 0) prof_begin(h, MGS_KC_SMOOTH, 2.0 * n3d(v.n) * 8.0);
                else prof_begin(h, MGS_KC_COARSE, 2.0 * n3d(v.n) * 8.0);
                launch_dgp_zpass64(l, colour ? B : A, A + half, B + (size_t)colour * half, v.rho, v.f, v.n, colour, c,
                                   h->s);
                prof_end(h);
            }
            std::swap(A, B);
        }
        return;
    }
    for (int s = 0; s < 2; s++) {
        op_smooth(h, l, 1);
        op_smooth(h, l, 0);
    }
}

// level-l residual, stored in res_l when store; rms returned when rms != null
int op_residual(mgs_handle* h, int l, bool store, double* rms)
{
    LevelD& v = h->L[l];
    if (store && l == 0) {
        int rc = ensure_res0(h);
        if (rc) return rc;
    }
    double* res = store ? v.res : nullptr;
    const double s = 8.0;
    prof_begin(h, l == 0 ? MGS_KC_RESIDUAL : MGS_KC_COARSE, (store ? 3.0 : 2.0) * n3d(v.n) * s);
    int nb;
    if (h->p.model == MGS_MODEL_FR)
        nb = launch_fr_residual64(h->p.fr_n, l, v.u, v.rho, v.f, res, rms ? h->partials : nullptr, v.n,
                                  fr_coefs(h->p, l), h->s);
    else
        nb = launch_dgp_residual64(l, v.u, v.rho, v.f, res, rms ? h->partials : nullptr, v.n, dgp_coefs(h->p, l),
                                   h->dgp_rms, h->s);
    prof_end(h);
    if (rms) {
        launch_sum_partials(h->partials, nb, h->dsum, h->s);
        HIPCHK(h, hipMemcpyAsync(h->hsum, h->dsum, sizeof(double), hipMemcpyDeviceToHost, h->s));
        HIPCHK(h, hipStreamSynchronize(h->s));
        *rms = std::sqrt(h->hsum[0] / n3d(v.n));
    }
    return 0;
}

int op_restrict_residual(mgs_handle* h, int l)
{
    LevelD& c = h->L[l];
    const LevelD& f = h->L[l - 1];
    if (l == 1) {
        prof_begin(h, MGS_KC_RESTRICT, (2.0 * n3d(f.n) + n3d(c.n)) * 8.0);
        if (h->p.model == MGS_MODEL_FR)
            launch_fr_restrict_residual1_64(h->p.fr_n, f.u, f.rho, c.rres, nullptr, nullptr, f.n, fr_coefs(h->p, 0),
                                            h->s);
        else
            launch_dgp_restrict_residual1_64(f.u, f.rho, c.rres, f.n, dgp_coefs(h->p, 0), h->s);
    } else {
        prof_begin(h, MGS_KC_COARSE, (n3d(f.n) + n3d(c.n)) * 8.0);
        launch_restrict8_64(f.res, c.rres, nullptr, c.n, h->s);
    }
    prof_end(h);
    return 0;
}

int op_restrict_field(mgs_handle* h, int l)
{
    LevelD& c = h->L[l];
    const LevelD& f = h->L[l - 1];
    prof_begin(h, l == 1 ? MGS_KC_RESTRICT : MGS_KC_COARSE, (n3d(f.n) + 2.0 * n3d(c.n)) * 8.0);
    launch_restrict8_64(f.u, c.ru, c.u, c.n, h->s);
    prof_end(h);
    return 0;
}

// restrict_residual then restrict_scalar_field onto level l, as a cycle runs
// them; for f(R) onto level 1 one pass reads u once for both
static int op_restrict_both(mgs_handle* h, int l)
{
    if (l == 1 && h->p.model == MGS_MODEL_FR) {
        LevelD& c = h->L[l];
        const LevelD& f = h->L[l - 1];
        prof_begin(h, MGS_KC_RESTRICT, (2.0 * n3d(f.n) + 3.0 * n3d(c.n)) * 8.0);
        launch_fr_restrict_residual1_64(h->p.fr_n, f.u, f.rho, c.rres, c.ru, c.u, f.n, fr_coefs(h->p, 0), h->s);
        prof_end(h);
        return 0;
    }
    op_restrict_residual(h, l);
    return op_restrict_field(h, l);
}

int op_restrict_density(mgs_handle* h, int l)
{
    LevelD& c = h->L[l];
    const LevelD& f = h->L[l - 1];
    prof_begin(h, l == 1 ? MGS_KC_RESTRICT : MGS_KC_COARSE, (n3d(f.n) + n3d(c.n)) * 8.0);
    launch_restrict8_64(f.rho, c.rho, nullptr, c.n, h->s);
    prof_end(h);
    return 0;
}

int op_rhs(mgs_handle* h, int l)
{
    LevelD& v = h->L[l];
    prof_begin(h, MGS_KC_COARSE, 4.0 * n3d(v.n) * 8.0);
    if (h->p.model == MGS_MODEL_FR)
        launch_fr_rhs64(h->p.fr_n, v.ru, v.rho, v.rres, v.f, v.n, fr_coefs(h->p, l), h->s);
    else
        launch_dgp_rhs64(v.ru, v.rho, v.rres, v.f, v.n, dgp_coefs(h->p, l), h->s);
    prof_end(h);
    return 0;
}

// in_cycle: the correction of a V/F/W cycle, which a colour-0 relaxation
// pass of level l-1 follows; for f(R) only the colour-1 cells are corrected
// then (the pass overwrites colour 0 without reading it).  mgs_correct, the
// reference's standalone correct_solution, corrects every cell.
int op_correct(mgs_handle* h, int l, bool in_cycle)
{
    LevelD& c = h->L[l];
    LevelD& f = h->L[l - 1];
    const bool fr = h->p.model == MGS_MODEL_FR;
    const bool only1 = fr && in_cycle;
    prof_begin(h, l == 1 ? MGS_KC_CORRECT : MGS_KC_COARSE, ((only1 ? 1.0 : 2.0) * n3d(f.n) + 2.0 * n3d(c.n)) * 8.0);
    launch_correct64(c.u, c.ru, f.u, c.n, fr ? 1 : 0, only1, h->s);
    prof_end(h);
    return 0;
}

static int cache_densities(mgs_handle* h)
{
    if (h->rho_cached) return 0;
    for (int l = 1; l <= h->lc; l++) op_restrict_density(h, l);
    // the mixed nDGP kernels' FP32 densities
    for (size_t l = 0; l < h->M.size(); l++)
        if (h->M[l].rho) launch_round_f32(h->L[l].rho, h->M[l].rho, (size_t)n3d(h->M[l].n), h->s);
    h->rho_cached = true;
    return 0;
}

// Run `body` (launches only, no host synchronisation) from a hipGraph: captured
// on the first call, replayed afterwards.  The coarse levels of a V-cycle are
// ~10 small launches per level; one graph launch replaces them.
template <typename F>
static int run_captured(mgs_handle* h, hipGraphExec_t* exec, F body, int kc = MGS_KC_COARSE)
{
    if (!h->use_graphs) return body();
    if (!*exec) {
        const bool prof = h->prof;
        h->prof = false;
        hipGraph_t g = nullptr;
        HIPCHK(h, hipStreamBeginCapture(h->s, hipStreamCaptureModeThreadLocal));
        const int rc = body();
        hipError_t e = hipStreamEndCapture(h->s, &g);
        h->prof = prof;
        if (rc) return rc;
        if (e != hipSuccess) return set_err(h, MGS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_err(h, MGS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    prof_begin(h, kc, 0.0);
    HIPCHK(h, hipGraphLaunch(*exec, h->s));
    prof_end(h);
    return 0;
}

void drop_graphs(mgs_handle* h)
{
    if (h->g_plain) (void)hipGraphExecDestroy(h->g_plain);
    if (h->g_inner) (void)hipGraphExecDestroy(h->g_inner);
    if (h->g_rbase) (void)hipGraphExecDestroy(h->g_rbase);
    h->g_plain = nullptr;
    h->g_inner = nullptr;
    h->g_rbase = nullptr;
}

// V_cycle(ilevel), PMP2extradof.f90:352-555 (Nsmth = 2, :361)
int vcycle_plain(mgs_handle* h, int l)
{
    if (l >= h->lc) {
        op_smooth2(h, l);
        return 0;
    }
    op_smooth2(h, l);
    if (l > 0) op_residual(h, l, true, nullptr);
    op_restrict_both(h, l + 1);
    if (l == 0) {
        int rc = run_captured(h, &h->g_plain, [h]() {
            op_rhs(h, 1);
            return vcycle_plain(h, 1);
        });
        if (rc) return rc;
    } else {
        op_rhs(h, l + 1);
        vcycle_plain(h, l + 1);
    }
    op_correct(h, l + 1, true);
    op_smooth2(h, l);
    return 0;
}

// coarse-grid correction from level l with a `kind` cycle (0 V, 1 F, 2 W) on
// level l+1: the sequence V_cycle, F_cycle and W_cycle share
// (PMP2extradof.f90:418-545, 650-840, 935-1140)
static int cycle_plain(mgs_handle* h, int l, int kind);

static int cgc_plain(mgs_handle* h, int l, int kind)
{
    if (l > 0) op_residual(h, l, true, nullptr);
    op_restrict_both(h, l + 1);
    op_rhs(h, l + 1);
    int rc = cycle_plain(h, l + 1, kind);
    if (rc) return rc;
    op_correct(h, l + 1, true);
    return 0;
}

static void smooth2_plain(mgs_handle* h, int l) { op_smooth2(h, l); }

// F_cycle (PMP2extradof.f90:563-858): S2, CGC(F), S2, CGC(V), S2;
// W_cycle (:867-1157): S2, CGC(W), S2, CGC(W), S2; S2 only on the coarsest
static int cycle_plain(mgs_handle* h, int l, int kind)
{
    if (kind == 0) return vcycle_plain(h, l);
    int rc;
    smooth2_plain(h, l);
    if (l >= h->lc) return 0;
    if ((rc = cgc_plain(h, l, kind))) return rc;
    smooth2_plain(h, l);
    if ((rc = cgc_plain(h, l, kind == 1 ? 0 : 2))) return rc;
    smooth2_plain(h, l);
    return 0;
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 delta-form FAS V-cycle (inner) and FP64 outer loop
// ---------------------------------------------------------------------------
// One inner V-cycle at level l: V(sweeps0) on level 0, V(sweeps1) below.  On
// level 0, `first` means colour 0 of d was set by the fused outer residual
// (relaxed once from zero: that pass is skipped, and colour 1, not written
// there, is set by the colour-1 pass that follows) and `last` fuses the final
// colour-1 pass with the outer update of u.
static int vcycle_inner(mgs_handle* h, int l, bool first = false, bool last = false)
{
    const int frn = h->p.fr_n;
    LevelF& v = h->M[l];
    const DfCoef dc = df_coefs(h->p, l);
    const int cls = l == 0 ? MGS_KC_SMOOTH : MGS_KC_COARSE;
    if (l >= h->lc) {
        prof_begin(h, MGS_KC_COARSE, 3.0 * n3d(v.n) * 4.0);
        launch_df_coarsest(frn, v.d, v.xb, v.f, v.n, 8, dc, h->s);
        prof_end(h);
        return 0;
    }
    if (l > 0 && v.n <= h->tail_n && h->use_tail) {
        // the small levels in one launch
        TailLevelIn lv[16];
        const int nlev = h->lc - l + 1;
        for (int i = 0; i < nlev && i < 16; i++) {
            const LevelF& m = h->M[l + i];
            lv[i] = TailLevelIn{m.n, m.d, m.xb, m.f, m.rres, m.rd, df_coefs(h->p, l + i)};
        }
        prof_begin(h, MGS_KC_COARSE, 12.0 * n3d(v.n) * 4.0);
        const bool ok = launch_df_tail(frn, nlev, lv, h->s);
        prof_end(h);
        if (ok) return 0;
    }
    const double s = 4.0;
    const int npre = l == 0 ? h->sweeps0[0] : h->sweeps1[0];
    for (int it = (l == 0 && first) ? 1 : 0; it < 2 * npre; it++) {
        prof_begin(h, cls, 2.0 * n3d(v.n) * s);
        launch_df_smooth(frn, v.d, v.xb, v.f, v.n, (it & 1) ? 1 : 0, dc, h->s);
        prof_end(h);
    }
    LevelF& c = h->M[l + 1];
    prof_begin(h, l == 0 ? MGS_KC_RESTRICT : MGS_KC_COARSE, (3.0 * n3d(v.n) + 3.0 * n3d(c.n)) * s);
    launch_df_restrict_residual(frn, v.d, v.xb, v.f, c.rres, c.rd, c.d, v.n, dc, h->s);
    prof_end(h);
    auto coarse = [h, l, frn, &c]() {
        prof_begin(h, MGS_KC_COARSE, 4.0 * n3d(c.n) * 4.0);
        launch_df_rhs(frn, c.rd, c.xb, c.rres, c.f, c.n, df_coefs(h->p, l + 1), h->s);
        prof_end(h);
        return vcycle_inner(h, l + 1);
    };
    if (l == 0) {
        int rc = run_captured(h, &h->g_inner, coarse);
        if (rc) return rc;
    } else {
        coarse();
    }
    prof_begin(h, l == 0 ? MGS_KC_CORRECT : MGS_KC_COARSE, (3.0 * n3d(v.n) + 2.0 * n3d(c.n)) * s);
    launch_df_correct(frn, c.d, c.rd, v.d, v.xb, c.n, h->s);
    prof_end(h);
    const int nps = l == 0 ? h->sweeps0[1] : h->sweeps1[1];
    const int npost = (l == 0 && last) ? 2 * nps - 1 : 2 * nps;
    for (int it = 0; it < npost; it++) {
        prof_begin(h, cls, 2.0 * n3d(v.n) * s);
        launch_df_smooth(frn, v.d, v.xb, v.f, v.n, (it & 1) ? 1 : 0, dc, h->s);
        prof_end(h);
    }
    if (l == 0 && last) {
        // reads d (colour 0), xb, g (colour 1), reads and writes u: 24 B/cell
        prof_begin(h, MGS_KC_UPDATE, 24.0 * n3d(v.n));
        launch_df_smooth_update(frn, v.d, v.xb, v.f, h->L[0].u, v.n, dc, fr_coefs(h->p, 0).ubg, h->s);
        prof_end(h);
    }
    return 0;
}

// FP64 outer residual; also writes the inner problem's g, xb and colour 0 of
// d after its first colour-0 pass (reads u, delta; writes g, xb, d: 26 B/cell)
// norm_only: the residual norm without the next inner problem's data
// The inner problem's base x-bar (f(R)) or phi_b (nDGP), written on level 0
// by the outer residual, restricted to every level: one captured graph
static int restrict_base_all(mgs_handle* h)
{
    const bool fr = h->p.model == MGS_MODEL_FR;
    return run_captured(
        h, &h->g_rbase,
        [h, fr]() {
            for (int l = 1; l <= h->lc; l++) {
                prof_begin(h, l == 1 ? MGS_KC_RESTRICT : MGS_KC_COARSE,
                           (n3d(h->M[l - 1].n) + n3d(h->M[l].n)) * 4.0);
                if (fr)
                    launch_restrict_base(h->p.fr_n, h->M[l - 1].xb, h->M[l].xb, h->M[l].n, h->s);
                else
                    launch_restrict8_32(h->M[l - 1].xb, h->M[l].xb, nullptr, h->M[l].n, h->s);
                prof_end(h);
            }
            return 0;
        },
        MGS_KC_RESTRICT);
}

// The outer norm's partials summed and copied to the host; with_base, the
// base restriction (which only needs the level-0 base just written) is
// enqueued behind the copy and the host waits for the copy alone, so that
// the restriction runs while the host reads the norm and picks the next
// inner count (an inner cycle, queued next, runs after it in stream order)
static int finish_outer_norm(mgs_handle* h, int nb, bool with_base, double* rms)
{
    launch_sum_partials(h->partials, nb, h->dsum, h->s);
    HIPCHK(h, hipMemcpyAsync(h->hsum, h->dsum, sizeof(double), hipMemcpyDeviceToHost, h->s));
    if (with_base && h->ev_norm) {
        HIPCHK(h, hipEventRecord(h->ev_norm, h->s));
        int rc = restrict_base_all(h);
        if (rc) return rc;
        HIPCHK(h, hipEventSynchronize(h->ev_norm));
    } else {
        if (with_base) {
            int rc = restrict_base_all(h);
            if (rc) return rc;
        }
        HIPCHK(h, hipStreamSynchronize(h->s));
    }
    *rms = std::sqrt(h->hsum[0] / n3d(h->L[0].n));
    return 0;
}

// f(R) outer residual; unless norm_only it also writes the inner problem's
// data and restricts its base to every level
static int outer_residual(mgs_handle* h, double* rms, bool norm_only = false)
{
    LevelD& v = h->L[0];
    LevelF& m = h->M[0];
    prof_begin(h, MGS_KC_RESIDUAL, (norm_only ? 16.0 : 26.0) * n3d(v.n));
    const int nb = launch_outer_residual(h->p.fr_n, v.u, v.rho, m.f, m.xb, m.d, h->partials, v.n, fr_coefs(h->p, 0),
                                         df_coefs(h->p, 0), norm_only, h->s);
    prof_end(h);
    return finish_outer_norm(h, nb, !norm_only, rms);
}

// the reference's exits, PMP2extradof.f90:107-122
bool reference_exit(const mgs_solve_opts& o, int iter, double res, double res1, double res2, bool* conv)
{
    bool can_exit = false;
    if (res <= o.res_conv) {
        can_exit = true;
        *conv = true;
    }
    if (iter > 2) {
        if (std::fabs(res / res1 - 1.0) <= 5.0e-3) can_exit = true;
        if (std::fabs(res / res1 - 1.0) <= 1.0e-2 && std::fabs(res1 / res2 - 1.0) <= 1.0e-2) can_exit = true;
    }
    if (iter >= o.max_cycles) can_exit = true;
    return can_exit;
}

static int solve_plain(mgs_handle* h, const mgs_solve_opts& o, mgs_solve_report* rep)
{
    int rc;
    if ((rc = cache_densities(h))) return rc;
    double res0 = 0.0;
    if ((rc = op_residual(h, 0, false, &res0))) return rc;
    rep->res0 = res0;
    double res_old_1 = 1.0e5, res_old_2 = 1.0e10;
    int iter = 1;
    bool conv = false;
    for (;;) {
        if ((rc = cycle_plain(h, 0, o.cycle))) return rc;
        double res = 0.0;
        if ((rc = op_residual(h, 0, false, &res))) return rc;
        if (iter <= 128) rep->history[iter - 1] = res;
        rep->cycles = iter;
        rep->outer = iter;
        rep->res_final = res;
        if (!std::isfinite(res)) return set_err(h, MGS_E_DOMAIN, "residual is not finite");
        bool can_exit;
        if (o.mode == 0) {
            can_exit = reference_exit(o, iter, res, res_old_1, res_old_2, &conv);
        } else {
            conv = res < o.rel_tol * res0;
            can_exit = conv || iter >= o.max_cycles;
        }
        if (can_exit) break;
        iter++;
        res_old_2 = res_old_1;
        res_old_1 = res;
    }
    rep->converged = conv ? 1 : 0;
    return 0;
}

// ---- nDGP: inner V-cycle on the FP32 increment e (see mgs_dgp.hip) ----------
// 2 x Nsmth colour passes (0, 1, 0, 1) of the inner smoother on level l;
// z-marching levels ping-pong between e and the scratch array and end in e
static void dgpm_smooth4(mgs_handle* h, int l, int cls)
{
    LevelF& v = h->M[l];
    LevelD& V = h->L[l];
    const DgpCoef dc = dgp_coefs(h->p, l);
    float* tmp = (float*)h->dgp_tmp;
    const double bytes = 3.0 * n3d(v.n) * 4.0;
    if (dgp_zmarch(v.n) && v.rho) {
        const size_t half = (size_t)v.n * v.n * (v.n / 2);
        float* A = v.d;
        float* B = tmp;  // holds a level-0 array of floats
        for (int s = 0; s < 2; s++) {
            for (int colour = 0; colour < 2; colour++) {
                prof_begin(h, cls, bytes);
                launch_dgpm_zpass(colour ? B : A, A + half, B + (size_t)colour * half, v.xb, v.rho, v.f, v.n, colour,
                                  dc, h->s);
                prof_end(h);
            }
            std::swap(A, B);
        }
        return;
    }
    for (int it = 0; it < 4; it++) {
        prof_begin(h, cls, bytes);
        launch_dgpm_smooth(v.d, v.xb, V.rho, v.f, tmp, v.n, (it & 1) ? 1 : 0, dc, h->s);
        prof_end(h);
    }
}

static int vcycle_inner_dgp(mgs_handle* h, int l)
{
    LevelF& v = h->M[l];
    LevelD& V = h->L[l];
    const DgpCoef dc = dgp_coefs(h->p, l);
    float* tmp = (float*)h->dgp_tmp;
    const double s = 4.0;
    const int cls = l == 0 ? MGS_KC_SMOOTH : MGS_KC_COARSE;
    if (l >= h->lc) {
        prof_begin(h, MGS_KC_COARSE, 4.0 * n3d(v.n) * s);
        launch_dgpm_coarsest(v.d, v.xb, V.rho, v.f, tmp, v.n, 8, dc, h->s);
        prof_end(h);
        return 0;
    }
    dgpm_smooth4(h, l, cls);
    LevelF& c = h->M[l + 1];
    prof_begin(h, l == 0 ? MGS_KC_RESTRICT : MGS_KC_COARSE, 6.0 * n3d(v.n) * s);
    launch_dgpm_residual(v.d, v.xb, V.rho, v.rho, v.f, v.res, v.n, dc, h->s);
    launch_restrict8_32(v.res, c.rres, nullptr, c.n, h->s);
    launch_restrict8_32(v.d, c.rd, c.d, c.n, h->s);
    prof_end(h);
    auto coarse = [h, l, &c]() {
        prof_begin(h, MGS_KC_COARSE, 5.0 * n3d(c.n) * 4.0);
        launch_dgpm_residual(c.rd, c.xb, h->L[l + 1].rho, c.rho, c.rres, c.f, c.n, dgp_coefs(h->p, l + 1), h->s);
        prof_end(h);
        return vcycle_inner_dgp(h, l + 1);
    };
    if (l == 0) {
        int rc = run_captured(h, &h->g_inner, coarse);
        if (rc) return rc;
    } else {
        coarse();
    }
    prof_begin(h, l == 0 ? MGS_KC_CORRECT : MGS_KC_COARSE, (3.0 * n3d(v.n) + 2.0 * n3d(c.n)) * s);
    launch_dgpm_correct(c.d, c.rd, v.d, c.n, h->s);
    prof_end(h);
    dgpm_smooth4(h, l, cls);
    return 0;
}

// ---------------------------------------------------------------------------
// mixed precision: inner F- and W-cycles (mgs_solve_opts.cycle 1 and 2) on
// the delta-form problem, in the order of the reference's F_cycle and
// W_cycle (PMP2extradof.f90:563-858, 867-1157): smoothing, a coarse-grid
// correction by an F- (W-) cycle, smoothing, a correction by a V- (W-)
// cycle, smoothing; the coarsest level only its solve.  f(R) smoothing
// blocks: the V-cycle's pre-sweeps first and its post-sweeps after each
// correction (level 0: sweeps0, the first colour-0 pass skipped after the
// outer residual and the last colour-1 pass fused with the update; below:
// sweeps1); nDGP: two pairs each, as its V-cycle.  The levels are launched
// one by one (the captured graph and the single-workgroup tail follow the
// V-cycle's shape); corrections below a V-correction use the V-cycle itself.
// ---------------------------------------------------------------------------
static int fw_inner(mgs_handle* h, int l, int kind, bool first, bool last);

// `pairs` colour-0/colour-1 pairs of the f(R) inner smoother on level l
static void df_smooth_pairs(mgs_handle* h, int l, int pairs, bool skip_first, bool fuse_last)
{
    const int frn = h->p.fr_n;
    LevelF& v = h->M[l];
    const DfCoef dc = df_coefs(h->p, l);
    const int cls = l == 0 ? MGS_KC_SMOOTH : MGS_KC_COARSE;
    const int n = 2 * pairs - (fuse_last ? 1 : 0);
    for (int it = skip_first ? 1 : 0; it < n; it++) {
        prof_begin(h, cls, 2.0 * n3d(v.n) * 4.0);
        launch_df_smooth(frn, v.d, v.xb, v.f, v.n, (it & 1) ? 1 : 0, dc, h->s);
        prof_end(h);
    }
    if (fuse_last) {
        prof_begin(h, MGS_KC_UPDATE, 24.0 * n3d(v.n));
        launch_df_smooth_update(frn, v.d, v.xb, v.f, h->L[0].u, v.n, dc, fr_coefs(h->p, 0).ubg, h->s);
        prof_end(h);
    }
}

// f(R) coarse-grid correction of level l by a `kind` cycle on level l+1
static int df_cgc(mgs_handle* h, int l, int kind)
{
    const int frn = h->p.fr_n;
    LevelF& v = h->M[l];
    LevelF& c = h->M[l + 1];
    const int cls = l == 0 ? MGS_KC_RESTRICT : MGS_KC_COARSE;
    prof_begin(h, cls, (3.0 * n3d(v.n) + 3.0 * n3d(c.n)) * 4.0);
    launch_df_restrict_residual(frn, v.d, v.xb, v.f, c.rres, c.rd, c.d, v.n, df_coefs(h->p, l), h->s);
    prof_end(h);
    prof_begin(h, MGS_KC_COARSE, 4.0 * n3d(c.n) * 4.0);
    launch_df_rhs(frn, c.rd, c.xb, c.rres, c.f, c.n, df_coefs(h->p, l + 1), h->s);
    prof_end(h);
    const int rc = kind == 0 ? vcycle_inner(h, l + 1) : fw_inner(h, l + 1, kind, false, false);
    if (rc) return rc;
    prof_begin(h, l == 0 ? MGS_KC_CORRECT : MGS_KC_COARSE, (3.0 * n3d(v.n) + 2.0 * n3d(c.n)) * 4.0);
    launch_df_correct(frn, c.d, c.rd, v.d, v.xb, c.n, h->s);
    prof_end(h);
    return 0;
}

static int fw_inner(mgs_handle* h, int l, int kind, bool first, bool last)
{
    if (l >= h->lc) return vcycle_inner(h, l);  // the coarsest solve
    const int pre = l == 0 ? h->sweeps0[0] : h->sweeps1[0];
    const int post = l == 0 ? h->sweeps0[1] : h->sweeps1[1];
    int rc;
    df_smooth_pairs(h, l, pre, l == 0 && first, false);
    if ((rc = df_cgc(h, l, kind == 1 ? 1 : 2))) return rc;
    df_smooth_pairs(h, l, post, false, false);
    if ((rc = df_cgc(h, l, kind == 1 ? 0 : 2))) return rc;
    df_smooth_pairs(h, l, post, false, l == 0 && last);
    return 0;
}

static int fw_inner_dgp(mgs_handle* h, int l, int kind);

// nDGP coarse-grid correction of level l by a `kind` cycle on level l+1
// (vcycle_inner_dgp's transfers)
static int dgp_cgc(mgs_handle* h, int l, int kind)
{
    LevelF& v = h->M[l];
    LevelD& V = h->L[l];
    LevelF& c = h->M[l + 1];
    prof_begin(h, l == 0 ? MGS_KC_RESTRICT : MGS_KC_COARSE, 6.0 * n3d(v.n) * 4.0);
    launch_dgpm_residual(v.d, v.xb, V.rho, v.rho, v.f, v.res, v.n, dgp_coefs(h->p, l), h->s);
    launch_restrict8_32(v.res, c.rres, nullptr, c.n, h->s);
    launch_restrict8_32(v.d, c.rd, c.d, c.n, h->s);
    prof_end(h);
    prof_begin(h, MGS_KC_COARSE, 5.0 * n3d(c.n) * 4.0);
    launch_dgpm_residual(c.rd, c.xb, h->L[l + 1].rho, c.rho, c.rres, c.f, c.n, dgp_coefs(h->p, l + 1), h->s);
    prof_end(h);
    const int rc = kind == 0 ? vcycle_inner_dgp(h, l + 1) : fw_inner_dgp(h, l + 1, kind);
    if (rc) return rc;
    prof_begin(h, l == 0 ? MGS_KC_CORRECT : MGS_KC_COARSE, (3.0 * n3d(v.n) + 2.0 * n3d(c.n)) * 4.0);
    launch_dgpm_correct(c.d, c.rd, v.d, c.n, h->s);
    prof_end(h);
    return 0;
}

static int fw_inner_dgp(mgs_handle* h, int l, int kind)
{
    if (l >= h->lc) return vcycle_inner_dgp(h, l);  // the coarsest solve
    const int cls = l == 0 ? MGS_KC_SMOOTH : MGS_KC_COARSE;
    int rc;
    dgpm_smooth4(h, l, cls);
    if ((rc = dgp_cgc(h, l, kind == 1 ? 1 : 2))) return rc;
    dgpm_smooth4(h, l, cls);
    if ((rc = dgp_cgc(h, l, kind == 1 ? 0 : 2))) return rc;
    dgpm_smooth4(h, l, cls);
    return 0;
}

static int outer_residual_dgp(mgs_handle* h, double* rms)
{
    LevelD& v = h->L[0];
    LevelF& m = h->M[0];
    prof_begin(h, MGS_KC_RESIDUAL, 28.0 * n3d(v.n));
    const int nb = launch_dgpm_outer(v.u, v.rho, m.f, m.xb, m.d, h->partials, v.n, dgp_coefs(h->p, 0), h->s);
    prof_end(h);
    return finish_outer_norm(h, nb, true, rms);
}

// Mode 1 with inner_cycles 0: the inner cycles of the next outer iteration.
// An inner cycle contracts the FP64 residual by a nearly constant factor rho
// until the FP32 increment's accuracy is reached (a gain of 1e-5 to 1e-6 per
// outer iteration, hence the cap), so the count is the one that reaches 0.7
// of the target at the rho measured over the previous outer iteration (the
// next one contracts up to 1.5 times less per cycle), between 1 and the cap;
// the first outer iteration runs the cap.  Fewer outer iterations save their
// FP64 residual, update and base restriction, about half an inner cycle each.
// *reaches: the count is expected to reach the target (not capped).
int auto_inner_count(int inner_max, double res, double target, double res_prev, int m_prev, bool* reaches)
{
    *reaches = false;
    if (m_prev <= 0 || !(res_prev > 0.0) || !(res > 0.0) || !(target > 0.0)) return inner_max;
    const double rho = std::pow(res / res_prev, 1.0 / m_prev);
    if (!(rho > 0.0) || !(rho < 0.9)) return inner_max;
    const double m = std::ceil(std::log(0.7 * target / res) / std::log(rho));
    if (m > (double)inner_max) return inner_max;
    *reaches = true;
    return m >= 1.0 ? (int)m : 1;
}

static int auto_inner(const mgs_handle* h, double res, double target, double res_prev, int m_prev, bool* reaches)
{
    return auto_inner_count(h->inner_max, res, target, res_prev, m_prev, reaches);
}

// FP64 defect correction around the FP32 inner FAS (f(R) and nDGP)
static int solve_mixed(mgs_handle* h, const mgs_solve_opts& o, mgs_solve_report* rep)
{
    int rc;
    const bool fr = h->p.model == MGS_MODEL_FR;
    const bool automatic = o.inner_cycles == 0 && o.mode == 1;
    int inner = o.inner_cycles > 0 ? o.inner_cycles : 1;
    if (!fr && (rc = cache_densities(h))) return rc;
    double res_old_1 = 1.0e5, res_old_2 = 1.0e10;
    double res_prev = 0.0;
    int m_prev = 0;
    bool conv = false;
    int cycles = 0;
    // f(R), mode 1: when this outer residual is expected to end the solve
    // (automatic inner count predicted to reach the target, or the cycle
    // cap), take its norm alone; should the solve go on after all, the full
    // kernel recomputes the same residual with the inner problem's data
    bool last_expected = false;
    for (int it = 1;; it++) {
        double res = 0.0;
        if ((rc = fr ? outer_residual(h, &res, last_expected) : outer_residual_dgp(h, &res))) return rc;
        if (!std::isfinite(res)) return set_err(h, MGS_E_DOMAIN, "residual is not finite");
        if (it == 1) rep->res0 = res;
        rep->res_final = res;
        if (it > 1 && it - 1 <= 128) rep->history[it - 2] = res;
        rep->outer = it - 1;
        bool can_exit = false;
        if (it > 1) {
            if (o.mode == 0) {
                can_exit = reference_exit(o, it - 1, res, res_old_1, res_old_2, &conv);
            } else {
                conv = res < o.rel_tol * rep->res0;
                can_exit = conv || (it - 1) >= o.max_cycles;
            }
            res_old_2 = res_old_1;
            res_old_1 = res;
        } else if (o.mode == 1 && res == 0.0) {
            conv = true;
            can_exit = true;
        }
        if (can_exit) break;
        if (last_expected && (rc = outer_residual(h, &res))) return rc;
        bool reaches = false;
        if (automatic) {
            inner = auto_inner(h, res, o.rel_tol * rep->res0, res_prev, m_prev, &reaches);
            res_prev = res;
            m_prev = inner;
        }
        last_expected = fr && o.mode == 1 && (reaches || it >= o.max_cycles);
        // inner problem: the base, restricted to every level with the outer
        // residual, and the increment from zero
        if (fr) {
            // d = 0 with its first colour-0 pass came from the outer residual;
            // the last inner pass also applies u += u_bg (x(d) - xb)
            for (int ic = 0; ic < inner; ic++) {
                if ((rc = o.cycle == 0 ? vcycle_inner(h, 0, ic == 0, ic == inner - 1)
                                       : fw_inner(h, 0, o.cycle, ic == 0, ic == inner - 1)))
                    return rc;
                cycles++;
            }
        } else {
            for (int ic = 0; ic < inner; ic++) {
                if ((rc = o.cycle == 0 ? vcycle_inner_dgp(h, 0) : fw_inner_dgp(h, 0, o.cycle))) return rc;
                cycles++;
            }
            prof_begin(h, MGS_KC_UPDATE, 20.0 * n3d(h->L[0].n));
            launch_dgpm_update(h->L[0].u, h->M[0].d, h->L[0].n, h->s);
            prof_end(h);
        }
    }
    rep->cycles = cycles;
    rep->converged = conv ? 1 : 0;
    return 0;
}

int init_handle(mgs_handle* h, const mgs_params* params, bool use_arena, unsigned char* arena, size_t arena_cap)
{
    h->p = *params;
    h->use_arena = use_arena;
    h->arena = arena;
    h->arena_cap = arena_cap;
    h->arena_used = 0;
#ifdef MGS_TUNING
    // tuning builds only (make EXTRA=-DMGS_TUNING): overrides of the mixed
    // inner cycle's shape; the shipped library's numerics depend on no
    // environment variable
    if (!use_arena) {
        const char* e = std::getenv("MGS_NO_TAIL");
        h->use_tail = !(e && e[0] == '1');
        const char* tn = std::getenv("MGS_TAIL_N");
        const int tv = tn ? std::atoi(tn) : 0;
        if (tv == 8 || tv == 16 || tv == 32) h->tail_n = tv;
        const char* w = std::getenv("MGS_INNER_SWEEPS");
        int a = 0, b = 0;
        if (w && std::sscanf(w, "%d,%d", &a, &b) == 2 && a >= 1 && b >= 1 && a <= 8 && b <= 8) {
            h->sweeps0[0] = a;
            h->sweeps0[1] = b;
        }
        const char* w1 = std::getenv("MGS_COARSE_SWEEPS");
        if (w1 && std::sscanf(w1, "%d,%d", &a, &b) == 2 && a >= 1 && b >= 1 && a <= 8 && b <= 8) {
            h->sweeps1[0] = a;
            h->sweeps1[1] = b;
        }
        const char* im = std::getenv("MGS_INNER_MAX");
        if (im && std::atoi(im) >= 1 && std::atoi(im) <= 64) h->inner_max = std::atoi(im);
    }
#endif
    h->lmax = levelmax_of(params->ngrid);
    h->lmin = levelmin_of(params->ngrid);
    h->lc = h->lmax - h->lmin < 0 ? 0 : h->lmax - h->lmin;
    int rc = alloc_hierarchy(h);
    if (rc) return rc;
    if (h->p.model == MGS_MODEL_DGP) rc = dgp_alloc(h);
    return rc;
}

void release_handle(mgs_handle* h)
{
    prof_flush(h);
    drop_graphs(h);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->staging) (void)hipFree(h->staging);
    h->staging = nullptr;
    if (h->hsum && h->hsum != h->hsum_host) (void)hipHostFree(h->hsum);
    h->hsum = nullptr;
    if (h->ev_norm) (void)hipEventDestroy(h->ev_norm);
    h->ev_norm = nullptr;
    fft_release(h);
    pm_release(h);
}

int solve_on_handle(mgs_handle* h, const mgs_solve_opts& o, mgs_solve_report* rep)
{
    if (o.init_guess) {
        double ubg = 0.0;
        fr_background(h->p, &ubg);
        launch_fill_f64(h->L[0].u, (size_t)n3d(h->p.ngrid), h->p.model == MGS_MODEL_FR ? ubg : 0.0, h->s);
    }
    return h->p.precision == MGS_PREC_MIXED ? solve_mixed(h, o, rep) : solve_plain(h, o, rep);
}

}  // namespace mgs

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
static int check_params(const mgs_params* p, std::string* why)
{
    if (!p) {
        *why = "params is NULL";
        return MGS_E_ARG;
    }
    if (p->ngrid < 2 || (p->ngrid & 1)) {
        *why = "ngrid must be even and >= 2";
        return MGS_E_ARG;
    }
    if (p->ngrid > 1024) {
        // the kernels index a level with 32-bit offsets (at most 2^30 cells)
        *why = "ngrid must be <= 1024";
        return MGS_E_ARG;
    }
    if (p->model != MGS_MODEL_FR && p->model != MGS_MODEL_DGP) {
        *why = "Model currently unsupported (MG_model must be 1 = f(R) or 2 = DGP)";
        return MGS_E_ARG;
    }
    if (p->model == MGS_MODEL_FR && (p->fr_n < 0 || p->fr_n > 2)) {
        *why = "f(R) n must be 0, 1 or 2";
        return MGS_E_ARG;
    }
    if (p->precision != MGS_PREC_FP64 && p->precision != MGS_PREC_MIXED) {
        *why = "precision must be MGS_PREC_FP64 or MGS_PREC_MIXED";
        return MGS_E_ARG;
    }
    if (p->precision == MGS_PREC_MIXED && p->model == MGS_MODEL_FR && !(p->fr0 < 0.0)) {
        *why = "the mixed-precision f(R) solve requires fR0 < 0";
        return MGS_E_ARG;
    }
    if (!(p->box > 0.0) || !(p->aexpn > 0.0)) {
        *why = "box and aexpn must be positive";
        return MGS_E_ARG;
    }
    if (p->model == MGS_MODEL_DGP && !(p->h0rc > 0.0)) {
        *why = "DGP needs H0rc > 0";
        return MGS_E_ARG;
    }
    const int lmax = levelmax_of(p->ngrid), lmin = levelmin_of(p->ngrid);
    const int lc = lmax - lmin < 0 ? 0 : lmax - lmin;
    if (((p->ngrid >> lc) & 1) || (p->ngrid >> lc) < 2) {
        *why = "grid hierarchy: every level must have an even number of cells";
        return MGS_E_ARG;
    }
    if (p->precision == MGS_PREC_MIXED && lc < 1) {
        *why = "the mixed-precision solve needs at least two grid levels";
        return MGS_E_ARG;
    }
    return 0;
}

namespace mgs {
int validate_params(const mgs_params* p, std::string* why) { return check_params(p, why); }
}  // namespace mgs

extern "C" {

int mgs_create(const mgs_params* params, mgs_handle** out)
{
    if (!out) return set_err(nullptr, MGS_E_ARG, "out is NULL");
    *out = nullptr;
    std::string why;
    int rc = check_params(params, &why);
    if (rc) return set_err(nullptr, rc, why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(nullptr, MGS_E_HIP, "no HIP device available: libmgsolve has no CPU fallback");
    mgs_handle* h = new mgs_handle();
    if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return set_err(nullptr, MGS_E_HIP, "hipStreamCreate failed");
    }
    rc = mgs::init_handle(h, params, false, nullptr, 0);
    if (rc) {
        g_create_error = h->err;
        mgs_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

int mgs_destroy(mgs_handle* h)
{
    if (!h) return 0;
    if (h->s) (void)hipStreamSynchronize(h->s);
    mgs::release_handle(h);
    if (h->s) (void)hipStreamDestroy(h->s);
    delete h;
    return 0;
}

const char* mgs_last_error(const mgs_handle* h)
{
    return h ? h->err.c_str() : g_create_error.c_str();
}

int mgs_levels(const mgs_handle* h, int32_t* levelmax, int32_t* levelmin, int32_t* coarsest)
{
    if (!h) return MGS_E_ARG;
    if (levelmax) *levelmax = h->lmax;
    if (levelmin) *levelmin = h->lmin;
    if (coarsest) *coarsest = h->lc;
    return 0;
}

int mgs_set_params(mgs_handle* h, const mgs_params* params)
{
    if (!h) return MGS_E_ARG;
    std::string why;
    int rc = check_params(params, &why);
    if (rc) return set_err(h, rc, why);
    if (params->ngrid != h->p.ngrid || params->model != h->p.model || params->precision != h->p.precision)
        return set_err(h, MGS_E_ARG, "mgs_set_params cannot change ngrid, model or precision");
    h->p = *params;
    // the captured coarse cycles hold the old coefficients
    HIPCHK(h, hipStreamSynchronize(h->s));
    drop_graphs(h);
    return 0;
}

static size_t esize(int32_t dtype) { return dtype == MGS_F32 ? 4 : 8; }

// copy a lexicographic array (host or device) into device memory; returns a
// device pointer (the caller's own when it is already on the device)
static int to_device(mgs_handle* h, const void* src, size_t bytes, int32_t where, const void** dev)
{
    if (where == MGS_DEVICE) {
        *dev = src;
        return 0;
    }
    int rc = ensure_staging(h, bytes);
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(h->staging, src, bytes, hipMemcpyHostToDevice, h->s));
    *dev = h->staging;
    return 0;
}

int mgs_set_density(mgs_handle* h, const void* delta, int32_t dtype, int32_t where)
{
    if (!h || !delta) return set_err(h, MGS_E_ARG, "NULL argument");
    if (dtype != MGS_F32 && dtype != MGS_F64) return set_err(h, MGS_E_ARG, "dtype");
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    if (where == MGS_HOST) {
        // reject NaN/Inf before upload (the operators accept any finite delta,
        // as the reference's do)
        for (size_t q = 0; q < n3; q++) {
            const double v = dtype == MGS_F32 ? (double)((const float*)delta)[q] : ((const double*)delta)[q];
            if (!std::isfinite(v))
                return set_err(h, MGS_E_DOMAIN, "density contrast must be finite (cell " + std::to_string(q) + ")");
        }
    }
    const void* dev = nullptr;
    int rc = to_device(h, delta, n3 * esize(dtype), where, &dev);
    if (rc) return rc;
    if (dtype == MGS_F32)
        launch_pack_f64_from_f32((const float*)dev, h->L[0].rho, N, h->s);
    else
        launch_pack_f64_from_f64((const double*)dev, h->L[0].rho, N, h->s);
    HIPCHK(h, hipStreamSynchronize(h->s));
    h->has_density = true;
    h->rho_cached = false;
    return 0;
}

int mgs_set_field(mgs_handle* h, const void* u, int32_t dtype, int32_t where)
{
    if (!h || !u) return set_err(h, MGS_E_ARG, "NULL argument");
    if (dtype != MGS_F32 && dtype != MGS_F64) return set_err(h, MGS_E_ARG, "dtype");
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    const void* dev = nullptr;
    int rc = to_device(h, u, n3 * esize(dtype), where, &dev);
    if (rc) return rc;
    if (dtype == MGS_F32)
        launch_pack_f64_from_f32((const float*)dev, h->L[0].u, N, h->s);
    else
        launch_pack_f64_from_f64((const double*)dev, h->L[0].u, N, h->s);
    HIPCHK(h, hipStreamSynchronize(h->s));
    return 0;
}

int mgs_get_field(mgs_handle* h, void* u, int32_t dtype, int32_t where, double scale)
{
    if (!h || !u) return set_err(h, MGS_E_ARG, "NULL argument");
    if (dtype != MGS_F32 && dtype != MGS_F64) return set_err(h, MGS_E_ARG, "dtype");
    const int N = h->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    const size_t bytes = n3 * esize(dtype);
    void* dst = u;
    if (where == MGS_HOST) {
        int rc = ensure_staging(h, bytes);
        if (rc) return rc;
        dst = h->staging;
    }
    if (dtype == MGS_F32)
        launch_unpack_f64_to_f32(h->L[0].u, (float*)dst, N, scale, h->s);
    else
        launch_unpack_f64_to_f64(h->L[0].u, (double*)dst, N, scale, h->s);
    if (where == MGS_HOST) HIPCHK(h, hipMemcpyAsync(u, dst, bytes, hipMemcpyDeviceToHost, h->s));
    HIPCHK(h, hipStreamSynchronize(h->s));
    return 0;
}

static double* level_array(mgs_handle* h, int level, int which)
{
    LevelD& v = h->L[level];
    switch (which) {
    case MGS_ARR_U: return v.u;
    case MGS_ARR_RHO: return v.rho;
    case MGS_ARR_F: return v.f;
    case MGS_ARR_RES:
        if (level == 0 && ensure_res0(h)) return nullptr;
        return h->L[level].res;
    case MGS_ARR_RRES: return v.rres;
    case MGS_ARR_RU: return v.ru;
    }
    return nullptr;
}

int mgs_set_level_array(mgs_handle* h, int32_t level, int32_t which, const void* src, int32_t dtype, int32_t where)
{
    if (!h || !src || level < 0 || level > h->lc) return set_err(h, MGS_E_ARG, "level out of range");
    double* a = level_array(h, level, which);
    if (!a) return set_err(h, MGS_E_ARG, "no such array on this level");
    const int n = h->L[level].n;
    const size_t n3 = (size_t)n * n * n;
    const void* dev = nullptr;
    int rc = to_device(h, src, n3 * esize(dtype), where, &dev);
    if (rc) return rc;
    if (dtype == MGS_F32)
        launch_pack_f64_from_f32((const float*)dev, a, n, h->s);
    else
        launch_pack_f64_from_f64((const double*)dev, a, n, h->s);
    HIPCHK(h, hipStreamSynchronize(h->s));
    if (which == MGS_ARR_RHO) h->rho_cached = false;
    return 0;
}

int mgs_get_level_array(mgs_handle* h, int32_t level, int32_t which, void* dst, int32_t dtype, int32_t where)
{
    if (!h || !dst || level < 0 || level > h->lc) return set_err(h, MGS_E_ARG, "level out of range");
    double* a = level_array(h, level, which);
    if (!a) return set_err(h, MGS_E_ARG, "no such array on this level");
    const int n = h->L[level].n;
    const size_t bytes = (size_t)n * n * n * esize(dtype);
    void* out = dst;
    if (where == MGS_HOST) {
        int rc = ensure_staging(h, bytes);
        if (rc) return rc;
        out = h->staging;
    }
    if (dtype == MGS_F32)
        launch_unpack_f64_to_f32(a, (float*)out, n, 1.0, h->s);
    else
        launch_unpack_f64_to_f64(a, (double*)out, n, 1.0, h->s);
    if (where == MGS_HOST) HIPCHK(h, hipMemcpyAsync(dst, out, bytes, hipMemcpyDeviceToHost, h->s));
    HIPCHK(h, hipStreamSynchronize(h->s));
    return 0;
}

int mgs_background(const mgs_handle* h, double* fr_bg, double* u_bg)
{
    if (!h) return MGS_E_ARG;
    double ub = 0.0;
    const double f = fr_background(h->p, &ub);
    if (fr_bg) *fr_bg = f;
    if (u_bg) *u_bg = ub;
    return 0;
}

int mgs_initialise(mgs_handle* h)
{
    if (!h) return MGS_E_ARG;
    double ubg = 0.0;
    fr_background(h->p, &ubg);
    const size_t n3 = (size_t)h->p.ngrid * h->p.ngrid * h->p.ngrid;
    launch_fill_f64(h->L[0].u, n3, h->p.model == MGS_MODEL_FR ? ubg : 0.0, h->s);
    for (int l = 1; l <= h->lc; l++) {
        const size_t m3 = (size_t)h->L[l].n * h->L[l].n * h->L[l].n;
        LevelD& v = h->L[l];
        for (double* a : {v.u, v.f, v.res, v.rres, v.ru}) launch_fill_f64(a, m3, 0.0, h->s);
    }
    if (h->L[0].res) launch_fill_f64(h->L[0].res, n3, 0.0, h->s);
    HIPCHK(h, hipStreamSynchronize(h->s));
    return 0;
}

static int check_level(mgs_handle* h, int level, int lo)
{
    if (level < lo || level > h->lc) return set_err(h, MGS_E_ARG, "level out of range");
    return 0;
}

#define SYNC_RET(h)                                 \
    do {                                            \
        HIPCHK(h, hipGetLastError());               \
        HIPCHK(h, hipStreamSynchronize((h)->s));    \
        return 0;                                   \
    } while (0)

int mgs_smooth(mgs_handle* h, int32_t level, int32_t red)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 0)) return MGS_E_ARG;
    op_smooth(h, level, red);
    SYNC_RET(h);
}

int mgs_residual(mgs_handle* h, int32_t level, double* rms)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 0)) return MGS_E_ARG;
    double r = 0.0;
    int rc = op_residual(h, level, true, &r);
    if (rc) return rc;
    if (rms) *rms = r;
    SYNC_RET(h);
}

int mgs_restrict_residual(mgs_handle* h, int32_t level)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 1)) return MGS_E_ARG;
    op_restrict_residual(h, level);
    SYNC_RET(h);
}

int mgs_restrict_field(mgs_handle* h, int32_t level)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 1)) return MGS_E_ARG;
    op_restrict_field(h, level);
    SYNC_RET(h);
}

int mgs_restrict_density(mgs_handle* h, int32_t level)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 1)) return MGS_E_ARG;
    op_restrict_density(h, level);
    SYNC_RET(h);
}

int mgs_rhs(mgs_handle* h, int32_t level)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 1)) return MGS_E_ARG;
    op_rhs(h, level);
    SYNC_RET(h);
}

int mgs_correct(mgs_handle* h, int32_t level)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 1)) return MGS_E_ARG;
    op_correct(h, level, false);
    SYNC_RET(h);
}

int mgs_vcycle(mgs_handle* h, int32_t level)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 0)) return MGS_E_ARG;
    if (!h->has_density) return set_err(h, MGS_E_STATE, "no density set");
    cache_densities(h);
    vcycle_plain(h, level);
    SYNC_RET(h);
}

static int cycle_entry(mgs_handle* h, int32_t level, int kind)
{
    if (!h) return MGS_E_ARG;
    if (check_level(h, level, 0)) return MGS_E_ARG;
    if (!h->has_density) return set_err(h, MGS_E_STATE, "no density set");
    cache_densities(h);
    cycle_plain(h, level, kind);
    SYNC_RET(h);
}

int mgs_fcycle(mgs_handle* h, int32_t level) { return cycle_entry(h, level, 1); }

int mgs_wcycle(mgs_handle* h, int32_t level) { return cycle_entry(h, level, 2); }

int mgs_solve(mgs_handle* h, const mgs_solve_opts* opts, mgs_solve_report* report)
{
    if (!h || !opts || !report) return set_err(h, MGS_E_ARG, "NULL argument");
    if (!h->has_density) return set_err(h, MGS_E_STATE, "no density set (mgs_set_density)");
    std::memset(report, 0, sizeof(*report));
    if (opts->max_cycles < 1) return set_err(h, MGS_E_ARG, "max_cycles must be >= 1");
    if (opts->cycle < 0 || opts->cycle > 2) return set_err(h, MGS_E_ARG, "cycle must be 0 (V), 1 (F) or 2 (W)");
    // the solve's wall time on the stream; destroyed on every return path
    struct EventPair {
        hipEvent_t a = nullptr, b = nullptr;
        ~EventPair()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    HIPCHK(h, hipEventCreate(&ev.a));
    HIPCHK(h, hipEventCreate(&ev.b));
    HIPCHK(h, hipEventRecord(ev.a, h->s));
    const int rc = mgs::solve_on_handle(h, *opts, report);
    HIPCHK(h, hipEventRecord(ev.b, h->s));
    HIPCHK(h, hipEventSynchronize(ev.b));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess) ms = 0.0f;
    report->seconds = ms * 1e-3;
    if (rc) return rc;
    HIPCHK(h, hipGetLastError());
    return 0;
}

int mgs_profile(mgs_handle* h, int32_t enable)
{
    if (!h) return MGS_E_ARG;
    prof_flush(h);
    h->prof = enable != 0;
    for (auto& c : h->pc) c = ProfClass();
    return 0;
}

int mgs_profile_get(mgs_handle* h, int32_t kclass, int64_t* launches, double* ms, double* bytes)
{
    if (!h || kclass < 0 || kclass >= MGS_KC_COUNT) return MGS_E_ARG;
    prof_flush(h);
    if (launches) *launches = h->pc[kclass].launches;
    if (ms) *ms = h->pc[kclass].ms;
    if (bytes) *bytes = h->pc[kclass].bytes;
    return 0;
}

int mgs_use_graphs(mgs_handle* h, int32_t enable)
{
    if (!h) return MGS_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->s));
    drop_graphs(h);
    h->use_graphs = enable != 0;
    return 0;
}

int mgs_synchronize(mgs_handle* h)
{
    if (!h) return MGS_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->s));
    return 0;
}

int mgs_device_field(mgs_handle* h, void** ptr, size_t* bytes)
{
    if (!h || !ptr) return MGS_E_ARG;
    *ptr = h->L[0].u;
    if (bytes) *bytes = (size_t)n3d(h->p.ngrid) * sizeof(double);
    return 0;
}

}  // extern "C"
