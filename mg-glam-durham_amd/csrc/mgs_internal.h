// mgs_internal.h -- the handle behind include/mgsolve.h and the host-side
// helpers shared by mgs_api.cpp, mgs_dgp.hip, mgs_fft.hip and mgs_serve.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mgsolve.h"
#include "mgs_common.h"

namespace mgs {

// one level of the FP64 hierarchy (the reference's FI/FI2/FI3 sections)
struct LevelD {
    int n = 0;
    double* u = nullptr;     // FI2 (level 0) / FI3 S4
    double* rho = nullptr;   // FI (level 0) / FI3 S2
    double* f = nullptr;     // FI3 S5
    double* res = nullptr;   // FI3 S6 (level 0: FI3 itself, allocated on demand)
    double* rres = nullptr;  // FI3 S1
    double* ru = nullptr;    // FI3 S3
};

// one level of the FP32 delta-form hierarchy (mixed precision)
struct LevelF {
    int n = 0;
    float* d = nullptr;      // increment of x^{n+1}
    float* xb = nullptr;     // base x = u / u_bg
    float* f = nullptr;      // right-hand side (level 0: -r of the FP64 residual)
    float* res = nullptr;
    float* rres = nullptr;
    float* rd = nullptr;     // frozen R d
    float* rho = nullptr;    // nDGP, z-marching levels: the level's density rounded to FP32
};

struct ProfRec {
    int kc;
    double bytes;
    hipEvent_t a, b;
    bool ok;  // both events recorded
};

struct ProfClass {
    long long launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

// Level-0 pre/post sweeps of the mixed f(R) inner V-cycle (coarser inner
// levels keep the reference's Nsmth = 2).  One pre-sweep is enough on the
// finest level of the increment's problem: its smooth error comes straight
// from the FP64 residual, and the coarse correction removes it.  The serving
// engine runs the same sweeps, so its results stay bit-identical to
// mgs_solve's.
constexpr int MIXED_PRE0 = 1;
constexpr int MIXED_POST0 = 2;
// mode 1 with inner_cycles 0: cap on the automatic inner cycles per outer
// iteration (the FP32 increment's accuracy bounds what one outer step gains)
constexpr int MIXED_INNER_MAX = 4;

// DGP constants of one level, PMP2MGsolver_DGP.f90:36-48
struct DgpCoef {
    double dx2, dx4, alpha, beta, fct1;
    double fct2_1, fct3_1;     // fused level-1 restriction
    double inv_dx2, inv_dx4;   // 1/dx2, 1/dx4: powers of two, so x * inv == x / dx exactly
};

}  // namespace mgs

struct mgs_handle {
    mgs_params p{};
    int lmax = 0, lmin = 0, lc = 0;
    hipStream_t s = nullptr;
    std::vector<mgs::LevelD> L;
    std::vector<mgs::LevelF> M;
    std::vector<void*> allocs;
    // arena mode (mgs_solve_noalloc): device memory comes from [arena,
    // arena + arena_cap) instead of hipMalloc; with arena null it only counts
    // the bytes a hierarchy needs (arena_used)
    bool use_arena = false;
    unsigned char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    double hsum_host[16] = {};  // arena mode: host copy of the norm (no pinned allocation)
    double* partials = nullptr;
    size_t npart = 0;
    double* dsum = nullptr;
    double* hsum = nullptr;          // pinned
    void* staging = nullptr;
    size_t staging_bytes = 0;
    bool has_density = false;
    bool rho_cached = false;
    // DGP: out-of-place colour pass target and per-block partials
    double* dgp_tmp = nullptr;     // a colour half of level 0 in FP64 (N^3 floats)
    double* dgp_rms = nullptr;
    double* dgp_spare64 = nullptr;  // FP64 z-marching passes: second level-0 array
    // FFT plans (density generator, power spectrum)
    void* fft = nullptr;
    // particles and the potential of the PM step (mgs_pm.hip)
    void* pm = nullptr;
    // captured coarse-level parts of the V-cycles (levels >= 1), replayed
    // every cycle; rebuilt when the model scalars change
    hipGraphExec_t g_plain = nullptr;
    hipGraphExec_t g_inner = nullptr;
    // mixed solve: the inner problem's base restricted to every level
    hipGraphExec_t g_rbase = nullptr;
    // recorded after the outer norm's copy to the host (not in arena mode):
    // the host waits for the norm while the base restriction runs
    hipEvent_t ev_norm = nullptr;
    bool use_graphs = true;
    // mixed f(R): small inner levels in one launch; tuning builds
    // (-DMGS_TUNING) read MGS_NO_TAIL=1 to turn it off
    bool use_tail = true;
    int tail_n = 16;  // the tail starts at the first level with n <= tail_n (tuning builds: MGS_TAIL_N 8, 16 or 32)
    // mixed f(R): level-0 pre/post sweeps of the inner cycle and the
    // automatic inner-cycle cap (tuning builds: MGS_INNER_SWEEPS "a,b" and
    // MGS_INNER_MAX override them for mgs_solve only)
    int sweeps0[2] = {mgs::MIXED_PRE0, mgs::MIXED_POST0};
    int sweeps1[2] = {2, 2};  // the inner cycle's pre / post pairs below level 0 (above the tail)
    int inner_max = mgs::MIXED_INNER_MAX;
    // profiling
    bool prof = false;
    std::vector<mgs::ProfRec> recs;
    mgs::ProfClass pc[MGS_KC_COUNT];
    std::string err;
};

namespace mgs {
// mgs_api.cpp
FrCoef fr_coefs(const mgs_params& p, int level);
DfCoef df_coefs(const mgs_params& p, int level);
double fr_background(const mgs_params& p, double* ubg);
int ensure_staging(mgs_handle* h, size_t bytes);
// device memory for the handle: hipMalloc, or the next 256-byte aligned piece
// of the handle's arena (use_arena); recorded for mgs_destroy when allocated
int dev_alloc(mgs_handle* h, size_t bytes, void** out, const char* what);
// mgs_create's set-up of a handle without a stream of its own; arena mode when
// use_arena (arena null: count the bytes only)
int init_handle(mgs_handle* h, const mgs_params* params, bool use_arena, unsigned char* arena, size_t arena_cap);
// mgs_solve's body on a prepared handle (no timing events)
int solve_on_handle(mgs_handle* h, const mgs_solve_opts& o, mgs_solve_report* rep);
// release what init_handle allocated (not the stream)
void release_handle(mgs_handle* h);
void prof_begin(mgs_handle* h, int kc, double bytes);
void prof_end(mgs_handle* h);
int vcycle_plain(mgs_handle* h, int l);
int op_residual(mgs_handle* h, int l, bool store, double* rms);
int levelmax_of(int N);
int levelmin_of(int N);
// the parameter checks of mgs_create (0 or an MGS_E_* code with a reason)
int validate_params(const mgs_params* p, std::string* why);
// mixed mode 1, automatic inner count: the inner cycles of the next outer
// iteration from its start residual res, the target, and the previous outer
// iteration's start residual and count (m_prev 0: none yet, inner_max);
// *reaches: the count is expected to reach the target
int auto_inner_count(int inner_max, double res, double target, double res_prev, int m_prev, bool* reaches);
// mode-0 stopping rule after outer/cycle `iter` with residual res and the two
// before it (res1, res2); sets *conv when res <= res_conv
bool reference_exit(const mgs_solve_opts& o, int iter, double res, double res1, double res2, bool* conv);

// mgs_dgp.hip
DgpCoef dgp_coefs(const mgs_params& p, int level);
int dgp_alloc(mgs_handle* h);
void launch_dgp_smooth64(int level, double* u, const double* rho, const double* f, double* tmp, int n, int colour,
                         const DgpCoef& c, hipStream_t s);
int launch_dgp_residual64(int level, const double* u, const double* rho, const double* f, double* res,
                          double* partials, int n, const DgpCoef& c, double* scratch, hipStream_t s);
void launch_dgp_restrict_residual1_64(const double* u0, const double* delta, double* rres1, int N, const DgpCoef& c0,
                                      hipStream_t s);
void launch_dgp_rhs64(const double* ru, const double* rho, const double* rres, double* f, int n, const DgpCoef& c,
                      hipStream_t s);
// mixed-precision DGP (FP32 delta-form inner problem)
int launch_dgpm_outer(const double* u, const double* rho, float* g32, float* pb, float* e, double* partials, int n,
                      const DgpCoef& c, hipStream_t s);
void launch_dgpm_smooth(float* e, const float* pb, const double* rho, const float* rhs, float* tmp, int n, int colour,
                        const DgpCoef& c, hipStream_t s);
// rho32: the density rounded to FP32 (used on z-marching levels)
void launch_dgpm_residual(const float* e, const float* pb, const double* rho, const float* rho32, const float* rhs,
                          float* out, int n, const DgpCoef& c, hipStream_t s);
void launch_dgpm_coarsest(float* e, const float* pb, const double* rho, const float* rhs, float* tmp, int n,
                          int sweeps, const DgpCoef& c, hipStream_t s);
void launch_dgpm_correct(const float* eH, const float* rd, float* ef, int nc, hipStream_t s);
void launch_dgpm_update(double* u, const float* e, int n, hipStream_t s);
// nDGP levels whose passes and residuals z-march with 19-point register
// windows (mgs_dgp_dev.h): whole waves along m, runs of DGP_KC planes
constexpr int DGP_KC = 8;
inline bool dgp_zmarch(int n)
{
    const int h = n / 2;
    return h >= 64 && h % 64 == 0 && n % DGP_KC == 0;
}
// z-marching colour passes (levels with dgp_zmarch(n)): read
// the field's colour halves h0/h1 (e0/e1), write the pass colour's new half
// to out
void launch_dgp_zpass64(int level, const double* h0, const double* h1, double* out, const double* rho,
                        const double* f, int n, int colour, const DgpCoef& c, hipStream_t s);
void launch_dgpm_zpass(const float* e0, const float* e1, float* out, const float* pb, const float* rho32,
                       const float* rhs, int n, int colour, const DgpCoef& c, hipStream_t s);
void launch_round_f32(const double* a, float* b, size_t n, hipStream_t s);

// mgs_fft.hip
void fft_release(mgs_handle* h);

// mgs_pm.hip
void pm_release(mgs_handle* h);
// the handle's N^3 REAL*4 potential (x fastest), allocated on first use; null on failure
float* pm_potential(mgs_handle* h);
}  // namespace mgs

int set_err(mgs_handle* h, int code, const std::string& msg);
