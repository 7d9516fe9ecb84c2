// mgs_serve.hip -- continuous-batching f(R) solve server (mgs_server_* C-ABI).
//
// Requests wait in a FIFO queue; a worker thread admits them into slots as
// slots and HBM budget free up, and advances every active request by one
// outer iteration of the mixed-precision solve per step.  A step is a fixed
// sequence of batched persistent launches: each launch walks the tiles of
// all participating requests (a prefix sum of their tile counts in a device
// plan), so one launch smooths, restricts or corrects the same level of every
// request at once, whatever their grid sizes and parameters.  The sequence
// depends only on the deepest hierarchy in the batch, so it is captured once
// per depth as a hipGraph and replayed; requests joining or leaving only
// change the plan in device memory.
//
// Per request the arithmetic, the tiling and the reduction order are those of
// the single-request solve (mgs_api.cpp, solve_mixed with one inner cycle),
// so a result is bit-identical alone, in any mix, and to mgs_solve.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mgsolve.h"
#include "mgs_fr_dev.h"
#include "mgs_internal.h"
#include "mgs_kernels.h"

namespace mgs {
namespace serve {

constexpr int MAXS = 64;    // slots
constexpr int MAXL = 11;    // levels of a hierarchy (ngrid <= 1024)
constexpr int GRID = 2048;  // blocks of a batched stencil/transfer launch
constexpr int SWEEPS = 8;   // coarsest-level sweeps (as solve_mixed)

struct Shape {
    int TX, TY, XB, JB, KC;
    unsigned NB;
};

static Shape shape_of(const ZLaunch& L) { return Shape{L.TX, L.TY, L.XB, L.JB, L.KC, L.NB}; }

struct LevelDesc {
    Geo g;
    float *d, *xb, *f, *rres, *rd;
    DfCoef dc;
    Shape z;   // stencil passes on this level
    Shape zb;  // both-colour passes on this level (residual-type)
    Shape zc;  // coarse-column transfers between this level and the next finer one (levels >= 1)
};

struct SlotDesc {
    int frn, lc, pad0, pad1;
    double* u;
    const double* rho;
    double* partials;
    FrCoef fc;
    LevelDesc L[MAXL];
};

// which requests take part in each batched launch, and prefix sums of their
// tiles; lists at level l hold the requests with l < lc (standard V-cycle
// step at l) or lc == l (coarsest solve at l)
struct Plan {
    int nres;
    int res_slot[MAXS];
    unsigned res_pre[MAXS + 1];  // level-0 stencil tiles
    int ninner[MAXL];
    int inner_slot[MAXL][MAXS];
    unsigned z_pre[MAXL][MAXS + 1];     // level-l stencil tiles
    unsigned zc_pre[MAXL][MAXS + 1];    // level-(l+1) coarse columns (transfers l <-> l+1)
    unsigned cell_pre[MAXL][MAXS + 1];  // level-(l+1) cells in blocks of 256
    unsigned zn_pre[MAXL][MAXS + 1];    // level-(l+1) stencil tiles
    int ncoarse[MAXL];
    int coarse_slot[MAXL][MAXS];
};

// ---------------------------------------------------------------------------
// batched kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_slot(const unsigned* pre, int n, unsigned t)
{
    int lo = 0, hi = n;  // pre[lo] <= t < pre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ bool tile_col(const Geo& g, int width, const Shape& z, unsigned lb, ZCol& c)
{
    const int tx = (int)threadIdx.x % z.TX, ty = (int)threadIdx.x / z.TX;
    if (ty >= z.TY) return false;
    return zcol_at(width, g.n, z.TX, z.TY, z.XB, z.JB, z.KC, lb, tx, ty, c);
}

#define FRN_SWITCH(frn, CALL)      \
    switch (frn) {                 \
    case 0: {                      \
        constexpr int FRN = 0;     \
        CALL;                      \
    } break;                       \
    case 1: {                      \
        constexpr int FRN = 1;     \
        CALL;                      \
    } break;                       \
    default: {                     \
        constexpr int FRN = 2;     \
        CALL;                      \
    } break;                       \
    }

// FP64 outer residual (and the inner problem's g, xb, d after its first
// colour-0 pass) on level 0; one block partial per tile, as k_both_colours
__global__ __launch_bounds__(256) void kb_outer_residual(const SlotDesc* __restrict__ S, const Plan* __restrict__ P)
{
    const unsigned total = P->res_pre[P->nres];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_slot(P->res_pre, P->nres, t);
        const SlotDesc& sd = S[P->res_slot[i]];
        const LevelDesc& L = sd.L[0];
        const unsigned lb = t - P->res_pre[i];
        ZCol c;
        double acc = 0.0;
        if (tile_col(L.g, L.g.h, L.zb, lb, c)) {
            FRN_SWITCH(sd.frn, acc = both_colours_col(FrOuterOp<FRN>{sd.u, sd.rho, L.f, L.xb, L.d, sd.fc, L.dc,
                                                                     (unsigned)L.g.half},
                                                      L.g, c));
        }
        const double bs = block_sum(acc);
        if (threadIdx.x == 0) sd.partials[lb] = bs;
        __syncthreads();
    }
}

// per request: the sum of its level-0 partials in launch_sum_partials' order
__global__ __launch_bounds__(1024) void kb_sums(const SlotDesc* __restrict__ S, const Plan* __restrict__ P,
                                                double* __restrict__ out)
{
    __shared__ double s[1024];
    __shared__ double mid[1024];
    for (int i = blockIdx.x; i < P->nres; i += gridDim.x) {
        const int slot = P->res_slot[i];
        const SlotDesc& sd = S[slot];
        const int count = (int)sd.L[0].zb.NB;
        double r;
        if (count <= 8192) {
            block_chunk_sum(sd.partials, 0, count, s);
            r = s[0];
        } else {
            const int chunk = (count + 1023) / 1024;
            for (int c = 0; c < 1024; c++) {
                const int lo = c * chunk;
                const int hi = lo + chunk < count ? lo + chunk : count;
                block_chunk_sum(sd.partials, lo, hi, s);
                if (threadIdx.x == 0) mid[c] = s[0];
                __syncthreads();
            }
            block_chunk_sum(mid, 0, 1024, s);
            r = s[0];
        }
        if (threadIdx.x == 0) out[slot] = r;
        __syncthreads();
    }
}

// xb of level l+1 from level l
__global__ __launch_bounds__(256) void kb_restrict_base(const SlotDesc* __restrict__ S, const Plan* __restrict__ P,
                                                        int l)
{
    const unsigned total = P->cell_pre[l][P->ninner[l]];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_slot(P->cell_pre[l], P->ninner[l], t);
        const SlotDesc& sd = S[P->inner_slot[l][i]];
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        const unsigned nc = (unsigned)C.g.n;
        const unsigned cell = (t - P->cell_pre[l][i]) * 256u + threadIdx.x;
        if (cell < nc * nc * nc) {
            const int I = (int)(cell % nc), J = (int)((cell / nc) % nc), K = (int)(cell / (nc * nc));
            FRN_SWITCH(sd.frn, restrict_base_cell<FRN>(F.xb, C.xb, F.g, C.g, I, J, K));
        }
    }
}

// one colour pass of the inner smoother on level l
__global__ __launch_bounds__(256) void kb_smooth(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l,
                                                 int colour)
{
    const unsigned total = P->z_pre[l][P->ninner[l]];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_slot(P->z_pre[l], P->ninner[l], t);
        const SlotDesc& sd = S[P->inner_slot[l][i]];
        const LevelDesc& L = sd.L[l];
        ZCol c;
        if (tile_col(L.g, L.g.h, L.z, t - P->z_pre[l][i], c)) {
            FRN_SWITCH(sd.frn, colour_pass_col(DfSmoothOp<FRN>{L.d, L.xb, L.f, L.dc}, L.g, colour, c));
        }
    }
}

// the last colour-1 pass on level 0 fused with the outer update of u
__global__ __launch_bounds__(256) void kb_smooth_update(const SlotDesc* __restrict__ S, const Plan* __restrict__ P)
{
    const unsigned total = P->z_pre[0][P->ninner[0]];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_slot(P->z_pre[0], P->ninner[0], t);
        const SlotDesc& sd = S[P->inner_slot[0][i]];
        const LevelDesc& L = sd.L[0];
        ZCol c;
        if (tile_col(L.g, L.g.h, L.z, t - P->z_pre[0][i], c)) {
            FRN_SWITCH(sd.frn,
                       colour_pass_col(DfSmoothUpdateOp<FRN>{L.d, L.xb, L.f, sd.u, L.dc, sd.fc.ubg}, L.g, 1, c));
        }
    }
}

// residual of level l restricted to level l+1, with R d (frozen copy rd and
// the start value d of level l+1)
__global__ __launch_bounds__(256) void kb_restrict_residual(const SlotDesc* __restrict__ S,
                                                            const Plan* __restrict__ P, int l)
{
    const unsigned total = P->zc_pre[l][P->ninner[l]];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_slot(P->zc_pre[l], P->ninner[l], t);
        const SlotDesc& sd = S[P->inner_slot[l][i]];
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        ZCol c;
        if (tile_col(C.g, C.g.n, C.zc, t - P->zc_pre[l][i], c)) {
            FRN_SWITCH(sd.frn,
                       df_restrict_residual_col<FRN>(F.d, F.xb, F.f, C.rres, C.rd, C.d, F.g, C.g, F.dc, c));
        }
    }
}

// FAS right-hand side of level l+1: f = D(R d) - R r
__global__ __launch_bounds__(256) void kb_rhs(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    const unsigned total = P->zn_pre[l][P->ninner[l]];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_slot(P->zn_pre[l], P->ninner[l], t);
        const SlotDesc& sd = S[P->inner_slot[l][i]];
        const LevelDesc& C = sd.L[l + 1];
        ZCol c;
        if (tile_col(C.g, C.g.h, C.zb, t - P->zn_pre[l][i], c)) {
            FRN_SWITCH(sd.frn, both_colours_col(DfResOp<FRN>{C.rd, C.xb, C.rres, C.f, C.dc}, C.g, c));
        }
    }
}

// coarse-grid correction of level l from level l+1
__global__ __launch_bounds__(256) void kb_correct(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    const unsigned total = P->zc_pre[l][P->ninner[l]];
    for (unsigned t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_slot(P->zc_pre[l], P->ninner[l], t);
        const SlotDesc& sd = S[P->inner_slot[l][i]];
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        ZCol c;
        if (tile_col(C.g, C.g.n, C.zc, t - P->zc_pre[l][i], c)) {
            FRN_SWITCH(sd.frn, df_correct_col<FRN>(C.d, C.rd, F.d, F.xb, F.g, C.g, c));
        }
    }
}

// coarsest-level sweeps of the requests whose coarsest level is l, one
// workgroup each
__global__ __launch_bounds__(1024) void kb_coarsest(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    for (int i = blockIdx.x; i < P->ncoarse[l]; i += gridDim.x) {
        const SlotDesc& sd = S[P->coarse_slot[l][i]];
        const LevelDesc& L = sd.L[l];
        FRN_SWITCH(sd.frn, df_coarsest_block<FRN>(L.d, L.xb, L.f, L.g, SWEEPS, L.dc));
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

static double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// one request's device hierarchy: level-0 u and delta in FP64, the FP32
// inner hierarchy, and the level-0 block partials, carved out of one
// allocation
struct Hier {
    int N = 0, lc = 0;
    size_t bytes = 0;
    void* base = nullptr;
    double *u = nullptr, *rho = nullptr, *partials = nullptr;
    float *d[MAXL] = {}, *xb[MAXL] = {}, *f[MAXL] = {}, *rres[MAXL] = {}, *rd[MAXL] = {};
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static int coarsest_of(int N)
{
    const int lc = levelmax_of(N) - levelmin_of(N);
    return lc < 0 ? 0 : lc;
}

// carve (or just measure, when base is null) a hierarchy for ngrid N
static size_t layout(Hier& h, int N, char* base)
{
    h.N = N;
    h.lc = coarsest_of(N);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    const size_t n3 = (size_t)N * N * N;
    h.u = (double*)take(n3 * 8);
    h.rho = (double*)take(n3 * 8);
    h.partials = (double*)take(((size_t)make_zlaunch_both(make_geo(N)).NB + 64) * 8);
    for (int l = 0; l <= h.lc; l++) {
        const size_t m = (size_t)(N >> l) * (N >> l) * (N >> l);
        h.d[l] = (float*)take(m * 4);
        h.xb[l] = (float*)take(m * 4);
        h.f[l] = (float*)take(m * 4);
        if (l > 0) {
            h.rres[l] = (float*)take(m * 4);
            h.rd[l] = (float*)take(m * 4);
        }
    }
    h.bytes = off;
    return off;
}

struct Request {
    int64_t ticket = 0;
    mgs_params p{};
    std::vector<char> delta;
    int32_t dtype = MGS_F32;
    int32_t max_outer = 60;
    double rel_tol = 1e-8;
    Clock::time_point t_submit, t_admit, t_done;
    bool done = false;
    mgs_request_result res{};
    std::vector<double> u;  // raw field, reference order
};

struct Slot {
    Request* r = nullptr;
    Hier* h = nullptr;
    bool fresh = true;  // outer residual of the start field not yet taken
    int it = 0;         // outer residuals taken so far
    double res0 = 0.0;
};

}  // namespace serve
}  // namespace mgs

using namespace mgs;
using namespace mgs::serve;

struct mgs_server {
    mgs_server_opts o{};
    int device = 0;
    hipStream_t s = nullptr;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::deque<Request*> waiting;
    std::map<int64_t, Request*> reqs;
    int64_t next_ticket = 1;
    Slot slots[MAXS];
    std::map<int, std::vector<Hier*>> pool;  // idle hierarchies by ngrid
    size_t used = 0, budget = 0;
    SlotDesc *dS = nullptr, *hS = nullptr;
    Plan *dP = nullptr, *hP = nullptr;
    double *dsum = nullptr, *hsum = nullptr;
    std::map<int, hipGraphExec_t> graphs;  // by batch depth
    bool desc_dirty = true;
    // statistics
    std::vector<double> lat;
    int64_t submitted = 0, rejected = 0, steps = 0;
    int max_batch = 0;
    bool any_submit = false;
    Clock::time_point first_submit, last_done;
    std::string err;
};

static thread_local std::string g_server_create_error;

static int srv_err(mgs_server* S, int code, const std::string& msg)
{
    if (S)
        S->err = msg;
    else
        g_server_create_error = msg;
    return code;
}

namespace {

// free idle pooled hierarchies (other sizes first) until `need` more bytes fit
void trim_pool(mgs_server* S, size_t need)
{
    for (auto it = S->pool.begin(); it != S->pool.end() && S->used + need > S->budget;) {
        auto& v = it->second;
        while (!v.empty() && S->used + need > S->budget) {
            Hier* h = v.back();
            v.pop_back();
            (void)hipFree(h->base);
            S->used -= h->bytes;
            delete h;
        }
        it = v.empty() ? S->pool.erase(it) : std::next(it);
    }
}

Hier* take_hier(mgs_server* S, int N)
{
    auto it = S->pool.find(N);
    if (it != S->pool.end() && !it->second.empty()) {
        Hier* h = it->second.back();
        it->second.pop_back();
        return h;
    }
    Hier probe;
    const size_t bytes = layout(probe, N, nullptr);
    if (S->used + bytes > S->budget) trim_pool(S, bytes);
    if (S->used + bytes > S->budget) return nullptr;
    void* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) return nullptr;
    Hier* h = new Hier();
    layout(*h, N, (char*)base);
    h->base = base;
    S->used += bytes;
    return h;
}

size_t footprint(int N)
{
    Hier probe;
    return layout(probe, N, nullptr);
}

void fill_desc(SlotDesc& sd, const Hier& h, const mgs_params& p)
{
    std::memset(&sd, 0, sizeof(sd));
    sd.frn = p.fr_n;
    sd.lc = h.lc;
    sd.u = h.u;
    sd.rho = h.rho;
    sd.partials = h.partials;
    sd.fc = fr_coefs(p, 0);
    for (int l = 0; l <= h.lc; l++) {
        LevelDesc& L = sd.L[l];
        L.g = make_geo(h.N >> l);
        L.d = h.d[l];
        L.xb = h.xb[l];
        L.f = h.f[l];
        L.rres = h.rres[l];
        L.rd = h.rd[l];
        L.dc = df_coefs(p, l);
        L.z = shape_of(make_zlaunch(L.g));
        L.zb = shape_of(make_zlaunch_both(L.g));
        if (l > 0) L.zc = shape_of(make_zlaunch(L.g.n, L.g.n));
    }
}

// upload the density of a freshly admitted request and start from u = u_bg
int start_request(mgs_server* S, Slot& sl)
{
    Request* r = sl.r;
    Hier* h = sl.h;
    const int N = r->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    if (r->dtype == MGS_F32) {
        if (hipMemcpyAsync(h->d[0], r->delta.data(), n3 * 4, hipMemcpyHostToDevice, S->s) != hipSuccess)
            return MGS_E_HIP;
        launch_pack_f64_from_f32(h->d[0], h->rho, N, S->s);
    } else {
        if (hipMemcpyAsync(h->u, r->delta.data(), n3 * 8, hipMemcpyHostToDevice, S->s) != hipSuccess)
            return MGS_E_HIP;
        launch_pack_f64_from_f64(h->u, h->rho, N, S->s);
    }
    double ubg = 1.0;
    fr_background(r->p, &ubg);
    launch_fill_f64(h->u, n3, ubg, S->s);
    if (hipStreamSynchronize(S->s) != hipSuccess) return MGS_E_HIP;
    std::vector<char>().swap(r->delta);
    return 0;
}

// copy the converged field out and hand the hierarchy back to the pool
int finish_request(mgs_server* S, Slot& sl, int status)
{
    Request* r = sl.r;
    Hier* h = sl.h;
    int rc = status;
    if (!rc) {
        const int N = r->p.ngrid;
        const size_t n3 = (size_t)N * N * N;
        r->u.resize(n3);
        launch_unpack_f64_to_f64(h->u, h->rho, N, 1.0, S->s);
        if (hipMemcpyAsync(r->u.data(), h->rho, n3 * 8, hipMemcpyDeviceToHost, S->s) != hipSuccess ||
            hipStreamSynchronize(S->s) != hipSuccess)
            rc = MGS_E_HIP;
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->t_done = Clock::now();
        r->res.status = rc;
        r->res.queued_ms = ms_between(r->t_submit, r->t_admit);
        r->res.latency_ms = ms_between(r->t_submit, r->t_done);
        r->done = true;
        S->pool[h->N].push_back(h);
        if (!rc) {
            S->lat.push_back(r->res.latency_ms);
            S->last_done = r->t_done;
        }
    }
    sl = Slot();
    S->desc_dirty = true;
    S->cv_done.notify_all();
    return rc;
}

// move waiting requests into free slots, FIFO, while their hierarchies fit
int admit(mgs_server* S)
{
    int admitted = 0;
    for (;;) {
        Request* r = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->waiting.empty()) break;
            r = S->waiting.front();
        }
        int free_slot = -1;
        int active = 0;
        for (int i = 0; i < S->o.max_active; i++) {
            if (S->slots[i].r)
                active++;
            else if (free_slot < 0)
                free_slot = i;
        }
        if (free_slot < 0) break;
        Hier* h = take_hier(S, r->p.ngrid);
        if (!h) {
            if (active == 0) {
                // it cannot fit even alone: fail it rather than wait forever
                std::lock_guard<std::mutex> lk(S->mu);
                S->waiting.pop_front();
                r->t_admit = r->t_done = Clock::now();
                r->res.status = MGS_E_NOMEM;
                r->done = true;
                S->cv_done.notify_all();
                continue;
            }
            break;
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->waiting.pop_front();
            r->t_admit = Clock::now();
        }
        Slot& sl = S->slots[free_slot];
        sl.r = r;
        sl.h = h;
        sl.fresh = true;
        sl.it = 0;
        sl.res0 = 0.0;
        const int rc = start_request(S, sl);
        if (rc) {
            finish_request(S, sl, rc);
            continue;
        }
        S->desc_dirty = true;
        admitted++;
    }
    return admitted;
}

void build_plan(mgs_server* S, int* depth, bool* any_inner)
{
    Plan& P = *S->hP;
    std::memset(&P, 0, sizeof(P));
    *depth = 0;
    *any_inner = false;
    for (int i = 0; i < S->o.max_active; i++) {
        Slot& sl = S->slots[i];
        if (!sl.r) continue;
        const SlotDesc& sd = S->hS[i];
        P.res_slot[P.nres] = i;
        P.res_pre[P.nres + 1] = P.res_pre[P.nres] + sd.L[0].zb.NB;
        P.nres++;
        if (sl.fresh) continue;
        *any_inner = true;
        if (sd.lc > *depth) *depth = sd.lc;
        for (int l = 0; l < sd.lc; l++) {
            const int k = P.ninner[l];
            const LevelDesc& C = sd.L[l + 1];
            const unsigned nc = (unsigned)C.g.n;
            P.inner_slot[l][k] = i;
            P.z_pre[l][k + 1] = P.z_pre[l][k] + sd.L[l].z.NB;
            P.zc_pre[l][k + 1] = P.zc_pre[l][k] + C.zc.NB;
            P.cell_pre[l][k + 1] = P.cell_pre[l][k] + (nc * nc * nc + 255u) / 256u;
            P.zn_pre[l][k + 1] = P.zn_pre[l][k] + C.zb.NB;
            P.ninner[l]++;
        }
        P.coarse_slot[sd.lc][P.ncoarse[sd.lc]++] = i;
    }
}

// the inner V-cycle of every non-fresh request, recursing to `depth`
void vcycle_batched(mgs_server* S, int l, int depth)
{
    hipStream_t s = S->s;
    const dim3 G(GRID), B(256);
    hipLaunchKernelGGL(kb_coarsest, dim3(MAXS), dim3(1024), 0, s, S->dS, S->dP, l);
    if (l >= depth) return;
    // pre-smoothing (level 0: the first colour-0 pass came with the outer residual)
    const int npre = 2 * (l == 0 ? MIXED_PRE0 : 2);
    for (int it = l == 0 ? 1 : 0; it < npre; it++) hipLaunchKernelGGL(kb_smooth, G, B, 0, s, S->dS, S->dP, l, it & 1);
    hipLaunchKernelGGL(kb_restrict_residual, G, B, 0, s, S->dS, S->dP, l);
    hipLaunchKernelGGL(kb_rhs, G, B, 0, s, S->dS, S->dP, l);
    vcycle_batched(S, l + 1, depth);
    hipLaunchKernelGGL(kb_correct, G, B, 0, s, S->dS, S->dP, l);
    // post-smoothing (level 0: the last colour-1 pass is fused with the update)
    const int npost = l == 0 ? 2 * MIXED_POST0 - 1 : 4;
    for (int it = 0; it < npost; it++) hipLaunchKernelGGL(kb_smooth, G, B, 0, s, S->dS, S->dP, l, it & 1);
    if (l == 0) hipLaunchKernelGGL(kb_smooth_update, G, B, 0, s, S->dS, S->dP);
}

// one step: inner cycle + update of the non-fresh requests, then the outer
// residual and its norm for all; depth < 0: no inner cycle
void enqueue_step(mgs_server* S, int depth)
{
    hipStream_t s = S->s;
    if (depth >= 0) {
        for (int l = 0; l < depth; l++)
            hipLaunchKernelGGL(kb_restrict_base, dim3(GRID), dim3(256), 0, s, S->dS, S->dP, l);
        vcycle_batched(S, 0, depth);
    }
    hipLaunchKernelGGL(kb_outer_residual, dim3(GRID), dim3(256), 0, s, S->dS, S->dP);
    hipLaunchKernelGGL(kb_sums, dim3(MAXS), dim3(1024), 0, s, S->dS, S->dP, S->dsum);
}

int run_step(mgs_server* S, int depth)
{
    if (!S->o.use_graphs) {
        enqueue_step(S, depth);
        return hipGetLastError() == hipSuccess ? 0 : MGS_E_HIP;
    }
    auto it = S->graphs.find(depth);
    if (it == S->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        if (hipStreamBeginCapture(S->s, hipStreamCaptureModeThreadLocal) != hipSuccess) return MGS_E_HIP;
        enqueue_step(S, depth);
        if (hipStreamEndCapture(S->s, &g) != hipSuccess) return MGS_E_HIP;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return MGS_E_HIP;
        it = S->graphs.emplace(depth, ex).first;
    }
    return hipGraphLaunch(it->second, S->s) == hipSuccess ? 0 : MGS_E_HIP;
}

void worker_main(mgs_server* S)
{
    (void)hipSetDevice(S->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            bool any_active = false;
            for (int i = 0; i < S->o.max_active; i++) any_active = any_active || S->slots[i].r;
            S->cv_work.wait(lk, [&] {
                bool act = false;
                for (int i = 0; i < S->o.max_active; i++) act = act || S->slots[i].r;
                return S->stop || act || !S->waiting.empty();
            });
            if (S->stop) break;
            (void)any_active;
        }
        admit(S);
        int nact = 0;
        for (int i = 0; i < S->o.max_active; i++) nact += S->slots[i].r ? 1 : 0;
        if (!nact) {
            // the head of the queue is waiting for memory held by nobody: it
            // was failed by admit(); sleep until new work arrives
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            continue;
        }
        // descriptors and plan for this batch
        for (int i = 0; i < S->o.max_active; i++)
            if (S->slots[i].r && S->desc_dirty) fill_desc(S->hS[i], *S->slots[i].h, S->slots[i].r->p);
        int depth = 0;
        bool any_inner = false;
        build_plan(S, &depth, &any_inner);
        int rc = 0;
        if (hipMemcpyAsync(S->dS, S->hS, sizeof(SlotDesc) * MAXS, hipMemcpyHostToDevice, S->s) != hipSuccess ||
            hipMemcpyAsync(S->dP, S->hP, sizeof(Plan), hipMemcpyHostToDevice, S->s) != hipSuccess)
            rc = MGS_E_HIP;
        S->desc_dirty = false;
        if (!rc) rc = run_step(S, any_inner ? depth : -1);
        if (!rc && (hipMemcpyAsync(S->hsum, S->dsum, sizeof(double) * MAXS, hipMemcpyDeviceToHost, S->s) != hipSuccess ||
                    hipStreamSynchronize(S->s) != hipSuccess))
            rc = MGS_E_HIP;
        S->steps++;
        S->max_batch = std::max(S->max_batch, nact);
        // per request: the stopping rule of solve_mixed (mode 1)
        for (int i = 0; i < S->o.max_active; i++) {
            Slot& sl = S->slots[i];
            if (!sl.r) continue;
            if (rc) {
                finish_request(S, sl, rc);
                continue;
            }
            Request* r = sl.r;
            const int N = r->p.ngrid;
            const double res = std::sqrt(S->hsum[i] / ((double)N * N * N));
            sl.it++;
            if (!std::isfinite(res)) {
                finish_request(S, sl, MGS_E_DOMAIN);
                continue;
            }
            bool stop = false, conv = false;
            if (sl.it == 1) {
                sl.res0 = res;
                r->res.res0 = res;
                if (res == 0.0) stop = conv = true;
            } else {
                conv = res < r->rel_tol * sl.res0;
                stop = conv || (sl.it - 1) >= r->max_outer;
            }
            r->res.res_final = res;
            r->res.outer = sl.it - 1;
            r->res.converged = conv ? 1 : 0;
            if (sl.fresh) {
                sl.fresh = false;
                S->desc_dirty = true;  // joins the inner cycle: new plan
            }
            if (stop) finish_request(S, sl, 0);
        }
    }
}

}  // namespace

extern "C" {

int mgs_server_create(const mgs_server_opts* opts, mgs_server** out)
{
    if (!out) return srv_err(nullptr, MGS_E_ARG, "out is NULL");
    *out = nullptr;
    mgs_server_opts o{};
    if (opts) o = *opts;
    if (o.max_active <= 0) o.max_active = 16;
    if (o.max_active > MAXS) return srv_err(nullptr, MGS_E_ARG, "max_active must be <= 64");
    if (o.queue_capacity <= 0) o.queue_capacity = 1 << 16;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return srv_err(nullptr, MGS_E_HIP, "no HIP device available: libmgsolve has no CPU fallback");
    mgs_server* S = new mgs_server();
    S->o = o;
    (void)hipGetDevice(&S->device);
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
        delete S;
        return srv_err(nullptr, MGS_E_HIP, "hipMemGetInfo");
    }
    S->budget = o.hbm_budget ? (size_t)o.hbm_budget : (size_t)(0.6 * (double)freeb);
    if (hipStreamCreate(&S->s) != hipSuccess || hipMalloc((void**)&S->dS, sizeof(SlotDesc) * MAXS) != hipSuccess ||
        hipMalloc((void**)&S->dP, sizeof(Plan)) != hipSuccess ||
        hipMalloc((void**)&S->dsum, sizeof(double) * MAXS) != hipSuccess ||
        hipHostMalloc((void**)&S->hS, sizeof(SlotDesc) * MAXS, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&S->hP, sizeof(Plan), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&S->hsum, sizeof(double) * MAXS, hipHostMallocDefault) != hipSuccess) {
        mgs_server_destroy(S);
        return srv_err(nullptr, MGS_E_NOMEM, "server buffers");
    }
    std::memset(S->hS, 0, sizeof(SlotDesc) * MAXS);
    S->worker = std::thread(worker_main, S);
    *out = S;
    return 0;
}

int mgs_server_destroy(mgs_server* S)
{
    if (!S) return 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    if (S->s) (void)hipStreamSynchronize(S->s);
    for (auto& kv : S->graphs) (void)hipGraphExecDestroy(kv.second);
    for (int i = 0; i < MAXS; i++)
        if (S->slots[i].h) {
            (void)hipFree(S->slots[i].h->base);
            delete S->slots[i].h;
        }
    for (auto& kv : S->pool)
        for (Hier* h : kv.second) {
            (void)hipFree(h->base);
            delete h;
        }
    for (auto& kv : S->reqs) delete kv.second;
    if (S->dS) (void)hipFree(S->dS);
    if (S->dP) (void)hipFree(S->dP);
    if (S->dsum) (void)hipFree(S->dsum);
    if (S->hS) (void)hipHostFree(S->hS);
    if (S->hP) (void)hipHostFree(S->hP);
    if (S->hsum) (void)hipHostFree(S->hsum);
    if (S->s) (void)hipStreamDestroy(S->s);
    delete S;
    return 0;
}

const char* mgs_server_last_error(const mgs_server* S)
{
    return S ? S->err.c_str() : g_server_create_error.c_str();
}

int mgs_server_submit(mgs_server* S, const mgs_request* req, int64_t* ticket)
{
    if (!S || !req || !ticket || !req->delta) return srv_err(S, MGS_E_ARG, "NULL argument");
    std::string why;
    int rc = validate_params(&req->params, &why);
    if (rc) return srv_err(S, rc, why);
    const mgs_params& p = req->params;
    if (p.model != MGS_MODEL_FR || p.precision != MGS_PREC_MIXED)
        return srv_err(S, MGS_E_ARG, "the server solves f(R) requests in MGS_PREC_MIXED");
    if (p.ngrid < 16 || p.ngrid > 1024) return srv_err(S, MGS_E_ARG, "the server takes 16 <= ngrid <= 1024");
    if (coarsest_of(p.ngrid) + 1 > MAXL) return srv_err(S, MGS_E_ARG, "too many levels");
    if (req->dtype != MGS_F32 && req->dtype != MGS_F64) return srv_err(S, MGS_E_ARG, "dtype");
    if (!(req->rel_tol > 0.0) || req->max_outer < 1) return srv_err(S, MGS_E_ARG, "rel_tol > 0 and max_outer >= 1");
    if (footprint(p.ngrid) > S->budget) return srv_err(S, MGS_E_NOMEM, "request larger than the server's HBM budget");
    const size_t n3 = (size_t)p.ngrid * p.ngrid * p.ngrid;
    const size_t es = req->dtype == MGS_F32 ? 4 : 8;
    for (size_t q = 0; q < n3; q++) {
        const double v = req->dtype == MGS_F32 ? (double)((const float*)req->delta)[q] : ((const double*)req->delta)[q];
        if (!std::isfinite(v))
            return srv_err(S, MGS_E_DOMAIN, "density contrast must be finite (cell " + std::to_string(q) + ")");
    }
    Request* r = new Request();
    r->p = p;
    r->dtype = req->dtype;
    r->max_outer = req->max_outer;
    r->rel_tol = req->rel_tol;
    r->delta.assign((const char*)req->delta, (const char*)req->delta + n3 * es);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if ((int)S->waiting.size() >= S->o.queue_capacity) {
            S->rejected++;
            delete r;
            S->err = "queue full";
            return MGS_E_BUSY;
        }
        r->ticket = S->next_ticket++;
        r->t_submit = Clock::now();
        if (!S->any_submit) {
            S->any_submit = true;
            S->first_submit = r->t_submit;
        }
        S->submitted++;
        S->reqs[r->ticket] = r;
        S->waiting.push_back(r);
        *ticket = r->ticket;
    }
    S->cv_work.notify_all();
    return 0;
}

int mgs_server_wait(mgs_server* S, int64_t ticket, void* u, int32_t dtype, double scale, mgs_request_result* res)
{
    if (!S) return MGS_E_ARG;
    if (u && dtype != MGS_F32 && dtype != MGS_F64) return srv_err(S, MGS_E_ARG, "dtype");
    Request* r = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end()) return srv_err(S, MGS_E_ARG, "unknown ticket");
        r = it->second;
        S->cv_done.wait(lk, [&] { return r->done; });
        S->reqs.erase(it);
    }
    int rc = r->res.status;
    if (res) *res = r->res;
    if (!rc && u) {
        const size_t n3 = r->u.size();
        for (size_t q = 0; q < n3; q++) {
            const double v = scale == 1.0 ? r->u[q] : r->u[q] * scale;
            if (dtype == MGS_F32)
                ((float*)u)[q] = (float)v;
            else
                ((double*)u)[q] = v;
        }
    }
    if (rc) S->err = "request failed (code " + std::to_string(rc) + ")";
    delete r;
    return rc;
}

int mgs_server_get_stats(mgs_server* S, mgs_server_stats* st)
{
    if (!S || !st) return MGS_E_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = S->submitted;
    st->completed = (int64_t)S->lat.size();
    st->rejected = S->rejected;
    st->steps = S->steps;
    st->max_batch = S->max_batch;
    if (!S->lat.empty()) {
        std::vector<double> v = S->lat;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            const double pos = q * (double)(v.size() - 1);
            const size_t lo = (size_t)pos;
            const size_t hi = std::min(lo + 1, v.size() - 1);
            return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        double sum = 0.0;
        for (double x : v) sum += x;
        st->mean_ms = sum / (double)v.size();
        const double span = ms_between(S->first_submit, S->last_done) * 1e-3;
        st->solves_per_s = span > 0.0 ? (double)v.size() / span : 0.0;
    }
    return 0;
}

}  // extern "C"
