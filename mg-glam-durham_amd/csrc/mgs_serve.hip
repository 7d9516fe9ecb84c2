// mgs_serve.hip -- continuous-batching f(R) / nDGP solve server (mgs_server_* C-ABI).
//
// Requests wait in a FIFO queue; a worker thread admits them into slots as
// slots and HBM budget free up, and advances every active request by one
// outer iteration of the mixed-precision solve per step.  A step is a fixed
// sequence of batched persistent launches: each launch walks the tiles of
// all participating requests (a prefix sum of their tile counts in a device
// plan), so one launch smooths, restricts or corrects the same level of every
// request at once, whatever their grid sizes and parameters.  The sequence
// depends only on the deepest f(R) and nDGP hierarchies in the batch, so it
// is captured once per pair of depths as a hipGraph and replayed; requests joining or leaving only
// change the plan in device memory.
//
// Per request the arithmetic, the tiling and the reduction order are those of
// the single-request solve (mgs_api.cpp, solve_mixed with one inner cycle),
// so a result is bit-identical alone, in any mix, and to mgs_solve.  f(R)
// and nDGP requests share a step: each model has its own participation lists
// and kernels (the nDGP inner cycle is vcycle_inner_dgp's: out-of-place
// colour passes through a per-request scratch array, 19-point operators on
// FP64 densities cached on every level); the outer residual and its norm are
// one launch over both.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <sys/mman.h>

#include "../../include/mgsolve.h"
#include "mgs_dgp_dev.h"
#include "mgs_fr_dev.h"
#include "mgs_internal.h"
#include "mgs_kernels.h"

namespace mgs {
namespace serve {

constexpr int MAXS = 64;    // slots
constexpr int MAXL = 11;    // levels of a hierarchy (ngrid <= 1024)
constexpr int GRID = 8192;  // blocks of a batched stencil/transfer launch (tiles are walked, XCD-contiguous)
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
    // nDGP
    const double* rhol;  // FP64 density of this level
    DgpF gf;
};

struct SlotDesc {
    int frn, lc, model, zdgp;  // zdgp: nDGP outer residual on z-marching tiles (dgp_zmarch(ngrid))
    double* u;
    const double* rho;
    double* partials;
    FrCoef fc;
    // nDGP: FP64 operator coefficients of level 0, the tiles of its outer
    // residual (as launch_dgpm_outer: z-marching runs, else make_launch(h,
    // 2 n^2)), and scratch of ngrid^3 floats (the colour passes'
    // out-of-place half, then the residual before restriction)
    DgpCoef gc0;
    Shape zo;
    float* scr;
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
    // nDGP requests, in blocks of 256 cells
    int nd[MAXL];
    int d_slot[MAXL][MAXS];
    unsigned dh_pre[MAXL][MAXS + 1];  // level-l cells of one colour
    unsigned df_pre[MAXL][MAXS + 1];  // level-l cells
    unsigned dc_pre[MAXL][MAXS + 1];  // level-(l+1) cells
    int ndcoarse[MAXL];
    int dcoarse_slot[MAXL][MAXS];
    int nupd;
    int upd_slot[MAXS];
    unsigned upd_pre[MAXS + 1];  // level-0 cells (phi += e)
    // f(R) requests by where their inner cycle stands in its outer
    // iteration (level-0 stencil tiles): not the first cycle (their level-0
    // pre-smoothing starts with a colour-0 pass), the last one (its final
    // colour-1 pass applies the outer update), not the last one (a plain
    // final colour-1 pass)
    int nnf, nlast, nnl;
    int nf_slot[MAXS], last_slot[MAXS], nl_slot[MAXS];
    unsigned nf_pre[MAXS + 1], last_pre[MAXS + 1], nl_pre[MAXS + 1];
    // f(R) requests in the first cycle of an outer iteration: the base xb
    // restricted from level l to l+1 (level-(l+1) cells in blocks of 256)
    int nrb[MAXL];
    int rb_slot[MAXL][MAXS];
    unsigned rb_pre[MAXL][MAXS + 1];
};

// the level-0 lists of the f(R) inner cycle
enum { LIST_NOT_FIRST = 0, LIST_LAST = 1, LIST_NOT_LAST = 2 };

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

// The tiles 0 .. total-1 of a batched launch, XCD-contiguous: the blocks of
// XCD x (blockIdx.x % 8 == x, the hardware's round-robin placement of
// blocks) take the x-th eighth of the tile list, in order.  Neighbouring
// tiles (a plane run's halo planes, the rows beside a slab) then share an
// XCD's L2, as map_block arranges for a single-request launch.  A tile's
// values do not depend on which block runs it.
struct XcdTiles {
    unsigned t, end, stride;
    __device__ __forceinline__ explicit XcdTiles(unsigned total)
    {
        const unsigned per = (total + 7u) / 8u, x = blockIdx.x & 7u;
        t = x * per + (blockIdx.x >> 3);
        end = (x + 1u) * per < total ? (x + 1u) * per : total;
        stride = gridDim.x >> 3;
    }
};

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

// thread (x, row) of tile lb of a make_launch(width, rows) grid, as map_xy
// (the tile is block lb of that grid, thread threadIdx.x = tx + TX ty)
__device__ __forceinline__ bool tile_xy(int width, unsigned rows, const Shape& z, unsigned lb, int& x, unsigned& row)
{
    const int tx = (int)threadIdx.x % z.TX, ty = (int)threadIdx.x / z.TX;
    if (ty >= z.TY) return false;
    const unsigned xb = z.XB == 1 ? 0u : lb % (unsigned)z.XB, rb = z.XB == 1 ? lb : lb / (unsigned)z.XB;
    x = (int)(xb * (unsigned)z.TX) + tx;
    row = rb * (unsigned)z.TY + (unsigned)ty;
    return x < width && row < rows;
}

// FP64 outer residual (and the inner problem's g, xb, d: f(R) after its
// first colour-0 pass, nDGP d = 0) on level 0; one block partial per tile, as
// k_both_colours / k_dgpm_outer
__global__ __launch_bounds__(256) void kb_outer_residual(const SlotDesc* __restrict__ S, const Plan* __restrict__ P)
{
    const unsigned total = P->res_pre[P->nres];
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
        const int i = find_slot(P->res_pre, P->nres, t);
        const SlotDesc& sd = S[P->res_slot[i]];
        const LevelDesc& L = sd.L[0];
        const unsigned lb = t - P->res_pre[i];
        double acc = 0.0;
        if (sd.model == MGS_MODEL_DGP && sd.zdgp) {
            ZCol c;
            if (tile_col(L.g, L.g.h, sd.zo, lb, c))
                acc = dgp_both_zcol(DgpmOuterOp<true>{sd.u, sd.rho, L.f, L.xb, L.d, sd.gc0, (unsigned)L.g.half}, L.g,
                                    c.m, c.j, c.k0);
        } else if (sd.model == MGS_MODEL_DGP) {
            int m;
            unsigned row;
            if (tile_xy(L.g.h, 2u * (unsigned)L.g.n * (unsigned)L.g.n, sd.zo, lb, m, row))
                acc = dgpm_outer_cell(sd.u, sd.rho, L.f, L.xb, L.d, L.g, sd.gc0, m, row);
        } else {
            ZCol c;
            if (tile_col(L.g, L.g.h, L.zb, lb, c)) {
                FRN_SWITCH(sd.frn,
                           acc = both_colours_col(FrOuterOp<FRN, false, true>{sd.u, sd.rho, L.f, L.xb, L.d, sd.fc,
                                                                              L.dc, (unsigned)L.g.half},
                                                  L.g, c));
            }
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
        const int count = (int)(P->res_pre[i + 1] - P->res_pre[i]);
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

// xb of level l+1 from level l, for the requests starting an outer iteration
__global__ __launch_bounds__(256) void kb_restrict_base(const SlotDesc* __restrict__ S, const Plan* __restrict__ P,
                                                        int l)
{
    const unsigned total = P->rb_pre[l][P->nrb[l]];
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
        const int i = find_slot(P->rb_pre[l], P->nrb[l], t);
        const SlotDesc& sd = S[P->rb_slot[l][i]];
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        const unsigned nc = (unsigned)C.g.n;
        const unsigned cell = (t - P->rb_pre[l][i]) * 256u + threadIdx.x;
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
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
        const int i = find_slot(P->z_pre[l], P->ninner[l], t);
        const SlotDesc& sd = S[P->inner_slot[l][i]];
        const LevelDesc& L = sd.L[l];
        ZCol c;
        if (tile_col(L.g, L.g.h, L.z, t - P->z_pre[l][i], c)) {
            FRN_SWITCH(sd.frn, colour_pass_col(DfSmoothOp<FRN>{L.d, L.xb, L.f, L.dc}, L.g, colour, c));
        }
    }
}

// one level-0 list of the f(R) inner cycle (LIST_*)
__device__ __forceinline__ void level0_list(const Plan* __restrict__ P, int list, int& n, const int*& slot,
                                            const unsigned*& pre)
{
    if (list == LIST_NOT_FIRST) {
        n = P->nnf;
        slot = P->nf_slot;
        pre = P->nf_pre;
    } else if (list == LIST_LAST) {
        n = P->nlast;
        slot = P->last_slot;
        pre = P->last_pre;
    } else {
        n = P->nnl;
        slot = P->nl_slot;
        pre = P->nl_pre;
    }
}

// one colour pass of the inner smoother on level 0 for the requests of a list
__global__ __launch_bounds__(256) void kb_smooth0(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int list,
                                                  int colour)
{
    int n;
    const int* slot;
    const unsigned* pre;
    level0_list(P, list, n, slot, pre);
    const unsigned total = pre[n];
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
        const int i = find_slot(pre, n, t);
        const SlotDesc& sd = S[slot[i]];
        const LevelDesc& L = sd.L[0];
        ZCol c;
        if (tile_col(L.g, L.g.h, L.z, t - pre[i], c)) {
            FRN_SWITCH(sd.frn, colour_pass_col(DfSmoothOp<FRN, true>{L.d, L.xb, L.f, L.dc}, L.g, colour, c));
        }
    }
}

// the last colour-1 pass on level 0 fused with the outer update of u, for
// the requests in the last inner cycle of their outer iteration
__global__ __launch_bounds__(256) void kb_smooth_update(const SlotDesc* __restrict__ S, const Plan* __restrict__ P)
{
    const unsigned total = P->last_pre[P->nlast];
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
        const int i = find_slot(P->last_pre, P->nlast, t);
        const SlotDesc& sd = S[P->last_slot[i]];
        const LevelDesc& L = sd.L[0];
        ZCol c;
        if (tile_col(L.g, L.g.h, L.z, t - P->last_pre[i], c)) {
            FRN_SWITCH(sd.frn,
                       colour_pass_col(DfSmoothUpdateOp<FRN, true>{L.d, L.xb, L.f, sd.u, L.dc, sd.fc.ubg}, L.g, 1, c));
        }
    }
}

// residual of level l restricted to level l+1, with R d (frozen copy rd and
// the start value d of level l+1)
__global__ __launch_bounds__(256) void kb_restrict_residual(const SlotDesc* __restrict__ S,
                                                            const Plan* __restrict__ P, int l)
{
    const unsigned total = P->zc_pre[l][P->ninner[l]];
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
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
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
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
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
        const int i = find_slot(P->zc_pre[l], P->ninner[l], t);
        const SlotDesc& sd = S[P->inner_slot[l][i]];
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        ZCol c;
        if (tile_col(C.g, C.g.n, C.zc, t - P->zc_pre[l][i], c)) {
            if (l == 0) {
                FRN_SWITCH(sd.frn, (df_correct_col<FRN, false, true>(C.d, C.rd, F.d, F.xb, F.g, C.g, c)));
            } else {
                FRN_SWITCH(sd.frn, df_correct_col<FRN>(C.d, C.rd, F.d, F.xb, F.g, C.g, c));
            }
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

// ---- nDGP (vcycle_inner_dgp's steps) ----------------------------------------
// Each launch walks blocks of 256 cells of every participating request; the
// cell bodies are those of the single-request kernels, so the order cells are
// visited in does not change a value.

// body(request, cell) for every cell of the blocks of 256 cells listed by
// the prefix sums `pre` of the requests `slots`
template <typename Body>
__device__ __forceinline__ void dgp_cells(const SlotDesc* __restrict__ S, const unsigned* pre, int count,
                                          const int* slots, Body body)
{
    const unsigned total = pre[count];
    for (XcdTiles it(total); it.t < it.end; it.t += it.stride) {
        const unsigned t = it.t;
        const int i = find_slot(pre, count, t);
        body(S[slots[i]], (t - pre[i]) * 256u + threadIdx.x);
    }
}

// (I, J, K) of cell `cell` of an nc^3 level
struct Ijk {
    int I, J, K;
};
__device__ __forceinline__ Ijk ijk_of(int nc, unsigned cell)
{
    const unsigned n = (unsigned)nc;
    return Ijk{(int)(cell % n), (int)((cell / n) % n), (int)(cell / (n * n))};
}

// xb of level l+1 from level l
__global__ __launch_bounds__(256) void kbd_restrict_base(const SlotDesc* __restrict__ S, const Plan* __restrict__ P,
                                                         int l)
{
    dgp_cells(S, P->dc_pre[l], P->nd[l], P->d_slot[l], [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        if (cell < (unsigned)(2 * C.g.half)) {
            const Ijk c = ijk_of(C.g.n, cell);
            restrict8_cell<float>(F.xb, C.xb, nullptr, F.g, C.g, c.I, c.J, c.K);
        }
    });
}

// one colour pass of the inner smoother on level l, out of place into scr
__global__ __launch_bounds__(256) void kbd_smooth(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l,
                                                  int colour)
{
    dgp_cells(S, P->dh_pre[l], P->nd[l], P->d_slot[l], [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc& L = sd.L[l];
        if (cell < (unsigned)L.g.half)
            dgpm_smooth_cell(L.d, L.xb, L.rhol, L.f, sd.scr, L.g, colour, L.gf, (int)(cell % (unsigned)L.g.h),
                             cell / (unsigned)L.g.h);
    });
}

// the pass's colour half back from scr
__global__ __launch_bounds__(256) void kbd_copy(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l,
                                                int colour)
{
    dgp_cells(S, P->dh_pre[l], P->nd[l], P->d_slot[l], [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc& L = sd.L[l];
        if (cell < (unsigned)L.g.half) L.d[(size_t)colour * L.g.half + cell] = sd.scr[cell];
    });
}

// inner residual D(e) - f of level l into scr
__global__ __launch_bounds__(256) void kbd_residual(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    dgp_cells(S, P->df_pre[l], P->nd[l], P->d_slot[l], [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc& L = sd.L[l];
        if (cell < (unsigned)(2 * L.g.half))
            dgpm_residual_cell(L.d, L.xb, L.rhol, L.f, sd.scr, L.g, L.gf, (int)(cell % (unsigned)L.g.h),
                               cell / (unsigned)L.g.h);
    });
}

// R r and R e of level l: rres, and rd and the start value d of level l+1
__global__ __launch_bounds__(256) void kbd_restrict(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    dgp_cells(S, P->dc_pre[l], P->nd[l], P->d_slot[l], [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        if (cell < (unsigned)(2 * C.g.half)) {
            const Ijk c = ijk_of(C.g.n, cell);
            restrict8_cell<float>(sd.scr, C.rres, nullptr, F.g, C.g, c.I, c.J, c.K);
            restrict8_cell<float>(F.d, C.rd, C.d, F.g, C.g, c.I, c.J, c.K);
        }
    });
}

// FAS right-hand side of level l+1: f = D(R e) - R r
__global__ __launch_bounds__(256) void kbd_rhs(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    dgp_cells(S, P->dc_pre[l], P->nd[l], P->d_slot[l], [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc& C = sd.L[l + 1];
        if (cell < (unsigned)(2 * C.g.half))
            dgpm_residual_cell(C.rd, C.xb, C.rhol, C.rres, C.f, C.g, C.gf, (int)(cell % (unsigned)C.g.h),
                               cell / (unsigned)C.g.h);
    });
}

// coarse-grid correction of level l from level l+1
__global__ __launch_bounds__(256) void kbd_correct(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    dgp_cells(S, P->dc_pre[l], P->nd[l], P->d_slot[l], [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc &F = sd.L[l], &C = sd.L[l + 1];
        if (cell < (unsigned)(2 * C.g.half)) {
            const Ijk c = ijk_of(C.g.n, cell);
            dgpm_correct_cell(C.d, C.rd, F.d, F.g, C.g, c.I, c.J, c.K);
        }
    });
}

// phi += e on level 0
__global__ __launch_bounds__(256) void kbd_update(const SlotDesc* __restrict__ S, const Plan* __restrict__ P)
{
    dgp_cells(S, P->upd_pre, P->nupd, P->upd_slot, [&](const SlotDesc& sd, unsigned cell) {
        const LevelDesc& L = sd.L[0];
        if (cell < (unsigned)(2 * L.g.half)) sd.u[cell] = sd.u[cell] + (double)L.d[cell];
    });
}

// coarsest-level sweeps of the nDGP requests whose coarsest level is l
__global__ __launch_bounds__(1024) void kbd_coarsest(const SlotDesc* __restrict__ S, const Plan* __restrict__ P, int l)
{
    for (int i = blockIdx.x; i < P->ndcoarse[l]; i += gridDim.x) {
        const SlotDesc& sd = S[P->dcoarse_slot[l][i]];
        const LevelDesc& L = sd.L[l];
        dgpm_coarsest_block(L.d, L.xb, L.rhol, L.f, sd.scr, L.g, SWEEPS, L.gf);
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
// inner hierarchy, and the level-0 block partials (nDGP: also the FP64
// density of every coarser level and the scratch array), carved out of one
// allocation
struct Hier {
    int N = 0, lc = 0, model = MGS_MODEL_FR;
    size_t bytes = 0;
    void* base = nullptr;
    double *u = nullptr, *rho = nullptr, *partials = nullptr;
    float *d[MAXL] = {}, *xb[MAXL] = {}, *f[MAXL] = {}, *rres[MAXL] = {}, *rd[MAXL] = {};
    double* rhol[MAXL] = {};  // nDGP: rhol[0] = rho
    float* scr = nullptr;     // nDGP
};

// idle hierarchies are pooled by (ngrid, model)
static int pool_key(int N, int model) { return N * 4 + model; }

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static int coarsest_of(int N)
{
    const int lc = levelmax_of(N) - levelmin_of(N);
    return lc < 0 ? 0 : lc;
}

// carve (or just measure, when base is null) a hierarchy for ngrid N
static size_t layout(Hier& h, int N, int model, char* base)
{
    h.N = N;
    h.model = model;
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
    const bool dgp = model == MGS_MODEL_DGP;
    const Geo g0 = make_geo(N);
    // nDGP: the per-cell tiling's partial count bounds the z-marching one
    const size_t nb = dgp ? (size_t)make_launch(g0.h, 2L * N * N).NB : (size_t)make_zlaunch_both(g0).NB;
    // the partials, launch_sum_partials' chunk sums after them, the norm
    h.partials = (double*)take((nb + 1024 + 64) * 8);
    if (dgp) {
        h.rhol[0] = h.rho;
        for (int l = 1; l <= h.lc; l++) h.rhol[l] = (double*)take((size_t)(N >> l) * (N >> l) * (N >> l) * 8);
        h.scr = (float*)take(n3 * 4);
    }
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
    char* delta = nullptr;  // host staging (take_stage), until admission
    size_t delta_bytes = 0;
    int32_t dtype = MGS_F32;
    int32_t max_outer = 60;
    double rel_tol = 1e-8;
    int32_t exit_rule = MGS_EXIT_REL_TOL;
    double res_conv = 0.0;
    bool verify = false;    // keep delta until wait, which runs the FP64 comparison solve
    bool waited = false;    // a waiter has taken the ticket (under mu)
    Clock::time_point t_submit, t_admit, t_done;
    bool done = false;
    mgs_request_result res{};
    // The result stays in HBM, in the request's hierarchy (hold), until
    // mgs_server_wait copies it out; when admission needs the memory first,
    // the raw field (reference order) is spilled to u.
    Hier* hold = nullptr;
    bool spilling = false;
    std::vector<double> u;
    ~Request() { std::free(delta); }
};

struct Slot {
    Request* r = nullptr;
    Hier* h = nullptr;
    bool pending = false;  // admitted while a step ran: joins the next step
    bool fresh = true;     // outer residual of the start field not yet taken
    int it = 0;         // outer residuals taken so far
    double res0 = 0.0;
    double res1 = 1.0e5, res2 = 1.0e10;  // the two residuals before the last (mode-0 rule)
    // inner cycles of the current outer iteration and those done; f(R)
    // requests with MGS_EXIT_REL_TOL choose the count as mgs_solve's mode 1
    // with inner_cycles 0 (auto_inner_count from the previous outer
    // iteration's start residual and count), others run one
    int inner = 1, ic = 0;
    double res_prev = 0.0;
    int m_prev = 0;
    bool outer_now = false;  // this step ends with the request's outer residual (set by build_plan)
    bool starting = false;   // upload and start residual queued on the admission stream (finish_starts)
    void* upin = nullptr;    // pinned bounce buffer of the upload, returned by finish_starts
};

}  // namespace serve
}  // namespace mgs

using namespace mgs;
using namespace mgs::serve;

struct mgs_server {
    mgs_server_opts o{};
    int device = 0;
    hipStream_t s = nullptr;   // the batched steps
    hipStream_t cs = nullptr;  // admission uploads and spills, beside the steps
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::deque<Request*> waiting;
    std::map<int64_t, Request*> reqs;
    int64_t next_ticket = 1;
    Slot slots[MAXS];
    std::map<int, std::vector<Hier*>> pool;  // idle hierarchies by pool_key
    size_t used = 0, budget = 0;
    SlotDesc *dS = nullptr, *hS = nullptr;
    Plan *dP = nullptr, *hP = nullptr;
    double *dsum = nullptr, *hsum = nullptr;
    double* hstart = nullptr;  // pinned: each slot's start-residual sum, written by the admission stream
    std::map<int, hipGraphExec_t> graphs;  // by the batch's f(R) and nDGP depths
    bool desc_dirty = true;
    // done requests whose result is held in HBM, oldest first, and the
    // hierarchies waiters are copying out (both under mu)
    std::deque<Request*> held;
    int claimed = 0;
    int waiters = 0;  // threads inside mgs_server_wait (under mu); destroy waits for none

    // idle host staging buffers by size (under mu)
    std::map<size_t, std::vector<char*>> stages;
    size_t stage_idle = 0;
    // pinned copy-out staging: idle buffers by size, bytes pinned in all (under mu)
    std::map<size_t, std::vector<void*>> pins;
    size_t pinned = 0;
    // idle copy-out streams of the waiters (under mu), kept until destroy
    std::vector<hipStream_t> wstreams;
    // statistics
    std::vector<double> lat;
    int64_t submitted = 0, rejected = 0, steps = 0, spilled = 0;
    double step_ms = 0.0, admit_ms = 0.0, finish_ms = 0.0;  // worker time
    int max_batch = 0;
    bool any_submit = false;
    Clock::time_point first_submit, last_done;
};

// the calling thread's last server error (submit and wait may run on
// different client threads; a message per thread keeps them apart)
static thread_local std::string g_server_error;

// no NaN or Inf among n values: a branch-free reduction the compiler
// vectorizes (|v| <= max is false for both)
static bool all_finite(const void* a, int32_t dtype, size_t n)
{
    unsigned bad = 0;
    if (dtype == MGS_F32) {
        const float* v = (const float*)a;
        for (size_t q = 0; q < n; q++) bad |= !(std::fabs(v[q]) <= 3.4028234663852886e38f);
    } else {
        const double* v = (const double*)a;
        for (size_t q = 0; q < n; q++) bad |= !(std::fabs(v[q]) <= 1.7976931348623157e308);
    }
    return bad == 0;
}

// Host staging for submitted densities, recycled by size so that steady
// serving copies into resident pages (a fresh buffer pays a page fault per
// 4 KiB on first touch): 2 MiB-aligned with transparent huge pages asked for.
constexpr size_t STAGE_IDLE_MAX = size_t(4) << 30;

static char* take_stage(mgs_server* S, size_t bytes)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->stages.find(bytes);
        if (it != S->stages.end() && !it->second.empty()) {
            char* p = it->second.back();
            it->second.pop_back();
            S->stage_idle -= bytes;
            return p;
        }
    }
    const size_t huge = size_t(2) << 20;
    const size_t len = (bytes + huge - 1) / huge * huge;
    char* p = (char*)std::aligned_alloc(huge, len);
    if (p) (void)madvise(p, len, MADV_HUGEPAGE);
    return p;
}

static void give_stage(mgs_server* S, char* p, size_t bytes)
{
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stage_idle + bytes <= STAGE_IDLE_MAX) {
            S->stages[bytes].push_back(p);
            S->stage_idle += bytes;
            return;
        }
    }
    std::free(p);
}

// Copy n values into dst and check them in the same pass: each 64 KiB chunk
// is checked right after it is copied, while it is still in cache; 4 threads
// on fields of 16 MiB or more.  False when a value is not finite.
static bool copy_finite(char* dst, const void* src, int32_t dtype, size_t n)
{
    const size_t es = dtype == MGS_F32 ? 4 : 8;
    auto part = [=](size_t lo, size_t hi) {
        const size_t ch = (size_t(64) << 10) / es;
        bool ok = true;
        for (size_t a = lo; a < hi; a += ch) {
            const size_t b = std::min(hi, a + ch);
            std::memcpy(dst + a * es, (const char*)src + a * es, (b - a) * es);
            ok = all_finite(dst + a * es, dtype, b - a) && ok;
        }
        return ok;
    };
    const int nt = n * es >= (size_t(16) << 20) ? 4 : 1;
    if (nt == 1) return part(0, n);
    const size_t slice = (n + nt - 1) / nt;
    bool ok[4] = {true, true, true, true};
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) {
        const size_t lo = slice * t, hi = std::min(n, lo + slice);
        if (lo >= hi) break;
        th.emplace_back([&ok, &part, t, lo, hi] { ok[t] = part(lo, hi); });
    }
    ok[0] = part(0, std::min(n, slice));
    for (auto& x : th) x.join();
    return ok[0] && ok[1] && ok[2] && ok[3];
}

static int srv_err(mgs_server* S, int code, const std::string& msg)
{
    (void)S;
    g_server_error = msg;
    return code;
}

namespace {

// free idle pooled hierarchies (other sizes first) until `need` more bytes fit
// free idle pooled hierarchies (other sizes first) until `need` more bytes
// fit.  The pool is shared with the waiters (under mu); S->used is the
// worker's.
void trim_pool(mgs_server* S, size_t need)
{
    std::vector<Hier*> victims;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        size_t freed = 0;
        for (auto it = S->pool.begin(); it != S->pool.end() && S->used - freed + need > S->budget;) {
            auto& v = it->second;
            while (!v.empty() && S->used - freed + need > S->budget) {
                victims.push_back(v.back());
                freed += v.back()->bytes;
                v.pop_back();
            }
            it = v.empty() ? S->pool.erase(it) : std::next(it);
        }
    }
    for (Hier* h : victims) {
        (void)hipFree(h->base);
        S->used -= h->bytes;
        delete h;
    }
}

Hier* take_hier(mgs_server* S, int N, int model)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->pool.find(pool_key(N, model));
        if (it != S->pool.end() && !it->second.empty()) {
            Hier* h = it->second.back();
            it->second.pop_back();
            return h;
        }
    }
    Hier probe;
    const size_t bytes = layout(probe, N, model, nullptr);
    if (S->used + bytes > S->budget) trim_pool(S, bytes);
    if (S->used + bytes > S->budget) return nullptr;
    void* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) return nullptr;
    Hier* h = new Hier();
    layout(*h, N, model, (char*)base);
    h->base = base;
    S->used += bytes;
    return h;
}

size_t footprint(int N, int model)
{
    Hier probe;
    return layout(probe, N, model, nullptr);
}

void fill_desc(SlotDesc& sd, const Hier& h, const mgs_params& p)
{
    std::memset(&sd, 0, sizeof(sd));
    const bool dgp = p.model == MGS_MODEL_DGP;
    sd.frn = p.fr_n;
    sd.lc = h.lc;
    sd.model = dgp ? MGS_MODEL_DGP : MGS_MODEL_FR;
    sd.u = h.u;
    sd.rho = h.rho;
    sd.partials = h.partials;
    if (dgp) {
        const Geo g0 = make_geo(h.N);
        sd.gc0 = dgp_coefs(p, 0);
        sd.zdgp = dgp_zmarch(h.N) ? 1 : 0;
        if (sd.zdgp) {
            sd.zo = shape_of(make_zlaunch(g0.h, h.N, DGP_KC));
        } else {
            const Launch lo = make_launch(g0.h, 2L * h.N * h.N);
            sd.zo = Shape{lo.TX, lo.TY, lo.XB, 0, 0, lo.NB};
        }
        sd.scr = h.scr;
    } else {
        sd.fc = fr_coefs(p, 0);
    }
    for (int l = 0; l <= h.lc; l++) {
        LevelDesc& L = sd.L[l];
        L.g = make_geo(h.N >> l);
        L.d = h.d[l];
        L.xb = h.xb[l];
        L.f = h.f[l];
        L.rres = h.rres[l];
        L.rd = h.rd[l];
        if (dgp) {
            L.rhol = h.rhol[l];
            L.gf = dgpf(dgp_coefs(p, l));
            continue;
        }
        L.dc = df_coefs(p, l);
        L.z = shape_of(make_zlaunch(L.g));
        L.zb = shape_of(make_zlaunch_both(L.g));
        if (l > 0) L.zc = shape_of(make_zlaunch(L.g.n, L.g.n));
    }
}

// Pinned staging for uploads and copy-outs, pooled by size and kept until the
// server is destroyed (a buffer per transfer would cost a pinning each); at
// most PIN_MAX bytes in all, beyond which a transfer goes through the
// runtime's pageable path.
constexpr size_t PIN_MAX = size_t(2) << 30;
constexpr size_t PIN_MIN = size_t(1) << 20;  // smaller fields go direct

void* take_pin(mgs_server* S, size_t bytes)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->pins.find(bytes);
        if (it != S->pins.end() && !it->second.empty()) {
            void* p = it->second.back();
            it->second.pop_back();
            return p;
        }
        if (S->pinned + bytes > PIN_MAX) return nullptr;
        S->pinned += bytes;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->pinned -= bytes;
        return nullptr;
    }
    return p;
}

void give_pin(mgs_server* S, void* p, size_t bytes)
{
    std::lock_guard<std::mutex> lk(S->mu);
    S->pins[bytes].push_back(p);
}

// host copy of a large field with a few threads (one slice each)
void copy_host(void* dst, const void* src, size_t bytes)
{
    const int nt = bytes >= (size_t(16) << 20) ? 4 : 1;
    if (nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t slice = (bytes / nt + 4095) & ~size_t(4095);
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) {
        const size_t o = slice * t;
        if (o >= bytes) break;
        const size_t len = std::min(slice, bytes - o);
        th.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, len); });
    }
    std::memcpy(dst, src, std::min(slice, bytes));
    for (auto& x : th) x.join();
}

// upload the density of a freshly admitted request (nDGP: and restrict it to
// every level, as cache_densities), start from u = u_bg (nDGP: phi = 0) and
// take the start field's outer residual.  Queued on the admission stream
// (the hierarchy is idle, from the pool or new, so this runs beside a step
// of the active requests); finish_starts waits for it and applies the norm.
int start_request(mgs_server* S, Slot& sl, int slot)
{
    Request* r = sl.r;
    Hier* h = sl.h;
    hipStream_t cs = S->cs;
    const int N = r->p.ngrid;
    const size_t n3 = (size_t)N * N * N;
    // through a pinned bounce buffer: a multi-threaded host copy into it, then
    // a DMA that runs on while the worker goes on (the runtime's pageable
    // path copies on one thread and returns only when it is done)
    const size_t bytes = n3 * (r->dtype == MGS_F32 ? 4 : 8);
    void* dst = r->dtype == MGS_F32 ? (void*)h->d[0] : (void*)h->u;
    const void* src = r->delta;
    sl.upin = bytes >= PIN_MIN ? take_pin(S, bytes) : nullptr;
    if (sl.upin) {
        copy_host(sl.upin, r->delta, bytes);
        src = sl.upin;
    }
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, cs) != hipSuccess) return MGS_E_HIP;
    if (r->dtype == MGS_F32)
        launch_pack_f64_from_f32(h->d[0], h->rho, N, cs);
    else
        launch_pack_f64_from_f64(h->u, h->rho, N, cs);
    double ubg = 0.0;
    if (h->model == MGS_MODEL_DGP) {
        for (int l = 1; l <= h->lc; l++) launch_restrict8_64(h->rhol[l - 1], h->rhol[l], nullptr, N >> l, cs);
    } else {
        fr_background(r->p, &ubg);
    }
    launch_fill_f64(h->u, n3, ubg, cs);
    // the start field's outer residual and its norm, as the single-request
    // solve takes them (same tiles, same partial sums, launch_sum_partials'
    // order); they also set the first inner problem's g, xb and d
    int nb;
    if (h->model == MGS_MODEL_DGP)
        nb = launch_dgpm_outer(h->u, h->rho, h->f[0], h->xb[0], h->d[0], h->partials, N, dgp_coefs(r->p, 0), cs);
    else
        nb = launch_outer_residual(r->p.fr_n, h->u, h->rho, h->f[0], h->xb[0], h->d[0], h->partials, N,
                                   fr_coefs(r->p, 0), df_coefs(r->p, 0), false, cs);
    double* dsum = h->partials + nb + 1024;
    launch_sum_partials(h->partials, nb, dsum, cs);
    if (hipMemcpyAsync(&S->hstart[slot], dsum, sizeof(double), hipMemcpyDeviceToHost, cs) != hipSuccess)
        return MGS_E_HIP;
    sl.starting = true;
    return 0;
}

// the request is done: record it; a failed request's hierarchy goes back to
// the pool, a solved one's stays with the request (its result)
void complete_request(mgs_server* S, Request* r, Hier* h, int rc)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->t_done = Clock::now();
        r->res.status = rc;
        r->res.queued_ms = ms_between(r->t_submit, r->t_admit);
        r->res.latency_ms = ms_between(r->t_submit, r->t_done);
        r->done = true;
        if (rc) {
            S->pool[pool_key(h->N, h->model)].push_back(h);
        } else {
            r->hold = h;
            S->held.push_back(r);
            S->lat.push_back(r->res.latency_ms);
            S->last_done = r->t_done;
        }
    }
    S->cv_done.notify_all();
}

// Hand a finished request back: its slot is free at once, and its field
// (complete: every step is synchronized before the stopping rule runs)
// stays in HBM until mgs_server_wait copies it out.
int finish_request(mgs_server* S, Slot& sl, int status)
{
    Request* r = sl.r;
    Hier* h = sl.h;
    sl = Slot();
    S->desc_dirty = true;
    complete_request(S, r, h, status);
    return status;
}

// Admission needs memory that unclaimed results hold: copy the oldest one's
// field to the host and pool its hierarchy.  False when none is held.
bool spill_oldest(mgs_server* S)
{
    Request* r = nullptr;
    Hier* h = nullptr;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->held.empty()) return false;
        r = S->held.front();
        S->held.pop_front();
        h = r->hold;
        r->hold = nullptr;
        r->spilling = true;
    }
    const int N = h->N;
    const size_t n3 = (size_t)N * N * N;
    r->u.resize(n3);
    // the held hierarchy is out of every step: the admission stream
    launch_unpack_f64_to_f64(h->u, h->rho, N, 1.0, S->cs);
    const bool ok = hipMemcpyAsync(r->u.data(), h->rho, n3 * 8, hipMemcpyDeviceToHost, S->cs) == hipSuccess &&
                    hipStreamSynchronize(S->cs) == hipSuccess;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!ok) r->res.status = MGS_E_HIP;
        r->spilling = false;
        S->spilled++;
        S->pool[pool_key(h->N, h->model)].push_back(h);
    }
    S->cv_done.notify_all();
    return true;
}

// The outer residual of a request's outer iteration has been taken (sum:
// its squared-residual sum): solve_mixed's stopping rule (mode 1 for
// MGS_EXIT_REL_TOL, mode 0 for MGS_EXIT_REFERENCE), then the next outer
// iteration's inner cycles.  True when the request finished (its slot is
// free).
bool take_outer(mgs_server* S, Slot& sl, double sum)
{
    Request* r = sl.r;
    const int N = r->p.ngrid;
    const double res = std::sqrt(sum / ((double)N * N * N));
    sl.it++;
    if (!std::isfinite(res)) {
        finish_request(S, sl, MGS_E_DOMAIN);
        return true;
    }
    bool stop = false, conv = false;
    if (sl.it == 1) {
        sl.res0 = res;
        r->res.res0 = res;
        if (res == 0.0 && r->exit_rule == MGS_EXIT_REL_TOL) stop = conv = true;
    } else if (r->exit_rule == MGS_EXIT_REFERENCE) {
        mgs_solve_opts o{};
        o.max_cycles = r->max_outer;
        o.res_conv = r->res_conv;
        stop = reference_exit(o, sl.it - 1, res, sl.res1, sl.res2, &conv);
        sl.res2 = sl.res1;
        sl.res1 = res;
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
    if (stop) {
        finish_request(S, sl, 0);
        return true;
    }
    // the next outer iteration's inner cycles (solve_mixed's order)
    if (r->p.model == MGS_MODEL_FR && r->exit_rule == MGS_EXIT_REL_TOL) {
        bool reaches = false;
        sl.inner = auto_inner_count(MIXED_INNER_MAX, res, r->rel_tol * sl.res0, sl.res_prev, sl.m_prev, &reaches);
        sl.res_prev = res;
        sl.m_prev = sl.inner;
    } else {
        sl.inner = 1;
    }
    sl.ic = 0;
    return false;
}

// move waiting requests into free slots, FIFO, while their hierarchies fit;
// during_step: a step of the active requests is running (the new slots join
// the next one)
int admit(mgs_server* S, bool during_step)
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
        Hier* h = take_hier(S, r->p.ngrid, r->p.model);
        while (!h && spill_oldest(S)) h = take_hier(S, r->p.ngrid, r->p.model);
        bool claimed = false;
        if (!h) {
            std::lock_guard<std::mutex> lk(S->mu);
            claimed = S->claimed > 0;
        }
        if (!h) {
            if (active == 0 && !claimed) {
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
        sl.pending = during_step;
        sl.fresh = true;
        sl.inner = 1;
        sl.ic = 0;
        sl.res_prev = 0.0;
        sl.m_prev = 0;
        sl.it = 0;
        sl.res0 = 0.0;
        sl.res1 = 1.0e5;
        sl.res2 = 1.0e10;
        const int rc = start_request(S, sl, free_slot);
        if (rc) {
            (void)hipStreamSynchronize(S->cs);
            sl.starting = false;
            if (sl.upin) give_pin(S, sl.upin, r->delta_bytes);
            sl.upin = nullptr;
            finish_request(S, sl, rc);
            continue;
        }
        S->desc_dirty = true;
        admitted++;
    }
    return admitted;
}

// Wait for the admission stream and apply each new request's start residual
// (its first outer iteration's norm: the request's first step is then
// already an inner cycle); the host staging of the density goes back to the
// pool.
void finish_starts(mgs_server* S)
{
    bool any = false;
    for (int i = 0; i < S->o.max_active; i++) any = any || S->slots[i].starting;
    if (!any) return;
    const bool ok = hipStreamSynchronize(S->cs) == hipSuccess;
    for (int i = 0; i < S->o.max_active; i++) {
        Slot& sl = S->slots[i];
        if (!sl.starting) continue;
        sl.starting = false;
        Request* r = sl.r;
        if (sl.upin) {
            give_pin(S, sl.upin, r->delta_bytes);
            sl.upin = nullptr;
        }
        if (!r->verify) {  // a verified request's density stays for the FP64 solve in wait
            give_stage(S, r->delta, r->delta_bytes);
            r->delta = nullptr;
        }
        if (!ok) {
            finish_request(S, sl, MGS_E_HIP);
            continue;
        }
        take_outer(S, sl, S->hstart[i]);
    }
}

static unsigned blocks256(size_t cells) { return (unsigned)((cells + 255) / 256); }

// participation lists of the next step; *dfr / *ddgp: the deepest f(R) /
// nDGP hierarchy taking an inner cycle, -1 when there is none
void build_plan(mgs_server* S, int* dfr, int* ddgp)
{
    Plan& P = *S->hP;
    std::memset(&P, 0, sizeof(P));
    *dfr = *ddgp = -1;
    auto add = [](int& n, int* slot, unsigned* pre, int i, unsigned tiles) {
        slot[n] = i;
        pre[n + 1] = pre[n] + tiles;
        n++;
    };
    for (int i = 0; i < S->o.max_active; i++) {
        Slot& sl = S->slots[i];
        sl.outer_now = false;
        if (!sl.r) continue;
        const SlotDesc& sd = S->hS[i];
        const bool dgp = sd.model == MGS_MODEL_DGP;
        // the outer residual: nDGP every step, f(R) after the last inner
        // cycle of an outer iteration (and for a fresh request's start field)
        sl.outer_now = dgp || sl.fresh || sl.ic == sl.inner - 1;
        if (sl.outer_now) add(P.nres, P.res_slot, P.res_pre, i, dgp ? sd.zo.NB : sd.L[0].zb.NB);
        if (sl.fresh) continue;
        if (dgp) {
            *ddgp = std::max(*ddgp, sd.lc);
            for (int l = 0; l < sd.lc; l++) {
                const int k = P.nd[l];
                P.d_slot[l][k] = i;
                P.dh_pre[l][k + 1] = P.dh_pre[l][k] + blocks256(sd.L[l].g.half);
                P.df_pre[l][k + 1] = P.df_pre[l][k] + blocks256(2 * sd.L[l].g.half);
                P.dc_pre[l][k + 1] = P.dc_pre[l][k] + blocks256(2 * sd.L[l + 1].g.half);
                P.nd[l]++;
            }
            P.dcoarse_slot[sd.lc][P.ndcoarse[sd.lc]++] = i;
            P.upd_slot[P.nupd] = i;
            P.upd_pre[P.nupd + 1] = P.upd_pre[P.nupd] + blocks256(2 * sd.L[0].g.half);
            P.nupd++;
            continue;
        }
        *dfr = std::max(*dfr, sd.lc);
        const unsigned z0 = sd.L[0].z.NB;
        if (sl.ic > 0) add(P.nnf, P.nf_slot, P.nf_pre, i, z0);
        if (sl.ic == sl.inner - 1)
            add(P.nlast, P.last_slot, P.last_pre, i, z0);
        else
            add(P.nnl, P.nl_slot, P.nl_pre, i, z0);
        if (sl.ic == 0)
            for (int l = 0; l < sd.lc; l++) {
                const unsigned nc = (unsigned)sd.L[l + 1].g.n;
                add(P.nrb[l], P.rb_slot[l], P.rb_pre[l], i, (nc * nc * nc + 255u) / 256u);
            }
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

// the inner V-cycle of every non-fresh f(R) request, recursing to `depth`
void vcycle_batched(mgs_server* S, int l, int depth)
{
    hipStream_t s = S->s;
    const dim3 G(GRID), B(256);
    hipLaunchKernelGGL(kb_coarsest, dim3(MAXS), dim3(1024), 0, s, S->dS, S->dP, l);
    if (l >= depth) return;
    // pre-smoothing (level 0: in the first inner cycle of an outer iteration
    // the first colour-0 pass came with the outer residual)
    const int npre = 2 * (l == 0 ? MIXED_PRE0 : 2);
    for (int it = 0; it < npre; it++) {
        if (l == 0 && it == 0)
            hipLaunchKernelGGL(kb_smooth0, G, B, 0, s, S->dS, S->dP, (int)LIST_NOT_FIRST, 0);
        else
            hipLaunchKernelGGL(kb_smooth, G, B, 0, s, S->dS, S->dP, l, it & 1);
    }
    hipLaunchKernelGGL(kb_restrict_residual, G, B, 0, s, S->dS, S->dP, l);
    hipLaunchKernelGGL(kb_rhs, G, B, 0, s, S->dS, S->dP, l);
    vcycle_batched(S, l + 1, depth);
    hipLaunchKernelGGL(kb_correct, G, B, 0, s, S->dS, S->dP, l);
    // post-smoothing (level 0: in the last inner cycle of an outer iteration
    // the last colour-1 pass is fused with the update)
    const int npost = l == 0 ? 2 * MIXED_POST0 - 1 : 4;
    for (int it = 0; it < npost; it++) hipLaunchKernelGGL(kb_smooth, G, B, 0, s, S->dS, S->dP, l, it & 1);
    if (l == 0) {
        hipLaunchKernelGGL(kb_smooth0, G, B, 0, s, S->dS, S->dP, (int)LIST_NOT_LAST, 1);
        hipLaunchKernelGGL(kb_smooth_update, G, B, 0, s, S->dS, S->dP);
    }
}

// the nDGP inner V-cycle of every non-fresh nDGP request, recursing to
// `depth` (vcycle_inner_dgp: 2 x 2 colour passes before and after)
void vcycle_dgp(mgs_server* S, int l, int depth)
{
    hipStream_t s = S->s;
    const dim3 G(GRID), B(256);
    hipLaunchKernelGGL(kbd_coarsest, dim3(MAXS), dim3(1024), 0, s, S->dS, S->dP, l);
    if (l >= depth) return;
    for (int it = 0; it < 4; it++) {
        hipLaunchKernelGGL(kbd_smooth, G, B, 0, s, S->dS, S->dP, l, it & 1);
        hipLaunchKernelGGL(kbd_copy, G, B, 0, s, S->dS, S->dP, l, it & 1);
    }
    hipLaunchKernelGGL(kbd_residual, G, B, 0, s, S->dS, S->dP, l);
    hipLaunchKernelGGL(kbd_restrict, G, B, 0, s, S->dS, S->dP, l);
    hipLaunchKernelGGL(kbd_rhs, G, B, 0, s, S->dS, S->dP, l);
    vcycle_dgp(S, l + 1, depth);
    hipLaunchKernelGGL(kbd_correct, G, B, 0, s, S->dS, S->dP, l);
    for (int it = 0; it < 4; it++) {
        hipLaunchKernelGGL(kbd_smooth, G, B, 0, s, S->dS, S->dP, l, it & 1);
        hipLaunchKernelGGL(kbd_copy, G, B, 0, s, S->dS, S->dP, l, it & 1);
    }
}

// one step: inner cycle + update of the non-fresh requests, then the outer
// residual and its norm for all; dfr / ddgp < 0: no f(R) / nDGP inner cycle
void enqueue_step(mgs_server* S, int dfr, int ddgp)
{
    hipStream_t s = S->s;
    if (dfr >= 0) {
        for (int l = 0; l < dfr; l++)
            hipLaunchKernelGGL(kb_restrict_base, dim3(GRID), dim3(256), 0, s, S->dS, S->dP, l);
        vcycle_batched(S, 0, dfr);
    }
    if (ddgp >= 0) {
        for (int l = 0; l < ddgp; l++)
            hipLaunchKernelGGL(kbd_restrict_base, dim3(GRID), dim3(256), 0, s, S->dS, S->dP, l);
        vcycle_dgp(S, 0, ddgp);
        hipLaunchKernelGGL(kbd_update, dim3(GRID), dim3(256), 0, s, S->dS, S->dP);
    }
    hipLaunchKernelGGL(kb_outer_residual, dim3(GRID), dim3(256), 0, s, S->dS, S->dP);
    hipLaunchKernelGGL(kb_sums, dim3(MAXS), dim3(1024), 0, s, S->dS, S->dP, S->dsum);
}

int run_step(mgs_server* S, int dfr, int ddgp)
{
    if (!S->o.use_graphs) {
        enqueue_step(S, dfr, ddgp);
        return hipGetLastError() == hipSuccess ? 0 : MGS_E_HIP;
    }
    const int key = (dfr + 1) * 64 + (ddgp + 1);
    auto it = S->graphs.find(key);
    if (it == S->graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        if (hipStreamBeginCapture(S->s, hipStreamCaptureModeThreadLocal) != hipSuccess) return MGS_E_HIP;
        enqueue_step(S, dfr, ddgp);
        if (hipStreamEndCapture(S->s, &g) != hipSuccess) return MGS_E_HIP;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return MGS_E_HIP;
        it = S->graphs.emplace(key, ex).first;
    }
    return hipGraphLaunch(it->second, S->s) == hipSuccess ? 0 : MGS_E_HIP;
}

// One iteration: launch a step of the active requests, admit waiting
// requests while it runs (their uploads go on the admission stream, beside
// the step; they join the next step), wait for the step's norms, apply the
// stopping rule.  With no request active, admission runs alone first.
void worker_main(mgs_server* S)
{
    (void)hipSetDevice(S->device);
    auto count_active = [S] {
        int n = 0;
        for (int i = 0; i < S->o.max_active; i++) n += S->slots[i].r ? 1 : 0;
        return n;
    };
    auto timed_admit = [S](bool during_step) {
        const Clock::time_point ta = Clock::now();
        admit(S, during_step);
        const double dt = ms_between(ta, Clock::now());
        std::lock_guard<std::mutex> lk(S->mu);  // statistics are read under mu
        S->admit_ms += dt;
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] {
                bool act = false;
                for (int i = 0; i < S->o.max_active; i++) act = act || S->slots[i].r;
                return S->stop || act || !S->waiting.empty();
            });
            if (S->stop) break;
        }
        if (!count_active()) {
            timed_admit(false);
            finish_starts(S);
        }
        const int nact = count_active();
        if (!nact) {
            // the head of the queue waits for a hierarchy a waiter is copying
            // out (or was failed by admit()): sleep until one is returned
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            continue;
        }
        // descriptors and plan for this batch
        for (int i = 0; i < S->o.max_active; i++)
            if (S->slots[i].r && S->desc_dirty) fill_desc(S->hS[i], *S->slots[i].h, S->slots[i].r->p);
        const Clock::time_point ts = Clock::now();
        int dfr = -1, ddgp = -1;
        build_plan(S, &dfr, &ddgp);
        int rc = 0;
        if (hipMemcpyAsync(S->dS, S->hS, sizeof(SlotDesc) * MAXS, hipMemcpyHostToDevice, S->s) != hipSuccess ||
            hipMemcpyAsync(S->dP, S->hP, sizeof(Plan), hipMemcpyHostToDevice, S->s) != hipSuccess)
            rc = MGS_E_HIP;
        S->desc_dirty = false;
        if (!rc) rc = run_step(S, dfr, ddgp);
        if (!rc && hipMemcpyAsync(S->hsum, S->dsum, sizeof(double) * MAXS, hipMemcpyDeviceToHost, S->s) != hipSuccess)
            rc = MGS_E_HIP;
        // hS and hP are not touched again before the synchronization below
        // (admission only marks descriptors dirty)
        if (!rc) timed_admit(true);
        if (hipStreamSynchronize(S->s) != hipSuccess) rc = rc ? rc : MGS_E_HIP;
        const Clock::time_point tf = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->steps++;
            S->max_batch = std::max(S->max_batch, nact);
            S->step_ms += ms_between(ts, tf);
        }
        // per request: the stopping rule of solve_mixed (mode 1 or 0)
        for (int i = 0; i < S->o.max_active; i++) {
            Slot& sl = S->slots[i];
            if (!sl.r) continue;
            if (sl.pending) {  // admitted during this step: no norm of its own yet
                sl.pending = false;
                continue;
            }
            if (rc) {
                finish_request(S, sl, rc);
                continue;
            }
            if (!sl.outer_now) {  // an inner cycle inside its outer iteration: no norm
                sl.ic++;
                continue;
            }
            take_outer(S, sl, S->hsum[i]);
        }
        // the requests admitted during the step
        finish_starts(S);
        {
            const double dt = ms_between(tf, Clock::now());
            std::lock_guard<std::mutex> lk(S->mu);
            S->finish_ms += dt;
        }
    }
}

// A waiter's copy-out stream: an idle one of the server's, or a new
// non-blocking one; returned to the server after the copy (a stream per
// claim would cost a creation and a destruction each).
hipStream_t take_wstream(mgs_server* S)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->wstreams.empty()) {
            hipStream_t ws = S->wstreams.back();
            S->wstreams.pop_back();
            return ws;
        }
    }
    hipStream_t ws = nullptr;
    if (hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) != hipSuccess) return nullptr;
    return ws;
}

void give_wstream(mgs_server* S, hipStream_t ws)
{
    std::lock_guard<std::mutex> lk(S->mu);
    S->wstreams.push_back(ws);
}

// a result held in HBM into the caller's host array: unpacked into the
// reference order in the requested type and scale (into the hierarchy's
// level-0 scratch), copied to pinned staging and from there into the array,
// on a copy-out stream of the server's
int copy_out(mgs_server* S, Hier* h, void* u, int32_t dtype, double scale)
{
    if (hipSetDevice(S->device) != hipSuccess) return MGS_E_HIP;
    hipStream_t ws = take_wstream(S);
    if (!ws) return MGS_E_HIP;
    const int N = h->N;
    const size_t n3 = (size_t)N * N * N;
    const size_t bytes = n3 * (dtype == MGS_F32 ? sizeof(float) : sizeof(double));
    const void* dev;
    if (dtype == MGS_F32) {
        launch_unpack_f64_to_f32(h->u, h->d[0], N, scale, ws);
        dev = h->d[0];
    } else {
        launch_unpack_f64_to_f64(h->u, h->rho, N, scale, ws);
        dev = h->rho;
    }
    void* pin = bytes >= PIN_MIN ? take_pin(S, bytes) : nullptr;
    const hipError_t e = hipMemcpyAsync(pin ? pin : u, dev, bytes, hipMemcpyDeviceToHost, ws);
    const hipError_t es = hipStreamSynchronize(ws);
    give_wstream(S, ws);
    const bool ok = e == hipSuccess && es == hipSuccess;
    if (pin) {
        if (ok) copy_host(u, pin, bytes);
        give_pin(S, pin, bytes);
    }
    return ok ? 0 : MGS_E_HIP;
}

}  // namespace

// The served field's difference from the plain FP64 method on the same
// request: a handle of its own (mgs_create, FP64), the request's density and
// exit rule, mgs_solve from the background, the raw fields compared on the
// host.  Runs in the waiting client's thread, beside the server's stream.
static int verify_fp64(const Request* r, const std::vector<double>& served, mgs_request_result* res)
{
    mgs_params p = r->p;
    p.precision = MGS_PREC_FP64;
    mgs_handle* h = nullptr;
    int rc = mgs_create(&p, &h);
    if (rc) return rc;
    const size_t n3 = served.size();
    std::vector<double> u64(n3);
    mgs_solve_opts o{};
    o.mode = r->exit_rule == MGS_EXIT_REL_TOL ? 1 : 0;
    o.max_cycles = r->exit_rule == MGS_EXIT_REL_TOL ? std::max(100, r->max_outer) : r->max_outer;
    o.rel_tol = r->rel_tol;
    o.res_conv = r->res_conv;
    o.init_guess = 1;
    mgs_solve_report rep{};
    rc = mgs_set_density(h, r->delta, r->dtype, MGS_HOST);
    if (!rc) rc = mgs_solve(h, &o, &rep);
    if (!rc) rc = mgs_get_field(h, u64.data(), MGS_F64, MGS_HOST, 1.0);
    mgs_destroy(h);
    if (rc) return rc;
    double dmax = 0.0, umax = 0.0, d2 = 0.0, u2 = 0.0;
    for (size_t q = 0; q < n3; q++) {
        const double d = served[q] - u64[q];
        dmax = std::max(dmax, std::fabs(d));
        umax = std::max(umax, std::fabs(u64[q]));
        d2 += d * d;
        u2 += u64[q] * u64[q];
    }
    res->verified = 1;
    res->fp64_rel_linf = umax > 0.0 ? dmax / umax : dmax;
    res->fp64_rel_l2 = u2 > 0.0 ? std::sqrt(d2 / u2) : std::sqrt(d2);
    res->fp64_res_final = rep.res_final;
    return 0;
}

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
    // a non-blocking stream: no null-stream call of any other thread (a
    // client's, or a waiter's FP64 comparison solve) synchronizes with the
    // worker's stream, which is being captured into graphs between steps
    if (hipStreamCreateWithFlags(&S->s, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&S->cs, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&S->dS, sizeof(SlotDesc) * MAXS) != hipSuccess ||
        hipMalloc((void**)&S->dP, sizeof(Plan)) != hipSuccess ||
        hipMalloc((void**)&S->dsum, sizeof(double) * MAXS) != hipSuccess ||
        hipHostMalloc((void**)&S->hS, sizeof(SlotDesc) * MAXS, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&S->hP, sizeof(Plan), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&S->hsum, sizeof(double) * MAXS, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&S->hstart, sizeof(double) * MAXS, hipHostMallocDefault) != hipSuccess) {
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
    {
        // blocked waiters wake, see stop and leave with MGS_E_STATE; a waiter
        // copying a result out finishes first
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.notify_all();
        S->cv_done.wait(lk, [&] { return S->waiters == 0; });
    }
    if (S->s) (void)hipStreamSynchronize(S->s);
    if (S->cs) (void)hipStreamSynchronize(S->cs);
    S->held.clear();  // their hierarchies are freed with the requests below
    for (auto& kv : S->graphs) (void)hipGraphExecDestroy(kv.second);
    for (int i = 0; i < MAXS; i++) {
        if (S->slots[i].h) {
            (void)hipFree(S->slots[i].h->base);
            delete S->slots[i].h;
        }
        if (S->slots[i].upin) (void)hipHostFree(S->slots[i].upin);
    }
    for (auto& kv : S->pool)
        for (Hier* h : kv.second) {
            (void)hipFree(h->base);
            delete h;
        }
    for (auto& kv : S->reqs) {
        if (kv.second->hold) {
            (void)hipFree(kv.second->hold->base);
            delete kv.second->hold;
        }
        delete kv.second;
    }
    for (auto& kv : S->stages)
        for (char* p : kv.second) std::free(p);
    for (auto& kv : S->pins)
        for (void* p : kv.second) (void)hipHostFree(p);
    if (S->dS) (void)hipFree(S->dS);
    if (S->dP) (void)hipFree(S->dP);
    if (S->dsum) (void)hipFree(S->dsum);
    if (S->hS) (void)hipHostFree(S->hS);
    if (S->hP) (void)hipHostFree(S->hP);
    if (S->hsum) (void)hipHostFree(S->hsum);
    if (S->hstart) (void)hipHostFree(S->hstart);
    for (hipStream_t ws : S->wstreams) (void)hipStreamDestroy(ws);
    if (S->s) (void)hipStreamDestroy(S->s);
    if (S->cs) (void)hipStreamDestroy(S->cs);
    delete S;
    return 0;
}

const char* mgs_server_last_error(const mgs_server* S)
{
    (void)S;
    return g_server_error.c_str();
}

int mgs_server_submit(mgs_server* S, const mgs_request* req, int64_t* ticket)
{
    if (!S || !req || !ticket || !req->delta) return srv_err(S, MGS_E_ARG, "NULL argument");
    std::string why;
    int rc = validate_params(&req->params, &why);
    if (rc) return srv_err(S, rc, why);
    const mgs_params& p = req->params;
    if ((p.model != MGS_MODEL_FR && p.model != MGS_MODEL_DGP) || p.precision != MGS_PREC_MIXED)
        return srv_err(S, MGS_E_ARG, "the server solves f(R) and nDGP requests in MGS_PREC_MIXED");
    if (p.ngrid < 16 || p.ngrid > 1024) return srv_err(S, MGS_E_ARG, "the server takes 16 <= ngrid <= 1024");
    if (p.model == MGS_MODEL_DGP && (p.ngrid & (p.ngrid - 1)))
        return srv_err(S, MGS_E_ARG, "the server takes nDGP requests with a power-of-two ngrid");
    if (coarsest_of(p.ngrid) + 1 > MAXL) return srv_err(S, MGS_E_ARG, "too many levels");
    if (req->dtype != MGS_F32 && req->dtype != MGS_F64) return srv_err(S, MGS_E_ARG, "dtype");
    if (req->max_outer < 1) return srv_err(S, MGS_E_ARG, "max_outer >= 1");
    if (req->exit_rule == MGS_EXIT_REL_TOL) {
        if (!(req->rel_tol > 0.0)) return srv_err(S, MGS_E_ARG, "rel_tol > 0");
    } else if (req->exit_rule == MGS_EXIT_REFERENCE) {
        if (!(req->res_conv >= 0.0) || !std::isfinite(req->res_conv))
            return srv_err(S, MGS_E_ARG, "res_conv must be finite and >= 0");
    } else {
        return srv_err(S, MGS_E_ARG, "exit_rule must be MGS_EXIT_REL_TOL or MGS_EXIT_REFERENCE");
    }
    if (footprint(p.ngrid, p.model) > S->budget)
        return srv_err(S, MGS_E_NOMEM, "request larger than the server's HBM budget");
    const size_t n3 = (size_t)p.ngrid * p.ngrid * p.ngrid;
    const size_t es = req->dtype == MGS_F32 ? 4 : 8;
    Request* r = new Request();
    r->p = p;
    r->dtype = req->dtype;
    r->max_outer = req->max_outer;
    r->rel_tol = req->rel_tol;
    r->exit_rule = req->exit_rule;
    r->res_conv = req->res_conv;
    r->verify = req->verify_fp64 != 0;
    r->delta_bytes = n3 * es;
    r->delta = take_stage(S, r->delta_bytes);
    if (!r->delta) {
        delete r;
        return srv_err(S, MGS_E_NOMEM, "host staging buffer");
    }
    if (!copy_finite(r->delta, req->delta, req->dtype, n3)) {
        give_stage(S, r->delta, r->delta_bytes);
        r->delta = nullptr;
        delete r;
        for (size_t q = 0; q < n3; q++) {
            const double v =
                req->dtype == MGS_F32 ? (double)((const float*)req->delta)[q] : ((const double*)req->delta)[q];
            if (!std::isfinite(v))
                return srv_err(S, MGS_E_DOMAIN, "density contrast must be finite (cell " + std::to_string(q) + ")");
        }
        return srv_err(S, MGS_E_DOMAIN, "density contrast must be finite");
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if ((int)S->waiting.size() >= S->o.queue_capacity) {
            S->rejected++;
            // back to the idle staging (mu is held: the pool is appended to here)
            if (S->stage_idle + r->delta_bytes <= STAGE_IDLE_MAX) {
                S->stages[r->delta_bytes].push_back(r->delta);
                S->stage_idle += r->delta_bytes;
                r->delta = nullptr;
            }
            delete r;
            g_server_error = "queue full";
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

int mgs_server_test(mgs_server* S, int64_t ticket, int32_t* done)
{
    if (!S || !done) return MGS_E_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return srv_err(S, MGS_E_ARG, "unknown ticket");
    *done = it->second->done ? 1 : 0;
    return 0;
}

int mgs_server_wait(mgs_server* S, int64_t ticket, void* u, size_t u_count, int32_t dtype, double scale,
                    mgs_request_result* res)
{
    if (!S) return MGS_E_ARG;
    if (u && dtype != MGS_F32 && dtype != MGS_F64) return srv_err(S, MGS_E_ARG, "dtype");
    Request* r = nullptr;
    Hier* h = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        if (S->stop) return srv_err(S, MGS_E_STATE, "the server is being destroyed");
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end()) return srv_err(S, MGS_E_ARG, "unknown ticket");
        r = it->second;
        if (r->waited) return srv_err(S, MGS_E_STATE, "ticket already waited for");
        const size_t n3 = (size_t)r->p.ngrid * r->p.ngrid * r->p.ngrid;
        if (u && u_count < n3)
            return srv_err(S, MGS_E_ARG, "output array of " + std::to_string(u_count) + " elements for a field of " +
                                             std::to_string(n3));
        r->waited = true;
        S->waiters++;
        S->cv_done.wait(lk, [&] { return (r->done && !r->spilling) || S->stop; });
        if (!(r->done && !r->spilling)) {
            // destroy began first: the request (and its memory) is destroy's
            S->waiters--;
            S->cv_done.notify_all();
            return srv_err(S, MGS_E_STATE, "the server was destroyed before the request finished");
        }
        S->reqs.erase(it);
        if (r->hold) {  // the result is still in HBM: claim it
            h = r->hold;
            r->hold = nullptr;
            S->held.erase(std::find(S->held.begin(), S->held.end(), r));
            S->claimed++;
        }
    }
    int rc = r->res.status;
    const size_t n3 = (size_t)r->p.ngrid * r->p.ngrid * r->p.ngrid;
    // a verified request's raw field in FP64, from HBM or its spill
    std::vector<double> raw;
    if (h) {
        if (!rc && r->verify) {
            raw.resize(n3);
            rc = copy_out(S, h, raw.data(), MGS_F64, 1.0);
        } else if (!rc && u) {
            rc = copy_out(S, h, u, dtype, scale);
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->pool[pool_key(h->N, h->model)].push_back(h);
            S->claimed--;
        }
        S->cv_work.notify_all();  // its memory is free for admission
    } else if (!rc) {
        raw.swap(r->u);
    }
    if (!rc && u && !raw.empty()) {
        for (size_t q = 0; q < n3; q++) {
            const double v = scale == 1.0 ? raw[q] : raw[q] * scale;
            if (dtype == MGS_F32)
                ((float*)u)[q] = (float)v;
            else
                ((double*)u)[q] = v;
        }
    }
    if (res) *res = r->res;
    if (!rc && r->verify && res) rc = verify_fp64(r, raw, res);
    if (res) res->status = rc;
    if (rc) g_server_error = "request failed (code " + std::to_string(rc) + ")";
    if (r->delta) give_stage(S, r->delta, r->delta_bytes);
    r->delta = nullptr;
    delete r;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->waiters--;
    }
    S->cv_done.notify_all();
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
    st->step_ms = S->step_ms;
    st->admit_ms = S->admit_ms;
    st->finish_ms = S->finish_ms;
    st->spilled = S->spilled;
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
