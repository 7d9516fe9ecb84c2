// mgs_fr.hip -- f(R) multigrid kernels for gfx950 (CDNA4).
//
// Plain FP64 path: the reference's operators (PMP2MGsolver_fR.f90,
// PMP2extradof.f90:1168-1578) evaluated in FP64 with the reference's operand
// order, on the colour-split layout of mgs_common.h.  Built with
// -ffp-contract=off, so the residual, the restrictions, the right-hand side
// and the prolongation round every operation as the CPU oracle does and are
// bit-identical to it.  The smoother solves the same local polynomial
// u^{n+2} + P u + Q = 0; by default its root comes from a cancellation-free
// FP32 seed refined by FP64 Newton steps (MGS_FLAG_REFERENCE_ROOTS selects
// the reference's closed forms evaluated literally with the device libm).
//
// Mixed path: the FP32 delta-form FAS used inside the FP64 defect-correction
// loop (mgs_api.cpp, solve_mixed).  Unknown d = x^{n+1} - xb^{n+1} with
// x = u / u_bg and xb an FP32 copy of the current FP64 iterate.  The inner
// operator
//     D(d) = A (sum_nb d_nb - 6 d) + C (1/xb - 1/x)
// is N(xb + .) - N(xb) written so that every term is proportional to d: FP32
// rounding then scales with the correction and the outer FP64 residual
// converges to FP64 accuracy.  Each cell is solved from the FP32 closed-form
// seed followed by a Newton step on D(d) = rhs in d.
//
// Stencil kernels march over z (mgs_stencil.h); transfer kernels map one
// thread to one coarse cell; all are memory-bound, with 32-bit index math.

#include <math.h>

#include "mgs_fr_dev.h"
#include "mgs_kernels.h"

namespace mgs {

// restrict_residual_fR(1): the 8 children's residuals averaged, recomputed
// from u and delta, PMP2MGsolver_fR.f90:607-749.  A thread owns coarse column
// (I, J) and marches over K; fine plane 2K+2 (the upper z faces of K) is the
// lower child plane of K+1 and plane 2K-1 is carried from K-1, so u and delta
// are read once.  The operation order is the reference's.
//
// With ru1 non-null the same pass also restricts u (restrict_scalar_field,
// the 8-child average in restrict8_cell's order, so bit-identical to it) into
// ru1 and the coarse start value u1.
template <int FRN>
__global__ __launch_bounds__(256) void k_fr_restrict_residual1_64(const double* __restrict__ u0,
                                                                  const double* __restrict__ delta,
                                                                  double* __restrict__ rres1,
                                                                  double* __restrict__ ru1, double* __restrict__ u1,
                                                                  Geo gf, Geo gc, FrCoef c, int TX, int TY, int XB,
                                                                  int JB, int KC, unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    const int I = cc.m, J = cc.j, N = gf.n;
    Pts12 F;
    F.init(gf, I, J);
    // per fine plane: u^{n+1}, fct1 delta and fct4/u of the 4 children, and
    // u^{n+1} of the 8 face neighbours; us: running sum of the children's u
    // (x fastest), continued from `base` for the upper plane of a coarse cell
    struct Pl {
        double w[4], t1[4], t2[4], f[8], us;
    };
    auto load = [&](int z, Pl& p, bool cont, double base) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const unsigned o = F.at(q, z);
            const double u = u0[o];
            p.w[q] = wpow<FRN>(u);
            p.t1[q] = c.fct1r * delta[o];
            p.t2[q] = c.fct4_1 / u;
            if (q == 0)
                p.us = cont ? base + u : u;
            else
                p.us = p.us + u;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) p.f[q] = wpow<FRN>(u0[F.at(4 + q, z)]);
    };
    Pl p0, p1, p2;
    double wl[4];
    const int zs = 2 * cc.k0;
    load(zs, p0, false, 0.0);
    {
        const int zl = zs == 0 ? N - 1 : zs - 1;
#pragma unroll
        for (int q = 0; q < 4; q++) wl[q] = wpow<FRN>(u0[F.at(q, zl)]);
    }
    for (int K = cc.k0; K < cc.k1; K++) {
        load(2 * K + 1, p1, true, p0.us);
        load(2 * K + 2 == N ? 0 : 2 * K + 2, p2, false, 0.0);
        double Q = -8.0 * c.fct2r;
#pragma unroll
        for (int q = 0; q < 4; q++) Q = Q + p0.t1[q] + p0.t2[q];
#pragma unroll
        for (int q = 0; q < 4; q++) Q = Q + p1.t1[q] + p1.t2[q];
        double s8 = p0.w[0];
        s8 = s8 + p0.w[1];
        s8 = s8 + p0.w[2];
        s8 = s8 + p0.w[3];
#pragma unroll
        for (int q = 0; q < 4; q++) s8 = s8 + p1.w[q];
        double P = -3.0 * s8;
        P = P + p0.f[0] + p0.f[1] + p1.f[0] + p1.f[1];  // x-1 faces
        P = P + p0.f[2] + p0.f[3] + p1.f[2] + p1.f[3];  // x+2 faces
        P = P + p0.f[4] + p0.f[5] + p1.f[4] + p1.f[5];  // y-1 faces
        P = P + p0.f[6] + p0.f[7] + p1.f[6] + p1.f[7];  // y+2 faces
        P = P + wl[0] + wl[1] + wl[2] + wl[3];          // z-1 faces
        P = P + p2.w[0] + p2.w[1] + p2.w[2] + p2.w[3];  // z+2 faces
        const unsigned oc = sidx(gc, I, J, K);
        rres1[oc] = (P * c.fct3_1 + Q) / 8.0;
        if (ru1) {
            const double v = p1.us / 8.0;
            ru1[oc] = v;
            u1[oc] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) wl[q] = p1.w[q];
        p0 = p2;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_restrict8(const T* __restrict__ fine, T* __restrict__ out1,
                                                   T* __restrict__ out2, Geo gf, Geo gc, int TX, int TY, int XB,
                                                   unsigned NB)
{
    int I;
    unsigned row;
    if (!map_xy(gc.n, (long)gc.n * gc.n, TX, TY, XB, NB, I, row)) return;
    restrict8_cell<T>(fine, out1, out2, gf, gc, I, (int)(row % (unsigned)gc.n), (int)(row / (unsigned)gc.n));
}

// correct_solution (PMP2extradof.f90:1361-1578), one coarse column per thread;
// ONLY1: colour-1 sons only (son_corrected); GUARD: keep a son's old value
// unless the new one is > 0 (f(R)).  Fine plane 2K takes coarse
// planes K-1 (beside) and K (centre), plane 2K+1 planes K+1 and K, so fine
// planes 2K+1 and 2K+2 both take coarse planes K and K+1: the column marches
// in that pair order with two coarse planes of uH and R u in registers (a
// third, prefetched, would not fit four waves per SIMD).  Each son is
// written once, from values nothing in the launch writes, so the order of
// the updates does not change the results.
template <bool ONLY1, bool GUARD, bool NT = false>
__global__ __launch_bounds__(256) void k_correct64(const double* __restrict__ uH, const double* __restrict__ ru,
                                                   double* __restrict__ uf, Geo gf, Geo gc, int TX, int TY, int XB,
                                                   int JB, int KC, unsigned NB)
{
    // coarse arrays: elem() offsets (a coarse level holds at most 2^27 cells)
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    const int I = cc.m, J = cc.j;
    Block9 B;
    B.init(gc, I, J);
    double ha[9], ra[9], hb[9], rb[9];  // coarse planes K and K+1
    {
        const int Km = wrap_dec(cc.k0, gc.n);
        double o[4];
        correct64_sons_load<0, ONLY1, NT>(uf, gf, I, J, cc.k0, o);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            ha[q] = elem(uH, B.at(q, Km));
            ra[q] = elem(ru, B.at(q, Km));
            hb[q] = elem(uH, B.at(q, cc.k0));
            rb[q] = elem(ru, B.at(q, cc.k0));
        }
        correct64_sons_apply<0, ONLY1, GUARD, NT>(ha, ra, hb, rb, uf, gf, I, J, cc.k0, o);  // fine plane 2 k0
    }
    for (int K = cc.k0; K < cc.k1; K++) {
        const bool more = K + 1 < cc.k1;
        double o1[4], o0[4];
        correct64_sons_load<1, ONLY1, NT>(uf, gf, I, J, K, o1);
        if (more) correct64_sons_load<0, ONLY1, NT>(uf, gf, I, J, K + 1, o0);
        const int Kp = wrap_inc(K, gc.n);
#pragma unroll
        for (int q = 0; q < 9; q++) {
            ha[q] = hb[q];
            ra[q] = rb[q];
            hb[q] = elem(uH, B.at(q, Kp));
            rb[q] = elem(ru, B.at(q, Kp));
        }
        correct64_sons_apply<1, ONLY1, GUARD, NT>(hb, rb, ha, ra, uf, gf, I, J, K, o1);  // fine plane 2K+1
        if (more) correct64_sons_apply<0, ONLY1, GUARD, NT>(ha, ra, hb, rb, uf, gf, I, J, K + 1, o0);  // 2K+2
    }
}

// deterministic sum of block partials: stage 1 folds contiguous chunks, stage
// 2 (one block) the chunk sums; the order depends only on the count
__global__ __launch_bounds__(1024) void k_sum_chunks(const double* __restrict__ p, int count, int chunk,
                                                     double* __restrict__ out)
{
    __shared__ double s[1024];
    const int lo = blockIdx.x * chunk;
    const int hi = lo + chunk < count ? lo + chunk : count;
    block_chunk_sum(p, lo, hi, s);
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void k_pack(const TI* __restrict__ lex, TO* __restrict__ split, Geo g, int TX,
                                              int TY, int XB, unsigned NB)
{
    int i;
    unsigned row;
    if (!map_xy(g.n, (long)g.n * g.n, TX, TY, XB, NB, i, row)) return;
    const int j = (int)(row % (unsigned)g.n), k = (int)(row / (unsigned)g.n);
    split[sidx(g, i, j, k)] = (TO)lex[(size_t)row * g.n + i];
}

template <typename TO>
__global__ __launch_bounds__(256) void k_unpack(const double* __restrict__ split, TO* __restrict__ lex, Geo g,
                                                double scale, int TX, int TY, int XB, unsigned NB)
{
    int i;
    unsigned row;
    if (!map_xy(g.n, (long)g.n * g.n, TX, TY, XB, NB, i, row)) return;
    const int j = (int)(row % (unsigned)g.n), k = (int)(row / (unsigned)g.n);
    const double v = split[sidx(g, i, j, k)];
    lex[(size_t)row * g.n + i] = (TO)(scale == 1.0 ? v : v * scale);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ p, size_t count, T v)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy(const T* __restrict__ s, T* __restrict__ d, size_t count)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) d[i] = s[i];
}

// ===========================================================================
// Mixed precision: FP32 delta-form FAS
// ===========================================================================

template <int FRN>
__global__ __launch_bounds__(256) void k_df_restrict_residual(const float* __restrict__ d0,
                                                              const float* __restrict__ xb0,
                                                              const float* __restrict__ g0, float* __restrict__ rres1,
                                                              float* __restrict__ rd1, float* __restrict__ d1, Geo gf,
                                                              Geo gc, DfCoef c, int TX, int TY, int XB, int JB, int KC,
                                                              unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    // lanes t-1 / t+1 hold columns I-1 / I+1 of the same row and plane run
    // when they are in the same block row, in this wave and inside the level
    const unsigned t = threadIdx.y * (unsigned)TX + threadIdx.x;
    const bool lok = threadIdx.x > 0 && (t & 63u) != 0;
    const bool rok = (int)threadIdx.x + 1 < TX && cc.m + 1 < gc.n && ((t + 1) & 63u) != 0;
    df_restrict_residual_col<FRN, true>(d0, xb0, g0, rres1, rd1, d1, gf, gc, c, cc, lok, rok);
}

template <int FRN>
__global__ __launch_bounds__(256) void k_restrict_base(const float* __restrict__ xf, float* __restrict__ xc, Geo gf,
                                                       Geo gc, int TX, int TY, int XB, unsigned NB)
{
    int I;
    unsigned row;
    if (!map_xy(gc.n, (long)gc.n * gc.n, TX, TY, XB, NB, I, row)) return;
    restrict_base_cell<FRN>(xf, xc, gf, gc, I, (int)(row % (unsigned)gc.n), (int)(row / (unsigned)gc.n));
}

template <int FRN, bool NT = false>
__global__ __launch_bounds__(256) void k_df_correct(const float* __restrict__ dH, const float* __restrict__ rd,
                                                    float* __restrict__ df, const float* __restrict__ xbf, Geo gf,
                                                    Geo gc, int TX, int TY, int XB, int JB, int KC, unsigned NB)
{
    ZCol cc;
    if (!zcol(gc.n, gc.n, TX, TY, XB, JB, KC, NB, cc)) return;
    // lanes t-1 / t+1 hold columns I-1 / I+1 of the same row and plane run
    // when they are in the same block row, in this wave and inside the level
    const unsigned t = threadIdx.y * (unsigned)TX + threadIdx.x;
    const bool lok = threadIdx.x > 0 && (t & 63u) != 0;
    const bool rok = (int)threadIdx.x + 1 < TX && cc.m + 1 < gc.n && ((t + 1) & 63u) != 0;
    df_correct_col<FRN, true, NT>(dH, rd, df, xbf, gf, gc, cc, lok, rok);
}

template <int FRN>
__global__ __launch_bounds__(1024) void k_df_coarsest(float* __restrict__ d, const float* __restrict__ xb,
                                                      const float* __restrict__ rhs, Geo g, int sweeps, DfCoef c)
{
    df_coarsest_block<FRN>(d, xb, rhs, g, sweeps, c);
}

template <int FRN>
__global__ __launch_bounds__(1024) void k_df_tail(TailDesc T)
{
    df_tail_vcycle<FRN>(T);
}

// ===========================================================================
// launchers
// ===========================================================================
#define LAUNCH_ARGS L.TX, L.TY, L.XB, L.NB
#define ZARGS L.TX, L.TY, L.XB, L.JB, L.KC, L.NB

#define FRN_DISPATCH(frn, KERNEL, ...)                                                   \
    do {                                                                                 \
        if ((frn) == 0)                                                                  \
            hipLaunchKernelGGL((KERNEL<0>), L.grid, L.block, 0, s, __VA_ARGS__);         \
        else if ((frn) == 1)                                                             \
            hipLaunchKernelGGL((KERNEL<1>), L.grid, L.block, 0, s, __VA_ARGS__);         \
        else                                                                             \
            hipLaunchKernelGGL((KERNEL<2>), L.grid, L.block, 0, s, __VA_ARGS__);         \
    } while (0)

// the same with the kernel's one-touch accesses non-temporal where nt
#define FRN_DISPATCH_NT(frn, nt, KERNEL, ...)                                               \
    do {                                                                                    \
        if (nt) {                                                                           \
            if ((frn) == 0)                                                                 \
                hipLaunchKernelGGL((KERNEL<0, true>), L.grid, L.block, 0, s, __VA_ARGS__);  \
            else if ((frn) == 1)                                                            \
                hipLaunchKernelGGL((KERNEL<1, true>), L.grid, L.block, 0, s, __VA_ARGS__);  \
            else                                                                            \
                hipLaunchKernelGGL((KERNEL<2, true>), L.grid, L.block, 0, s, __VA_ARGS__);  \
        } else {                                                                            \
            FRN_DISPATCH(frn, KERNEL, __VA_ARGS__);                                         \
        }                                                                                   \
    } while (0)

template <class OP>
static void colour_pass(const OP& op, const Geo& g, int colour, hipStream_t s)
{
    const ZLaunch L = make_zlaunch(g);
    hipLaunchKernelGGL((k_colour_pass<OP>), L.grid, L.block, 0, s, op, g, colour, ZARGS);
}

template <class OP>
static int both_colours(const OP& op, double* partials, const Geo& g, hipStream_t s)
{
    const ZLaunch L = make_zlaunch_both(g);
    hipLaunchKernelGGL((k_both_colours<OP>), L.grid, L.block, 0, s, op, partials, g, ZARGS);
    return (int)L.NB;
}

void launch_pack_f64_from_f64(const double* lex, double* split, int n, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_pack<double, double>), L.grid, L.block, 0, s, lex, split, g, LAUNCH_ARGS);
}

void launch_pack_f64_from_f32(const float* lex, double* split, int n, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_pack<float, double>), L.grid, L.block, 0, s, lex, split, g, LAUNCH_ARGS);
}

void launch_unpack_f64_to_f64(const double* split, double* lex, int n, double scale, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_unpack<double>), L.grid, L.block, 0, s, split, lex, g, scale, LAUNCH_ARGS);
}

void launch_unpack_f64_to_f32(const double* split, float* lex, int n, double scale, hipStream_t s)
{
    const Geo g = make_geo(n);
    const Launch L = make_launch(n, (long)n * n);
    hipLaunchKernelGGL((k_unpack<float>), L.grid, L.block, 0, s, split, lex, g, scale, LAUNCH_ARGS);
}

static dim3 flat_grid(size_t count)
{
    size_t b = (count + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

void launch_fill_f64(double* p, size_t count, double v, hipStream_t s)
{
    hipLaunchKernelGGL((k_fill<double>), flat_grid(count), dim3(256), 0, s, p, count, v);
}

void launch_fill_f32(float* p, size_t count, float v, hipStream_t s)
{
    hipLaunchKernelGGL((k_fill<float>), flat_grid(count), dim3(256), 0, s, p, count, v);
}

void launch_copy_f64(const double* src, double* dst, size_t count, hipStream_t s)
{
    hipLaunchKernelGGL((k_copy<double>), flat_grid(count), dim3(256), 0, s, src, dst, count);
}

void launch_copy_f32(const float* src, float* dst, size_t count, hipStream_t s)
{
    hipLaunchKernelGGL((k_copy<float>), flat_grid(count), dim3(256), 0, s, src, dst, count);
}

void launch_fr_smooth64(int frn, int level, double* u, const double* rho, const double* f, int n, int colour,
                        const FrCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const bool nt = n >= NT_MIN_N;
#define SM(FRN, CO, RR)                                                             \
    do {                                                                            \
        if (nt)                                                                     \
            colour_pass(FrSmoothOp<FRN, CO, RR, true>{u, rho, f, c}, g, colour, s); \
        else                                                                        \
            colour_pass(FrSmoothOp<FRN, CO, RR>{u, rho, f, c}, g, colour, s);       \
    } while (0)
#define SM2(FRN, CO)            \
    do {                        \
        if (c.reference_roots)  \
            SM(FRN, CO, true);  \
        else                    \
            SM(FRN, CO, false); \
    } while (0)
    if (level == 0) {
        if (frn == 0) SM2(0, false);
        else if (frn == 1) SM2(1, false);
        else SM2(2, false);
    } else {
        if (frn == 0) SM2(0, true);
        else if (frn == 1) SM2(1, true);
        else SM2(2, true);
    }
#undef SM2
#undef SM
}

int launch_fr_residual64(int frn, int level, const double* u, const double* rho, const double* f, double* res,
                         double* partials, int n, const FrCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
    const bool nt = n >= NT_MIN_N;
#define RS(FRN, CO)                                                                      \
    (nt ? both_colours(FrResOp<FRN, CO, true>{u, rho, f, res, c}, partials, g, s) \
        : both_colours(FrResOp<FRN, CO>{u, rho, f, res, c}, partials, g, s))
    if (level == 0) {
        if (frn == 0) return RS(0, false);
        if (frn == 1) return RS(1, false);
        return RS(2, false);
    }
    if (frn == 0) return RS(0, true);
    if (frn == 1) return RS(1, true);
    return RS(2, true);
#undef RS
}

void launch_fr_restrict_residual1_64(int frn, const double* u0, const double* delta, double* rres1, double* ru1,
                                     double* u1, int N, const FrCoef& c0, hipStream_t s)
{
    const Geo gf = make_geo(N), gc = make_geo(N / 2);
    const ZLaunch L = make_zlaunch(gc.n, gc.n);
    FRN_DISPATCH(frn, k_fr_restrict_residual1_64, u0, delta, rres1, ru1, u1, gf, gc, c0, ZARGS);
}

void launch_restrict8_64(const double* fine, double* out1, double* out2, int nc, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const Launch L = make_launch(nc, (long)nc * nc);
    hipLaunchKernelGGL((k_restrict8<double>), L.grid, L.block, 0, s, fine, out1, out2, gf, gc, LAUNCH_ARGS);
}

void launch_restrict8_32(const float* fine, float* out1, float* out2, int nc, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const Launch L = make_launch(nc, (long)nc * nc);
    hipLaunchKernelGGL((k_restrict8<float>), L.grid, L.block, 0, s, fine, out1, out2, gf, gc, LAUNCH_ARGS);
}

void launch_fr_rhs64(int frn, const double* ru, const double* rho, const double* rres, double* f, int n,
                     const FrCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
#define RHS(FRN)                                                                              \
    do {                                                                                      \
        if (n >= NT_MIN_N)                                                                    \
            both_colours(FrResOp<FRN, true, true>{ru, rho, rres, f, c}, nullptr, g, s);       \
        else                                                                                  \
            both_colours(FrResOp<FRN, true>{ru, rho, rres, f, c}, nullptr, g, s);             \
    } while (0)
    if (frn == 0) RHS(0);
    else if (frn == 1) RHS(1);
    else RHS(2);
#undef RHS
}

void launch_correct64(const double* uH, const double* ru, double* ufine, int nc, int guard, bool only1,
                      hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const ZLaunch L = make_zlaunch(nc, nc);
#define C64(O, G)                                                                                          \
    do {                                                                                                   \
        if (2 * nc >= NT_MIN_N)                                                                            \
            hipLaunchKernelGGL((k_correct64<O, G, true>), L.grid, L.block, 0, s, uH, ru, ufine, gf, gc, ZARGS); \
        else                                                                                               \
            hipLaunchKernelGGL((k_correct64<O, G>), L.grid, L.block, 0, s, uH, ru, ufine, gf, gc, ZARGS);   \
    } while (0)
    if (only1 && guard)
        C64(true, true);
    else if (only1)
        C64(true, false);
    else if (guard)
        C64(false, true);
    else
        C64(false, false);
#undef C64
}

// partials[0..count) -> out[0]; stage 1 writes its chunk sums after the
// partials (the buffer holds count + 1024 doubles)
void launch_sum_partials(const double* partials, int count, double* out, hipStream_t s)
{
    if (count <= 8192) {
        hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(1024), 0, s, partials, count, count, out);
        return;
    }
    const int nchunk = 1024;
    const int chunk = (count + nchunk - 1) / nchunk;
    double* mid = const_cast<double*>(partials) + count;
    hipLaunchKernelGGL(k_sum_chunks, dim3(nchunk), dim3(1024), 0, s, partials, count, chunk, mid);
    hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(1024), 0, s, (const double*)mid, nchunk, nchunk, out);
}

int launch_outer_residual(int frn, const double* u, const double* delta, float* g32, float* xb, float* d,
                          double* partials, int n, const FrCoef& c, const DfCoef& dc, bool norm_only, hipStream_t s)
{
    const Geo g = make_geo(n);
    const unsigned half = (unsigned)g.half;
    const bool nt = n >= NT_MIN_N;
#define OUTER1(FRN, NO)                                                                                   \
    return nt ? both_colours(FrOuterOp<FRN, NO, true>{u, delta, g32, xb, d, c, dc, half}, partials, g, s) \
              : both_colours(FrOuterOp<FRN, NO>{u, delta, g32, xb, d, c, dc, half}, partials, g, s)
#define OUTER(FRN)        \
    do {                  \
        if (norm_only)    \
            OUTER1(FRN, true); \
        OUTER1(FRN, false);    \
    } while (0)
    if (frn == 0) OUTER(0);
    if (frn == 1) OUTER(1);
    OUTER(2);
#undef OUTER
#undef OUTER1
}

void launch_df_smooth_update(int frn, const float* d, const float* xb, const float* rhs, double* u, int n,
                             const DfCoef& c, double ubg, hipStream_t s)
{
    const Geo g = make_geo(n);
#define UPD(FRN)                                                                        \
    do {                                                                                \
        if (n >= NT_MIN_N)                                                              \
            colour_pass(DfSmoothUpdateOp<FRN, true>{d, xb, rhs, u, c, ubg}, g, 1, s);   \
        else                                                                            \
            colour_pass(DfSmoothUpdateOp<FRN>{d, xb, rhs, u, c, ubg}, g, 1, s);         \
    } while (0)
    if (frn == 0) UPD(0);
    else if (frn == 1) UPD(1);
    else UPD(2);
#undef UPD
}

void launch_restrict_base(int frn, const float* xb_f, float* xb_c, int nc, hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const Launch L = make_launch(nc, (long)nc * nc);
    FRN_DISPATCH(frn, k_restrict_base, xb_f, xb_c, gf, gc, LAUNCH_ARGS);
}

void launch_df_smooth(int frn, float* d, const float* xb, const float* rhs, int n, int colour, const DfCoef& c,
                      hipStream_t s)
{
    const Geo g = make_geo(n);
#define SMD(FRN)                                                             \
    do {                                                                     \
        if (n >= NT_MIN_N)                                                   \
            colour_pass(DfSmoothOp<FRN, true>{d, xb, rhs, c}, g, colour, s); \
        else                                                                 \
            colour_pass(DfSmoothOp<FRN>{d, xb, rhs, c}, g, colour, s);       \
    } while (0)
    if (frn == 0) SMD(0);
    else if (frn == 1) SMD(1);
    else SMD(2);
#undef SMD
}

void launch_df_residual(int frn, const float* d, const float* xb, const float* rhs, float* res, int n,
                        const DfCoef& c, hipStream_t s)
{
    const Geo g = make_geo(n);
#define RSD(FRN)                                                                      \
    do {                                                                              \
        if (n >= NT_MIN_N)                                                            \
            both_colours(DfResOp<FRN, true>{d, xb, rhs, res, c}, nullptr, g, s);      \
        else                                                                          \
            both_colours(DfResOp<FRN>{d, xb, rhs, res, c}, nullptr, g, s);            \
    } while (0)
    if (frn == 0) RSD(0);
    else if (frn == 1) RSD(1);
    else RSD(2);
#undef RSD
}

void launch_df_restrict_residual(int frn, const float* d0, const float* xb0, const float* g0, float* rres1,
                                 float* rd1, float* d1, int N, const DfCoef& c0, hipStream_t s)
{
    const Geo gf = make_geo(N), gc = make_geo(N / 2);
    int kc = zrun_length(gc.n);
#ifdef MGS_TUNING
    // tuning builds: MGS_ZRUN_RR sets this kernel's coarse planes per thread
    static const int rr = [] {
        const char* e = std::getenv("MGS_ZRUN_RR");
        return e ? std::atoi(e) : 0;
    }();
    if (rr > 0) kc = rr < gc.n ? rr : gc.n;
    // and MGS_RR_TX (64 or 128) its block width in coarse columns
    static const int rr_tx = [] {
        const char* e = std::getenv("MGS_RR_TX");
        const int v = e ? std::atoi(e) : 0;
        return (v == 64 || v == 128) ? v : 0;
    }();
    const ZLaunch L = make_zlaunch(gc.n, gc.n, kc, rr_tx);
#else
    const ZLaunch L = make_zlaunch(gc.n, gc.n, kc);
#endif
    FRN_DISPATCH(frn, k_df_restrict_residual, d0, xb0, g0, rres1, rd1, d1, gf, gc, c0, ZARGS);
}

void launch_df_rhs(int frn, const float* rd, const float* xb, const float* rres, float* f, int n, const DfCoef& c,
                   hipStream_t s)
{
    launch_df_residual(frn, rd, xb, rres, f, n, c, s);
}

void launch_df_correct(int frn, const float* dH, const float* rd, float* dfine, const float* xbfine, int nc,
                       hipStream_t s)
{
    const Geo gf = make_geo(2 * nc), gc = make_geo(nc);
    const ZLaunch L = make_zlaunch(nc, nc);
    FRN_DISPATCH_NT(frn, 2 * nc >= NT_MIN_N, k_df_correct, dH, rd, dfine, xbfine, gf, gc, ZARGS);
}

bool launch_df_tail(int frn, int nlev, const TailLevelIn* lv, hipStream_t s)
{
    if (nlev < 2 || nlev > TAIL_MAXL) return false;
    long lds = 0;  // the LDS copies: d, xb, g of the first level, d, xb, g, rres, rd below
    for (int i = 0; i < nlev; i++) lds += (i == 0 ? 3L : 5L) * lv[i].n * lv[i].n * lv[i].n;
    if (lds > TAIL_LDS_FLOATS) return false;
    TailDesc T;
    T.nl = nlev - 1;
    for (int i = 0; i < nlev; i++) {
        T.L[i].g = make_geo(lv[i].n);
        T.L[i].d = lv[i].d;
        T.L[i].xb = lv[i].xb;
        T.L[i].f = lv[i].f;
        T.L[i].rres = lv[i].rres;
        T.L[i].rd = lv[i].rd;
        T.L[i].dc = lv[i].dc;
    }
    if (frn == 0) hipLaunchKernelGGL((k_df_tail<0>), dim3(1), dim3(1024), 0, s, T);
    else if (frn == 1) hipLaunchKernelGGL((k_df_tail<1>), dim3(1), dim3(1024), 0, s, T);
    else hipLaunchKernelGGL((k_df_tail<2>), dim3(1), dim3(1024), 0, s, T);
    return true;
}

void launch_df_coarsest(int frn, float* d, const float* xb, const float* rhs, int n, int sweeps, const DfCoef& c,
                        hipStream_t s)
{
    const Geo g = make_geo(n);
    int threads = (int)g.half;
    if (threads > 1024) threads = 1024;
    threads = ((threads + 63) / 64) * 64;
    if (frn == 0) hipLaunchKernelGGL((k_df_coarsest<0>), dim3(1), dim3(threads), 0, s, d, xb, rhs, g, sweeps, c);
    else if (frn == 1) hipLaunchKernelGGL((k_df_coarsest<1>), dim3(1), dim3(threads), 0, s, d, xb, rhs, g, sweeps, c);
    else hipLaunchKernelGGL((k_df_coarsest<2>), dim3(1), dim3(threads), 0, s, d, xb, rhs, g, sweeps, c);
}

}  // namespace mgs
