// mgs_pm.hip -- the particle side of a PM step around the scalar-field solve:
// CIC density deposit (DENSIT) and the particle push with the CIC-gathered
// potential gradient (MOVE), on particles resident in HBM.
//
// Particles are the reference's REAL*4 arrays XPAR..VZ (PMP2mod_tools.f90:46)
// in its convention: positions in grid units in [1, N+1), cell I = INT(X).
// They live on the device in six float arrays owned by the handle.
//
// Deposit into the solver's level-0 density (colour-split FP64, the array
// mgs_set_density fills), so a solve can follow without a host trip.  The
// particles are binned by the 16^3-cell tile of their CIC base cell, in runs
// of consecutive particles of one tile (a counting sort of the runs: per-tile
// counts whose atomics also rank the runs, an exclusive scan, a scatter of
// (first particle, length) into tile order); one workgroup per tile then
// accumulates its particles' 8 weights in a 17^3 fixed-point buffer in LDS
// and adds each touched cell to the grid once, instead of eight global
// atomics per particle.  Integer accumulation makes the sum independent of
// the order of the additions (the reference adds into REAL*4 FI).
//
// Move: one thread per particle gathers the 32 potential values of its 2x2x2
// node block and their outer neighbours (PMP2main.f90:496-558), forms the
// doubled central-difference gradients at the 8 nodes, interpolates them
// with the CIC weights, kicks, drifts, wraps, and adds v^2 WPAR to a block
// partial of SVEL.  The potential (mgs_potential, + mgs_potential_add_fifth_force)
// is REAL*4 in x-fastest order, as the reference's FI; the gradients and
// weights are formed in FP64 from it, as there.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <string>

#include "../../include/mgsolve.h"
#include "mgs_device.h"
#include "mgs_internal.h"
#include "mgs_kernels.h"

namespace mgs {

struct PmState {
    int64_t np = 0;
    float* buf = nullptr;  // x, y, z, vx, vy, vz: 6 * np floats
    float* phi = nullptr;  // N^3 potential, x fastest, REAL*4 like the reference's FI
    double* part = nullptr;
    int* bad = nullptr;    // particles outside [1, N+1) seen by the last kernel
    int npart_blocks = 0;
    // deposit binning (runs of particles, see k_bin_count): per particle its
    // key and rank, the run descriptors in tile order; per tile its count of
    // runs and first run
    int64_t nbin = 0;
    int* key = nullptr;
    int* rank = nullptr;
    int64_t* runs = nullptr;  // nbin (at most one run per particle)
    int ntiles = 0;
    int* count = nullptr;
    int64_t* first = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
};

static PmState* pm_of(mgs_handle* h) { return (PmState*)h->pm; }

constexpr int PM_BLOCK = 256;
constexpr int PM_GRID_MAX = 8192;  // persistent blocks of the particle kernels

static int pm_grid(int64_t np)
{
    const int64_t b = (np + PM_BLOCK - 1) / PM_BLOCK;
    return (int)(b < PM_GRID_MAX ? (b < 1 ? 1 : b) : PM_GRID_MAX);
}

void pm_release(mgs_handle* h)
{
    PmState* p = pm_of(h);
    if (!p) return;
    if (p->buf) (void)hipFree(p->buf);
    if (p->phi) (void)hipFree(p->phi);
    if (p->part) (void)hipFree(p->part);
    if (p->bad) (void)hipFree(p->bad);
    for (void* q : {(void*)p->key, (void*)p->rank, (void*)p->runs, (void*)p->count, (void*)p->first, p->scan_tmp})
        if (q) (void)hipFree(q);
    delete p;
    h->pm = nullptr;
}

float* pm_potential(mgs_handle* h)
{
    PmState* p = pm_of(h);
    if (!p) {
        p = new PmState();
        h->pm = p;
    }
    if (!p->phi) {
        const size_t n3 = (size_t)h->p.ngrid * h->p.ngrid * h->p.ngrid;
        if (hipMalloc((void**)&p->phi, n3 * sizeof(float)) != hipSuccess) {
            p->phi = nullptr;
            return nullptr;
        }
    }
    return p->phi;
}

// INT(X) of the reference for X in [1, N+1): the 1-based cell; -1 outside
__device__ __forceinline__ int cell_of(double X, int N)
{
    if (!(X >= 1.0 && X < (double)N + 1.0)) return -1;
    return (int)X;
}

// DENSIT's CIC loop, PMP2mod_density.f90:235-281 (0-based cells here), over
// the particles binned by tile.
constexpr int PM_TILE = 16;                                            // cells per tile side
constexpr int PM_TBUF = PM_TILE + 1;                                   // with the +1 CIC reach
constexpr int PM_TCELLS = PM_TBUF * PM_TBUF * PM_TBUF;                 // 4913 (38.4 KiB of LDS)

// tile of 0-based base cell (i0, j0, k0); TT tiles per side
__device__ __forceinline__ int tile_of(int i0, int j0, int k0, int TT)
{
    return ((k0 / PM_TILE) * TT + j0 / PM_TILE) * TT + i0 / PM_TILE;
}

// The binning passes take PM_RUN consecutive particles per thread (particles
// in array order are usually neighbours in space, so consecutive ones usually
// share a tile) and bin *runs*: maximal sequences of a thread's particles in
// one tile, described by (first particle, length).
constexpr int PM_RUN = 8;
constexpr int KEY_OUTSIDE = -1;  // particle outside [1, N+1)
constexpr int KEY_CONT = -2;     // particle continuing the run before it

// per run: its tile (in key[] at the run's first particle; the others KEY_CONT)
// and its rank among the runs of that tile (in rank[] there); per tile: the
// count of runs
__global__ __launch_bounds__(PM_BLOCK) void k_bin_count(const float* __restrict__ xs, const float* __restrict__ ys,
                                                        const float* __restrict__ zs, int64_t np, int N, int TT,
                                                        int* __restrict__ key, int* __restrict__ rank,
                                                        int* __restrict__ count, int* __restrict__ bad)
{
    const int64_t nchunk = (np + PM_RUN - 1) / PM_RUN;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t ch = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ch < nchunk; ch += stride) {
        const int64_t p0 = ch * PM_RUN;
        // a whole chunk with 16-byte aligned rows: two 16-byte loads per array
        const bool vec = (np & 3) == 0 && p0 + PM_RUN <= np;
        float x[PM_RUN], y[PM_RUN], z[PM_RUN];
        if (vec) {
#pragma unroll
            for (int v = 0; v < PM_RUN / 4; v++) {
                const float4 a = ((const float4*)(xs + p0))[v], b = ((const float4*)(ys + p0))[v],
                             c = ((const float4*)(zs + p0))[v];
                x[4 * v] = a.x, x[4 * v + 1] = a.y, x[4 * v + 2] = a.z, x[4 * v + 3] = a.w;
                y[4 * v] = b.x, y[4 * v + 1] = b.y, y[4 * v + 2] = b.z, y[4 * v + 3] = b.w;
                z[4 * v] = c.x, z[4 * v + 1] = c.y, z[4 * v + 2] = c.z, z[4 * v + 3] = c.w;
            }
        } else {
#pragma unroll
            for (int r = 0; r < PM_RUN; r++) {
                const int64_t p = p0 + r < np ? p0 + r : np - 1;
                x[r] = xs[p];
                y[r] = ys[p];
                z[r] = zs[p];
            }
        }
        int t[PM_RUN];
#pragma unroll
        for (int r = 0; r < PM_RUN; r++) {
            t[r] = KEY_OUTSIDE;
            if (p0 + r < np) {
                const int I = cell_of(x[r], N), J = cell_of(y[r], N), K = cell_of(z[r], N);
                if (I < 0 || J < 0 || K < 0)
                    atomicAdd(bad, 1);
                else
                    t[r] = tile_of(I - 1, J - 1, K - 1, TT);
            }
        }
        int k[PM_RUN], rk[PM_RUN];
#pragma unroll
        for (int r = 0; r < PM_RUN; r++) {
            const bool start = t[r] >= 0 && (r == 0 || t[r] != t[r - 1]);
            k[r] = start || t[r] < 0 ? t[r] : KEY_CONT;
            rk[r] = 0;
            if (start) rk[r] = atomicAdd(&count[t[r]], 1);
        }
        if (p0 + PM_RUN <= np) {  // key + p0 is 32-byte aligned (p0 a multiple of 8)
#pragma unroll
            for (int v = 0; v < PM_RUN / 4; v++) {
                ((int4*)(key + p0))[v] = make_int4(k[4 * v], k[4 * v + 1], k[4 * v + 2], k[4 * v + 3]);
                ((int4*)(rank + p0))[v] = make_int4(rk[4 * v], rk[4 * v + 1], rk[4 * v + 2], rk[4 * v + 3]);
            }
        } else {
#pragma unroll
            for (int r = 0; r < PM_RUN; r++)
                if (p0 + r < np) {
                    key[p0 + r] = k[r];
                    rank[p0 + r] = rk[r];
                }
        }
    }
}

// run descriptors into tile order: (first particle) * PM_RUN + (length - 1)
__global__ __launch_bounds__(PM_BLOCK) void k_bin_scatter(int64_t np, const int* __restrict__ key,
                                                          const int* __restrict__ rank,
                                                          const int64_t* __restrict__ first,
                                                          int64_t* __restrict__ runs)
{
    const int64_t nchunk = (np + PM_RUN - 1) / PM_RUN;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t ch = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ch < nchunk; ch += stride) {
        const int64_t p0 = ch * PM_RUN;
        int k[PM_RUN];
        if (p0 + PM_RUN <= np) {
#pragma unroll
            for (int v = 0; v < PM_RUN / 4; v++) {
                const int4 a = ((const int4*)(key + p0))[v];
                k[4 * v] = a.x, k[4 * v + 1] = a.y, k[4 * v + 2] = a.z, k[4 * v + 3] = a.w;
            }
        } else {
#pragma unroll
            for (int r = 0; r < PM_RUN; r++) k[r] = p0 + r < np ? key[p0 + r] : KEY_OUTSIDE;
        }
#pragma unroll
        for (int r = 0; r < PM_RUN; r++) {
            if (k[r] < 0) continue;
            int len = 1;
#pragma unroll
            for (int q = r + 1; q < PM_RUN; q++) {
                bool cont = true;
#pragma unroll
                for (int u = r + 1; u <= q; u++) cont = cont && k[u] == KEY_CONT;
                len += cont ? 1 : 0;
            }
            runs[first[k[r]] + rank[p0 + r]] = (p0 + r) * PM_RUN + (len - 1);
        }
    }
}

// The deposit sums in fixed point so that its result does not depend on the
// order in which particles meet a cell: every CIC contribution (DENSIT's
// arithmetic in FP64, a value in [0, WPAR]) is rounded once to an integer
// multiple of 2^-40 and the multiples are added as 64-bit integers, which is
// exact and associative; a cell then holds -1 + (its integer) * 2^-40.  The
// result is bitwise the same run to run, within 2^-41 per contribution of
// the FP64 sum; a cell's integer holds 1 + delta up to 2^24 and converts
// exactly up to 2^13.
constexpr double PM_FIX = 1099511627776.0;          // 2^40
constexpr double PM_UNFIX = 1.0 / 1099511627776.0;  // 2^-40

__device__ __forceinline__ unsigned long long pm_fix(double v) { return __double2ull_rn(v * PM_FIX); }

// one workgroup per tile: the CIC weights of the particles of its runs
// accumulated in LDS over the tile's cells and their +1 reach, then added
// per touched cell into the level's density storage, which holds the
// fixed-point sums until k_deposit_finish converts them.
//
// PHASED: the launch covers the tiles of one parity class (tile coordinates
// (2a + px, 2b + py, 2c + pz), phase = px + 2 py + 4 pz) and adds with plain
// loads and stores.  Two tiles of a class are at least two tiles apart, so
// their 17^3 boxes are disjoint (with TT even also across the periodic wrap,
// whose reach lands in tile 0's box only), and the eight classes run as
// eight launches in order.  Otherwise every tile in one launch adds with
// device-scope 64-bit atomics.  Integer sums make both bitwise the same.
template <bool PHASED>
__global__ __launch_bounds__(PM_BLOCK) void k_tile_deposit(const float* __restrict__ xs, const float* __restrict__ ys,
                                                           const float* __restrict__ zs,
                                                           const int64_t* __restrict__ runs,
                                                           const int* __restrict__ count,
                                                           const int64_t* __restrict__ first, int nlaunch, int TT,
                                                           int phase, unsigned long long* __restrict__ sums, Geo g,
                                                           double wpar)
{
    __shared__ unsigned long long acc[PM_TCELLS];
    const int N = g.n;
    for (int bt = blockIdx.x; bt < nlaunch; bt += gridDim.x) {
        int t = bt;
        if constexpr (PHASED) {
            const int H = TT >> 1;
            const int a = bt % H, b2 = (bt / H) % H, c2 = bt / (H * H);
            t = ((2 * c2 + ((phase >> 2) & 1)) * TT + 2 * b2 + ((phase >> 1) & 1)) * TT + 2 * a + (phase & 1);
        }
        const int c = count[t];
        if (c == 0) continue;  // uniform over the block
        for (int i = threadIdx.x; i < PM_TCELLS; i += blockDim.x) acc[i] = 0ull;
        __syncthreads();
        const int i00 = (t % TT) * PM_TILE, j00 = ((t / TT) % TT) * PM_TILE, k00 = (t / (TT * TT)) * PM_TILE;
        const int64_t o = first[t];
        // thread q: particle q % PM_RUN of run q / PM_RUN (idle past its length)
        for (int64_t q = threadIdx.x; q < (int64_t)c * PM_RUN; q += blockDim.x) {
            const int64_t d = runs[o + q / PM_RUN];
            const int u = (int)(q % PM_RUN);
            if (u > (int)(d % PM_RUN)) continue;
            const int64_t pr = d / PM_RUN;
            const double X = xs[pr + u], Y = ys[pr + u], Z = zs[pr + u];
            const int I = (int)X, J = (int)Y, K = (int)Z;  // in [1, N]: checked when binned
            const double D1 = X - (double)I, D2 = Y - (double)J, D3 = Z - (double)K;
            const double T1 = 1.0 - D1, T2 = 1.0 - D2, T3 = 1.0 - D3;
            const double T2W = T2 * wpar, D2W = D2 * wpar;
            const int b = ((K - 1 - k00) * PM_TBUF + (J - 1 - j00)) * PM_TBUF + (I - 1 - i00);
            atomicAdd(&acc[b], pm_fix(T3 * T1 * T2W));
            atomicAdd(&acc[b + 1], pm_fix(T3 * D1 * T2W));
            atomicAdd(&acc[b + PM_TBUF], pm_fix(T3 * T1 * D2W));
            atomicAdd(&acc[b + PM_TBUF + 1], pm_fix(T3 * D1 * D2W));
            atomicAdd(&acc[b + PM_TBUF * PM_TBUF], pm_fix(D3 * T1 * T2W));
            atomicAdd(&acc[b + PM_TBUF * PM_TBUF + 1], pm_fix(D3 * D1 * T2W));
            atomicAdd(&acc[b + PM_TBUF * PM_TBUF + PM_TBUF], pm_fix(D3 * T1 * D2W));
            atomicAdd(&acc[b + PM_TBUF * PM_TBUF + PM_TBUF + 1], pm_fix(D3 * D1 * D2W));
        }
        __syncthreads();
        for (int i = threadIdx.x; i < PM_TCELLS; i += blockDim.x) {
            const unsigned long long v = acc[i];
            if (v == 0ull) continue;
            int gi = i00 + i % PM_TBUF, gj = j00 + (i / PM_TBUF) % PM_TBUF, gk = k00 + i / (PM_TBUF * PM_TBUF);
            gi = gi >= N ? gi - N : gi;
            gj = gj >= N ? gj - N : gj;
            gk = gk >= N ? gk - N : gk;
            if constexpr (PHASED)
                sums[split_index(g, gi, gj, gk)] += v;
            else
                atomicAdd(&sums[split_index(g, gi, gj, gk)], v);
        }
        __syncthreads();
    }
}

// the fixed-point sums (in place) -> delta = -1 + sum * 2^-40
__global__ __launch_bounds__(256) void k_deposit_finish(double* __restrict__ rho, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned long long v = reinterpret_cast<const unsigned long long*>(rho)[i];
        rho[i] = -1.0 + (double)v * PM_UNFIX;
    }
}

// MOVE's particle loop for the f(R) and nDGP models (no friction or mass
// terms, PMP2main.f90:488-684): kick with the CIC-interpolated gradient of
// the potential, drift, periodic wrap, SVEL partial per block
__global__ __launch_bounds__(PM_BLOCK) void k_move(float* __restrict__ xs, float* __restrict__ ys,
                                                   float* __restrict__ zs, float* __restrict__ vxs,
                                                   float* __restrict__ vys, float* __restrict__ vzs, int64_t np,
                                                   const float* __restrict__ phi, int N, double pconst,
                                                   double xconst, double wpar, double* __restrict__ part,
                                                   int* __restrict__ bad)
{
    // XN = FLOAT(NGRID)+1.-1.E-8 and YN = FLOAT(NGRID), evaluated as the
    // reference does (default REAL: the 1e-8 is below float resolution)
    const double XN = (double)((float)N + 1.0f - 1.0e-8f);
    const double YN = (double)N;
    const size_t n = (size_t)N;
    double svel = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += stride) {
        double X = xs[p], Y = ys[p], Z = zs[p];
        double VVX = vxs[p], VVY = vys[p], VVZ = vzs[p];
        const int I = cell_of(X, N), J = cell_of(Y, N), K = cell_of(Z, N);
        if (I < 0 || J < 0 || K < 0) {
            atomicAdd(bad, 1);
            continue;
        }
        const double D1 = X - (double)I, D2 = Y - (double)J, D3 = Z - (double)K;
        const double T1 = 1.0 - D1, T2 = 1.0 - D2, T3 = 1.0 - D3;
        const double T2T1 = T2 * T1, T2D1 = T2 * D1, D2T1 = D2 * T1, D2D1 = D2 * D1;
        // 0-based indices of I-1 .. I+2 (x), J-1 .. J+2 (y), K-1 .. K+2 (z)
        const int x1 = I - 1, x2 = I == N ? 0 : I, x3 = I + 1 >= N ? I + 1 - N : I + 1, x0 = I == 1 ? N - 1 : I - 2;
        const int y1 = J - 1, y2 = J == N ? 0 : J, y3 = J + 1 >= N ? J + 1 - N : J + 1, y0 = J == 1 ? N - 1 : J - 2;
        const int z1 = K - 1, z2 = K == N ? 0 : K, z3 = K + 1 >= N ? K + 1 - N : K + 1, z0 = K == 1 ? N - 1 : K - 2;
#define PHI(a, b, c) phi[(size_t)(a) + n * ((size_t)(b) + n * (size_t)(c))]
        const float F111 = PHI(x1, y1, z1), F211 = PHI(x2, y1, z1), F121 = PHI(x1, y2, z1), F221 = PHI(x2, y2, z1);
        const float F112 = PHI(x1, y1, z2), F212 = PHI(x2, y1, z2), F122 = PHI(x1, y2, z2), F222 = PHI(x2, y2, z2);
        const float F113 = PHI(x1, y1, z3), F213 = PHI(x2, y1, z3), F123 = PHI(x1, y2, z3), F223 = PHI(x2, y2, z3);
        const float F110 = PHI(x1, y1, z0), F210 = PHI(x2, y1, z0), F120 = PHI(x1, y2, z0), F220 = PHI(x2, y2, z0);
        const float F311 = PHI(x3, y1, z1), F321 = PHI(x3, y2, z1), F131 = PHI(x1, y3, z1), F231 = PHI(x2, y3, z1);
        const float F312 = PHI(x3, y1, z2), F322 = PHI(x3, y2, z2), F132 = PHI(x1, y3, z2), F232 = PHI(x2, y3, z2);
        const float F011 = PHI(x0, y1, z1), F021 = PHI(x0, y2, z1), F101 = PHI(x1, y0, z1), F201 = PHI(x2, y0, z1);
        const float F012 = PHI(x0, y1, z2), F022 = PHI(x0, y2, z2), F102 = PHI(x1, y0, z2), F202 = PHI(x2, y0, z2);
#undef PHI
        // doubled gradients at the 8 nodes (the 1/2 is in pconst); the
        // reference's F's are default REAL, so each difference is a REAL*4
        // subtraction (PMP2main.f90:580-602), widened afterwards
#define G2(a, b) (double)((a) - (b))
        const double GX111 = G2(F211, F011), GX211 = G2(F311, F111), GX121 = G2(F221, F021), GX221 = G2(F321, F121);
        const double GX112 = G2(F212, F012), GX212 = G2(F312, F112), GX122 = G2(F222, F022), GX222 = G2(F322, F122);
        const double GY111 = G2(F121, F101), GY211 = G2(F221, F201), GY121 = G2(F131, F111), GY221 = G2(F231, F211);
        const double GY112 = G2(F122, F102), GY212 = G2(F222, F202), GY122 = G2(F132, F112), GY222 = G2(F232, F212);
        const double GZ111 = G2(F112, F110), GZ211 = G2(F212, F210), GZ121 = G2(F122, F120), GZ221 = G2(F222, F220);
        const double GZ112 = G2(F113, F111), GZ212 = G2(F213, F211), GZ122 = G2(F123, F121), GZ222 = G2(F223, F221);
#undef G2
        const double GX = pconst * (T3 * (T2T1 * GX111 + T2D1 * GX211 + D2T1 * GX121 + D2D1 * GX221) +
                                    D3 * (T2T1 * GX112 + T2D1 * GX212 + D2T1 * GX122 + D2D1 * GX222));
        const double GY = pconst * (T3 * (T2T1 * GY111 + T2D1 * GY211 + D2T1 * GY121 + D2D1 * GY221) +
                                    D3 * (T2T1 * GY112 + T2D1 * GY212 + D2T1 * GY122 + D2D1 * GY222));
        const double GZ = pconst * (T3 * (T2T1 * GZ111 + T2D1 * GZ211 + D2T1 * GZ121 + D2D1 * GZ221) +
                                    D3 * (T2T1 * GZ112 + T2D1 * GZ212 + D2T1 * GZ122 + D2D1 * GZ222));
        VVX = VVX + GX;
        VVY = VVY + GY;
        VVZ = VVZ + GZ;
        X = X + VVX * xconst;
        Y = Y + VVY * xconst;
        Z = Z + VVZ * xconst;
        if (X < 1.0) X = X + YN;
        if (X >= XN) X = X - YN;
        if (Y < 1.0) Y = Y + YN;
        if (Y >= XN) Y = Y - YN;
        if (Z < 1.0) Z = Z + YN;
        if (Z >= XN) Z = Z - YN;
        svel = svel + (VVX * VVX + VVY * VVY + VVZ * VVZ) * wpar;
        float xf = (float)X, yf = (float)Y, zf = (float)Z;
        // a coordinate that rounds to N+1 in REAL*4 is pulled back inside
        if ((int)xf == N + 1) xf = xf - 1.0e-3f;
        if ((int)yf == N + 1) yf = yf - 1.0e-3f;
        if ((int)zf == N + 1) zf = zf - 1.0e-3f;
        xs[p] = xf;
        ys[p] = yf;
        zs[p] = zf;
        vxs[p] = (float)VVX;
        vys[p] = (float)VVY;
        vzs[p] = (float)VVZ;
    }
    const double t = block_sum(svel);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// WPAR = YN**3/FLOAT(Nparticles) (PMP2mod_density.f90:40, PMP2main.f90:470):
// the particle count is converted to default REAL first, as there
// WPAR = YN**3/FLOAT(Nparticles) (PMP2mod_density.f90:40, PMP2main.f90:470):
// the count through default REAL, the quotient in REAL*8, stored in WPAR, an
// implicitly typed default REAL
static double wpar_of(int N, int64_t np)
{
    const double yn = (double)N;
    return (double)(float)(yn * yn * yn / (double)(float)np);
}

static int pm_ensure_particles(mgs_handle* h, int64_t np)
{
    PmState* p = pm_of(h);
    if (!p) {
        p = new PmState();
        h->pm = p;
    }
    if (!p->bad && hipMalloc((void**)&p->bad, sizeof(int)) != hipSuccess) {
        p->bad = nullptr;
        return set_err(h, MGS_E_NOMEM, "particle counters");
    }
    if (!p->part) {
        if (hipMalloc((void**)&p->part, (PM_GRID_MAX + 1024) * sizeof(double)) != hipSuccess) {
            p->part = nullptr;
            return set_err(h, MGS_E_NOMEM, "particle partials");
        }
    }
    if (p->np == np && p->buf) return 0;
    if (p->buf) (void)hipFree(p->buf);
    p->buf = nullptr;
    p->np = 0;
    if (hipMalloc((void**)&p->buf, (size_t)np * 6 * sizeof(float)) != hipSuccess) {
        p->buf = nullptr;
        return set_err(h, MGS_E_NOMEM, "particle arrays (" + std::to_string(np) + " particles)");
    }
    p->np = np;
    return 0;
}

// out-of-range particles counted by the last kernel (synchronises)
static int pm_check_bad(mgs_handle* h, const char* what)
{
    int bad = 0;
    if (hipMemcpyAsync(&bad, pm_of(h)->bad, sizeof(int), hipMemcpyDeviceToHost, h->s) != hipSuccess ||
        hipStreamSynchronize(h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, std::string(what) + ": HIP error");
    if (bad)
        return set_err(h, MGS_E_DOMAIN, std::string(what) + ": " + std::to_string(bad) +
                                            " particle(s) outside [1, N+1) or not finite were skipped");
    return 0;
}

}  // namespace mgs

using namespace mgs;

extern "C" {

int mgs_particles_set(mgs_handle* h, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                      const float* vz, int64_t npart, int32_t where)
{
    if (!h) return MGS_E_ARG;
    if (!x || !y || !z) return set_err(h, MGS_E_ARG, "positions are NULL");
    if (npart < 1) return set_err(h, MGS_E_ARG, "npart < 1");
    if (where != MGS_HOST && where != MGS_DEVICE) return set_err(h, MGS_E_ARG, "where");
    int rc = pm_ensure_particles(h, npart);
    if (rc) return rc;
    PmState* p = pm_of(h);
    const size_t bytes = (size_t)npart * sizeof(float);
    const hipMemcpyKind kind = where == MGS_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    const float* src[6] = {x, y, z, vx, vy, vz};
    for (int a = 0; a < 6; a++) {
        float* dst = p->buf + (size_t)a * npart;
        const hipError_t e = src[a] ? hipMemcpyAsync(dst, src[a], bytes, kind, h->s) : hipMemsetAsync(dst, 0, bytes, h->s);
        if (e != hipSuccess) return set_err(h, MGS_E_HIP, "particle upload");
    }
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "particle upload");
    return 0;
}

int mgs_particles_get(mgs_handle* h, float* x, float* y, float* z, float* vx, float* vy, float* vz, int32_t where)
{
    if (!h) return MGS_E_ARG;
    PmState* p = pm_of(h);
    if (!p || !p->buf) return set_err(h, MGS_E_STATE, "no particles set");
    if (where != MGS_HOST && where != MGS_DEVICE) return set_err(h, MGS_E_ARG, "where");
    const size_t bytes = (size_t)p->np * sizeof(float);
    const hipMemcpyKind kind = where == MGS_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    float* dst[6] = {x, y, z, vx, vy, vz};
    for (int a = 0; a < 6; a++)
        if (dst[a] && hipMemcpyAsync(dst[a], p->buf + (size_t)a * p->np, bytes, kind, h->s) != hipSuccess)
            return set_err(h, MGS_E_HIP, "particle download");
    if (hipStreamSynchronize(h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "particle download");
    return 0;
}

// the deposit's binning buffers for the current particle count and ntiles
static int pm_ensure_bins(mgs_handle* h, PmState* p, int ntiles)
{
    if (p->nbin != p->np) {
        for (void* q : {(void*)p->key, (void*)p->rank, (void*)p->runs})
            if (q) (void)hipFree(q);
        p->key = p->rank = nullptr;
        p->runs = nullptr;
        p->nbin = 0;
        const size_t n = (size_t)p->np;
        if (hipMalloc((void**)&p->key, n * sizeof(int)) != hipSuccess ||
            hipMalloc((void**)&p->rank, n * sizeof(int)) != hipSuccess ||
            hipMalloc((void**)&p->runs, n * sizeof(int64_t)) != hipSuccess)
            return set_err(h, MGS_E_NOMEM, "deposit binning arrays");
        p->nbin = p->np;
    }
    if (p->ntiles != ntiles) {
        for (void* q : {(void*)p->count, (void*)p->first, p->scan_tmp})
            if (q) (void)hipFree(q);
        p->count = nullptr;
        p->first = nullptr;
        p->scan_tmp = nullptr;
        p->ntiles = 0;
        size_t bytes = 0;
        if (hipMalloc((void**)&p->count, (size_t)ntiles * sizeof(int)) != hipSuccess ||
            hipMalloc((void**)&p->first, (size_t)ntiles * sizeof(int64_t)) != hipSuccess ||
            hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, p->count, p->first, ntiles, h->s) != hipSuccess ||
            hipMalloc(&p->scan_tmp, bytes + 256) != hipSuccess)
            return set_err(h, MGS_E_NOMEM, "deposit tile arrays");
        p->scan_bytes = bytes + 256;
        p->ntiles = ntiles;
    }
    return 0;
}

int mgs_cic_deposit(mgs_handle* h)
{
    if (!h) return MGS_E_ARG;
    PmState* p = pm_of(h);
    if (!p || !p->buf) return set_err(h, MGS_E_STATE, "no particles set");
    if (p->np >= ((int64_t)1 << 31)) return set_err(h, MGS_E_ARG, "the deposit takes fewer than 2^31 particles");
    const int N = h->p.ngrid;
    const Geo g = make_geo(N);
    const int TT = (N + PM_TILE - 1) / PM_TILE;
    int rc = pm_ensure_bins(h, p, TT * TT * TT);
    if (rc) return rc;
    const int64_t np = p->np;
    const float* b = p->buf;
    prof_begin(h, MGS_KC_PM, (double)np * 56.0 + 32.0 * (double)N * N * N);
    // the density storage holds the fixed-point sums (zero) until k_deposit_finish
    launch_fill_f64(h->L[0].rho, (size_t)N * N * N, 0.0, h->s);
    if (hipMemsetAsync(p->bad, 0, sizeof(int), h->s) != hipSuccess ||
        hipMemsetAsync(p->count, 0, (size_t)p->ntiles * sizeof(int), h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, "deposit");
    const int gbin = pm_grid((np + PM_RUN - 1) / PM_RUN);
    hipLaunchKernelGGL(k_bin_count, dim3(gbin), dim3(PM_BLOCK), 0, h->s, b, b + np, b + 2 * np, np, N, TT,
                       p->key, p->rank, p->count, p->bad);
    if (hipcub::DeviceScan::ExclusiveSum(p->scan_tmp, p->scan_bytes, p->count, p->first, p->ntiles, h->s) !=
        hipSuccess)
        return set_err(h, MGS_E_HIP, "deposit scan");
    hipLaunchKernelGGL(k_bin_scatter, dim3(gbin), dim3(PM_BLOCK), 0, h->s, np, p->key, p->rank, p->first, p->runs);
    unsigned long long* sums = reinterpret_cast<unsigned long long*>(h->L[0].rho);
    if (TT % 2 == 0) {
        const int per = p->ntiles / 8;  // tiles of one parity class
        for (int ph = 0; ph < 8; ph++)
            hipLaunchKernelGGL(k_tile_deposit<true>, dim3((unsigned)per), dim3(PM_BLOCK), 0, h->s, b, b + np,
                               b + 2 * np, p->runs, p->count, p->first, per, TT, ph, sums, g, wpar_of(N, np));
    } else {
        hipLaunchKernelGGL(k_tile_deposit<false>, dim3((unsigned)p->ntiles), dim3(PM_BLOCK), 0, h->s, b, b + np,
                           b + 2 * np, p->runs, p->count, p->first, p->ntiles, TT, 0, sums, g, wpar_of(N, np));
    }
    {
        const size_t n3 = (size_t)N * N * N;
        const unsigned gfin = (unsigned)std::min<size_t>((n3 + 255) / 256, 65536);
        hipLaunchKernelGGL(k_deposit_finish, dim3(gfin), dim3(256), 0, h->s, h->L[0].rho, n3);
    }
    prof_end(h);
    rc = pm_check_bad(h, "CIC deposit");
    h->has_density = true;
    h->rho_cached = false;
    return rc;
}

int mgs_move(mgs_handle* h, double pconst, double xconst, double* svel)
{
    if (!h) return MGS_E_ARG;
    PmState* p = pm_of(h);
    if (!p || !p->buf) return set_err(h, MGS_E_STATE, "no particles set");
    if (!p->phi) return set_err(h, MGS_E_STATE, "no potential: call mgs_potential first");
    if (!std::isfinite(pconst) || !std::isfinite(xconst)) return set_err(h, MGS_E_ARG, "pconst/xconst not finite");
    const int N = h->p.ngrid;
    const int grid = pm_grid(p->np);
    prof_begin(h, MGS_KC_PM, (double)p->np * 48.0);
    if (hipMemsetAsync(p->bad, 0, sizeof(int), h->s) != hipSuccess) return set_err(h, MGS_E_HIP, "move");
    const int64_t np = p->np;
    float* b = p->buf;
    hipLaunchKernelGGL(k_move, dim3(grid), dim3(PM_BLOCK), 0, h->s, b, b + np, b + 2 * np, b + 3 * np, b + 4 * np,
                       b + 5 * np, np, p->phi, N, pconst, xconst, wpar_of(N, np), p->part, p->bad);
    launch_sum_partials(p->part, grid, h->dsum, h->s);
    prof_end(h);
    double s = 0.0;
    if (hipMemcpyAsync(&s, h->dsum, sizeof(double), hipMemcpyDeviceToHost, h->s) != hipSuccess)
        return set_err(h, MGS_E_HIP, "move");
    int rc = pm_check_bad(h, "move");
    if (svel) *svel = s;
    return rc;
}

}  // extern "C"
