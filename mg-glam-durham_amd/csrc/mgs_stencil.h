// mgs_stencil.h -- z-marching 7-point stencil drivers on the colour-split
// layout (see mgs_common.h).
//
// A thread owns one (m, j) column of a level and marches over a run of KC
// planes k.  The z-neighbours of a cell are the other colour's values at the
// same (m, j) in planes k-1 and k+1, and the other colour's value at
// (m, j, k) is always one of the cell's two x-neighbours, so a register
// window (prev, cur, next) over the other colour's column supplies three of
// the six neighbours: each step loads only next, the second x-neighbour and
// the two y-neighbours.  Blocks are (TX m-values) x (TY rows j) x (KC planes),
// ordered k-run-major and remapped per XCD so that one XCD sweeps a
// contiguous slab of planes (its L2 then holds the rows j+-1 and the planes
// k+-1 a block needs).
#pragma once

#include <cstdlib>

#include "mgs_device.h"

namespace mgs {

// Block order flag in ZLaunch::JB (row-block counts stay below 2^16): each
// XCD owns a slab of JB/8 block rows and marches it plane run by plane run,
// rows fastest.  A plane run's halo planes are then the neighbouring runs of
// the same slab, which the same XCD's L2 still holds, and the rows above and
// below are its own (but for the slab edges): the z-halo re-reads of a
// rows-fastest order over the whole grid (whose plane-run footprint exceeds
// an XCD's L2) stay on chip.
constexpr int ZSLAB = 1 << 20;
// bits 21-22 of JB: log2 of (number of slabs / 8).  A slab is marched plane run
// by plane run, so its rows times one run's planes must stay in an XCD's 4 MB
// L2 for the next run to find its halo plane there: from n = 1024 (one XCD's
// eighth of the rows holds about 6 MB per run of 4 planes) there are 16
// slabs, each XCD marching two in turn.
constexpr int ZSLAB_SHIFT = 21;

struct ZLaunch {
    dim3 grid, block;
    int TX, TY, XB, JB, KC;
    unsigned NB;
};

#ifdef MGS_TUNING
// tuning builds: the launch-shape overrides below apply to levels with n <= MGS_ZMAXN (default: all)
inline int tuning_max_n()
{
    static const int v = [] {
        const char* e = std::getenv("MGS_ZMAXN");
        return e ? std::atoi(e) : 1 << 30;
    }();
    return v;
}
#endif

// planes marched per thread; in tuning builds (-DMGS_TUNING) MGS_ZRUN overrides it
inline int zrun_length(int n)
{
#ifdef MGS_TUNING
    static const int forced = [] {
        const char* e = std::getenv("MGS_ZRUN");
        return e ? std::atoi(e) : 0;
    }();
    if (forced > 0 && n <= tuning_max_n()) return forced < n ? forced : n;
#endif
    // measured at 512^3 (runs of 2, 4, 6, 8, 16, 32 planes): 4 is best; the
    // window's two prologue loads per run cost less than the lost
    // parallelism of longer runs.  At 256^3 (the 512^3 solve's level 1) runs
    // of 2, 8 and 16 planes and 64-wide blocks are slower as well.
    return n >= 4 ? 4 : n;
}

// (TX x TY)-thread blocks over `width` columns x n rows, each thread marching
// a run of KC planes; k-run-major block order.  tx_cap > 0 caps the block
// width (tuning builds)
inline ZLaunch make_zlaunch(int width, int n, int kc, int tx_cap = 0)
{
    // widest block row; in tuning builds MGS_ZTX overrides it
#ifdef MGS_TUNING
    static const int max_tx_env = [] {
        const char* e = std::getenv("MGS_ZTX");
        const int v = e ? std::atoi(e) : 0;
        return (v == 32 || v == 64 || v == 128) ? v : 256;
    }();
    const int max_tx = n <= tuning_max_n() ? max_tx_env : 256;
#else
    constexpr int max_tx = 256;
#endif
    ZLaunch L;
    L.TX = width < max_tx ? width : max_tx;
    if (tx_cap > 0 && L.TX > tx_cap) L.TX = tx_cap;
    int ty = 256 / L.TX;
    if (ty > n) ty = n;
    L.TY = ty < 1 ? 1 : ty;
    L.XB = (width + L.TX - 1) / L.TX;
    L.JB = (n + L.TY - 1) / L.TY;
    L.KC = kc < 1 ? 1 : (kc > n ? n : kc);
    // block order: in tuning builds MGS_ZORDER=k makes the plane runs of one column block
    // consecutive (flagged to the kernels by a negative JB)
#ifdef MGS_TUNING
    static const bool kfast = [] {
        const char* e = std::getenv("MGS_ZORDER");
        return e && e[0] == 'k';
    }();
#else
    constexpr bool kfast = false;
#endif
    const int kb = (n + L.KC - 1) / L.KC;
    L.NB = (unsigned)(L.XB * L.JB * kb);
    const unsigned per = (L.NB + 7) / 8;
    L.grid = dim3(per * 8);
    L.block = dim3(L.TX, L.TY);
    const int slab_log = n >= 1024 ? 1 : 0;  // 8 << slab_log slabs
    if (kfast)
        L.JB = -L.JB;
    else if (L.JB % (8 << slab_log) == 0 && kb > 1)
        L.JB |= ZSLAB | (slab_log << ZSLAB_SHIFT);  // row slabs, one or more per XCD (zcol_at)
    return L;
}

inline ZLaunch make_zlaunch(int width, int n) { return make_zlaunch(width, n, zrun_length(n)); }

// stencil passes over one colour half: columns are the n/2 values of m
inline ZLaunch make_zlaunch(const Geo& g) { return make_zlaunch(g.h, g.n); }

// passes over both colours (residual-type kernels, two windows per column);
// in tuning builds MGS_ZRUN_BOTH overrides their run length
inline ZLaunch make_zlaunch_both(const Geo& g)
{
#ifdef MGS_TUNING
    static const int forced = [] {
        const char* e = std::getenv("MGS_ZRUN_BOTH");
        return e ? std::atoi(e) : 0;
    }();
#else
    constexpr int forced = 0;
#endif
    return make_zlaunch(g.h, g.n, forced > 0 ? forced : zrun_length(g.n));
}

struct ZCol {
    int m, j, k0, k1;
};

// column and plane run of thread (tx, ty) of logical block lb; false for
// idle threads
__device__ __forceinline__ bool zcol_at(int width, int n, int TX, int TY, int XB, int JB, int KC, unsigned lb, int tx,
                                        int ty, ZCol& c)
{
    unsigned xb, jb, kb;
    if (JB > 0 && (JB & ZSLAB)) {  // x blocks fastest, then the slab's rows, then plane runs, then slabs
        const unsigned JS = (unsigned)(JB & (ZSLAB - 1)) >> (3 + ((JB >> ZSLAB_SHIFT) & 3));
        const unsigned KB = (unsigned)((n + KC - 1) / KC);
        xb = XB == 1 ? 0u : lb % (unsigned)XB;
        unsigned t = XB == 1 ? lb : lb / (unsigned)XB;
        const unsigned jl = t % JS;
        t /= JS;
        kb = t % KB;
        jb = (t / KB) * JS + jl;
    } else if (JB > 0) {  // x blocks fastest, then rows, then plane runs
        xb = XB == 1 ? 0u : lb % (unsigned)XB;
        const unsigned t = XB == 1 ? lb : lb / (unsigned)XB;
        jb = t % (unsigned)JB;
        kb = t / (unsigned)JB;
    } else {  // plane runs fastest
        const unsigned KB = (unsigned)((n + KC - 1) / KC);
        kb = lb % KB;
        const unsigned t = lb / KB;
        jb = t % (unsigned)(-JB);
        xb = t / (unsigned)(-JB);
    }
    c.m = (int)(xb * TX + tx);
    c.j = (int)(jb * TY + ty);
    c.k0 = (int)kb * KC;
    c.k1 = c.k0 + KC < n ? c.k0 + KC : n;
    return c.m < width && c.j < n;
}

// the same for this thread of a standalone launch (XCD-remapped block id)
__device__ __forceinline__ bool zcol(int width, int n, int TX, int TY, int XB, int JB, int KC, unsigned NB, ZCol& c)
{
    unsigned lb;
    if (!map_block(NB, lb)) return false;
    return zcol_at(width, n, TX, TY, XB, JB, KC, lb, threadIdx.x, threadIdx.y, c);
}

__device__ __forceinline__ bool zcol(const Geo& g, int TX, int TY, int XB, int JB, int KC, unsigned NB, ZCol& c)
{
    return zcol(g.h, g.n, TX, TY, XB, JB, KC, NB, c);
}

// Column window over one colour half `a` (32-bit offsets; a level holds at
// most 2^30 cells).  fetch(k) issues the loads that give the six neighbour
// values of the opposite-colour cell at (m, j, k), whose x-parity is s;
// apply() turns them into the neighbours and advances the window to k+1.
// Splitting the two lets a loop issue the loads of plane k+1 before it
// computes and stores plane k.
template <typename T>
struct ZPre {
    T next, xo, yp, ym;
};

template <typename T>
struct ZWin {
    const T* __restrict__ a;
    unsigned n, h, m, jrow, jprow, jmrow;  // row offsets are j*h etc. within a plane
    unsigned plane;                         // n * h
    T prev, cur;

    __device__ __forceinline__ void init(const T* base, const Geo& g, int m_, int j, int k0)
    {
        a = base;
        n = (unsigned)g.n;
        h = (unsigned)g.h;
        m = (unsigned)m_;
        plane = n * h;
        jrow = (unsigned)j * h;
        jprow = (unsigned)wrap_inc(j, g.n) * h;
        jmrow = (unsigned)wrap_dec(j, g.n) * h;
        prev = elem(a, (unsigned)wrap_dec(k0, g.n) * plane + jrow + m);
        cur = elem(a, (unsigned)k0 * plane + jrow + m);
    }

    __device__ __forceinline__ ZPre<T> fetch(int k, int s) const
    {
        const unsigned kp = (unsigned)(k + 1 == (int)n ? 0 : k + 1);
        const unsigned pk = (unsigned)k * plane;
        unsigned mo;
        if (s) mo = (m + 1 == h) ? 0u : m + 1;
        else mo = (m == 0) ? h - 1 : m - 1;
        ZPre<T> p;
        p.next = elem(a, kp * plane + jrow + m);
        p.xo = elem(a, pk + jrow + mo);
        p.yp = elem(a, pk + jprow + m);
        p.ym = elem(a, pk + jmrow + m);
        return p;
    }

    __device__ __forceinline__ void apply(const ZPre<T>& p, int s, T& xp, T& xm, T& yp, T& ym, T& zp, T& zm)
    {
        xp = s ? p.xo : cur;
        xm = s ? cur : p.xo;
        yp = p.yp;
        ym = p.ym;
        zp = p.next;
        zm = prev;
        prev = cur;
        cur = p.next;
    }
};

// One colour pass.  OP::load(own offset, other offset) loads the cell's own
// data (everything but the marched field); OP::update(own offset, other
// offset, own data, value of the other-colour cell at the same (m, j, k),
// xp, xm, yp, ym, zp, zm) applies the update, with neighbours in the
// reference's order M1u, M1l, M2u, M2l, M3u, M3l.  Offsets index the full
// level array.  The loads of plane k+1 are issued before plane k is updated.
// A run of exactly R planes, unrolled: every load of the run is issued
// before the first update (R + 2 window values, R own records), and the
// window needs no register rotation.
template <int R, class OP>
__device__ __forceinline__ void colour_pass_run(const OP& op, ZWin<typename OP::T>& w, unsigned own, unsigned oth,
                                                unsigned col, int s0, int k0)
{
    using T = typename OP::T;
    using Own = typename OP::Own;
    ZPre<T> p[R];
    Own o[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned rel = (unsigned)(k0 + r) * w.plane + col;
        p[r] = w.fetch(k0 + r, s0 ^ (r & 1));
        o[r] = op.load(own + rel, oth + rel);
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned rel = (unsigned)(k0 + r) * w.plane + col;
        const T other = w.cur;
        T xp, xm, yp, ym, zp, zm;
        w.apply(p[r], s0 ^ (r & 1), xp, xm, yp, ym, zp, zm);
        op.update(own + rel, oth + rel, o[r], other, xp, xm, yp, ym, zp, zm);
    }
}

// plane-run length of the unrolled path (the default run, zrun_length)
constexpr int ZRUN_UNROLLED = 4;

template <class OP>
__device__ __forceinline__ void colour_pass_col(const OP& op, const Geo& g, int colour, const ZCol& c)
{
    using T = typename OP::T;
    using Own = typename OP::Own;
    ZWin<T> w;
    w.init(op.field() + (size_t)(1 - colour) * g.half, g, c.m, c.j, c.k0);
    const unsigned own = (unsigned)colour * (unsigned)g.half;
    const unsigned oth = (unsigned)(1 - colour) * (unsigned)g.half;
    const unsigned col = w.jrow + w.m;
    if (c.k1 - c.k0 == ZRUN_UNROLLED) {
        colour_pass_run<ZRUN_UNROLLED>(op, w, own, oth, col, (colour + c.j + c.k0) & 1, c.k0);
        return;
    }
    int k = c.k0;
    ZPre<T> pc = w.fetch(k, (colour + c.j + k) & 1);
    Own oc = op.load(own + (unsigned)k * w.plane + col, oth + (unsigned)k * w.plane + col);
    for (; k < c.k1; k++) {
        ZPre<T> pn = pc;
        Own on = oc;
        if (k + 1 < c.k1) {
            const unsigned rel = (unsigned)(k + 1) * w.plane + col;
            pn = w.fetch(k + 1, (colour + c.j + k + 1) & 1);
            on = op.load(own + rel, oth + rel);
        }
        const unsigned rel = (unsigned)k * w.plane + col;
        const T other = w.cur;
        T xp, xm, yp, ym, zp, zm;
        w.apply(pc, (colour + c.j + k) & 1, xp, xm, yp, ym, zp, zm);
        op.update(own + rel, oth + rel, oc, other, xp, xm, yp, ym, zp, zm);
        pc = pn;
        oc = on;
    }
}

template <class OP>
__global__ __launch_bounds__(256) void k_colour_pass(OP op, Geo g, int colour, int TX, int TY, int XB, int JB, int KC,
                                                     unsigned NB)
{
    ZCol c;
    if (!zcol(g, TX, TY, XB, JB, KC, NB, c)) return;
    colour_pass_col(op, g, colour, c);
}

// Both colours per column: OP::load(offset) loads a cell's own data;
// OP::cell(offset, own data, centre, xp, xm, yp, ym, zp, zm) returns a value
// whose square is accumulated; the block sum goes to partials[logical block]
// when partials is non-null.  Loads of plane k+1 are issued before plane k
// is evaluated.
// A run of exactly R planes of both colours, unrolled like colour_pass_run;
// the squares are accumulated in the same order as the loop below.
template <int R, class OP>
__device__ __forceinline__ double both_colours_run(const OP& op, ZWin<typename OP::T>& w0, ZWin<typename OP::T>& w1,
                                                   unsigned half, unsigned col, int j, int k0)
{
    using T = typename OP::T;
    using Own = typename OP::Own;
    ZPre<T> p0[R], p1[R];
    Own o0[R], o1[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned rel = (unsigned)(k0 + r) * w0.plane + col;
        p1[r] = w1.fetch(k0 + r, (j + k0 + r) & 1);
        p0[r] = w0.fetch(k0 + r, (1 + j + k0 + r) & 1);
        o0[r] = op.load(rel);
        o1[r] = op.load(half + rel);
    }
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const T c0 = w0.cur, c1 = w1.cur;
        const unsigned off = (unsigned)(k0 + r) * w0.plane + col;
        T xp, xm, yp, ym, zp, zm;
        w1.apply(p1[r], (j + k0 + r) & 1, xp, xm, yp, ym, zp, zm);
        const double r0 = op.cell(off, o0[r], c0, xp, xm, yp, ym, zp, zm);
        w0.apply(p0[r], (1 + j + k0 + r) & 1, xp, xm, yp, ym, zp, zm);
        const double r1 = op.cell(half + off, o1[r], c1, xp, xm, yp, ym, zp, zm);
        acc += r0 * r0 + r1 * r1;
    }
    return acc;
}

template <class OP>
__device__ __forceinline__ double both_colours_col(const OP& op, const Geo& g, const ZCol& c)
{
    using T = typename OP::T;
    using Own = typename OP::Own;
    double acc = 0.0;
    ZWin<T> w0, w1;  // windows over the red (0) and black (1) halves
    w0.init(op.field(), g, c.m, c.j, c.k0);
    w1.init(op.field() + g.half, g, c.m, c.j, c.k0);
    const unsigned half = (unsigned)g.half;
    const unsigned col = w0.jrow + w0.m;
    if (c.k1 - c.k0 == ZRUN_UNROLLED) return both_colours_run<ZRUN_UNROLLED>(op, w0, w1, half, col, c.j, c.k0);
    int k = c.k0;
    // red cell: neighbours in the black half, x-parity (j+k)&1;
    // black cell: neighbours in the red half, x-parity (1+j+k)&1
    ZPre<T> p1 = w1.fetch(k, (c.j + k) & 1);
    ZPre<T> p0 = w0.fetch(k, (1 + c.j + k) & 1);
    Own o0 = op.load((unsigned)k * w0.plane + col);
    Own o1 = op.load(half + (unsigned)k * w0.plane + col);
    for (; k < c.k1; k++) {
        ZPre<T> q1 = p1, q0 = p0;
        Own n0 = o0, n1 = o1;
        if (k + 1 < c.k1) {
            const unsigned rel = (unsigned)(k + 1) * w0.plane + col;
            q1 = w1.fetch(k + 1, (c.j + k + 1) & 1);
            q0 = w0.fetch(k + 1, (1 + c.j + k + 1) & 1);
            n0 = op.load(rel);
            n1 = op.load(half + rel);
        }
        const T c0 = w0.cur, c1 = w1.cur;
        const unsigned off = (unsigned)k * w0.plane + col;
        T xp, xm, yp, ym, zp, zm;
        w1.apply(p1, (c.j + k) & 1, xp, xm, yp, ym, zp, zm);
        const double r0 = op.cell(off, o0, c0, xp, xm, yp, ym, zp, zm);
        w0.apply(p0, (1 + c.j + k) & 1, xp, xm, yp, ym, zp, zm);
        const double r1 = op.cell(half + off, o1, c1, xp, xm, yp, ym, zp, zm);
        acc += r0 * r0 + r1 * r1;
        p1 = q1;
        p0 = q0;
        o0 = n0;
        o1 = n1;
    }
    return acc;
}

template <class OP>
__global__ __launch_bounds__(256) void k_both_colours(OP op, double* __restrict__ partials, Geo g, int TX, int TY,
                                                      int XB, int JB, int KC, unsigned NB)
{
    ZCol c;
    double acc = 0.0;
    if (zcol(g, TX, TY, XB, JB, KC, NB, c)) acc = both_colours_col(op, g, c);
    if (partials) store_partial(partials, NB, block_sum(acc));
}

}  // namespace mgs
