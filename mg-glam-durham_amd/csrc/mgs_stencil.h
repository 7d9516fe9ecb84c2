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

#include "mgs_device.h"

namespace mgs {

struct ZLaunch {
    dim3 grid, block;
    int TX, TY, XB, JB, KC;
    unsigned NB;
};

inline int zrun_length(int n)
{
    if (n >= 256) return 16;
    if (n >= 64) return 8;
    if (n >= 16) return 4;
    return n;
}

inline ZLaunch make_zlaunch(const Geo& g)
{
    ZLaunch L;
    L.TX = g.h < 256 ? g.h : 256;
    int ty = 256 / L.TX;
    if (ty > g.n) ty = g.n;
    L.TY = ty < 1 ? 1 : ty;
    L.XB = (g.h + L.TX - 1) / L.TX;
    L.JB = (g.n + L.TY - 1) / L.TY;
    L.KC = zrun_length(g.n);
    const int kb = (g.n + L.KC - 1) / L.KC;
    L.NB = (unsigned)(L.XB * L.JB * kb);
    const unsigned per = (L.NB + 7) / 8;
    L.grid = dim3(per * 8);
    L.block = dim3(L.TX, L.TY);
    return L;
}

struct ZCol {
    int m, j, k0, k1;
};

// column and plane run of this thread; false for idle threads
__device__ __forceinline__ bool zcol(const Geo& g, int TX, int TY, int XB, int JB, int KC, unsigned NB, ZCol& c)
{
    unsigned lb;
    if (!map_block(NB, lb)) return false;
    const unsigned xb = XB == 1 ? 0u : lb % (unsigned)XB;
    const unsigned t = XB == 1 ? lb : lb / (unsigned)XB;
    const unsigned jb = t % (unsigned)JB;
    const unsigned kb = t / (unsigned)JB;
    c.m = (int)(xb * TX + threadIdx.x);
    c.j = (int)(jb * TY + threadIdx.y);
    c.k0 = (int)kb * KC;
    c.k1 = c.k0 + KC < g.n ? c.k0 + KC : g.n;
    return c.m < g.h && c.j < g.n;
}

// Column window over one colour half `a` (32-bit offsets; a level holds at
// most 2^30 cells).  load() at step k gives the six neighbour values of the
// opposite-colour cell at (m, j, k), whose x-parity is s.
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
        prev = a[(unsigned)wrap_dec(k0, g.n) * plane + jrow + m];
        cur = a[(unsigned)k0 * plane + jrow + m];
    }

    // neighbours of the other colour's cell at plane k (x-parity s); advances
    // the window to k+1
    __device__ __forceinline__ void step(int k, int s, T& xp, T& xm, T& yp, T& ym, T& zp, T& zm)
    {
        const unsigned kp = (unsigned)(k + 1 == (int)n ? 0 : k + 1);
        const unsigned pk = (unsigned)k * plane;
        const T next = a[kp * plane + jrow + m];
        unsigned mo;
        if (s) mo = (m + 1 == h) ? 0u : m + 1;
        else mo = (m == 0) ? h - 1 : m - 1;
        const T xo = a[pk + jrow + mo];
        yp = a[pk + jprow + m];
        ym = a[pk + jmrow + m];
        xp = s ? xo : cur;
        xm = s ? cur : xo;
        zp = next;
        zm = prev;
        prev = cur;
        cur = next;
    }
};

// One colour pass: OP::update(offset of the cell in the full level array,
// xp, xm, yp, ym, zp, zm) with neighbours in the reference's order
// M1u, M1l, M2u, M2l, M3u, M3l.
template <class OP>
__global__ __launch_bounds__(256) void k_colour_pass(OP op, Geo g, int colour, int TX, int TY, int XB, int JB, int KC,
                                                     unsigned NB)
{
    ZCol c;
    if (!zcol(g, TX, TY, XB, JB, KC, NB, c)) return;
    using T = typename OP::T;
    ZWin<T> w;
    w.init(op.field() + (size_t)(1 - colour) * g.half, g, c.m, c.j, c.k0);
    const unsigned own = (unsigned)colour * (unsigned)g.half;
    for (int k = c.k0; k < c.k1; k++) {
        T xp, xm, yp, ym, zp, zm;
        const int s = (colour + c.j + k) & 1;
        w.step(k, s, xp, xm, yp, ym, zp, zm);
        op.update(own + (unsigned)k * w.plane + w.jrow + w.m, xp, xm, yp, ym, zp, zm);
    }
}

// Both colours per column: OP::cell(offset, centre, xp, xm, yp, ym, zp, zm)
// returns a value whose square is accumulated; the block sum goes to
// partials[logical block] when partials is non-null.
template <class OP>
__global__ __launch_bounds__(256) void k_both_colours(OP op, double* __restrict__ partials, Geo g, int TX, int TY,
                                                      int XB, int JB, int KC, unsigned NB)
{
    ZCol c;
    double acc = 0.0;
    if (zcol(g, TX, TY, XB, JB, KC, NB, c)) {
        using T = typename OP::T;
        ZWin<T> w0, w1;  // windows over the red (0) and black (1) halves
        w0.init(op.field(), g, c.m, c.j, c.k0);
        w1.init(op.field() + g.half, g, c.m, c.j, c.k0);
        const unsigned half = (unsigned)g.half;
        for (int k = c.k0; k < c.k1; k++) {
            const T c0 = w0.cur, c1 = w1.cur;
            const unsigned off = (unsigned)k * w0.plane + w0.jrow + w0.m;
            T xp, xm, yp, ym, zp, zm;
            // red cell: neighbours in the black half, x-parity (j+k)&1
            w1.step(k, (c.j + k) & 1, xp, xm, yp, ym, zp, zm);
            const double r0 = op.cell(off, c0, xp, xm, yp, ym, zp, zm);
            // black cell: neighbours in the red half, x-parity (1+j+k)&1
            w0.step(k, (1 + c.j + k) & 1, xp, xm, yp, ym, zp, zm);
            const double r1 = op.cell(half + off, c1, xp, xm, yp, ym, zp, zm);
            acc += r0 * r0 + r1 * r1;
        }
    }
    if (partials) store_partial(partials, NB, block_sum(acc));
}

}  // namespace mgs
