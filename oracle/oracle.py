"""CPU oracle for the f(R)/nDGP multigrid solve -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product (libmgsolve) never calls it.

Contents
--------
* ctypes bindings to liboracle.so, the C restatement of the reference's
  multigrid kernels (mgs_oracle.h): every kernel in an FP32-storage variant
  (``prec='f32'``, bit-faithful to the reference) and an FP64 variant
  (``prec='f64'``, the oracle proper).
* ctypes bindings to oracle/_ref/libmgref.so, the reference's own Fortran
  routines compiled by oracle/Makefile (``RefLib``), used to pin the C
  restatement and to write golden vectors.  Present only where the reference
  sources are (the build container); tests that need it skip otherwise.
* numpy restatements of the build-only pieces: the seeded lognormal density
  (Philox-4x32-10 + FFT filter), the fifth-force potential and gradient
  (POTENT_FOR_MG / MOVE, PMP2main.f90:869-935, 565-603) and the power
  spectrum with the reference's NGP |k| binning (POWERfft5,
  PMP2mod_power.f90:209-463).

Arrays are numpy 3-D arrays indexed [k, j, i] (z, y, x) in C order, i.e. the
same flat memory as the reference's FI(M1,M2,M3) arrays.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OrcParams(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("model", C.c_int32), ("fr_n", C.c_int32), ("branch", C.c_int32),
        ("box", C.c_double), ("om", C.c_double), ("oml", C.c_double), ("a", C.c_double),
        ("fr0", C.c_double), ("h0rc", C.c_double), ("polish", C.c_int32), ("pad_", C.c_int32),
    ]


class OrcOpts(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("max_cycles", C.c_int32), ("res_conv", C.c_double),
        ("rel_tol", C.c_double), ("init_guess", C.c_int32), ("normalize", C.c_int32),
        ("dgp_lexi", C.c_int32), ("cycle", C.c_int32),
    ]


class OrcReport(C.Structure):
    _fields_ = [
        ("cycles", C.c_int32), ("converged", C.c_int32), ("res0", C.c_double),
        ("res_final", C.c_double), ("history", C.c_double * 256),
    ]


@dataclass
class Params:
    """Model parameters (meaning as in the reference's Tools module)."""
    N: int
    model: int = 1          # 1 f(R), 2 DGP
    fr_n: int = 1
    branch: int = 1         # DGP: +1 normal, -1 self-accelerating
    box: float = 256.0
    om: float = 0.3
    oml: float = 0.7
    a: float = 1.0
    fr0: float = -1.0e-5
    h0rc: float = 1.0
    polish: int = 0

    def c(self) -> OrcParams:
        return OrcParams(self.N, self.model, self.fr_n, self.branch, self.box, self.om,
                         self.oml, self.a, self.fr0, self.h0rc, self.polish, 0)

    def with_(self, **kw) -> "Params":
        d = dict(self.__dict__)
        d.update(kw)
        return Params(**d)


def f32(x: float) -> float:
    """Round a Python float to IEEE single, as the reference's Real*4 inputs are."""
    return float(np.float32(x))


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER(OrcParams)
        for prec, ct in (("f32", C.c_float), ("f64", C.c_double)):
            ptr = C.POINTER(ct)
            getattr(L, f"orc_fr_smooth_{prec}").argtypes = [P, C.c_int, C.c_int, ptr, ptr, ptr]
            getattr(L, f"orc_fr_residual_{prec}").argtypes = [P, C.c_int, ptr, ptr, ptr, ptr]
            getattr(L, f"orc_fr_residual_{prec}").restype = C.c_double
            getattr(L, f"orc_fr_restrict_residual1_{prec}").argtypes = [P, ptr, ptr, ptr]
            getattr(L, f"orc_restrict8_{prec}").argtypes = [C.c_int, ptr, ptr]
            getattr(L, f"orc_fr_rhs_{prec}").argtypes = [P, C.c_int, ptr, ptr, ptr, ptr]
            getattr(L, f"orc_correct_{prec}").argtypes = [C.c_int, ptr, ptr, ptr, C.c_int]
            getattr(L, f"orc_dgp_smooth_{prec}").argtypes = [P, C.c_int, C.c_int, C.c_int, ptr, ptr, ptr]
            getattr(L, f"orc_dgp_residual_{prec}").argtypes = [P, C.c_int, ptr, ptr, ptr, ptr]
            getattr(L, f"orc_dgp_residual_{prec}").restype = C.c_double
            getattr(L, f"orc_dgp_restrict_residual1_{prec}").argtypes = [P, ptr, ptr, ptr]
            getattr(L, f"orc_dgp_rhs_{prec}").argtypes = [P, C.c_int, ptr, ptr, ptr, ptr]
            getattr(L, f"orc_solve_{prec}").argtypes = [P, ptr, ptr, C.POINTER(OrcOpts), C.POINTER(OrcReport)]
            getattr(L, f"orc_vcycles_{prec}").argtypes = [P, ptr, ptr, C.c_int, C.c_int, C.POINTER(C.c_double)]
            getattr(L, f"orc_cycles_{prec}").argtypes = [P, ptr, ptr, C.c_int, C.c_int, C.c_int,
                                                         C.POINTER(C.c_double)]
        L.orc_levelmax.argtypes = [C.c_int]
        L.orc_levelmin.argtypes = [C.c_int]
        L.orc_fr_ubg.argtypes = [P]
        L.orc_fr_ubg.restype = C.c_double
        L.orc_fr_fbg.argtypes = [P]
        L.orc_fr_fbg.restype = C.c_double
        _lib = L
    return _lib


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


def _ptr(a, prec):
    if a is None:
        return None
    assert a.dtype == _dt(prec) and a.flags.c_contiguous
    ct = C.c_float if prec == "f32" else C.c_double
    return a.ctypes.data_as(C.POINTER(ct))


def levelmax(N):
    return lib().orc_levelmax(N)


def levelmin(N):
    return lib().orc_levelmin(N)


def coarsest_level(N):
    return max(levelmax(N) - levelmin(N), 0)


def fr_fbg(p: Params) -> float:
    return lib().orc_fr_fbg(C.byref(p.c()))


def fr_ubg(p: Params) -> float:
    return lib().orc_fr_ubg(C.byref(p.c()))


# ---- single kernels (arrays are modified in place where the kernel writes) --

def fr_smooth(p, level, red, u, rho, f=None, prec="f64"):
    getattr(lib(), f"orc_fr_smooth_{prec}")(C.byref(p.c()), level, int(red), _ptr(u, prec), _ptr(rho, prec), _ptr(f, prec))


def fr_residual(p, level, u, rho, f=None, res=None, prec="f64"):
    return getattr(lib(), f"orc_fr_residual_{prec}")(C.byref(p.c()), level, _ptr(u, prec), _ptr(rho, prec),
                                                     _ptr(f, prec), _ptr(res, prec))


def fr_restrict_residual1(p, u0, delta, prec="f64"):
    nc = p.N // 2
    out = np.zeros((nc, nc, nc), _dt(prec))
    getattr(lib(), f"orc_fr_restrict_residual1_{prec}")(C.byref(p.c()), _ptr(u0, prec), _ptr(delta, prec), _ptr(out, prec))
    return out


def restrict8(fine, prec="f64"):
    nc = fine.shape[0] // 2
    out = np.zeros((nc, nc, nc), _dt(prec))
    getattr(lib(), f"orc_restrict8_{prec}")(nc, _ptr(fine, prec), _ptr(out, prec))
    return out


def fr_rhs(p, level, ru, rho, rres, prec="f64"):
    out = np.zeros_like(ru)
    getattr(lib(), f"orc_fr_rhs_{prec}")(C.byref(p.c()), level, _ptr(ru, prec), _ptr(rho, prec), _ptr(rres, prec), _ptr(out, prec))
    return out


def correct(uH, ru, ufine, guard, prec="f64"):
    getattr(lib(), f"orc_correct_{prec}")(uH.shape[0], _ptr(uH, prec), _ptr(ru, prec), _ptr(ufine, prec), int(guard))


def dgp_smooth(p, level, red, u, rho, f=None, lexi=False, prec="f64"):
    getattr(lib(), f"orc_dgp_smooth_{prec}")(C.byref(p.c()), level, int(red), int(lexi), _ptr(u, prec), _ptr(rho, prec), _ptr(f, prec))


def dgp_residual(p, level, u, rho, f=None, res=None, prec="f64"):
    return getattr(lib(), f"orc_dgp_residual_{prec}")(C.byref(p.c()), level, _ptr(u, prec), _ptr(rho, prec),
                                                      _ptr(f, prec), _ptr(res, prec))


def dgp_restrict_residual1(p, u0, delta, prec="f64"):
    nc = p.N // 2
    out = np.zeros((nc, nc, nc), _dt(prec))
    getattr(lib(), f"orc_dgp_restrict_residual1_{prec}")(C.byref(p.c()), _ptr(u0, prec), _ptr(delta, prec), _ptr(out, prec))
    return out


def dgp_rhs(p, level, ru, rho, rres, prec="f64"):
    out = np.zeros_like(ru)
    getattr(lib(), f"orc_dgp_rhs_{prec}")(C.byref(p.c()), level, _ptr(ru, prec), _ptr(rho, prec), _ptr(rres, prec), _ptr(out, prec))
    return out


def vcycles(p, u, delta, ncycles, dgp_lexi=False, prec="f64"):
    """ncycles V-cycles in place on u; returns the level-0 RMS residual after each."""
    hist = (C.c_double * max(ncycles, 1))()
    getattr(lib(), f"orc_vcycles_{prec}")(C.byref(p.c()), _ptr(u, prec), _ptr(delta, prec), ncycles, int(dgp_lexi), hist)
    return np.array(hist[:ncycles])


def cycles(p, u, delta, ncycles, cycle=0, dgp_lexi=False, prec="f64"):
    """ncycles V (0), F (1) or W (2) cycles in place on u; returns the level-0
    RMS residual after each."""
    hist = (C.c_double * max(ncycles, 1))()
    getattr(lib(), f"orc_cycles_{prec}")(C.byref(p.c()), _ptr(u, prec), _ptr(delta, prec), ncycles, int(cycle),
                                         int(dgp_lexi), hist)
    return np.array(hist[:ncycles])


def solve(p, delta, u=None, mode=1, rel_tol=1e-8, res_conv=1e-8, max_cycles=60, normalize=False,
          dgp_lexi=False, prec="f64", cycle=0):
    """scalar_field-style solve.  mode 0: the reference's exits; mode 1: relative residual.
    Returns (u, report dict)."""
    dt = _dt(prec)
    delta = np.ascontiguousarray(delta, dtype=dt)
    init = u is None
    u = np.zeros_like(delta) if init else np.ascontiguousarray(u, dtype=dt).copy()
    o = OrcOpts(mode, max_cycles, res_conv, rel_tol, 1 if init else 0, 1 if normalize else 0, int(dgp_lexi),
                int(cycle))
    r = OrcReport()
    getattr(lib(), f"orc_solve_{prec}")(C.byref(p.c()), _ptr(u, prec), _ptr(delta, prec), C.byref(o), C.byref(r))
    rep = dict(cycles=r.cycles, converged=bool(r.converged), res0=r.res0, res_final=r.res_final,
               history=np.array(r.history[:r.cycles]))
    return u, rep


# ---------------------------------------------------------------------------
# The reference's own routines (oracle/_ref/libmgref.so)
# ---------------------------------------------------------------------------

def fi3_sections(N, level):
    """(x0, y0, z0) offsets, 0-based, of the six level-l sections of the
    reference's packed FI3 scratch array (PMP2MGsolver_fR.f90:76-83,
    PMP2extradof.f90:1191-1196): keys rres, rho, ru, u, f, res."""
    n = N >> level
    koff = n * (2 ** level - 2)
    koff2 = n * (2 ** level - 1)
    return {
        "rres": (0, 0, koff), "rho": (n, 0, koff), "ru": (0, n, koff),
        "u": (n, n, koff), "f": (0, n, koff2), "res": (n, 0, koff2),
    }


class RefLib:
    """The reference's multigrid routines on the Tools module state."""

    def __init__(self, omp=False):
        name = "libmgref_omp.so" if omp else "libmgref.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        L.ref_init.argtypes = [C.c_int] * 5 + [C.c_double] * 7
        fp = C.POINTER(C.c_float)
        L.ref_put.argtypes = [C.c_int, fp]
        L.ref_get.argtypes = [C.c_int, fp]
        for nm in ("ref_smooth",):
            getattr(L, nm).argtypes = [C.c_int, C.c_int]
        for nm in ("ref_restrict_residual", "ref_rhs", "ref_restrict_density", "ref_restrict_field",
                   "ref_correct", "ref_scalar_field"):
            getattr(L, nm).argtypes = [C.c_int]
        L.ref_residual.argtypes = [C.c_int]
        L.ref_residual.restype = C.c_double
        L.ref_vcycles.argtypes = [C.c_int, C.POINTER(C.c_double)]
        L.ref_cycles.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double)]
        self.L = L
        self.N = None

    def init(self, p: Params, iter_max=4, res_conv=1e-8):
        self.N = p.N
        self.L.ref_init(p.N, p.model, p.fr_n, p.branch, iter_max, p.box, p.om, p.oml, p.a, p.fr0, p.h0rc, res_conv)

    def put(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        assert a.size == self.N ** 3
        self.L.ref_put(which, a.ctypes.data_as(C.POINTER(C.c_float)))

    def get(self, which):
        a = np.empty((self.N,) * 3, np.float32)
        self.L.ref_get(which, a.ctypes.data_as(C.POINTER(C.c_float)))
        return a

    # FI3 level sections
    def get_sec(self, level, name):
        n = self.N >> level
        x0, y0, z0 = fi3_sections(self.N, level)[name]
        return self.get(3)[z0:z0 + n, y0:y0 + n, x0:x0 + n].copy()

    def put_sec(self, level, name, arr):
        n = self.N >> level
        x0, y0, z0 = fi3_sections(self.N, level)[name]
        fi3 = self.get(3)
        fi3[z0:z0 + n, y0:y0 + n, x0:x0 + n] = arr
        self.put(3, fi3)

    def smooth(self, level, red):
        self.L.ref_smooth(level, int(red))

    def residual(self, level):
        return self.L.ref_residual(level)

    def restrict_residual(self, level):
        self.L.ref_restrict_residual(level)

    def rhs(self, level):
        self.L.ref_rhs(level)

    def restrict_density(self, level):
        self.L.ref_restrict_density(level)

    def restrict_field(self, level):
        self.L.ref_restrict_field(level)

    def correct(self, level):
        self.L.ref_correct(level)

    def vcycles(self, n):
        h = (C.c_double * n)()
        self.L.ref_vcycles(n, h)
        return np.array(h[:])

    def cycles(self, n, cycle):
        """n F_cycle(0) (cycle 1) or W_cycle(0) (cycle 2) calls, residual after each."""
        h = (C.c_double * n)()
        self.L.ref_cycles(n, cycle, h)
        return np.array(h[:])

    def scalar_field(self, istep=1):
        self.L.ref_scalar_field(istep)


# ---------------------------------------------------------------------------
# Seeded lognormal density (build-only component; no reference counterpart)
# ---------------------------------------------------------------------------

_PHILOX_M0 = np.uint64(0xD2511F53)
_PHILOX_M1 = np.uint64(0xCD9E8D57)
_PHILOX_W0 = np.uint32(0x9E3779B9)
_PHILOX_W1 = np.uint32(0xBB67AE85)


class RefPmLib:
    """The reference's PM-step routines (DENSIT, POTENTfft5, POTENT_FOR_MG,
    MOVE, POWERfft5) on the Tools module state, through
    oracle/ref_pm_harness.f90 (oracle/_ref/libmgref_pm.so).  Golden-vector
    generation only."""

    def __init__(self):
        path = os.path.join(HERE, "_ref", "libmgref_pm.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        L.ref_pm_init.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64] + [C.c_double] * 7
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        L.ref_pm_put_particles.argtypes = [fp] * 6
        L.ref_pm_get_particles.argtypes = [fp] * 6
        L.ref_pm_put.argtypes = [C.c_int, fp]
        L.ref_pm_get.argtypes = [C.c_int, fp]
        L.ref_pm_move.argtypes = [C.c_double]
        L.ref_pm_move.restype = C.c_double
        L.ref_pm_power.argtypes = [C.c_int, C.c_int, C.c_int] + [dp] * 6
        self.L = L
        self.N = None
        self.np = 0

    def init(self, N, model=1, fr_n=1, npart=1, box=64.0, om=0.3, oml=0.7, a=1.0, astep=0.01, fr0=-1e-5,
             h0rc=1.0):
        self.N, self.np = N, int(npart)
        self.L.ref_pm_init(N, model, fr_n, int(npart), box, om, oml, a, astep, fr0, h0rc)

    def put_particles(self, *arrs):
        bufs = [np.ascontiguousarray(a, np.float32) for a in arrs]
        assert all(b.size == self.np for b in bufs) and len(bufs) == 6
        self.L.ref_pm_put_particles(*[b.ctypes.data_as(C.POINTER(C.c_float)) for b in bufs])

    def get_particles(self):
        bufs = [np.empty(self.np, np.float32) for _ in range(6)]
        self.L.ref_pm_get_particles(*[b.ctypes.data_as(C.POINTER(C.c_float)) for b in bufs])
        return bufs

    def put(self, which, arr):
        a = np.ascontiguousarray(arr, np.float32)
        assert a.size == self.N ** 3
        self.L.ref_pm_put(which, a.ctypes.data_as(C.POINTER(C.c_float)))

    def get(self, which):
        a = np.empty((self.N,) * 3, np.float32)
        self.L.ref_pm_get(which, a.ctypes.data_as(C.POINTER(C.c_float)))
        return a

    def densit(self):
        self.L.ref_pm_densit()

    def potentfft5(self):
        self.L.ref_pm_potentfft5()

    def potent_for_mg(self):
        self.L.ref_pm_potent_for_mg()

    def move(self, factv=1.0):
        return self.L.ref_pm_move(factv)

    def power(self, ipol=0, iswitch=1, icic=0):
        h = self.N // 2
        out = [np.zeros(h) for _ in range(6)]
        self.L.ref_pm_power(ipol, iswitch, icic, *[o.ctypes.data_as(C.POINTER(C.c_double)) for o in out])
        return out


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) on uint32 arrays."""
    c0 = c0.astype(np.uint32); c1 = c1.astype(np.uint32)
    c2 = c2.astype(np.uint32); c3 = c3.astype(np.uint32)
    k0 = np.uint32(k0); k1 = np.uint32(k1)
    for r in range(rounds):
        p0 = c0.astype(np.uint64) * _PHILOX_M0
        p1 = c2.astype(np.uint64) * _PHILOX_M1
        hi0 = (p0 >> np.uint64(32)).astype(np.uint32); lo0 = (p0 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        hi1 = (p1 >> np.uint64(32)).astype(np.uint32); lo1 = (p1 & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        if r < rounds - 1:
            k0 = np.uint32((int(k0) + int(_PHILOX_W0)) & 0xFFFFFFFF)
            k1 = np.uint32((int(k1) + int(_PHILOX_W1)) & 0xFFFFFFFF)
    return c0, c1, c2, c3


def white_noise(N, seed):
    """One standard normal per cell: Philox counter = (cell index lo, hi, 0, 0),
    key = (seed lo, seed hi); Box-Muller on two 53-bit uniforms, cosine branch."""
    idx = np.arange(N ** 3, dtype=np.uint64)
    lo = (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    hi = (idx >> np.uint64(32)).astype(np.uint32)
    z = np.zeros_like(lo)
    x0, x1, x2, x3 = philox4x32(lo, hi, z, z, seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    u1 = ((x0 >> np.uint32(5)).astype(np.float64) * 67108864.0 + (x1 >> np.uint32(6)).astype(np.float64) + 1.0) / 9007199254740992.0
    u2 = ((x2 >> np.uint32(5)).astype(np.float64) * 67108864.0 + (x3 >> np.uint32(6)).astype(np.float64)) / 9007199254740992.0
    g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
    return g.reshape(N, N, N)


def kgrid(N):
    m = np.fft.fftfreq(N, d=1.0 / N)  # integer wavenumbers, -N/2..N/2-1
    return m


def lognormal_density(N, box, sigma_g, seed, slope=-1.5, kcut_frac=0.5):
    """delta = exp(g - sigma_g^2/2) - 1, g Gaussian with P_g(k) ~ k^slope
    exp(-k^2/(2 k_c^2)), k_c = kcut_frac * k_Nyq, k_Nyq = pi N / box,
    rescaled to real-space standard deviation sigma_g."""
    w = white_noise(N, seed)
    wk = np.fft.fftn(w)
    m = kgrid(N)
    kf = 2.0 * math.pi / box
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    k = kf * np.sqrt(kx * kx + ky * ky + kz * kz)
    kc = kcut_frac * math.pi * N / box
    with np.errstate(divide="ignore"):
        amp = np.where(k > 0, k ** (0.5 * slope) * np.exp(-k * k / (4.0 * kc * kc)), 0.0)
    g = np.fft.ifftn(wk * amp).real
    g -= g.mean()
    g *= sigma_g / math.sqrt(np.mean(g * g))
    return np.exp(g - 0.5 * sigma_g ** 2) - 1.0


# ---------------------------------------------------------------------------
# Fifth force and power spectrum
# ---------------------------------------------------------------------------

def fifth_force_potential(p: Params, u_norm):
    """psi added to the Newtonian potential by POTENT_FOR_MG
    (PMP2main.f90:884-950): f(R) psi = -1/2 c~^2 fR_bg u^(n+1) with the
    normalised u of scalar_field's exit; DGP psi = phi/2."""
    if p.model == 1:
        ctilde = 2.99792458e3 * p.N / p.box
        fct1 = ctilde ** 2 * 0.5 * fr_fbg(p)
        return -fct1 * np.asarray(u_norm, np.float64) ** (p.fr_n + 1)
    return 0.5 * np.asarray(u_norm, np.float64)


def central_gradient(psi):
    """(d/dx, d/dy, d/dz) psi = (psi[i+1] - psi[i-1]) / 2 in grid units,
    periodic: the on-grid form of MOVE's 2x-gradient (PMP2main.f90:565-594)."""
    gx = 0.5 * (np.roll(psi, -1, axis=2) - np.roll(psi, 1, axis=2))
    gy = 0.5 * (np.roll(psi, -1, axis=1) - np.roll(psi, 1, axis=1))
    gz = 0.5 * (np.roll(psi, -1, axis=0) - np.roll(psi, 1, axis=0))
    return gx, gy, gz


# dLog of POWERfft5's log bins, a REAL*4 parameter (PMP2mod_power.f90:12)
DLOG32 = np.float32(0.02)
_LIBM = None


def log_bins(n2max):
    """MPlog for every integer |m|^2 = 0 .. n2max (PMP2mod_power.f90:396):
    INT(LOG10(RK)/dLog + 1000.) - 1000 clamped below at 1, evaluated as the
    compiled reference does, in REAL*4 with the C library's log10f (RK =
    SQRT(REAL(MI+MJ+MK)), a REAL*4 square root); entry 0 is unused."""
    global _LIBM
    if _LIBM is None:
        import ctypes.util

        _LIBM = C.CDLL(ctypes.util.find_library("m"))
        _LIBM.log10f.restype = C.c_float
        _LIBM.log10f.argtypes = [C.c_float]
    rk = np.sqrt(np.arange(n2max + 1, dtype=np.float32))
    out = np.ones(n2max + 1, np.int64)
    for n2 in range(1, n2max + 1):
        v = np.float32(np.float32(_LIBM.log10f(float(rk[n2]))) / DLOG32) + np.float32(1000.0)
        out[n2] = max(1, int(v) - 1000)
    return out


def power_spectrum_ex(field, box, ipol=0, icic=0):
    """POWERfft5(iPol, .) (PMP2mod_power.f90:209-463) over the complex
    spectrum c_m = (1/N^3) sum f e^{-i k.x} of the integer wavevectors m,
    k = 0 excluded, with the reference's REAL*4 |m| (RK = sqrt(REAL(|m|^2))):

    * icic = 0 (:394-410): bin MP = max(INT(RK + 0.5), 1); log bin MPlog
      (log_bins), whose power is summed only for ipol = 0;
    * icic = 1 (:373-393): MP = max(INT(RK + 0.0001), 1), MP2 = MP + 1, weights
      w1 = MP - RK + 1 and w2 = 1 - w1 on the power, the mode count and the
      bin's k (MP w1, (MP + 1) w2); no log bins;
    * ipol = 1: the power is weighted by P2 = 1.5 mz^2/RK^2 - 0.5 (z the line
      of sight: MI of :371 is the first index after the transposition) and
      multiplied by 2.5 at the end (:420).

    Bins 1..N/2 (:440-456): k = <RK> 2 pi/box, P = box^3 sum(power)/count,
    zero where the count is not positive.  Returns (k, P, count, klog, Plog,
    countlog)."""
    f = np.asarray(field, np.float64)
    N = f.shape[0]
    ck = np.fft.fftn(f) / N ** 3
    pw = (ck.real ** 2 + ck.imag ** 2).ravel()
    m = kgrid(N).astype(np.int64)
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    n2 = (kx * kx + ky * ky + kz * kz).ravel()
    # the first index of FI when the bins are filled is z: the third
    # transform runs after the I <-> K transposition (:307-316)
    mz2 = (kz * kz).ravel()
    sel = n2 > 0
    n2, mz2, pw = n2[sel], mz2[sel], pw[sel]
    rk = np.sqrt(n2.astype(np.float32))
    if ipol:
        cosmu2 = mz2.astype(np.float32) / (rk * rk)
        amp = pw * (np.float32(1.5) * cosmu2 - np.float32(0.5)).astype(np.float64)
    else:
        amp = pw
    nb = N + 2
    if icic:
        mp = np.maximum((rk + np.float32(0.0001)).astype(np.int64), 1)
        mp2 = np.minimum(mp + 1, N)
        w1 = (mp.astype(np.float32) - rk + np.float32(1.0)).astype(np.float32)
        w2 = (np.float32(1.0) - w1).astype(np.float64)
        w1 = w1.astype(np.float64)
        d1 = mp.astype(np.float64)
        cnt = np.bincount(mp, weights=w1, minlength=nb) + np.bincount(mp2, weights=w2, minlength=nb)
        psum = np.bincount(mp, weights=amp * w1, minlength=nb) + np.bincount(mp2, weights=amp * w2, minlength=nb)
        ksum = np.bincount(mp, weights=d1 * w1, minlength=nb) + np.bincount(mp2, weights=(d1 + 1.0) * w2, minlength=nb)
        lcnt = lpsum = lksum = np.zeros(nb)
    else:
        mp = np.maximum((rk + np.float32(0.5)).astype(np.int64), 1)
        cnt = np.bincount(mp, minlength=nb).astype(np.float64)
        psum = np.bincount(mp, weights=amp, minlength=nb)
        ksum = np.bincount(mp, weights=rk.astype(np.float64), minlength=nb)
        lb = np.minimum(log_bins(int(n2.max()))[n2], N)
        lcnt = np.bincount(lb, minlength=nb).astype(np.float64)
        lpsum = np.bincount(lb, weights=pw, minlength=nb) if not ipol else np.zeros(nb)
        lksum = np.bincount(lb, weights=rk.astype(np.float64), minlength=nb)
    if ipol:
        psum = 2.5 * psum
    bins = np.arange(1, N // 2 + 1)

    def finish(c, ps, ks):
        c = c[bins]
        with np.errstate(invalid="ignore", divide="ignore"):
            P = np.where(c > 0, box ** 3 * ps[bins] / c, 0.0)
            k = np.where(c > 0, ks[bins] / c * 2.0 * math.pi / box, 0.0)
        return k, P, np.where(c > 0, c, 0.0)

    return finish(cnt, psum, ksum) + finish(lcnt, lpsum, lksum)


def power_spectrum(field, box):
    """The NGP monopole of POWERfft5 (power_spectrum_ex with ipol = icic = 0):
    bin MP = max(INT(RK + 0.5), 1), P = box^3 <|c_m|^2>, k_bin = <RK> 2 pi/box.
    Returns (k, P, count) for bins 1..N/2."""
    return power_spectrum_ex(field, box)[:3]
