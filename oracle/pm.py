"""CPU restatement (numpy) of the reference's PM-step routines around the
scalar-field solve: the CIC density deposit, the FFT Poisson solve, the
fifth-force potential addition and the particle push.

Test infrastructure only: imported by tests/ and bench.py's checks, never by
the product.  Arrays are indexed [k, j, i] (x fastest in memory, the
reference's FI order); particle positions are REAL*4 in grid units in
[1, N+1) with cell I = INT(X), as in the reference.

Parity pins:
* golden vectors from the reference's own DENSIT, POTENTfft5, POTENT_FOR_MG
  and MOVE, compiled from its sources by oracle/Makefile (libmgref_pm.so)
  and written by tests/golden/make_golden.py (tests/test_oracle_pm_golden.py):
  MOVE and POTENT_FOR_MG bit for bit, the deposit and the Poisson solve to
  the REAL*4 rounding the reference's FI storage adds;
* invariants (tests/test_oracle_pm.py): mass conservation and exact node /
  lattice / wrap cases of the deposit, the single-Fourier-mode closed form
  and the discrete-Laplacian inverse of the Poisson solve, exact node
  gradients of a sinusoidal potential, the wrap and the REAL*4 edge rule of
  the push.
"""
import numpy as np

import oracle as O


def wpar(N, npart):
    """WPAR = YN**3/FLOAT(Nparticles) (PMP2mod_density.f90:40,
    PMP2main.f90:470): the count goes through default REAL, the quotient is
    formed in REAL*8 (YN is REAL*8) and stored in WPAR, an implicitly typed
    default REAL."""
    return float(np.float32(float(N) ** 3 / float(np.float32(npart))))


def _cells(X, N):
    X = np.asarray(X, np.float64)
    if not (np.all(np.isfinite(X)) and np.all(X >= 1.0) and np.all(X < N + 1.0)):
        raise ValueError("particle outside [1, N+1)")
    I = X.astype(np.int64)  # INT(X), X >= 1
    return I, X - I.astype(np.float64)


def cic_deposit(x, y, z, N):
    """DENSIT's CIC branch, PMP2mod_density.f90:235-281: FI = -1, then every
    particle adds WPAR times its CIC weights to its 8 cells (FP64 sums)."""
    I, D1 = _cells(np.asarray(x, np.float32), N)
    J, D2 = _cells(np.asarray(y, np.float32), N)
    K, D3 = _cells(np.asarray(z, np.float32), N)
    w = wpar(N, len(I))
    T1, T2, T3 = 1.0 - D1, 1.0 - D2, 1.0 - D3
    T2W, D2W = T2 * w, D2 * w
    i0, j0, k0 = I - 1, J - 1, K - 1
    i1, j1, k1 = np.where(I == N, 0, I), np.where(J == N, 0, J), np.where(K == N, 0, K)
    FI = -np.ones((N, N, N))
    for (kk, jj, ii, v) in ((k0, j0, i0, T3 * T1 * T2W), (k0, j0, i1, T3 * D1 * T2W), (k0, j1, i0, T3 * T1 * D2W),
                            (k0, j1, i1, T3 * D1 * D2W), (k1, j0, i0, D3 * T1 * T2W), (k1, j0, i1, D3 * D1 * T2W),
                            (k1, j1, i0, D3 * T1 * D2W), (k1, j1, i1, D3 * D1 * D2W)):
        np.add.at(FI, (kk, jj, ii), v)
    return FI


# PMP2main.f90:711: real*8, parameter :: P16 = 6.28318530718 -- a default
# REAL literal, so the parameter holds its REAL*4 rounding
P16 = float(np.float32(6.28318530718))


def green_table(N):
    """POTENTfft5's GREENf (PMP2main.f90:736-742) indexed by the complex
    wavenumber k = 0 .. N-1: FFTPACK5's half-complex wavenumber is
    i = min(k, N - k), G(0) = 2, G(N/2) = -2, G(i) = 2 cos(P16 i / N) between
    (even N)."""
    k = np.arange(N)
    i = np.minimum(k, N - k)
    G = 2.0 * np.cos(P16 * i / N)
    G[0] = 2.0
    G[N // 2] = -2.0
    return G


def green_eigen(N):
    """POTENTfft5's discrete-Laplacian eigenvalues on the half spectrum
    [kz, ky, kx] (PMP2main.f90:736-742, 805-816): A = (G(ky) - 6 + G(kx)) +
    G(kz) with the reference's GREENf table (green_table) and A = 1 where
    |A| < 1e-7 (the mean mode)."""
    G = green_table(N)
    Gh = G[: N // 2 + 1]
    A = ((G[None, :, None] - 6.0) + Gh[None, None, :]) + G[:, None, None]
    return np.where(np.abs(A) < 1.0e-7, 1.0, A)


def poisson_fft(delta, trfi):
    """POTENTfft5, PMP2main.f90:705-867: phi with hat(phi) = trfi hat(delta)
    / A(k) (A = 1 for the mean mode); FFTPACK5's normalised forward and
    unnormalised backward transforms compose to numpy's rfftn/irfftn.  FP64
    result (the REAL*4 FI it is stored in: potential_fi)."""
    d = np.asarray(delta, np.float64)
    N = d.shape[0]
    F = np.fft.rfftn(d)
    return np.fft.irfftn(F * trfi / green_eigen(N), s=d.shape, axes=(0, 1, 2))


def potential_fi(delta, trfi, psi=None):
    """The REAL*4 potential FI after POTENTfft5 and, with psi, after
    POTENT_FOR_MG's FI = FI + sf (PMP2main.f90:884-935: the sum in FP64,
    rounded once)."""
    fi = poisson_fft(delta, trfi).astype(np.float32)
    if psi is not None:
        fi = (fi.astype(np.float64) + np.asarray(psi, np.float64)).astype(np.float32)
    return fi


def trfi_of(om, aexpn):
    """POTENTfft5's trfi = 1.5*Om/aexpn (PMP2main.f90:727-731, standard LCDM
    branch): implicitly typed default REAL, evaluated in REAL*4"""
    return float(np.float32(np.float32(1.5) * np.float32(om)) / np.float32(aexpn))


def potent_for_mg(p, fi, fi2):
    """POTENT_FOR_MG, PMP2main.f90:869-935 (f(R) and nDGP branches): FI = FI
    + sf with sf = -fct1 FI2**(n+1) (f(R), fct1 = c~^2 fR_bg / 2) or
    0.5 FI2 (nDGP); FI2 is REAL*4, so FI2**2 and FI2**3 are REAL*4 products,
    and the sum is formed in REAL*8 and stored in the REAL*4 FI."""
    fi2 = np.asarray(fi2, np.float32)
    if p.model == 1:
        ctilde = 2.99792458e3 * float(p.N) / float(np.float32(p.box))
        fct1 = ctilde ** 2 * 0.5 * O.fr_fbg(p)
        w = fi2 if p.fr_n == 0 else (fi2 * fi2 if p.fr_n == 1 else (fi2 * fi2) * fi2)
        sf = -fct1 * w.astype(np.float64)
    else:
        sf = 0.5 * fi2.astype(np.float64)
    return (np.asarray(fi, np.float32).astype(np.float64) + sf).astype(np.float32)


def laplacian7(phi):
    """7-point discrete Laplacian in grid units, periodic"""
    out = -6.0 * phi
    for ax in range(3):
        out = out + np.roll(phi, 1, axis=ax) + np.roll(phi, -1, axis=ax)
    return out


def move(phi, x, y, z, vx, vy, vz, pconst, xconst):
    """MOVE's particle loop for the f(R) and nDGP models, PMP2main.f90:
    488-684 (no friction or varying-mass terms, Linear_nDGP false): REAL*8
    arithmetic on REAL*4 particles, node gradients of the REAL*4 potential
    phi [k, j, i] differenced in REAL*4.
    Returns (x, y, z, vx, vy, vz) as float32 and SVEL."""
    N = np.asarray(phi).shape[0]
    X = np.asarray(x, np.float32).astype(np.float64)
    Y = np.asarray(y, np.float32).astype(np.float64)
    Z = np.asarray(z, np.float32).astype(np.float64)
    VVX = np.asarray(vx, np.float32).astype(np.float64)
    VVY = np.asarray(vy, np.float32).astype(np.float64)
    VVZ = np.asarray(vz, np.float32).astype(np.float64)
    I, D1 = _cells(X, N)
    J, D2 = _cells(Y, N)
    K, D3 = _cells(Z, N)
    T1, T2, T3 = 1.0 - D1, 1.0 - D2, 1.0 - D3
    T2T1, T2D1, D2T1, D2D1 = T2 * T1, T2 * D1, D2 * T1, D2 * D1

    def idx(C):  # 0-based indices of C-1, C, C+1, C+2 (1-based C, periodic)
        return (C - 2) % N, (C - 1) % N, C % N, (C + 1) % N

    x0, x1, x2, x3 = idx(I)
    y0, y1, y2, y3 = idx(J)
    z0, z1, z2, z3 = idx(K)

    phi32 = np.asarray(phi, np.float32)

    def F(a, b, c):
        return phi32[c, b, a]

    F111, F211, F121, F221 = F(x1, y1, z1), F(x2, y1, z1), F(x1, y2, z1), F(x2, y2, z1)
    F112, F212, F122, F222 = F(x1, y1, z2), F(x2, y1, z2), F(x1, y2, z2), F(x2, y2, z2)
    F113, F213, F123, F223 = F(x1, y1, z3), F(x2, y1, z3), F(x1, y2, z3), F(x2, y2, z3)
    F110, F210, F120, F220 = F(x1, y1, z0), F(x2, y1, z0), F(x1, y2, z0), F(x2, y2, z0)
    F311, F321, F131, F231 = F(x3, y1, z1), F(x3, y2, z1), F(x1, y3, z1), F(x2, y3, z1)
    F312, F322, F132, F232 = F(x3, y1, z2), F(x3, y2, z2), F(x1, y3, z2), F(x2, y3, z2)
    F011, F021, F101, F201 = F(x0, y1, z1), F(x0, y2, z1), F(x1, y0, z1), F(x2, y0, z1)
    F012, F022, F102, F202 = F(x0, y1, z2), F(x0, y2, z2), F(x1, y0, z2), F(x2, y0, z2)
    # the doubled node gradients GX111 = F211 - F011, ... (:580-602): the
    # F's are implicitly typed default REAL, so each difference is rounded
    # to REAL*4 before it meets the REAL*8 weights
    def G(a, b):
        return (a - b).astype(np.float64)

    GX = pconst * (T3 * (T2T1 * G(F211, F011) + T2D1 * G(F311, F111) + D2T1 * G(F221, F021) + D2D1 * G(F321, F121)) +
                   D3 * (T2T1 * G(F212, F012) + T2D1 * G(F312, F112) + D2T1 * G(F222, F022) + D2D1 * G(F322, F122)))
    GY = pconst * (T3 * (T2T1 * G(F121, F101) + T2D1 * G(F221, F201) + D2T1 * G(F131, F111) + D2D1 * G(F231, F211)) +
                   D3 * (T2T1 * G(F122, F102) + T2D1 * G(F222, F202) + D2T1 * G(F132, F112) + D2D1 * G(F232, F212)))
    GZ = pconst * (T3 * (T2T1 * G(F112, F110) + T2D1 * G(F212, F210) + D2T1 * G(F122, F120) + D2D1 * G(F222, F220)) +
                   D3 * (T2T1 * G(F113, F111) + T2D1 * G(F213, F211) + D2T1 * G(F123, F121) + D2D1 * G(F223, F221)))
    VVX, VVY, VVZ = VVX + GX, VVY + GY, VVZ + GZ
    XN = float(np.float32(np.float32(N) + np.float32(1.0) - np.float32(1.0e-8)))
    YN = float(N)
    out = []
    for P, V in ((X, VVX), (Y, VVY), (Z, VVZ)):
        P = P + V * xconst
        P = np.where(P < 1.0, P + YN, P)
        P = np.where(P >= XN, P - YN, P)
        Pf = P.astype(np.float32)
        Pf = np.where(Pf.astype(np.int64) == N + 1, Pf - np.float32(1.0e-3), Pf).astype(np.float32)
        out.append(Pf)
    svel = float(np.sum((VVX ** 2 + VVY ** 2 + VVZ ** 2) * wpar(N, len(X))))
    return out[0], out[1], out[2], VVX.astype(np.float32), VVY.astype(np.float32), VVZ.astype(np.float32), svel


def pconst_xconst(aexpn, astep, om, oml, factv=1.0):
    """MOVE's PCONST and XCONST for the LCDM background (PMP2main.f90:333-335):
    every operand is default REAL (AEXPN, ASTEP, Om, OmL, FactV and the
    implicitly typed Ahalf), so both are evaluated in REAL*4, left to right,
    x**3 as (x*x)*x, and only stored in REAL*8."""
    f = np.float32
    a, st, om, oml, fv = f(aexpn), f(astep), f(om), f(oml), f(factv)
    a3 = (a * a) * a
    pconst = ((-np.sqrt(a / (om + oml * a3))) * st) * f(0.5) / fv
    ahalf = a + st / f(2.0)
    h3 = (ahalf * ahalf) * ahalf
    xconst = st / np.sqrt(ahalf * (om + oml * h3)) / ahalf
    return float(f(pconst)), float(f(xconst))
