"""Checks of the FP64 oracle that do not rely on the reference's outputs.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, coefficient or index, a transposed operand) fails it:
  * residual of a sinusoid in u^{n+1} against the discrete-Laplacian
    eigenvalue 2(cos(2 pi k/N) - 1) (pins c~^2 s/dx^2, the density and the
    background terms, the neighbour set);
  * after a colour pass of the smoother, the residual of every updated cell
    vanishes (the closed-form smoother and the residual are independent code;
    pins P, Q, fct1, fct2 and the FAS term at level 1);
  * the fused level-1 restriction equals the average of the 8 children's
    level-0 residuals computed by the residual routine (brute force);
  * restriction / prolongation against an independent numpy trilinear
    interpolation, constants and linear fields;
  * the whole solve in the linear regime against an FFT solve of the
    linearised equation with the same discrete operator;
  * DGP: phi = 0 for delta = 0, and the linear limit = FFT Poisson solve;
  * Philox-4x32-10 known-answer vectors, the power-spectrum estimator on a
    single mode, the central gradient of a sinusoid.
"""
import math

import numpy as np
import pytest

import oracle as O


def P(N=16, **kw):
    d = dict(N=N, model=1, fr_n=1, box=64.0, om=0.3, oml=0.7, a=0.7, fr0=-1e-5)
    d.update(kw)
    return O.Params(**d)


def consts(p):
    """Model constants from their definitions: c~ = c N / Box
    (PMP2MGsolver_fR.f90:54), the background curvatures R and R0, f_R of
    the background and u_bg = |f_R|^(1/(n+1)) (PMP2MGsolver_fR.f90:59-65)."""
    ctilde = 2.99792458e3 * p.N / p.box
    Rbg = 3.0 * (p.om / p.a ** 3 + 4.0 * p.oml)
    R0 = 3.0 * (p.om + 4.0 * p.oml)
    fbg = p.fr0 * (R0 / Rbg) ** (p.fr_n + 1)
    ubg = abs(fbg) ** (1.0 / (p.fr_n + 1))
    return ctilde, Rbg, fbg, ubg


@pytest.mark.parametrize("n", [0, 1, 2])
def test_residual_of_a_sinusoid(n):
    p = P(fr_n=n)
    N = p.N
    ctilde, Rbg, fbg, ubg = consts(p)
    k = 3
    x = np.arange(N)
    w = ubg ** (n + 1) * (1.0 + 0.2 * np.sin(2 * np.pi * k * x / N))[None, None, :] * np.ones((N, N, N))
    u = w ** (1.0 / (n + 1))
    rng = np.random.default_rng(3)
    delta = rng.standard_normal((N, N, N))
    res = np.zeros_like(u)
    O.fr_residual(p, 0, u, delta, None, res, "f64")
    lam = 2.0 * (np.cos(2 * np.pi * k / N) - 1.0)
    lap_w = lam * (w - ubg ** (n + 1))
    expect = ctilde ** 2 * np.sign(fbg) * lap_w + (p.om / p.a) * delta - (Rbg * p.a ** 2 / 3.0) * (ubg / u - 1.0)
    scale = ctilde ** 2 * ubg ** (n + 1)
    assert np.max(np.abs(res - expect)) <= 1e-11 * scale


@pytest.mark.parametrize("n", [0, 1, 2])
@pytest.mark.parametrize("level", [0, 1])
def test_smoother_zeroes_the_local_residual(n, level):
    p = P(N=32, fr_n=n)
    ctilde, Rbg, fbg, ubg = consts(p)
    nl = p.N >> level
    rng = np.random.default_rng(10 + n)
    rho = rng.lognormal(0, 1, (nl,) * 3) - 1.0
    u = ubg * (1.0 + 0.3 * rng.random((nl,) * 3))
    f = 0.01 * rng.standard_normal((nl,) * 3) if level else None
    for red in (1, 0):
        O.fr_smooth(p, level, red, u, rho, f, "f64")
        res = np.zeros_like(u)
        O.fr_residual(p, level, u, rho, f, res, "f64")
        i, j, k = np.meshgrid(np.arange(nl), np.arange(nl), np.arange(nl), indexing="ij")
        colour = ((i + j + k) % 2 == 0) if red else ((i + j + k) % 2 == 1)
        scale = ctilde ** 2 * ubg ** (n + 1) / 4 ** level
        # cells of the colour just updated satisfy their local equation
        assert np.max(np.abs(res[colour])) <= 1e-9 * scale
        # the other colour does not (the check is not vacuous)
        assert np.max(np.abs(res[~colour])) > 1e-6 * scale


@pytest.mark.parametrize("model", [1, 2])
def test_fused_restriction_is_mean_of_child_residuals(model):
    p = P(N=16, model=model, a=1.0)
    N = p.N
    rng = np.random.default_rng(5)
    delta = rng.lognormal(0, 0.8, (N,) * 3) - 1.2
    if model == 1:
        _, _, _, ubg = consts(p)
        u = ubg * (1.0 + 0.3 * rng.random((N,) * 3))
        res = np.zeros_like(u)
        O.fr_residual(p, 0, u, delta, None, res, "f64")
        fused = O.fr_restrict_residual1(p, u, delta, "f64")
    else:
        u = 0.05 * rng.standard_normal((N,) * 3)
        res = np.zeros_like(u)
        O.dgp_residual(p, 0, u, delta, None, res, "f64")
        fused = O.dgp_restrict_residual1(p, u, delta, "f64")
    mean8 = res.reshape(N // 2, 2, N // 2, 2, N // 2, 2).mean(axis=(1, 3, 5))
    scale = np.max(np.abs(res))
    assert np.max(np.abs(fused - mean8)) <= 1e-11 * scale


def test_restriction_averages_children():
    rng = np.random.default_rng(1)
    f = rng.standard_normal((8, 8, 8))
    c = O.restrict8(f, "f64")
    assert np.allclose(c, f.reshape(4, 2, 4, 2, 4, 2).mean(axis=(1, 3, 5)), rtol=0, atol=1e-15)
    assert np.all(O.restrict8(np.full((8, 8, 8), 2.5), "f64") == 2.5)


def trilinear_reference(e):
    """Cell-centred trilinear prolongation with periodic wrap (weights 3/4 and
    1/4 per axis), written independently of the oracle."""
    out = e
    for ax in range(3):
        up = np.roll(out, -1, axis=ax)
        dn = np.roll(out, 1, axis=ax)
        lo = 0.75 * out + 0.25 * dn   # son nearer the lower neighbour
        hi = 0.75 * out + 0.25 * up
        shp = list(out.shape)
        shp[ax] *= 2
        new = np.empty(shp)
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = slice(0, None, 2)
        sl_hi[ax] = slice(1, None, 2)
        new[tuple(sl_lo)] = lo
        new[tuple(sl_hi)] = hi
        out = new
    return out


def test_prolongation_is_trilinear():
    rng = np.random.default_rng(2)
    uH = rng.standard_normal((4, 4, 4))
    ru = rng.standard_normal((4, 4, 4))
    uf = rng.standard_normal((8, 8, 8))
    expect = uf + trilinear_reference(uH - ru)
    O.correct(uH, ru, uf, False, "f64")
    assert np.allclose(uf, expect, rtol=0, atol=1e-14)
    # constants prolong exactly
    uf = np.zeros((8, 8, 8))
    O.correct(np.full((4, 4, 4), 3.0), np.full((4, 4, 4), 1.0), uf, False, "f64")
    assert np.all(uf == 2.0)


def test_prolongation_guard_keeps_positive_u():
    uH = np.full((4, 4, 4), -5.0)
    ru = np.zeros((4, 4, 4))
    uf = np.full((8, 8, 8), 1.0)
    O.correct(uH, ru, uf, True, "f64")
    assert np.all(uf == 1.0)   # every corrected value would be <= 0: all skipped
    O.correct(uH, ru, uf, False, "f64")
    assert np.all(uf == -4.0)


@pytest.mark.parametrize("n", [0, 1, 2])
def test_linear_regime_matches_fft_solve(n):
    """sigma -> 0: u = u_bg (1 + e); with w = u^{n+1} ~ u_bg^{n+1} (1 + (n+1) e)
    and u_bg/u - 1 ~ -e the residual linearises to
    A (n+1) lap(e) + (Om/a) delta + C e = 0,  A = c~^2 s u_bg^{n+1}, C = Rbar a^2/3,
    solved in Fourier space with the discrete Laplacian's eigenvalues."""
    N = 32
    p = P(N=N, fr_n=n, a=1.0, box=64.0, polish=1)
    ctilde, Rbg, fbg, ubg = consts(p)
    eps = 1e-5
    delta = eps * O.lognormal_density(N, p.box, 1.0, seed=11)
    u, rep = O.solve(p, delta, mode=1, rel_tol=1e-12, max_cycles=60, prec="f64")
    A = ctilde ** 2 * np.sign(fbg) * ubg ** (n + 1)
    C = Rbg * p.a ** 2 / 3.0
    m = np.fft.fftfreq(N, 1.0 / N)
    lam = sum(np.meshgrid(*(2 * (np.cos(2 * np.pi * m / N) - 1),) * 3, indexing="ij"))
    ek = -np.fft.fftn((p.om / p.a) * delta) / (C + A * (n + 1) * lam)
    e = np.fft.ifftn(ek).real
    e_num = u / ubg - 1.0
    assert np.max(np.abs(e_num - e)) <= 1e-3 * np.max(np.abs(e))


def test_dgp_zero_density_gives_zero_field():
    p = P(N=16, model=2, a=1.0)
    u, rep = O.solve(p, np.zeros((16,) * 3), mode=1, rel_tol=1e-10, max_cycles=5, prec="f64")
    # sqrt(alpha*alpha) - alpha is 0 up to one rounding: phi may pick up a
    # uniform offset of that size (a null mode of the operator), never a gradient
    assert np.ptp(u) < 1e-25 and np.max(np.abs(u)) < 1e-10 and rep["res0"] < 1e-14


def test_dgp_linear_limit_is_poisson():
    """Small delta: L = lap(phi) - Om delta/(beta a) + O(delta^2)."""
    N = 32
    p = P(N=N, model=2, a=1.0, h0rc=1.0)
    eps = 1e-6
    delta = eps * O.lognormal_density(N, p.box, 1.0, seed=12)
    delta -= delta.mean()
    u = np.zeros_like(delta)
    O.vcycles(p, u, delta, 12, dgp_lexi=False, prec="f64")
    Orc = 1.0 / (4.0 * p.h0rc ** 2)
    beta = 1.0 + (0.5 * p.om + p.oml) / math.sqrt(Orc * (p.om + p.oml))
    m = np.fft.fftfreq(N, 1.0 / N)
    lam = sum(np.meshgrid(*(2 * (np.cos(2 * np.pi * m / N) - 1),) * 3, indexing="ij"))
    lam[0, 0, 0] = 1.0
    phik = np.fft.fftn(p.om * delta / (beta * p.a)) / lam
    phik[0, 0, 0] = 0.0
    phi = np.fft.ifftn(phik).real
    num = u - u.mean()
    assert np.max(np.abs(num - phi)) <= 1e-4 * np.max(np.abs(phi))


def test_philox_known_answers():
    """Philox-4x32-10 known-answer vectors of the Random123 distribution."""
    z = np.zeros(1, np.uint32)
    o = np.full(1, 0xFFFFFFFF, np.uint32)
    r = O.philox4x32(z, z, z, z, 0, 0)
    assert [int(v[0]) for v in r] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    r = O.philox4x32(o, o, o, o, 0xFFFFFFFF, 0xFFFFFFFF)
    assert [int(v[0]) for v in r] == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    c = [np.array([v], np.uint32) for v in (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344)]
    r = O.philox4x32(*c, 0xA4093822, 0x299F31D0)
    assert [int(v[0]) for v in r] == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_lognormal_density_statistics():
    N = 32
    d = O.lognormal_density(N, 64.0, 1.2, seed=5)
    g = np.log1p(d) + 0.5 * 1.2 ** 2
    assert abs(np.std(g) - 1.2) < 1e-12 and abs(np.mean(g)) < 1e-12
    assert d.min() > -1.0
    d2 = O.lognormal_density(N, 64.0, 1.2, seed=5)
    assert np.array_equal(d, d2)
    assert not np.array_equal(d, O.lognormal_density(N, 64.0, 1.2, seed=6))


def _binned_model(N, box, slope, cutoff_power):
    """per-|m|-bin average over the bin's modes of k^slope exp(-k^2/(cutoff_power k_c^2))"""
    m = O.kgrid(N)
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    rm = np.sqrt(kx * kx + ky * ky + kz * kz).ravel()
    sel = rm >= 0.5
    k = 2 * np.pi / box * rm[sel]
    kc = 0.5 * np.pi * N / box
    model = k ** slope * np.exp(-k * k / (cutoff_power * kc * kc))
    b = np.maximum((rm[sel] + 0.5).astype(np.int64), 1)
    s = np.bincount(b, weights=model, minlength=N + 1)
    c = np.bincount(b, minlength=N + 1)
    return (s / np.maximum(c, 1))[1: N // 2 + 1], c[1: N // 2 + 1]


def test_lognormal_gaussian_field_has_the_stated_spectrum():
    """g = log(1 + delta) + sigma^2/2 has P_g(k) proportional to
    k^-1.5 exp(-k^2 / (2 k_c^2)), k_c = k_Nyq / 2, bin by bin (6 seeds, bins
    of >= 50 modes, 20 % scatter bound); a filter of the power instead of its
    square root (slope -3) or a cutoff on the amplitude's scale
    (exp(-k^2 / k_c^2)) leaves ratios off by more than 50 %."""
    N, box, sg = 32, 64.0, 1.1
    acc = np.zeros(N // 2)
    seeds = range(40, 46)
    for sd in seeds:
        d = O.lognormal_density(N, box, sg, seed=sd)
        g = np.log1p(d) + 0.5 * sg ** 2
        _, Pk, cnt = O.power_spectrum(g, box)
        acc += Pk
    acc /= len(seeds)

    def flatness(slope, cutoff_power):
        model, c = _binned_model(N, box, slope, cutoff_power)
        use = c >= 50
        r = acc[use] / model[use]
        return np.max(np.abs(r / np.median(r) - 1.0))

    assert flatness(-1.5, 2.0) < 0.2
    assert flatness(-3.0, 2.0) > 0.5
    assert flatness(-1.5, 1.0) > 0.5


def test_power_spectrum_single_mode():
    N, box = 16, 100.0
    x = np.arange(N)
    A = 0.3
    f = A * np.cos(2 * np.pi * 3 * x / N)[None, None, :] * np.ones((N, N, N))
    k, Pk, cnt = O.power_spectrum(f, box)
    # two modes (+-3 along x) carry |A/2|^2 each; bin 3 holds cnt[2] modes
    assert np.isclose(Pk[2], box ** 3 * 2 * (A / 2) ** 2 / cnt[2], rtol=1e-12)
    assert np.all(Pk[np.arange(len(Pk)) != 2] < 1e-25)
    assert np.isclose(k[0], 2 * np.pi / box * np.mean([1.0, 1.0, 1.0, 1.0, 1.0, 1.0]) * 1.0, rtol=0.3)


def test_log_bins_follow_the_closed_form():
    """POWERfft5's log bin of |m| = sqrt(n2) is INT(log10|m|/0.02) =
    floor(25 log10 n2) (PMP2mod_power.f90:396), clamped below at 1.  The
    reference's REAL*4 evaluation (the sum with 1000. keeps 2^-14 or 2^-13)
    can move a value that lies within 1e-4 below an integer up to it;
    elsewhere the bin is the exact floor.  Checked for every n2 up to 3 (512/2)^2: a wrong base, a
    missing square root or a lost rounding step fails."""
    n2max = 3 * 256 ** 2
    got = O.log_bins(n2max)[1:]
    n2 = np.arange(1, n2max + 1)
    v = 25.0 * np.log10(n2.astype(np.float64))
    near = np.abs(v - np.rint(v)) < 1e-4
    exact = np.maximum(np.floor(v).astype(np.int64), 1)
    assert np.array_equal(got[~near], exact[~near])
    up = np.maximum(np.rint(v[near]).astype(np.int64), 1)
    assert np.all((got[near] == up) | (got[near] == exact[near]))
    assert np.any(got[near] != exact[near])  # some are rounded up, as the REAL*4 sum does
    assert near.sum() < 50  # the rounded cases are the exception (n2 = 10^j among them)
    assert got[9] == 25 and got[99] == 50  # n2 = 10, 100: log10|m| = 1/2, 1


def test_central_gradient_of_sinusoid():
    N = 16
    x = np.arange(N)
    psi = np.sin(2 * np.pi * 2 * x / N)[None, None, :] * np.ones((N, N, N))
    gx, gy, gz = O.central_gradient(psi)
    expect = np.sin(2 * np.pi * 2 / N) * np.cos(2 * np.pi * 2 * x / N)
    assert np.allclose(gx[0, 0], expect, atol=1e-14)
    assert np.max(np.abs(gy)) == 0.0 and np.max(np.abs(gz)) == 0.0
