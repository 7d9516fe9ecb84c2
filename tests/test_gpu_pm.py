"""PM step around the solve on the GPU (mgs_particles_*, mgs_cic_deposit,
mgs_potential, mgs_potential_add_fifth_force, mgs_move) against the numpy
restatement in oracle/pm.py.

Bars: the deposit to its fixed-point rounding (2^-41 per contribution),
bitwise reproducible in any particle order, the REAL*4 potential to
one REAL*4 ulp (its FP64 value differs from numpy's FFT only at rounding
level, so the stored values are equal but for ties); the push bit-exact
(the reference's operation order and REAL*4 typing on REAL*4 particles, no
contraction on either side); SVEL to FP64 rounding (block order).  The same
kernels also run on the reference's own golden vectors (pm_*.npz)."""
import os

import numpy as np
import pytest

import oracle as O
import pm as P
import mgsolve as M
from conftest import GOLDEN, gpu_available


def particles(N, n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(1.0, N + 1.0, n).astype(np.float32) for _ in range(3))
    # edge cases: on the first node, just below N+1, on the wrap face
    x[:3] = np.float32([1.0, np.nextafter(np.float32(N + 1), np.float32(0)), N + 0.5])
    v = [rng.normal(0.0, 0.3, n).astype(np.float32) for _ in range(3)]
    return x, y, z, v


def fr_cfg(N, prec=M.PREC_FP64):
    return M.Config(N=N, model=M.MODEL_FR, fr_n=1, box=2.0 * N, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=-1e-5,
                    precision=prec)


@pytest.mark.gpu
@pytest.mark.parametrize("N", [20, 40, 44])
def test_binned_deposit_clustered_and_partial_tiles(N):
    """The deposit bins particles by 16^3-cell tile: grids that are not a
    multiple of 16 (a partial last tile on each axis), particles piled into a
    few cells (long runs of one tile, LDS accumulation contended), positions
    on both edges of [1, N+1) (the +1 reach wraps), and a shuffled order.
    N = 20 (2 tiles a side, the last partial) takes the eight parity-class
    launches with plain adds, whose last tile's reach wraps into tile 0;
    N = 40 and 44 (3 tiles) the one-launch path with atomics."""
    rng = np.random.default_rng(N)
    n = 30000
    x, y, z, _ = particles(N, n, N + 1)
    # a third in 3 cells, a third in one tile near the upper corner
    k = n // 3
    for a in (x, y, z):
        a[3:k] = np.float32(rng.choice([1.25, 7.5, N + 0.75], k - 3))
        a[k:2 * k] = np.float32(rng.uniform(N - 6.0, N + 1.0, k))
        np.clip(a, 1.0, np.nextafter(np.float32(N + 1), np.float32(0)), out=a)
    perm = rng.permutation(n)
    x, y, z = x[perm], y[perm], z[perm]
    print(f"[pm] binned deposit N={N} np={n}", flush=True)
    with M.Solver(fr_cfg(N)) as S:
        S.set_particles(x, y, z)
        S.cic_deposit()
        d = S.get_density(np.float64)
        S.cic_deposit()  # reuses the binning buffers
        d2 = S.get_density(np.float64)
    d_o = P.cic_deposit(x, y, z, N)
    w = P.wpar(N, n)
    # sums of up to ~1e4 weights per cell, each rounded to a multiple of 2^-40
    assert np.allclose(d, d_o, rtol=0, atol=1e-11 * w * max(1.0, np.max(d_o + 1.0) / w)), np.max(np.abs(d - d_o))
    # fixed-point sums: bitwise the same deposit again, whatever the order
    assert np.array_equal(d, d2)
    # every particle deposits the REAL*4 WPAR
    assert abs(np.sum(d + 1.0) - n * w) <= 1e-9 * N ** 3


@pytest.mark.gpu
def test_deposit_and_potential_match_oracle():
    N = 32
    x, y, z, _ = particles(N, 40000, 7)
    print(f"[pm] deposit N={N} np={x.size}", flush=True)
    with M.Solver(fr_cfg(N)) as S:
        S.set_particles(x, y, z)
        S.cic_deposit()
        d = S.get_density(np.float64)
        d_o = P.cic_deposit(x, y, z, N)
        w = P.wpar(N, x.size)
        # each contribution rounded to a multiple of 2^-40 (error <= 2^-41),
        # a few dozen per cell at most here
        assert np.allclose(d, d_o, rtol=0, atol=64 * 2.0 ** -41), np.max(np.abs(d - d_o))
        assert abs(np.sum(d + 1.0) - x.size * w) <= 1e-9 * N ** 3
        # a shuffled particle order deposits bitwise the same density
        perm = np.random.default_rng(1).permutation(x.size)
        S.set_particles(x[perm], y[perm], z[perm])
        S.cic_deposit()
        assert np.array_equal(S.get_density(np.float64), d)
        trfi = 1.5 * 0.3 / 1.0
        S.potential(trfi)
        phi = S.get_potential()
        phi_o = P.poisson_fft(d_o, trfi)
        # REAL*4 storage of an FP64 result that agrees to ~1e-15: equal but
        # for ties at a rounding boundary (one REAL*4 ulp)
        assert np.allclose(phi, phi_o, rtol=0, atol=1.2e-7 * np.max(np.abs(phi_o))), np.max(np.abs(phi - phi_o))
        assert np.mean(phi.astype(np.float32) == P.potential_fi(d_o, trfi)) > 0.999
        phi32 = S.get_potential(np.float32)
        assert np.array_equal(phi32, phi.astype(np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("model", [M.MODEL_FR, M.MODEL_DGP])
def test_potential_plus_fifth_force_after_a_solve(model):
    N = 32
    x, y, z, _ = particles(N, 60000, 11)
    if model == M.MODEL_FR:
        c = fr_cfg(N, M.PREC_MIXED)
    else:
        c = M.Config(N=N, model=M.MODEL_DGP, branch=1, box=2.0 * N, omega_m=0.3, omega_l=0.7, aexpn=1.0, h0rc=1.0,
                     precision=M.PREC_FP64)
    print(f"[pm] deposit -> solve -> potential + fifth force, model={model}", flush=True)
    with M.Solver(c) as S:
        S.set_particles(x, y, z)
        S.cic_deposit()
        if model == M.MODEL_FR:
            rep = S.solve(mode=1, rel_tol=1e-8, max_cycles=60)
        else:
            rep = S.solve(mode=0, res_conv=1e-30, max_cycles=30)
        assert rep["res_final"] < rep["res0"]
        d = S.get_density(np.float64)
        u = S.get_field(np.float64)
        S.potential(0.45)
        S.potential_add_fifth_force()
        phi = S.get_potential()
    p = O.Params(N=N, model=model, fr_n=1, branch=1, box=2.0 * N, om=0.3, oml=0.7, a=1.0, fr0=-1e-5, h0rc=1.0)
    un = u / O.fr_ubg(p) if model == M.MODEL_FR else u
    expect = P.potential_fi(d, 0.45, O.fifth_force_potential(p, un))
    scale = np.max(np.abs(expect))
    assert np.allclose(phi, expect, rtol=0, atol=2.4e-7 * scale)
    assert np.mean(phi.astype(np.float32) == expect) > 0.999


@pytest.mark.gpu
def test_move_matches_oracle_bitwise():
    N = 32
    x, y, z, (vx, vy, vz) = particles(N, 50000, 13)
    rng = np.random.default_rng(2)
    phi = rng.standard_normal((N, N, N))
    pc, xc = P.pconst_xconst(0.5, 0.02, 0.3, 0.7)
    print(f"[pm] move N={N} np={x.size}", flush=True)
    with M.Solver(fr_cfg(N)) as S:
        # a rough potential from a random density; the oracle pushes with the
        # potential the GPU holds
        S.set_density(P.laplacian7(phi))
        S.potential(0.7)
        phi_g = S.get_potential()
        S.set_particles(x, y, z, vx, vy, vz)
        svel = S.move(pc, xc)
        got = S.get_particles()
    expect = P.move(phi_g, x, y, z, vx, vy, vz, pc, xc)
    for a, b, name in zip(got, expect[:6], ("x", "y", "z", "vx", "vy", "vz")):
        assert np.array_equal(a, b), (name, int(np.sum(a != b)))
    assert svel == pytest.approx(expect[6], rel=1e-12)
    assert np.all(got[0] >= 1.0) and np.all(got[0] < N + 1.0)


def _golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


@pytest.mark.gpu
def test_pm_kernels_match_the_reference_golden_vectors():
    """The device's deposit, Poisson solve, push and P(k) against the
    reference's own DENSIT, POTENTfft5, MOVE and POWERfft5 outputs
    (tests/golden/pm_*.npz): push bit for bit, the rest within the REAL*4
    rounding of the reference's FI storage (bars as in
    tests/test_oracle_pm_golden.py)."""
    f = _golden("pm_deposit")
    N = int(f["N"][0])
    print(f"[pm-golden] deposit N={N} np={f['x'].size}", flush=True)
    with M.Solver(fr_cfg(N)) as S:
        S.set_particles(f["x"], f["y"], f["z"])
        S.cic_deposit()
        d = S.get_density()
    assert np.max(np.abs(d - f["fi"].astype(np.float64))) < 2e-6
    for N in (16, 12):
        f = _golden(f"pm_potential_{N}")
        _, box, om, oml, a = [float(v) for v in f["params"]]
        print(f"[pm-golden] potential N={N}", flush=True)
        with M.Solver(fr_cfg(N)) as S:
            S.set_density(f["delta"])
            S.potential(P.trfi_of(om, a))
            phi = S.get_potential(np.float32).astype(np.float64)
        ref = f["phi"].astype(np.float64)
        assert np.max(np.abs(phi - ref)) < 2e-6 * np.max(np.abs(ref))
    for tag in ("f1", "f12"):
        f = _golden(f"pm_move_{tag}")
        N, box, om, oml, a, astep, factv = [float(v) for v in f["params"]]
        N = int(N)
        pc, xc = P.pconst_xconst(a, astep, om, oml, factv)
        print(f"[pm-golden] move {tag}", flush=True)
        with M.Solver(fr_cfg(N)) as S:
            S.set_potential(f["phi"])
            S.set_particles(f["x_in"], f["y_in"], f["z_in"], f["vx_in"], f["vy_in"], f["vz_in"])
            S.move(pc, xc)
            got = S.get_particles()
        for arr, nm in zip(got, ("x", "y", "z", "vx", "vy", "vz")):
            assert np.array_equal(arr.view(np.int32), f[nm].view(np.int32)), (tag, nm)
    for N in (16, 12):
        f = _golden(f"pm_power_{N}")
        box = float(f["params"][1])
        print(f"[pm-golden] power N={N}", flush=True)
        cfg = fr_cfg(N)
        cfg.box = box
        with M.Solver(cfg) as S:
            S.set_density(f["field"])
            k, Pk, cnt = S.power_spectrum(which=0)
            ex = {(ip, ic): S.power_spectrum_ex(0, ip, ic) for ip, ic in ((0, 0), (0, 1), (1, 0))}
        assert np.array_equal(cnt, f["cnt_p0_c0"])
        assert np.max(np.abs(Pk / f["pk_p0_c0"] - 1.0)) < 1e-6
        assert np.max(np.abs(k / f["k_p0_c0"] - 1.0)) < 1e-6
        # POWERfft5's other options and log bins (the reference stores the
        # counts in REAL*4; its quadrupole is a difference of REAL*4 terms)
        scale = float(np.max(f["pk_p0_c0"]))
        for (ip, ic), (k2, P2, c2, kl, Pl, cl) in ex.items():
            t = f"p{ip}_c{ic}"
            f32 = lambda a: np.asarray(a).astype(np.float32).astype(np.float64)  # noqa: E731
            assert np.array_equal(f32(c2), f["cnt_" + t]) and np.array_equal(f32(cl), f["cntlog_" + t]), t
            for a, b in ((k2, f["k_" + t]), (kl, f["klog_" + t])):
                m = b != 0
                assert np.array_equal(a == 0, b == 0) and np.max(np.abs(a[m] / b[m] - 1.0), initial=0.0) < 1e-6, t
            tol = 1e-5 * scale if ip else 1e-6 * np.abs(f["pk_" + t])
            assert np.all(np.abs(P2 - f["pk_" + t]) <= tol), t
            assert np.all(np.abs(Pl - f["pklog_" + t]) <= 1e-6 * np.abs(f["pklog_" + t])), t


@pytest.mark.gpu
def test_pm_errors():
    N = 16
    with M.Solver(fr_cfg(N)) as S:
        with pytest.raises(M.MgsError, match="no particles"):
            S.cic_deposit()
        S.set_particles(np.float32([1.5, 17.5]), np.float32([2.0, 2.0]), np.float32([3.0, 3.0]))
        with pytest.raises(M.MgsError, match="outside"):
            S.cic_deposit()
        with pytest.raises(M.MgsError, match="no potential"):
            S.move(-0.1, 0.1)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_pm_without_gpu_is_an_error():
    with pytest.raises(M.MgsError):
        with M.Solver(fr_cfg(16)) as S:
            S.cic_deposit()
