"""The headline workloads at their full sizes, checked through properties
that do not depend on size (the oracle cannot run whole solves at 512^3 in
seconds, but its FP64 operators can evaluate a 512^3 field), and the
reference's CPU case at 128^3, where the oracle's whole solve is quick:

  * f(R) 128^3 (n=1, |fR0|=1e-5, sigma_g=1): device FP64 and mixed solves
    against the oracle's FP64 solve;
  * f(R) 512^3 (n=1, |fR0|=1e-6, sigma_g=1.5, seed 1002): the mixed solve's
    field satisfies the oracle's FP64 operator to relative residual 1e-8,
    and agrees with the plain FP64 solve's field;
  * nDGP 512^3 (normal branch, H0rc=1, sigma_g=1.2, seed 1005): the mixed
    solve reaches the FP64 solve's residual floor, measured by the oracle's
    operator, and the same field up to the free constant;
  * f(R) 1024^3 (n=1, |fR0|=1e-5, sigma_g=1.2, seed 1003), the single
    large solve: mixed and FP64 fields both satisfy the oracle's FP64
    operator to relative residual 1e-8 and agree with each other;
  * PM step at 512^3 with 512^3 particles: the deposit conserves mass
    (sum of 1 + delta = N^3), the density is >= -1, the potential with the
    fifth force is finite, and the push keeps every particle inside [1, N+1).
"""
import numpy as np
import pytest

import oracle as O
import mgsolve as M

N = 512
BOX = 512.0


def rel_linf(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.gpu
def test_config1_fr128_solves_match_the_oracle_solve():
    """The reference's CPU case (128^3, box 256, n=1, |fR0|=1e-5, sigma_g=1,
    FP64 V(2,2) to relative residual 1e-8): the oracle's whole solve against
    the device's FP64 solve (same cycles, histories to 1e-6 per cycle, field
    to 1e-9) and mixed solve (field to 1e-6)."""
    n, box = 128, 256.0
    op = O.Params(N=n, model=1, fr_n=1, branch=1, box=box, om=0.3, oml=0.7, a=1.0, fr0=-1e-5, h0rc=1.0, polish=1)
    delta = O.lognormal_density(n, box, 1.0, seed=1001)
    u_o, rep_o = O.solve(op, delta, mode=1, rel_tol=1e-8, max_cycles=60, prec="f64")
    print(f"[fullsize] config 1: oracle {rep_o['cycles']} cycles, rel. residual "
          f"{rep_o['res_final'] / rep_o['res0']:.3e}", flush=True)
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        c = M.Config(N=n, model=M.MODEL_FR, fr_n=1, box=box, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=-1e-5,
                     precision=prec)
        with M.Solver(c) as S:
            S.set_density(delta)
            rep = S.solve(mode=1, rel_tol=1e-8, max_cycles=60)
            u = S.get_field()
        err = rel_linf(u, u_o)
        print(f"  precision={prec}: rel. residual {rep['res_final'] / rep['res0']:.3e}, field vs oracle {err:.2e}",
              flush=True)
        assert rep["converged"]
        if prec == M.PREC_FP64:
            assert rep["cycles"] == rep_o["cycles"]
            assert np.allclose(rep["history"], rep_o["history"], rtol=1e-6, atol=0)
            assert err <= 1e-9
        else:
            assert err <= 1e-6


@pytest.mark.gpu
def test_fr512_mixed_solve_meets_the_oracle_operator():
    fr0, sigma, seed = -1e-6, 1.5, 1002
    op = O.Params(N=N, model=1, fr_n=1, branch=1, box=BOX, om=0.3, oml=0.7, a=1.0, fr0=fr0, h0rc=1.0, polish=1)
    out = {}
    delta = None
    for prec in (M.PREC_MIXED, M.PREC_FP64):
        c = M.Config(N=N, model=M.MODEL_FR, fr_n=1, box=BOX, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=fr0,
                     precision=prec)
        print(f"[fullsize] f(R) 512^3 precision={prec}", flush=True)
        with M.Solver(c) as S:
            if delta is None:
                S.make_lognormal_density(seed, sigma)
                delta = S.get_density(np.float64)
            else:
                S.set_density(delta)
            rep = S.solve(mode=1, rel_tol=1e-8, max_cycles=60)
            assert rep["converged"], rep
            _, ubg = S.background()
            out[prec] = (S.get_field(), rep)
    u0 = np.full((N, N, N), ubg)
    r0 = O.fr_residual(op, 0, u0, delta, None, None, "f64")
    for prec, (u, rep) in out.items():
        r = O.fr_residual(op, 0, u, delta, None, None, "f64")
        print(f"  precision={prec}: oracle rel. residual {r / r0:.3e} (device {rep['res_final'] / rep['res0']:.3e})",
              flush=True)
        assert r < 1.05e-8 * r0
        assert abs(rep["res0"] - r0) <= 1e-12 * r0
    err = rel_linf(out[M.PREC_MIXED][0], out[M.PREC_FP64][0])
    print(f"  mixed vs FP64 field rel. L_inf {err:.2e}", flush=True)
    assert err < 1e-6


@pytest.mark.gpu
def test_config3_fr1024_mixed_and_fp64_meet_the_oracle_operator():
    n, box, fr0, sigma, seed = 1024, 1024.0, -1e-5, 1.2, 1003
    op = O.Params(N=n, model=1, fr_n=1, branch=1, box=box, om=0.3, oml=0.7, a=1.0, fr0=fr0, h0rc=1.0, polish=1)
    delta = None
    fields = {}
    for prec in (M.PREC_MIXED, M.PREC_FP64):
        c = M.Config(N=n, model=M.MODEL_FR, fr_n=1, box=box, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=fr0,
                     precision=prec)
        print(f"[fullsize] f(R) 1024^3 precision={prec}", flush=True)
        with M.Solver(c) as S:
            if delta is None:
                S.make_lognormal_density(seed, sigma)
                delta = S.get_density(np.float64)
            else:
                S.set_density(delta)
            rep = S.solve(mode=1, rel_tol=1e-8, max_cycles=60)
            assert rep["converged"], rep
            fields[prec] = S.get_field()
        # the residual of u = u_bg is (Om/a) delta exactly
        r0 = 0.3 * float(np.sqrt(np.mean(delta * delta)))
        assert abs(rep["res0"] / r0 - 1.0) < 1e-10
        r = O.fr_residual(op, 0, fields[prec], delta, None, None, "f64")
        print(f"  {rep['seconds'] * 1e3:.1f} ms, oracle rel. residual {r / r0:.3e}", flush=True)
        assert r < 1.05e-8 * r0
    err = rel_linf(fields[M.PREC_MIXED], fields[M.PREC_FP64])
    print(f"  mixed vs FP64 field rel. L_inf {err:.2e}", flush=True)
    assert err < 1e-6


@pytest.mark.gpu
def test_dgp512_mixed_reaches_the_fp64_floor():
    sigma, seed = 1.2, 1005
    op = O.Params(N=N, model=2, fr_n=1, branch=1, box=BOX, om=0.3, oml=0.7, a=1.0, fr0=-1e-5, h0rc=1.0, polish=0)
    fields = {}
    delta = None
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        c = M.Config(N=N, model=M.MODEL_DGP, branch=1, box=BOX, omega_m=0.3, omega_l=0.7, aexpn=1.0, h0rc=1.0,
                     precision=prec)
        print(f"[fullsize] nDGP 512^3 precision={prec}", flush=True)
        with M.Solver(c) as S:
            if delta is None:
                S.make_lognormal_density(seed, sigma)
                delta = S.get_density(np.float64)
            else:
                S.set_density(delta)
            S.solve(mode=0, res_conv=1e-30, max_cycles=60)
            fields[prec] = S.get_field()
    r0 = O.dgp_residual(op, 0, np.zeros((N, N, N)), delta, None, None, "f64")
    r64 = O.dgp_residual(op, 0, fields[M.PREC_FP64], delta, None, None, "f64")
    rmx = O.dgp_residual(op, 0, fields[M.PREC_MIXED], delta, None, None, "f64")
    print(f"  oracle residuals: start {r0:.3e}, FP64 {r64:.3e}, mixed {rmx:.3e}", flush=True)
    assert r64 < 0.02 * r0
    assert rmx <= 1.01 * r64
    u64 = fields[M.PREC_FP64] - fields[M.PREC_FP64].mean()
    um = fields[M.PREC_MIXED] - fields[M.PREC_MIXED].mean()
    err = rel_linf(um, u64)
    print(f"  mixed vs FP64 field rel. L_inf (mean removed) {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.gpu
def test_pm512_step_invariants():
    rng = np.random.default_rng(77)
    g = np.arange(1, N + 1, dtype=np.float32)
    pos = []
    for ax in range(3):
        shape = [1, 1, 1]
        shape[2 - ax] = N
        p = np.broadcast_to(g.reshape(shape), (N, N, N)).reshape(-1) + rng.uniform(-0.5, 0.5, N ** 3).astype(
            np.float32)
        p = np.where(p < 1.0, p + N, p)
        pos.append(np.where(p >= N + 1.0, p - N, p).astype(np.float32))
    vel = [rng.normal(0.0, 0.05, N ** 3).astype(np.float32) for _ in range(3)]
    c = M.Config(N=N, model=M.MODEL_FR, fr_n=1, box=BOX, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=-1e-6,
                 precision=M.PREC_MIXED)
    print("[fullsize] PM step 512^3 grid, 512^3 particles", flush=True)
    with M.Solver(c) as S:
        S.set_particles(*pos, *vel)
        del pos, vel
        S.cic_deposit()
        d = S.get_density(np.float64)
        mass = float(np.sum(d + 1.0))
        print(f"  deposit: sum(1 + delta) = {mass:.9e}, min delta = {d.min():.3e}", flush=True)
        assert abs(mass - float(N) ** 3) <= 1e-9 * float(N) ** 3
        assert d.min() >= -1.0
        del d
        rep = S.solve(mode=1, rel_tol=1e-8, max_cycles=60)
        assert rep["converged"], rep
        S.potential(1.5 * 0.3)
        S.potential_add_fifth_force()
        phi = S.get_potential(np.float32)
        assert np.all(np.isfinite(phi))
        print(f"  potential mean {float(np.mean(phi, dtype=np.float64)):.3e}, max |phi| "
              f"{float(np.max(np.abs(phi))):.3e}", flush=True)
        del phi
        S.move(-0.5 * 0.02, 0.02)  # raises if a particle leaves [1, N+1)
        x, y, z, vx, vy, vz = S.get_particles()
        for a in (x, y, z):
            assert a.min() >= 1.0 and a.max() < N + 1.0
        assert np.all(np.isfinite(vx)) and np.all(np.isfinite(vy)) and np.all(np.isfinite(vz))
