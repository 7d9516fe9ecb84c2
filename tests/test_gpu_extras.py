"""GPU checks of the pieces around the solve, through the C-ABI:
graph replay vs eager launches, the seeded density generator, the
fifth-force potential/gradient and the power spectrum (against the numpy
restatements in oracle.py), and the error paths."""
import numpy as np
import pytest

import oracle as O
import mgsolve as M


def cfg(N, precision=M.PREC_FP64, **kw):
    d = dict(N=N, model=M.MODEL_FR, fr_n=1, box=64.0, omega_m=0.3, omega_l=0.7, aexpn=1.0, fr0=-1e-5,
             precision=precision)
    d.update(kw)
    return M.Config(**d)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,model", [(M.PREC_FP64, 1), (M.PREC_MIXED, 1), (M.PREC_FP64, 2)])
def test_graph_replays_equal_eager_launches(precision, model):
    """The captured coarse cycle replays (with eager level-0 work and host
    synchronisations between the replays) give bit-identical fields."""
    N = 32
    print(f"[graphs] precision={precision} model={model}", flush=True)
    out = []
    for graphs in (True, False):
        with M.Solver(cfg(N, precision, model=model, h0rc=1.0)) as S:
            S.use_graphs(graphs)
            fields = []
            for seed in (21, 22):
                S.set_density(O.lognormal_density(N, 64.0, 1.0, seed=seed))
                rep = S.solve(mode=1, rel_tol=1e-9, max_cycles=12)
                fields.append(S.get_field())
                S.synchronize()
            out.append(fields)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_lognormal_density_matches_numpy():
    N, box, sigma, seed = 32, 64.0, 1.3, 4242
    with M.Solver(cfg(N, box=box)) as S:
        S.make_lognormal_density(seed, sigma)
        d = S.get_density()
    ref = O.lognormal_density(N, box, sigma, seed)
    assert np.max(np.abs(d - ref) / (1.0 + np.abs(ref))) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("model,frn", [(1, 0), (1, 1), (1, 2), (2, 1)])
def test_fifth_force_matches_numpy(model, frn):
    N = 16
    c = cfg(N, model=model, fr_n=frn, h0rc=1.0)
    p = O.Params(N=N, model=model, fr_n=frn, box=c.box, om=c.omega_m, oml=c.omega_l, a=c.aexpn, fr0=c.fr0,
                 h0rc=1.0)
    rng = np.random.default_rng(3)
    with M.Solver(c) as S:
        _, ubg = S.background()
        u = ubg * (1.0 + 0.2 * rng.random((N,) * 3)) if model == 1 else 0.01 * rng.standard_normal((N,) * 3)
        S.set_field(u)
        psi, gx, gy, gz = S.fifth_force()
        psi32 = S.fifth_force(np.float32)[0]
    ref = O.fifth_force_potential(p, u / ubg if model == 1 else u)
    rx, ry, rz = O.central_gradient(ref)
    scale = np.max(np.abs(ref))
    for a, b in ((psi, ref), (gx, rx), (gy, ry), (gz, rz)):
        assert np.max(np.abs(a - b)) <= 1e-14 * scale
    assert np.max(np.abs(psi32 - ref)) <= 1e-6 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1, 2])
def test_power_spectrum_matches_numpy(which):
    N, box = 32, 100.0
    c = cfg(N, box=box)
    p = O.Params(N=N, model=1, fr_n=1, box=box, om=c.omega_m, oml=c.omega_l, a=1.0, fr0=c.fr0)
    with M.Solver(c) as S:
        S.make_lognormal_density(77, 1.0)
        S.solve(mode=1, rel_tol=1e-6, max_cycles=20)
        d = S.get_density()
        u = S.get_field()
        _, ubg = S.background()
        k, Pk, cnt = S.power_spectrum(which)
    field = [d, u, O.fifth_force_potential(p, u / ubg)][which]
    kr, Pr, cr = O.power_spectrum(field, box)
    assert np.array_equal(cnt, cr)
    assert np.allclose(k, kr, rtol=1e-12)
    assert np.allclose(Pk, Pr, rtol=1e-9, atol=1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize("N", [64, 48])
@pytest.mark.parametrize("ipol,icic", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_power_spectrum_options_match_the_oracle(N, ipol, icic):
    """POWERfft5's NGP / linear k-space bins, monopole / quadrupole and log
    bins against the oracle on a lognormal density: the bin rules agree
    exactly (NGP counts equal; linear weights summed in another order, 1e-12),
    k to 1e-12, P to 1e-9 of the largest bin."""
    box = 2.0 * N
    with M.Solver(cfg(N, box=box)) as S:
        S.make_lognormal_density(91, 1.1)
        d = S.get_density()
        print(f"[power] N={N} ipol={ipol} icic={icic}", flush=True)
        got = S.power_spectrum_ex(0, ipol, icic)
    ref = O.power_spectrum_ex(d, box, ipol, icic)
    k, Pk, cnt, kl, Pl, cl = got
    kr, Pr, cr, klr, Plr, clr = ref
    if icic:
        assert np.allclose(cnt, cr, rtol=1e-12, atol=0)
    else:
        assert np.array_equal(cnt, cr) and np.array_equal(cl, clr)
    assert np.allclose(k, kr, rtol=1e-12, atol=0) and np.allclose(kl, klr, rtol=1e-12, atol=0)
    scale = float(np.max(np.abs(Pr)))
    assert np.max(np.abs(Pk - Pr)) <= 1e-9 * scale
    assert np.max(np.abs(Pl - Plr)) <= 1e-9 * max(float(np.max(np.abs(Plr))), 1e-300)
    if icic:
        assert not np.any(cl) and not np.any(Pl)
    elif ipol:
        assert not np.any(Pl) and np.any(cl)
    # the plain entry point is the NGP monopole
    if ipol == 0 and icic == 0:
        with M.Solver(cfg(N, box=box)) as S:
            S.set_density(d)
            k0, P0, c0 = S.power_spectrum(0)
        assert np.array_equal(c0, cnt) and np.allclose(P0, Pk, rtol=1e-12) and np.allclose(k0, k, rtol=1e-14)


@pytest.mark.gpu
def test_error_paths():
    with M.Solver(cfg(16)) as S:
        with pytest.raises(M.MgsError, match="no density"):
            S.solve()
        bad = np.zeros((16,) * 3)
        bad[3, 4, 5] = np.nan
        with pytest.raises(M.MgsError, match="finite"):
            S.set_density(bad)
        with pytest.raises(M.MgsError, match="level"):
            S.smooth(99, 1)
        with pytest.raises(M.MgsError, match="cannot change"):
            S.set_params(cfg(32))
        # iteration cap reached: reported, not an error (PMP2extradof.f90:117-121)
        S.set_density(O.lognormal_density(16, 64.0, 1.0, seed=3))
        rep = S.solve(mode=0, max_cycles=1, res_conv=1e-30)
        assert rep["cycles"] == 1 and not rep["converged"]
