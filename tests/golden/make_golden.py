"""Write the golden vectors in tests/golden/ from the reference's own routines.

Run in the build container, where the reference sources are present:

    make -C oracle ref && python tests/golden/make_golden.py

The reference's multigrid routines (PMP2MGsolver_fR.f90, PMP2MGsolver_DGP.f90,
PMP2extradof.f90) are compiled from where they lie by oracle/Makefile into
oracle/_ref/libmgref.so and driven one routine at a time through
oracle/ref_harness.f90.  Each fixture holds the seeded inputs and the outputs
the reference computed on them (FP32 storage, as the reference stores FI,
FI2, FI3).  The reference's packed FI3 scratch array is unpacked into one
array per level and quantity with oracle.fi3_sections.

Fixtures (all little-endian numpy .npz, float32 fields):
  kernels_fr_n{0,1,2}_{a}.npz   one application of every f(R) routine, 16^3
  kernels_dgp_{N,S}_{a}.npz     one application of every DGP routine, 16^3
  vcycles_{fr,dgp}.npz          6 V-cycles from the reference's initial
                                field, 32^3: residual history + final field
  scalar_field_fr.npz           scalar_field(1) with the reference's own
                                exits, 32^3: final normalised field
  cycles_{F,W}_{fr,dgp}.npz     3 F_cycle / W_cycle calls, 16^3: residual
                                history + final field
  pm_*.npz                      the PM-step routines around the solve, 16^3
                                and 12^3: DENSIT's CIC deposit, POTENTfft5,
                                POTENT_FOR_MG (f(R) n = 0, 1, 2 and nDGP),
                                MOVE(FactV), POWERfft5 (NGP and CIC k-space
                                bins, monopole and quadrupole, linear and log
                                bins), from oracle/_ref/libmgref_pm.so
(`python make_golden.py cycles` or `... pm` writes only that group.)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

# Parameter sets whose float32 derived constants are exact (a = 1 or 1/2),
# so the FP32-storage oracle and the reference see identical coefficients.
PSETS = {
    "a1": dict(a=1.0, om=O.f32(0.3), oml=O.f32(0.7), box=256.0),
    "a05": dict(a=0.5, om=0.25, oml=0.75, box=64.0),
}


def fr_inputs(rng, p, N):
    ubg = O.fr_ubg(p)
    delta = (rng.lognormal(0.0, 1.0, (N, N, N)) - 1.6).astype(np.float32)
    # a spread of u around u_bg, with a tail to large values, so both closed-form
    # root branches of each n are exercised
    u = ubg * np.abs(1.0 + 0.4 * rng.standard_normal((N, N, N)))
    big = rng.random((N, N, N)) < 0.25
    u = np.where(big, u * 3.0, u).astype(np.float32)
    return delta, u


def dgp_inputs(rng, N):
    delta = (rng.lognormal(0.0, 1.0, (N, N, N)) - 1.6).astype(np.float32)
    u = (0.05 * rng.standard_normal((N, N, N))).astype(np.float32)
    return delta, u


def kernel_fixture(R, p, delta, u):
    """Apply every routine once, in V-cycle order, and record inputs/outputs."""
    out = {}
    R.init(p)
    R.put(1, delta)
    R.put(2, u)
    out["delta"] = delta
    out["u_in"] = u
    R.smooth(0, 1)
    out["u_red"] = R.get(2)
    R.smooth(0, 0)
    out["u_black"] = R.get(2)
    out["res0_rms"] = np.float64(R.residual(0))
    out["res0"] = R.get(3)
    R.restrict_residual(1)
    out["rres1"] = R.get_sec(1, "rres")
    R.restrict_field(1)
    out["ru1"] = R.get_sec(1, "ru")
    out["u1_init"] = R.get_sec(1, "u")
    R.restrict_density(1)
    out["rho1"] = R.get_sec(1, "rho")
    R.rhs(1)
    out["f1"] = R.get_sec(1, "f")
    R.smooth(1, 1)
    out["u1_red"] = R.get_sec(1, "u")
    R.smooth(1, 0)
    out["u1_black"] = R.get_sec(1, "u")
    R.residual(1)
    out["res1"] = R.get_sec(1, "res")
    R.restrict_residual(2)
    out["rres2"] = R.get_sec(2, "rres")
    R.restrict_field(2)
    out["ru2"] = R.get_sec(2, "ru")
    R.restrict_density(2)
    out["rho2"] = R.get_sec(2, "rho")
    # prolongation: perturb the level-1 solution so the correction is nonzero
    rng = np.random.default_rng(99)
    u1 = R.get_sec(1, "u")
    if p.model == 1:
        u1p = (u1 * (1.0 + 0.05 * rng.standard_normal(u1.shape))).astype(np.float32)
        # a few cells driven strongly negative exercise the positivity guard
        neg = rng.random(u1.shape) < 0.05
        u1p = np.where(neg, -3.0 * np.abs(u1), u1p).astype(np.float32)
    else:
        u1p = (u1 + 0.01 * rng.standard_normal(u1.shape)).astype(np.float32)
    R.put_sec(1, "u", u1p)
    out["u1_corr_in"] = u1p
    out["u_before_corr"] = R.get(2)
    R.correct(1)
    out["u_after_corr"] = R.get(2)
    return out


def branch_stats(p, u, delta):
    """Fraction of cells per closed-form branch of the level-0 smoother."""
    N = p.N
    n = p.fr_n
    ctilde = 2.99792458e3 * N / p.box
    a = p.a
    R_bg = 3.0 * (p.om / a ** 3 + 4.0 * p.oml)
    fbg = O.fr_fbg(p)
    fct2 = 1.0 / (ctilde * ctilde * np.sign(fbg)) / 6.0
    fct1 = -(p.om / a) * fct2
    Q2 = R_bg * a * a * fct2 / 3.0
    Q = Q2 * abs(fbg) ** (1.0 / (n + 1))
    w = u.astype(np.float64) ** (n + 1)
    s = sum(np.roll(w, sh, ax) for ax in range(3) for sh in (1, -1))
    P = fct1 * delta - Q2 - s / 6.0
    if n == 1:
        CC = (27 * Q) ** 2 - 4 * (-3 * P) ** 3
        return {"cardano": float(np.mean(CC >= 0)), "trig": float(np.mean(CC < 0))}
    if n == 2:
        return {"P>0": float(np.mean(P > 0)), "P<0": float(np.mean(P < 0))}
    return {}


def main():
    R = O.RefLib()
    rng = np.random.default_rng(20261202)
    N = 16
    for n in (0, 1, 2):
        for tag, ps in PSETS.items():
            p = O.Params(N=N, model=1, fr_n=n, fr0=-1.0e-5, **ps)
            delta, u = fr_inputs(rng, p, N)
            fx = kernel_fixture(R, p, delta, u)
            fx["params"] = np.array([p.N, p.model, p.fr_n, p.branch, p.box, p.om, p.oml, p.a, p.fr0, p.h0rc])
            np.savez_compressed(os.path.join(HERE, f"kernels_fr_n{n}_{tag}.npz"), **fx)
            print(f"kernels_fr_n{n}_{tag}", branch_stats(p, u, delta))
    for br, bn in ((1, "N"), (-1, "S")):
        for tag, ps in PSETS.items():
            p = O.Params(N=N, model=2, branch=br, h0rc=1.0, **ps)
            delta, u = dgp_inputs(rng, N)
            fx = kernel_fixture(R, p, delta, u)
            fx["params"] = np.array([p.N, p.model, p.fr_n, p.branch, p.box, p.om, p.oml, p.a, p.fr0, p.h0rc])
            np.savez_compressed(os.path.join(HERE, f"kernels_dgp_{bn}_{tag}.npz"), **fx)
            print(f"kernels_dgp_{bn}_{tag}")
    # full V-cycles at 32^3 from the reference's own initial field
    N = 32
    for model, name in ((1, "fr"), (2, "dgp")):
        p = O.Params(N=N, model=model, fr_n=1, fr0=-1.0e-5, h0rc=1.0, **PSETS["a05"])
        R.init(p, iter_max=6)
        delta = (rng.lognormal(0.0, 1.0, (N, N, N)) - np.exp(0.5)).astype(np.float32)
        R.put(1, delta)
        u0 = R.get(2)
        hist = R.vcycles(6)
        np.savez_compressed(os.path.join(HERE, f"vcycles_{name}.npz"), delta=delta, u_in=u0, hist=hist,
                            u_out=R.get(2),
                            params=np.array([p.N, p.model, p.fr_n, p.branch, p.box, p.om, p.oml, p.a, p.fr0, p.h0rc]))
        print(f"vcycles_{name}", hist)
    # scalar_field(1) with the reference's exits (iteration cap 8)
    p = O.Params(N=N, model=1, fr_n=1, fr0=-1.0e-5, **PSETS["a1"])
    R.init(p, iter_max=8, res_conv=1.0e-8)
    delta = (rng.lognormal(0.0, 0.8, (N, N, N)) - np.exp(0.32)).astype(np.float32)
    R.put(1, delta)
    R.scalar_field(1)
    np.savez_compressed(os.path.join(HERE, "scalar_field_fr.npz"), delta=delta, u_out=R.get(2),
                        params=np.array([p.N, p.model, p.fr_n, p.branch, p.box, p.om, p.oml, p.a, p.fr0, p.h0rc]))
    print("scalar_field_fr done")


def cycle_fixtures():
    """F_cycle and W_cycle (PMP2extradof.f90:563-1157), 3 calls each from
    the reference's initial field, 16^3, both models."""
    R = O.RefLib()
    rng = np.random.default_rng(20261202)
    N = 16
    for model, name in ((1, "fr"), (2, "dgp")):
        for cyc, cname in ((1, "F"), (2, "W")):
            p = O.Params(N=N, model=model, fr_n=1, fr0=-1.0e-5, h0rc=1.0, **PSETS["a05"])
            R.init(p, iter_max=3)
            delta = (rng.lognormal(0.0, 1.0, (N, N, N)) - np.exp(0.5)).astype(np.float32)
            R.put(1, delta)
            u0 = R.get(2)
            hist = R.cycles(3, cyc)
            np.savez_compressed(os.path.join(HERE, f"cycles_{cname}_{name}.npz"), delta=delta, u_in=u0, hist=hist,
                                u_out=R.get(2), cycle=np.array([cyc]),
                                params=np.array([p.N, p.model, p.fr_n, p.branch, p.box, p.om, p.oml, p.a, p.fr0,
                                                 p.h0rc]))
            print(f"cycles_{cname}_{name}", hist)


def _particles(rng, N, npart):
    """REAL*4 positions in [1, N+1): uniform, plus exact nodes, the last
    cell (I = N, wrapping to cell 1) and values just below N+1 and at 1"""
    pos = rng.uniform(1.0, N + 1.0, (3, npart)).astype(np.float32)
    top = np.nextafter(np.float32(N + 1), np.float32(0))
    special = np.array([1.0, float(N), float(N) + 0.5, top, 1.0 + 1e-6, N / 2 + 1.0], np.float32)
    for ax in range(3):
        pos[ax, : len(special)] = np.roll(special, ax)
        pos[ax, len(special): 2 * len(special)] = special[::-1]
    pos = np.clip(pos, np.float32(1.0), top)
    return pos


def pm_fixtures():
    """The reference's DENSIT (CIC), POTENTfft5, POTENT_FOR_MG, MOVE and
    POWERfft5 on seeded inputs (PMP2mod_density.f90:235-281,
    PMP2main.f90:304-935, PMP2mod_power.f90:209-463)."""
    R = O.RefPmLib()
    rng = np.random.default_rng(20261203)
    par = np.array
    # DENSIT: a particle count that makes WPAR inexact
    N, npart = 16, 3001
    R.init(N, npart=npart, box=64.0, om=0.3, oml=0.7, a=0.5)
    pos = _particles(rng, N, npart)
    zero = np.zeros(npart, np.float32)
    R.put_particles(pos[0], pos[1], pos[2], zero, zero, zero)
    R.densit()
    np.savez_compressed(os.path.join(HERE, "pm_deposit.npz"), x=pos[0], y=pos[1], z=pos[2], fi=R.get(1),
                        N=par([N]))
    print("pm_deposit", float(R.get(1).sum()))
    # POTENTfft5 (trfi = 1.5 Om / AEXPN in default REAL), two grid sizes
    for N, a in ((16, 0.5), (12, 1.0)):
        R.init(N, box=64.0, om=0.3, oml=0.7, a=a)
        delta = (rng.lognormal(0.0, 1.0, (N, N, N)) - 1.4).astype(np.float32)
        R.put(1, delta)
        R.potentfft5()
        np.savez_compressed(os.path.join(HERE, f"pm_potential_{N}.npz"), delta=delta, phi=R.get(1),
                            params=par([N, 64.0, O.f32(0.3), O.f32(0.7), a]))
        print(f"pm_potential_{N}", float(np.abs(R.get(1)).max()))
    # POTENT_FOR_MG: f(R) n = 0, 1, 2 and nDGP
    N = 16
    for model, frn, tag in ((1, 0, "fr_n0"), (1, 1, "fr_n1"), (1, 2, "fr_n2"), (2, 1, "dgp")):
        R.init(N, model=model, fr_n=frn, box=64.0, om=0.3, oml=0.7, a=0.5, fr0=-1e-5, h0rc=1.0)
        fi = (50.0 * rng.standard_normal((N, N, N))).astype(np.float32)
        fi2 = (1.0 + 0.3 * rng.standard_normal((N, N, N))).astype(np.float32)
        if model == 1:
            fi2 = np.abs(fi2).astype(np.float32)
        R.put(1, fi)
        R.put(2, fi2)
        R.potent_for_mg()
        np.savez_compressed(os.path.join(HERE, f"pm_potent_for_mg_{tag}.npz"), fi_in=fi, fi2=fi2, fi=R.get(1),
                            params=par([N, model, frn, 64.0, O.f32(0.3), O.f32(0.7), 0.5, -1e-5, 1.0]))
        print(f"pm_potent_for_mg_{tag}", float(np.abs(R.get(1) - fi).max()))
    # MOVE(FactV): a smooth potential, particles with velocities
    N, npart = 16, 2001
    for factv, tag in ((1.0, "f1"), (1.2, "f12")):
        R.init(N, npart=npart, box=64.0, om=0.3, oml=0.7, a=0.5, astep=0.02)
        kz, ky, kx = np.meshgrid(*(np.arange(N),) * 3, indexing="ij")
        phi = (40.0 * np.sin(2 * np.pi * (kx + 2 * ky) / N + 0.3) * np.cos(2 * np.pi * kz / N)
               + 5.0 * rng.standard_normal((N, N, N))).astype(np.float32)
        pos = _particles(rng, N, npart)
        vel = (3.0 * rng.standard_normal((3, npart))).astype(np.float32)
        R.put(1, phi)
        R.put_particles(pos[0], pos[1], pos[2], vel[0], vel[1], vel[2])
        enkin = R.move(factv)
        x, y, z, vx, vy, vz = R.get_particles()
        np.savez_compressed(os.path.join(HERE, f"pm_move_{tag}.npz"), phi=phi, x_in=pos[0], y_in=pos[1],
                            z_in=pos[2], vx_in=vel[0], vy_in=vel[1], vz_in=vel[2], x=x, y=y, z=z, vx=vx, vy=vy,
                            vz=vz, enkin=par([enkin]),
                            params=par([N, 64.0, O.f32(0.3), O.f32(0.7), 0.5, O.f32(0.02), O.f32(factv)]))
        print(f"pm_move_{tag}", enkin)
    # POWERfft5: NGP (iCIC = 0) and CIC (iCIC = 1) k-space bins, monopole
    # and quadrupole, on one random field per grid size
    for N in (16, 12):
        R.init(N, npart=N ** 3, box=100.0)
        f = rng.standard_normal((N, N, N)).astype(np.float32)
        out = {"field": f, "params": par([N, 100.0])}
        for ipol, icic in ((0, 0), (0, 1), (1, 0)):
            R.put(1, f)
            k, P, cnt, klog, plog, cntlog = R.power(ipol=ipol, iswitch=1, icic=icic)
            tag = f"p{ipol}_c{icic}"
            out.update({f"k_{tag}": k, f"pk_{tag}": P, f"cnt_{tag}": cnt, f"klog_{tag}": klog,
                        f"pklog_{tag}": plog, f"cntlog_{tag}": cntlog})
        np.savez_compressed(os.path.join(HERE, f"pm_power_{N}.npz"), **out)
        print(f"pm_power_{N}", out["cnt_p0_c0"][:4], out["pk_p0_c0"][:3])


if __name__ == "__main__":
    if sys.argv[1:] == ["cycles"]:
        cycle_fixtures()
    elif sys.argv[1:] == ["pm"]:
        pm_fixtures()
    else:
        main()
        cycle_fixtures()
        pm_fixtures()
