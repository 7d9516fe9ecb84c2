"""The oracle's PM-step restatements against the reference's own routines.

The fixtures tests/golden/pm_*.npz hold seeded inputs and the outputs of the
reference's DENSIT (CIC branch), POTENTfft5, POTENT_FOR_MG, MOVE(FactV) and
POWERfft5, compiled from its sources by oracle/Makefile (libmgref_pm.so) and
driven by tests/golden/make_golden.py.  Bars:
  * MOVE and POTENT_FOR_MG: bit for bit (the oracle restates the reference's
    REAL*4 typing: node gradients differenced in REAL*4, PCONST/XCONST/WPAR
    in REAL*4, FI2**(n+1) as REAL*4 products);
  * DENSIT: the reference accumulates every particle's weight into the REAL*4
    FI one rounding at a time; the oracle sums in FP64: 2e-6 absolute on
    fields of O(1) values (the last-bit spread of that accumulation);
  * POTENTfft5: the reference rounds FI to REAL*4 after each axis pass; the
    oracle transforms in FP64: 1e-6 of the field's maximum;
  * POWERfft5 (NGP bins): mode counts exact, P(k) and k to 1e-6 relative (the
    reference takes |k| with a REAL*4 sqrt and squares REAL*4 amplitudes);
    with the CIC k-space bins and the log bins as well: counts equal after
    the reference's REAL*4 storage of them; the quadrupole to 1e-5 of the
    largest monopole bin (it is a difference of REAL*4-rounded terms).
Each bar is orders of magnitude below what a wrong weight, index, sign,
Green's function, normalisation or bin rule produces (checked below by
planting such mistakes in the oracle's inputs).
"""
import os

import numpy as np
import pytest

import oracle as O
import pm as P
from conftest import GOLDEN


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def test_deposit_matches_reference_densit():
    f = load("pm_deposit")
    N = int(f["N"][0])
    fi = P.cic_deposit(f["x"], f["y"], f["z"], N)
    ref = f["fi"].astype(np.float64)
    assert np.max(np.abs(fi - ref)) < 2e-6
    # the fixture's particle count makes WPAR inexact: the REAL*4 WPAR is
    # what the reference used
    assert P.wpar(N, f["x"].size) == float(np.float32(N ** 3 / f["x"].size))
    # a planted mistake (x and y weights swapped) is far outside the bar
    bad = P.cic_deposit(f["y"], f["x"], f["z"], N)
    assert np.max(np.abs(bad - ref)) > 1e-2


@pytest.mark.parametrize("N", [16, 12])
def test_poisson_matches_reference_potentfft5(N):
    f = load(f"pm_potential_{N}")
    _, box, om, oml, a = [float(v) for v in f["params"]]
    trfi = P.trfi_of(om, a)
    phi = P.poisson_fft(f["delta"], trfi)
    ref = f["phi"].astype(np.float64)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(phi - ref)) < 1e-6 * scale
    # the REAL*4 result the device stores: within one REAL*4 rounding of the
    # reference's per-pass rounding
    assert np.max(np.abs(P.potential_fi(f["delta"], trfi).astype(np.float64) - ref)) < 2e-6 * scale
    # planted mistake: the continuous -k^2 Green's function (the commented-out
    # alternative at PMP2main.f90:817) instead of the discrete Laplacian
    m = np.fft.fftfreq(N, 1.0 / N)
    kz, ky, kx = np.meshgrid(m, m, m[: N // 2 + 1] if N % 2 == 0 else m, indexing="ij")
    k2 = -(2 * np.pi / N) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2)
    k2[0, 0, 0] = 1.0
    Fk = np.fft.rfftn(np.asarray(f["delta"], np.float64)) * trfi / k2[:, :, : N // 2 + 1]
    wrong = np.fft.irfftn(Fk, s=(N, N, N), axes=(0, 1, 2))
    assert np.max(np.abs(wrong - ref)) > 1e-2 * scale


@pytest.mark.parametrize("tag", ["fr_n0", "fr_n1", "fr_n2", "dgp"])
def test_potent_for_mg_matches_reference_bitwise(tag):
    f = load(f"pm_potent_for_mg_{tag}")
    N, model, frn, box, om, oml, a, fr0, h0rc = [float(v) for v in f["params"]]
    p = O.Params(N=int(N), model=int(model), fr_n=int(frn), box=box, om=om, oml=oml, a=a, fr0=fr0, h0rc=h0rc)
    out = P.potent_for_mg(p, f["fi_in"], f["fi2"])
    assert np.array_equal(out.view(np.int32), f["fi"].view(np.int32))
    # fifth_force_potential (psi from an FP64 field, the device's input)
    # agrees with the reference's sf to the REAL*4 power's rounding
    psi = O.fifth_force_potential(p, f["fi2"].astype(np.float64))
    sf = f["fi"].astype(np.float64) - f["fi_in"].astype(np.float64)
    tol = 4 * np.finfo(np.float32).eps * (np.abs(f["fi"]).astype(np.float64) + np.abs(psi))
    assert np.all(np.abs(psi - sf) <= tol)
    # planted mistake: the exponent n instead of n + 1 (f(R)) / phi instead of phi/2
    wrong = p.with_(fr_n=(p.fr_n + 1) % 3) if p.model == 1 else p
    psi_w = O.fifth_force_potential(wrong, f["fi2"].astype(np.float64))
    if p.model == 2:
        psi_w = 2.0 * psi_w
    assert np.max(np.abs(psi_w - sf)) > 1e-3 * np.max(np.abs(sf))


@pytest.mark.parametrize("tag", ["f1", "f12"])
def test_move_matches_reference_bitwise(tag):
    f = load(f"pm_move_{tag}")
    N, box, om, oml, a, astep, factv = [float(v) for v in f["params"]]
    pc, xc = P.pconst_xconst(a, astep, om, oml, factv)
    r = P.move(f["phi"], f["x_in"], f["y_in"], f["z_in"], f["vx_in"], f["vy_in"], f["vz_in"], pc, xc)
    for i, nm in enumerate(("x", "y", "z", "vx", "vy", "vz")):
        assert np.array_equal(r[i].view(np.int32), f[nm].view(np.int32)), nm
    # ENKIN = SVEL/2./(AEXPN+ASTEP/2.)**2, stored in REAL*4 (PMP2main.f90:689)
    ah = np.float32(np.float32(a) + np.float32(astep) / np.float32(2.0))
    enkin = np.float32(r[6] / 2.0 / float(ah * ah))
    assert enkin == np.float32(f["enkin"][0])
    # planted mistake: gradients differenced in FP64 change some last bits;
    # PCONST from FP64 arithmetic as well
    pc64 = -np.sqrt(a / (om + oml * a ** 3)) * astep * 0.5 / factv
    r64 = P.move(f["phi"].astype(np.float32), f["x_in"], f["y_in"], f["z_in"], f["vx_in"], f["vy_in"], f["vz_in"],
                 pc64, xc)
    assert not np.array_equal(r64[3], f["vx"])


@pytest.mark.parametrize("N", [16, 12])
def test_power_spectrum_matches_reference_powerfft5(N):
    f = load(f"pm_power_{N}")
    box = float(f["params"][1])
    k, Pk, cnt = O.power_spectrum(f["field"], box)
    assert np.array_equal(cnt, f["cnt_p0_c0"])
    assert np.max(np.abs(Pk / f["pk_p0_c0"] - 1.0)) < 1e-6
    assert np.max(np.abs(k / f["k_p0_c0"] - 1.0)) < 1e-6
    # planted mistake: floor instead of round for the bin (INT(RK) for
    # INT(RK + 0.5)) changes the mode counts of the fixture's bins
    m = O.kgrid(N)
    kz, ky, kx = np.meshgrid(m, m, m, indexing="ij")
    rk = np.sqrt(kx * kx + ky * ky + kz * kz).ravel()
    floor_bins = np.bincount(np.maximum(rk[rk >= 0.5].astype(np.int64), 1), minlength=N + 1)[1: N // 2 + 1]
    assert not np.array_equal(floor_bins.astype(np.float64), f["cnt_p0_c0"])


def _f32(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def _rel(a, b):
    m = b != 0
    return float(np.max(np.abs(a[m] / b[m] - 1.0))) if m.any() else 0.0


@pytest.mark.parametrize("N", [16, 12])
@pytest.mark.parametrize("ipol,icic", [(0, 0), (0, 1), (1, 0)])
def test_power_spectrum_options_match_reference_powerfft5(N, ipol, icic):
    """POWERfft5's NGP and CIC k-space bins, monopole and quadrupole, and the
    log bins (PMP2mod_power.f90:366-456)."""
    f = load(f"pm_power_{N}")
    box = float(f["params"][1])
    t = f"p{ipol}_c{icic}"
    k, Pk, cnt, klog, plog, cntlog = O.power_spectrum_ex(f["field"], box, ipol, icic)
    assert np.array_equal(_f32(cnt), f["cnt_" + t])
    assert np.array_equal(_f32(cntlog), f["cntlog_" + t])
    assert _rel(k, f["k_" + t]) < 1e-6 and _rel(klog, f["klog_" + t]) < 1e-6
    assert np.array_equal(Pk == 0, f["pk_" + t] == 0) and np.array_equal(plog == 0, f["pklog_" + t] == 0)
    if ipol == 0:
        assert _rel(Pk, f["pk_" + t]) < 1e-6 and _rel(plog, f["pklog_" + t]) < 1e-6
    else:
        scale = float(np.max(f["pk_p0_c0"]))
        assert np.max(np.abs(Pk - f["pk_" + t])) < 1e-5 * scale
        assert not np.any(plog)  # the log bins take no quadrupole power (:398-402)


def test_power_spectrum_planted_mistakes_fail_the_reference():
    f = load("pm_power_16")
    box = float(f["params"][1])
    fld = f["field"].astype(np.float64)
    scale = float(np.max(f["pk_p0_c0"]))
    # the quadrupole about x (the original first axis) instead of z
    _, Pq, *_ = O.power_spectrum_ex(np.transpose(fld, (2, 1, 0)), box, 1, 0)
    assert np.max(np.abs(Pq - f["pk_p1_c0"])) > 1e-2 * scale
