"""Pins of the PM-step oracle (oracle/pm.py) by what the mathematics fixes:
each check is chosen so that a plausible slip (a swapped weight, a shifted
neighbour, a dropped wrap, a wrong eigenvalue) fails it."""
import numpy as np
import pytest

import pm as P


def test_deposit_conserves_mass_and_is_zero_on_a_lattice():
    N = 8
    rng = np.random.default_rng(3)
    x, y, z = (rng.uniform(1.0, N + 1.0, 500).astype(np.float32) for _ in range(3))
    d = P.cic_deposit(x, y, z, N)
    # each particle deposits WPAR, which the reference holds in REAL*4: the
    # total is npart * WPAR, within a REAL*4 rounding of N^3
    assert abs(np.sum(d + 1.0) - x.size * P.wpar(N, x.size)) <= 1e-9 * N ** 3
    assert abs(x.size * P.wpar(N, x.size) - N ** 3) <= 6e-8 * N ** 3
    # one particle on every node: delta = 0 exactly
    g = np.arange(1, N + 1, dtype=np.float32)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    d = P.cic_deposit(X.ravel(), Y.ravel(), Z.ravel(), N)
    assert np.array_equal(d, np.zeros((N, N, N)))


def test_deposit_weights_of_one_particle():
    N = 6
    # X = 3.25 -> cell 3 (0-based 2) 0.75, cell 4 0.25; Y = 2.5 -> 1/2, 1/2;
    # Z = 5.75 -> cell 5 (0-based 4) 0.25, cell 6 (0-based 5) 0.75
    d = P.cic_deposit([3.25], [2.5], [5.75], N) + 1.0
    w = N ** 3
    expect = np.zeros((N, N, N))
    for (k, wz) in ((4, 0.25), (5, 0.75)):
        for (j, wy) in ((1, 0.5), (2, 0.5)):
            for (i, wx) in ((2, 0.75), (3, 0.25)):
                expect[k, j, i] = w * wx * wy * wz
    assert np.allclose(d, expect, rtol=1e-15, atol=0)


def test_deposit_wraps_at_the_upper_edge():
    N = 4
    d = P.cic_deposit([4.5], [1.0], [1.0], N) + 1.0
    assert d[0, 0, 3] == pytest.approx(0.5 * N ** 3) and d[0, 0, 0] == pytest.approx(0.5 * N ** 3)
    assert np.count_nonzero(d) == 2
    with pytest.raises(ValueError):
        P.cic_deposit([5.0], [1.0], [1.0], N)


def test_poisson_single_mode_closed_form():
    N = 16
    i = np.arange(N)
    K, J, I = np.meshgrid(i, i, i, indexing="ij")
    for m in ((1, 0, 0), (2, 3, 1), (0, 5, 7)):
        delta = np.cos(2 * np.pi * (m[0] * I + m[1] * J + m[2] * K) / N)
        # the eigenvalue with the reference's P16 (the REAL*4 value of 2 pi)
        A = sum(2 * np.cos(P.P16 * mm / N) - 2.0 for mm in m)
        phi = P.poisson_fft(delta, 1.7)
        assert np.allclose(phi, 1.7 * delta / A, rtol=0, atol=1e-10 * np.max(np.abs(delta / A)))


def test_poisson_inverts_the_discrete_laplacian():
    N = 16
    rng = np.random.default_rng(5)
    delta = rng.standard_normal((N, N, N))
    trfi = 0.45
    phi = P.poisson_fft(delta, trfi)
    lap = P.laplacian7(phi)
    # exact up to the reference's REAL*4 P16 in the Green's function (a
    # relative 3e-8 change of 2 pi)
    assert np.allclose(lap, trfi * (delta - delta.mean()), rtol=0, atol=1e-6)
    assert P.P16 == float(np.float32(2 * np.pi)) and abs(P.P16 / (2 * np.pi) - 1) < 5e-8
    assert phi.mean() == pytest.approx(trfi * delta.mean(), rel=1e-9, abs=1e-14)


def test_move_node_gradient_of_a_sinusoid():
    """phi = a sin(2 pi i / N): at a node the doubled central difference is
    2 a sin(2 pi / N) cos(2 pi i / N); halfway between two nodes the mean."""
    N = 16
    a = 0.3
    i = np.arange(N)
    phi = np.broadcast_to(a * np.sin(2 * np.pi * i / N), (N, N, N)).copy()
    pc, xc = -0.25, 0.0
    x = np.array([1.0, 5.0, 5.5, 16.0], np.float32)
    y = np.full(4, 3.0, np.float32)
    z = np.full(4, 7.0, np.float32)
    zero = np.zeros(4, np.float32)
    _, _, _, vx, vy, vz, _ = P.move(phi, x, y, z, zero, zero, zero, pc, xc)
    node = lambda c: 2 * a * np.sin(2 * np.pi / N) * np.cos(2 * np.pi * c / N)
    expect = pc * np.array([node(0), node(4), 0.5 * (node(4) + node(5)), node(15)])
    assert np.allclose(vx, expect.astype(np.float32), rtol=1e-6, atol=1e-9)
    assert np.all(vy == 0) and np.all(vz == 0)


def test_move_drift_wrap_edge_rule_and_svel():
    N = 8
    phi = np.zeros((N, N, N))
    x = np.array([1.5, 8.9, 2.0, 8.999999], np.float32)
    one = np.ones(4, np.float32)
    vx = np.array([-1.0, 0.5, 0.0, 0.0], np.float32)
    nx, ny, nz, nvx, _, _, svel = P.move(phi, x, one, one, vx, 0 * one, 0 * one, -0.5, 1.0)
    # 1.5 - 1 = 0.5 -> +8 = 8.5; REAL*4 8.9 + 0.5 -> -8
    assert nx[0] == np.float32(8.5) and nx[1] == np.float32(float(np.float32(8.9)) + 0.5 - 8.0)
    assert nx[2] == np.float32(2.0) and np.array_equal(nvx, vx)
    # 8.999999 in REAL*4 stays below 9 (N+1): no change
    assert nx[3] == np.float32(8.999999)
    assert svel == pytest.approx((1.0 + 0.25) * N ** 3 / 4)
    assert np.array_equal(ny, one) and np.array_equal(nz, one)
    # 8.5 + 0.5 * 0.9999998 = 8.9999999 < N+1 in FP64 rounds to 9 in REAL*4:
    # pulled back by 1e-3
    ex, _, _, _, _, _, _ = P.move(phi, np.float32([8.5]), np.float32([1.0]), np.float32([1.0]), np.float32([0.5]),
                                  np.float32([0.0]), np.float32([0.0]), 0.0, 0.9999998)
    assert ex[0] == np.float32(np.float32(9.0) - np.float32(1.0e-3))


def test_move_conserves_momentum_of_a_pair():
    """Two particles' gradients from the potential of their own deposit are
    equal and opposite (CIC gather is the transpose of CIC scatter)."""
    N = 16
    x = np.array([4.3, 9.8], np.float32)
    y = np.array([7.1, 6.2], np.float32)
    z = np.array([2.6, 3.9], np.float32)
    d = P.cic_deposit(x, y, z, N)
    phi = P.poisson_fft(d, 1.0)
    zero = np.zeros(2, np.float32)
    _, _, _, vx, vy, vz, _ = P.move(phi, x, y, z, zero, zero, zero, 1.0, 0.0)
    # self-forces vanish only on average for CIC; the pair sum is small against each force
    for v in (vx, vy, vz):
        assert abs(float(v[0]) + float(v[1])) <= 0.1 * max(abs(float(v[0])), abs(float(v[1])), 1e-12) + 1e-6
