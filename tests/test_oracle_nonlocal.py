"""CPU pins of the oracle's nonlocal KB projector operator
(PseudoPotOperator_mpi_v2.h:89-202, :336-570; GridSubgrid.h:381-492; SetBlockYlm.h:10-53).

1. Dense brute force: beta_j(r) = p(r) Y_lm(r^) is rebuilt for every grid point and every
   periodic image within the cutoff, with the real spherical harmonics taken from scipy's
   complex Y_l^m (standard real combination), not from the oracle's formulas.
2. Normalisation: a projector applied to itself gives int beta^2 = 1 (the radial part is
   normalised so that int p^2 r^2 dr = 1), up to grid quadrature.
3. The apply equals B diag(E) B^T psi dV, and the operator is symmetric.
4. GetValueByXi (RadialGrid2.h:81-136) is exact on polynomials of the degree each branch
   reproduces: 1 (outer linear intervals), 2 (cubic Hermite with centred slopes) and 4
   (quintic Hermite with 4th-order derivative estimates).
"""
import math

import numpy as np
import pytest
from scipy.special import sph_harm_y

import oracle as O
import qumasun_amd as Q


def real_ylm(l, m, theta, phi):
    """Real spherical harmonics from scipy's complex ones (no Condon-Shortley sign left)."""
    if m == 0:
        return np.real(sph_harm_y(l, 0, theta, phi))
    y = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return math.sqrt(2) * (-1) ** m * np.real(y)
    return math.sqrt(2) * (-1) ** m * np.imag(y)


def dense_beta_bruteforce(n, h, positions, projs):
    ng = n[0] * n[1] * n[2]
    iz, iy, ix = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    pts = np.stack([ix.ravel() * h[0], iy.ravel() * h[1], iz.ravel() * h[2]], axis=1)
    L = np.array([n[0] * h[0], n[1] * h[1], n[2] * h[2]])
    cols = []
    for R in positions:
        for pr in projs:
            l, rc, sig, amp = pr["l"], pr["cutoff"], pr["sigma"], pr["amplitude"]
            for m in ([0] if l == 0 else [-1, 0, 1]):
                col = np.zeros(ng)
                for sx in (-2, -1, 0, 1, 2):
                    for sy in (-2, -1, 0, 1, 2):
                        for sz in (-2, -1, 0, 1, 2):
                            d = pts - (R + np.array([sx, sy, sz]) * L)
                            r2 = (d ** 2).sum(1)
                            inside = r2 <= rc * rc
                            r = np.sqrt(r2[inside])
                            rad = amp * np.exp(-r2[inside] / (2 * sig * sig))
                            if l == 0:
                                y = np.full(r.shape, 1 / math.sqrt(4 * math.pi))
                            else:
                                with np.errstate(invalid="ignore", divide="ignore"):
                                    th = np.arccos(np.clip(d[inside, 2] / r, -1, 1))
                                    ph = np.arctan2(d[inside, 1], d[inside, 0])
                                y = real_ylm(1, m, th, ph)
                                y[r < 1e-7] = 0.0
                            col[inside] += rad * y
                cols.append(col)
    return np.array(cols)


@pytest.mark.parametrize("n,h,natom", [((14, 12, 13), (0.3, 0.32, 0.35), 3), ((10, 10, 10), (0.26, 0.26, 0.26), 2)])
def test_dense_beta_bruteforce(n, h, natom):
    rng = np.random.default_rng(7)
    L = np.array([n[0] * h[0], n[1] * h[1], n[2] * h[2]])
    pos = rng.random((natom, 3)) * L
    pos[0] = [2 * h[0], 3 * h[1], 1 * h[2]]      # one nucleus on a grid point (r = 0 case)
    projs = [Q.gaussian_projector(0, 1.0, 0.5, 1.9), Q.gaussian_projector(1, -0.5, 0.5, 1.9)]
    onl = O.OracleNL(n, h, pos, np.zeros(natom, np.int32), [projs])
    B = onl.dense_beta()
    Bref = dense_beta_bruteforce(n, h, pos, projs)
    assert B.shape == Bref.shape
    assert np.abs(B - Bref).max() < 1e-12 * np.abs(Bref).max()
    # apply = B diag(E) B^T psi dV, symmetric
    dV = h[0] * h[1] * h[2]
    E = onl.energies()
    assert np.allclose(E, np.tile([1.0, -0.5, -0.5, -0.5], natom))
    psi = rng.standard_normal(B.shape[1])
    hp = np.zeros_like(psi)
    onl.apply(psi, hp)
    ref = B.T @ (E * (B @ psi)) * dV
    assert np.abs(hp - ref).max() < 1e-12 * np.abs(ref).max()
    c = onl.project(psi)
    assert np.abs(c - B @ psi * dV).max() < 1e-12 * np.abs(c).max()
    u = rng.standard_normal(B.shape[1])
    hu = np.zeros_like(u)
    onl.apply(u, hu)
    assert abs(u @ hp - hu @ psi) < 1e-10 * abs(u @ hp)


def test_projector_normalisation():
    """Single nucleus, fine grid, cutoff far out: <beta_j | beta_j> = 1 within quadrature."""
    n = (48, 48, 48)
    h = (0.1,) * 3
    pos = np.array([[2.4, 2.4, 2.4]])
    projs = [Q.gaussian_projector(0, 1.0, 0.5, 2.3), Q.gaussian_projector(1, -0.5, 0.5, 2.3)]
    onl = O.OracleNL(n, h, pos, np.zeros(1, np.int32), [projs])
    B = onl.dense_beta()
    for j in range(4):
        c = onl.project(B[j].copy())
        assert abs(c[j] - 1.0) < 2e-3, (j, c[j])
        others = np.delete(c, j)
        assert np.abs(others).max() < 1e-10


@pytest.mark.parametrize("deg,branch_i", [(1, 0), (2, 1), (4, 5), (4, 9)])
def test_radial_by_xi_polynomial_exactness(deg, branch_i):
    nt = 16
    xi_min, dxi = -3.0, 0.25
    rng = np.random.default_rng(deg)
    coef = rng.standard_normal(deg + 1)
    poly = lambda x: sum(c * x ** k for k, c in enumerate(coef))
    table = np.array([poly(xi_min + k * dxi) for k in range(nt)])
    for t in (0.1, 0.5, 0.83):
        xi = xi_min + (branch_i + t) * dxi
        got = O.radial_by_xi(xi, table, xi_min, dxi)
        assert abs(got - poly(xi)) < 1e-10 * max(1.0, abs(poly(xi))), (deg, branch_i, t)
    # beyond the table: zero
    assert O.radial_by_xi(xi_min + (nt - 1 + 0.5) * dxi, table, xi_min, dxi) == 0.0
