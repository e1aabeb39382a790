"""Pins of the oracle's nonlocal energy and force restatement (oracle/force.py): the 8th-order
gradient's plane-wave symbol and antisymmetry, and the forces against minus the central
difference of the oracle's energy in each nucleus coordinate (smooth states and Gaussian
projectors whose cutoff truncates below 1e-7 of their peak and that are smooth at the nucleus,
so the energy is smooth in the positions)."""
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n,h", [((16, 12, 10), (0.3, 0.35, 0.4)), ((20, 20, 20), (0.25, 0.25, 0.25))])
def test_gradient8_symbol_and_antisymmetry(n, h):
    nx, ny, nz = n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    w = (672.0, -168.0, 32.0, -3.0)
    for m in ((1, 0, 0), (0, 2, 1), (3, -1, 2)):
        th = [2 * math.pi * m[d] / n[d] for d in range(3)]
        u = np.sin(th[0] * x + th[1] * y + th[2] * z).reshape(-1)
        cu = np.cos(th[0] * x + th[1] * y + th[2] * z).reshape(-1)
        g = O.gradient8(u, n, h, coef=1.5)
        for d in range(3):
            sym = sum(2 * wk * math.sin((k + 1) * th[d]) for k, wk in enumerate(w)) / (840.0 * h[d])
            assert np.allclose(g[d], 1.5 * sym * cu, atol=1e-12)
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal(nx * ny * nz), rng.standard_normal(nx * ny * nz)
    for ga, gb in zip(O.gradient8(a, n, h), O.gradient8(b, n, h)):
        assert abs(np.dot(ga, b) + np.dot(a, gb)) < 1e-9 * np.linalg.norm(ga) * np.linalg.norm(b)


def test_forces_are_minus_the_energy_derivative():
    n = (24, 24, 24)
    L = 6.0
    h = (L / n[0],) * 3
    sig, rc = 0.5, 3.0          # exp(-rc^2 / 2 sig^2) = 1.5e-8
    amp = math.sqrt(4.0 / (math.sqrt(math.pi) * sig ** 3))
    # the p channel as a table of r exp(-r^2 / 2 sig^2): the product with Y_1m is smooth at the
    # nucleus, so the grid sums behave like the integrals the identity F = -dE/dR rests on
    xi = np.linspace(math.log(1e-5), math.log(4.0), 4000)
    r = np.exp(xi)
    ptab = np.ascontiguousarray(2.0 * amp * r * np.exp(-r * r / (2 * sig * sig)))
    projs = [[dict(l=0, energy=1.0, cutoff=rc, kind="gaussian", sigma=sig, amplitude=amp),
              dict(l=1, energy=-0.5, cutoff=rc, kind="table", values=ptab, xi_min=xi[0],
                   xi_delta=xi[1] - xi[0], r_min=r[0])]]
    pos = np.array([[1.3, 2.1, 2.9], [4.2, 3.7, 1.1]])
    species = np.zeros(2, np.int32)
    nx, ny, nz = n
    zz, yy, xx = np.meshgrid(np.arange(nz) * h[2], np.arange(ny) * h[1], np.arange(nx) * h[0], indexing="ij")
    rng = np.random.default_rng(8)
    psi = []
    for _ in range(3):     # smooth states: plane waves with k h <= 0.27, where the 8th-order
                           # difference is exact to (k h)^8 / 630 < 1e-7
        f = np.zeros_like(xx)
        for _ in range(5):
            k = 2 * math.pi / L * rng.integers(-1, 2, 3)
            f += rng.standard_normal() * np.cos(k[0] * xx + k[1] * yy + k[2] * zz + rng.uniform(0, 6.28))
        psi.append(f.reshape(-1))
    psi = np.array(psi)
    occ = np.array([2.0, 1.2, 0.0])
    onl = O.OracleNL(n, h, pos, species, projs)
    e0, F = O.nonlocal_energy_force(n, h, onl, psi, occ, species, projs)
    assert e0 != 0.0
    delta = 1e-4
    for a in range(2):
        for ax in range(3):
            pe = []
            for sgn in (1, -1):
                p2 = pos.copy()
                p2[a, ax] += sgn * delta
                e, _ = O.nonlocal_energy_force(n, h, O.OracleNL(n, h, p2, species, projs), psi, occ, species, projs)
                pe.append(e)
            fd = -(pe[0] - pe[1]) / (2 * delta)
            assert abs(fd - F[a, ax]) <= 2e-6 * max(1.0, np.abs(F).max()), (a, ax, fd, F[a, ax])
