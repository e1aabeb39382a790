"""Pins of the oracle's SCF-step restatement (oracle/scf.py) by what the mathematics fixes:
Fermi-Dirac occupancies sum to N and reduce to the aufbau filling at small kbT; the density of
normalised states integrates to the occupancy sum; the FFT Poisson solve is exact on single
Fourier modes, linear, and blind to the mean; the LDA exchange is Dirac's, and V = d(rho e)/d rho
holds on both PZ81 correlation branches, which join continuously at rs = 1."""
import math

import numpy as np
import pytest

import oracle as O


def test_occupancy_sums_and_aufbau():
    rng = np.random.default_rng(3)
    eig = np.sort(rng.uniform(-1.0, 0.5, 24))
    for ne, kbT in ((16.0, 0.01), (31.0, 0.05), (8.0, 1e-3)):
        occ, mu = O.occupancy_fermi(eig, ne, kbT)
        assert abs(occ.sum() - ne) < 1e-6
        assert np.all(np.diff(occ) <= 1e-15) and occ.max() <= 2.0
    eig = np.arange(12) * 0.1
    occ, mu = O.occupancy_fermi(eig, 10.0, 1e-4)
    assert np.allclose(occ, [2] * 5 + [0] * 7, atol=1e-12)
    assert 0.4 < mu < 0.5


def test_density_integrates_to_occupancy():
    rng = np.random.default_rng(4)
    ng, dv = 1000, 0.01
    psi = rng.standard_normal((6, ng))
    psi /= np.sqrt(np.sum(psi * psi, axis=1, keepdims=True) * dv)
    occ = np.array([2.0, 2.0, 1.5, 0.5, 0.0, -0.0])
    rho = O.density(psi, occ)
    assert abs(rho.sum() * dv - occ.sum()) < 1e-12
    mixed, diff = O.density_mix(rho, np.zeros(ng), 0.3, dv)
    assert np.allclose(mixed, 0.3 * rho) and abs(diff - rho.sum() * dv) < 1e-12


@pytest.mark.parametrize("n,h", [((12, 10, 8), (0.3, 0.35, 0.4)), ((16, 16, 16), (0.25, 0.25, 0.25))])
def test_hartree_single_modes(n, h):
    nx, ny, nz = n
    L = [n[d] * h[d] for d in range(3)]
    z, y, x = np.meshgrid(np.arange(nz) * h[2], np.arange(ny) * h[1], np.arange(nx) * h[0], indexing="ij")
    for m in ((1, 0, 0), (0, 2, 1), (3, -1, 2), (nx // 2 - 1, 1, -2)):
        k = [2 * math.pi * m[d] / L[d] for d in range(3)]
        rho = np.cos(k[0] * x + k[1] * y + k[2] * z).reshape(-1)
        v = O.hartree(rho + 0.7, n, h)          # the mean does not enter
        assert np.allclose(v, 4 * math.pi / (k[0] ** 2 + k[1] ** 2 + k[2] ** 2) * rho, atol=1e-11)
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal(nx * ny * nz), rng.standard_normal(nx * ny * nz)
    assert np.allclose(O.hartree(2 * a - b, n, h), 2 * O.hartree(a, n, h) - O.hartree(b, n, h), atol=1e-10)


def test_lda_exchange_derivative_and_continuity():
    rho = np.logspace(-6, 2, 400)
    vx, vc, ex, ec = O.xc_lda(rho)
    dirac = -0.75 * (3.0 / math.pi) ** (1.0 / 3.0) * rho ** (1.0 / 3.0)
    assert np.allclose(ex, dirac, rtol=1e-13)
    # V = d(rho e)/d rho, central differences in log rho (both correlation branches)
    eps = 1e-5
    for e_fn, v in ((lambda r: O.xc_lda(r)[2], vx), (lambda r: O.xc_lda(r)[3], vc)):
        rp, rm = rho * (1 + eps), rho * (1 - eps)
        d = (rp * e_fn(rp) - rm * e_fn(rm)) / (rp - rm)
        sel = np.abs(np.log(0.6203504908994 / np.cbrt(rho))) > 1e-3     # away from the rs = 1 seam
        assert np.allclose(d[sel], v[sel], rtol=1e-7, atol=1e-12)
    # the rs >= 1 and rs < 1 parameter sets join at rs = 1
    rho1 = 3.0 / (4.0 * math.pi)
    lo = O.xc_lda(np.array([rho1 * (1 + 1e-9)]))
    hi = O.xc_lda(np.array([rho1 * (1 - 1e-9)]))
    assert abs(lo[3][0] - hi[3][0]) < 5e-5 and abs(lo[1][0] - hi[1][0]) < 5e-4
    # below 1e-8 everything is zero
    z = O.xc_lda(np.array([0.0, 5e-9]))
    assert all(np.all(a == 0.0) for a in z)
