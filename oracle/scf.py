"""CPU FP64 restatement of QUMASUN's SCF step around the eigensolver — TEST INFRASTRUCTURE ONLY
(the oracle; only tests/ may use it).

Each function cites the reference lines it restates.  What pins them (tests/test_oracle_scf.py):
  occupancy_fermi  sum = N to the bisection tolerance; kbT -> 0 gives the aufbau filling
  density          int rho dV = sum occ for normalised states
  hartree          closed form: a single Fourier mode rho = cos(k.r) gives V = 4 pi / |k|^2 rho;
                   linearity; V(k = 0) = 0
  xc_lda           e_x = -(3/4)(3/pi)^(1/3) rho^(1/3) (Dirac exchange); V = d(rho e)/d rho checked
                   by finite differences on both PZ81 branches; e_c and V_c continuous across rs = 1
                   (the two parameter sets were fitted to join there)
"""
import math

import numpy as np

__all__ = ["occupancy_fermi", "density", "density_mix", "hartree", "xc_lda"]


def occupancy_fermi(eig, num_electrons, kbT):
    """mSetOccupancySpinOff + SetOccupancy (qumasun_mpi.occupancy.h:61-108, SetOccupancy.h:7-18):
    bisection on mu; the occupancies of the last evaluated mu.  Returns (occ, mu)."""
    eig = np.asarray(eig, dtype=np.float64)
    n = len(eig)
    beta = 1.0 / kbT

    def fill(mu):
        occ = np.empty(n)
        s = 0.0
        for i in range(n):                         # SetOccupancy.h:12-16
            p = min(beta * (eig[i] - mu), 80.0)
            occ[i] = 2.0 / (math.exp(p) + 1.0)
            s += occ[i]
        return occ, s

    mu_min = eig[0]
    mu_max = eig[n - 1]
    mu_max = mu_max * 16.0 if mu_max > 0.0 else abs(mu_max) + 16.0      # :66-68
    if (int(round(num_electrons)) + 1) // 2 >= n:                       # :71-73
        mu_max *= 10.0
    mu = (mu_min + mu_max) / 2.0
    occ, _ = fill(mu)
    while abs(mu_max - mu_min) > 1.0e-9:                                 # :82-98
        occ, ne = fill(mu)
        if ne < num_electrons:
            mu_min = mu
        elif ne > num_electrons:
            mu_max = mu
        else:
            break
        mu = (mu_min + mu_max) / 2.0
    return occ, mu


def density(psi, occ):
    """mSetDensityOne (qumasun_mpi.density.h:10-32): rho = sum_{occ>0} occ psi^2, band order."""
    psi = np.asarray(psi, dtype=np.float64)
    rho = np.zeros(psi.shape[1])
    for b in range(psi.shape[0]):
        if occ[b] > 0.0:
            rho += (psi[b] * psi[b]) * occ[b]
    return rho


def density_mix(rho, prev, ratio, dvol):
    """mSetDensity's mixing (qumasun_mpi.density.h:46-64): returns (mixed, dV sum |rho - prev|)."""
    diff = rho - prev
    return diff * ratio + prev, float(np.sum(np.abs(diff)) * dvol)


def hartree(rho, n, h):
    """SetPotentialByPoissonFFT_keepRhok (poisson_fft.h:131-240) with FFTW's conventions
    (wrap_fft.h:14-38): forward FFT of rho (x fastest), times 4 pi / |k|^2 with k folded,
    k = 0 removed, inverse FFT / N, real part."""
    nx, ny, nz = n
    r = np.asarray(rho, dtype=np.float64).reshape(nz, ny, nx)
    rk = np.fft.fftn(r)
    def kaxis(m, hh):
        c = 2.0 * math.pi / (hh * m)
        idx = np.arange(m)
        return c * np.where(idx < m // 2, idx, idx - m)
    kz = kaxis(nz, h[2])[:, None, None]
    ky = kaxis(ny, h[1])[None, :, None]
    kx = kaxis(nx, h[0])[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    with np.errstate(divide="ignore"):
        f = np.where(k2 > 0.0, 4.0 * math.pi / np.where(k2 > 0.0, k2, 1.0), 0.0)
    return np.real(np.fft.ifftn(rk * f)).reshape(-1)


def xc_lda(rho, pcc=None):
    """Calc_XC_LDA (Vxc.h:46-75) pointwise on rho + pcc: returns (vx, vc, ex, ec)."""
    r = np.asarray(rho, dtype=np.float64) + (0.0 if pcc is None else np.asarray(pcc, dtype=np.float64))
    vx = np.zeros_like(r)
    vc = np.zeros_like(r)
    ex = np.zeros_like(r)
    ec = np.zeros_like(r)
    c2 = (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
    c3 = -(3.0 / (4.0 * math.pi)) * (9.0 * math.pi / 4.0) ** (1.0 / 3.0)
    on = r >= 1.0e-8
    rs = np.where(on, c2 / np.cbrt(np.where(on, r, 1.0)), 1.0)
    e_x = c3 / rs
    hi = rs >= 1.0
    # rs >= 1 (Vxc.h:17-21): gamma / (1 + b1 sqrt(rs) + b2 rs)
    g, b1, b2 = -0.1423, 1.0529, 0.3334
    sq = np.sqrt(rs)
    ec_hi = g / (1.0 + b1 * sq + b2 * rs)
    nd_hi = ec_hi * ec_hi * (b1 * 0.5 * sq + b2 * rs) / (3.0 * g)
    # rs < 1 (Vxc.h:23-27): B + A ln rs + D rs + C rs ln rs
    A, B, C, D = 0.0311, -0.0480, 0.0020, -0.0116
    lr = np.log(rs)
    ec_lo = B + A * lr + D * rs + C * rs * lr
    nd_lo = -(A + (D + C) * rs + C * rs * lr) / 3.0
    e_c = np.where(hi, ec_hi, ec_lo)
    nd = np.where(hi, nd_hi, nd_lo)
    ex[on] = e_x[on]
    ec[on] = e_c[on]
    vx[on] = (e_x + e_x / 3.0)[on]
    vc[on] = (e_c + nd)[on]
    return vx, vc, ex, ec
