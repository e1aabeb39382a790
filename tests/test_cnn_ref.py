"""CPU pins of the learned preconditioner's CPU reference (tests/cnn_ref.py): its convolution equals
a brute-force periodic sum on a tiny grid, its adjoint satisfies <N a, b> = <a, N^T b>, and
its kinetic operator matches the 8th-order stencil's plane-wave symbol."""
import numpy as np
import torch

import cnn_ref as R


def test_conv_brute_force():
    rng = np.random.default_rng(0)
    n = (5, 4, 6)   # (Nz, Ny, Nx)
    x = rng.standard_normal((1, 2) + n)
    W = rng.standard_normal((3, 2, 3, 3, 3))
    got = R.conv(torch.from_numpy(x), torch.from_numpy(W)).numpy()
    ref = np.zeros((1, 3) + n)
    for co in range(3):
        for z in range(n[0]):
            for y in range(n[1]):
                for xx in range(n[2]):
                    s = 0.0
                    for ci in range(2):
                        for t in range(27):
                            dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                            s += W[co, ci].reshape(27)[t] * x[0, ci, (z + dz) % n[0], (y + dy) % n[1], (xx + dx) % n[2]]
                    ref[0, co, z, y, xx] = s
    assert np.abs(got - ref).max() < 1e-12


def test_adjoint_identity():
    rng = np.random.default_rng(1)
    n = (6, 5, 7)
    P = R.split(torch.from_numpy(rng.standard_normal(R.NPARAM)))
    v = torch.from_numpy(rng.standard_normal((1, 1) + n))
    a = torch.from_numpy(rng.standard_normal((2, 1) + n))
    b = torch.from_numpy(rng.standard_normal((2, 1) + n))
    lhs = float((R.net(P, v, a) * b).sum())
    rhs = float((a * R.net_adj(P, v, b)).sum())
    assert abs(lhs - rhs) <= 1e-10 * max(1.0, abs(lhs))


def test_kinetic_symbol_plane_wave():
    """K = (T + s)^-1 on cos(k.r) is multiplication by 1/(T(k) + s), T from the stencil
    coefficients applied directly to the plane wave."""
    n = (8, 10, 12)
    h = (0.3, 0.25, 0.2)   # (hx, hy, hz) for (x, y, z)
    kv = (2, 3, 1)         # (kx, ky, kz)
    iz, iy, ix = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    f = np.cos(2 * np.pi * (kv[0] * ix / n[2] + kv[1] * iy / n[1] + kv[2] * iz / n[0]))
    # direct application of -1/2 Lap8 (periodic) to f
    c = [-14350 / 5040, 8064 / 5040, -1008 / 5040, 128 / 5040, -9 / 5040]
    lap = np.zeros_like(f)
    for ax, hh in ((2, h[0]), (1, h[1]), (0, h[2])):
        lap += c[0] * f / hh ** 2
        for m in range(1, 5):
            lap += c[m] * (np.roll(f, m, ax) + np.roll(f, -m, ax)) / hh ** 2
    Tf = -0.5 * lap
    Tk = float((Tf * f).sum() / (f * f).sum())
    T = torch.from_numpy(R.kinetic_symbol((n[2], n[1], n[0]), h))
    out = R.kinetic(torch.from_numpy(f[None, None]), T, 0.7).numpy()[0, 0]
    assert np.abs(out - f / (Tk + 0.7)).max() < 1e-12
