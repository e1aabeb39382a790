"""Pins of the oracle's k-point Hamiltonian (oracle/kpoint.py)."""
import math

import numpy as np
import pytest

import oracle as O
import qumasun_amd as Q

N = (12, 10, 14)
H = (0.35, 0.4, 0.3)


def _cell(seed=3):
    rng = np.random.default_rng(seed)
    L = [N[d] * H[d] for d in range(3)]
    pos = np.array([[0.3 * L[0], 0.4 * L[1], 0.2 * L[2]], [0.8 * L[0], 0.7 * L[1], 0.9 * L[2]]])
    projs = [[Q.gaussian_projector(0, 1.0, 0.5, 1.6), Q.gaussian_projector(1, -0.5, 0.5, 1.6)]]
    V = rng.uniform(-0.5, 0.5, N[0] * N[1] * N[2])
    return pos, projs, V


def test_stencil_plane_wave_symbol():
    k = (0.31, -0.2, 0.45)
    z, y, x = np.meshgrid(np.arange(N[2]), np.arange(N[1]), np.arange(N[0]), indexing="ij")
    for m in ((1, 0, 0), (2, -3, 1), (0, 4, -5)):
        th = [2 * math.pi * m[d] / N[d] for d in range(3)]
        u = np.exp(1j * (th[0] * x + th[1] * y + th[2] * z)).reshape(-1)
        s2 = sum((-14350.0 + 2 * sum(c * math.cos((i + 1) * th[d]) for i, c in enumerate(O.kpoint._C2)))
                 / (5040.0 * H[d] ** 2) for d in range(3))
        s1 = [2 * sum(w * math.sin((i + 1) * th[d]) for i, w in enumerate(O.kpoint._C1)) / (840.0 * H[d])
              for d in range(3)]
        sym = -0.5 * (s2 - sum(kk * kk for kk in k) - 2 * sum(k[d] * s1[d] for d in range(3)))
        assert np.allclose(O.kpoint_stencil(u, N, H, k), sym * u, atol=1e-10)


def test_gamma_point_is_the_real_hamiltonian():
    pos, projs, V = _cell()
    onl = O.OracleNL(N, H, pos, np.zeros(2, np.int32), projs)
    rng = np.random.default_rng(4)
    ur, ui = rng.standard_normal((2, 3, N[0] * N[1] * N[2]))
    out = O.kpoint_hamiltonian(N, H, V, onl, (0.0, 0.0, 0.0), ur + 1j * ui)
    ref = O.hamiltonian_block(N, H, V, onl, ur) + 1j * O.hamiltonian_block(N, H, V, onl, ui)
    assert np.abs(out - ref).max() <= 1e-10 * np.abs(ref).max()


def test_projector_columns_brute_force():
    """Every column against sum over images T of beta(x - T - R) exp(i k.(x - T)), with the
    Gaussian radial part and the real Ylm written out (s: 1/sqrt(4 pi); p: sqrt(3/4pi)(y, z, x)/r)."""
    pos, projs, _ = _cell()
    k = (0.4, -0.25, 0.3)
    onl = O.OracleNL(N, H, pos, np.zeros(2, np.int32), projs)
    B = onl.dense_beta_k(k)
    L = np.array([N[d] * H[d] for d in range(3)])
    z, y, x = np.meshgrid(np.arange(N[2]) * H[2], np.arange(N[1]) * H[1], np.arange(N[0]) * H[0], indexing="ij")
    X = np.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], 1)
    sig, rc = 0.5, 1.6
    amp = math.sqrt(4.0 / (math.sqrt(math.pi) * sig ** 3))
    j = 0
    for a in range(2):
        cols = np.zeros((4, X.shape[0]), np.complex128)
        for T in np.ndindex(5, 5, 5):
            t = (np.array(T) - 2) * L
            xu = X - t
            d = xu - pos[a]
            r2 = (d * d).sum(1)
            inside = r2 <= rc * rc
            r = np.sqrt(np.maximum(r2, 1e-300))
            g = amp * np.exp(-r2 / (2 * sig * sig)) * inside
            ph = np.exp(1j * (xu @ np.array(k)))
            c1 = math.sqrt(3.0 / (4 * math.pi))
            small = r2 < 1e-28
            cols[0] += g / math.sqrt(4 * math.pi) * ph
            for m, comp in enumerate((1, 2, 0)):
                yl = np.where(small, 0.0, c1 * d[:, comp] / r)
                cols[1 + m] += g * yl * ph
        for c in range(4):
            assert np.abs(B[j] - cols[c]).max() <= 1e-12 * max(1.0, np.abs(cols[c]).max()), (a, c)
            j += 1


def test_dense_hermitian_and_time_reversal():
    pos, projs, V = _cell()
    onl = O.OracleNL(N, H, pos, np.zeros(2, np.int32), projs)
    k = (0.5, 0.2, -0.35)
    Hk = O.kpoint_dense(N, H, V, onl, k)
    Hm = O.kpoint_dense(N, H, V, onl, tuple(-c for c in k))
    assert np.abs(Hk - Hk.conj().T).max() <= 1e-10 * np.abs(Hk).max()
    assert np.abs(Hm - Hk.conj()).max() <= 1e-10 * np.abs(Hk).max()
    e1 = np.linalg.eigvalsh(Hk)[:8]
    e2 = np.linalg.eigvalsh(Hm)[:8]
    assert np.allclose(e1, e2, atol=1e-9)
