"""CPU pins of the oracle's 8th-order Laplacian (GridDifference8th.h:20-185).

1. Bit-for-bit against the reference's own Laplasian8th<double>, compiled from its header
   into oracle/_ref (built when the reference sources are present).
2. Consistency: on a plane wave the periodic operator is diagonal with symbol s(k); the
   8th-order weights must reproduce -k^2 up to O((kh)^8) (only Fornberg's weights do),
   and s(k) must equal the oracle's action on cos(k.r) on odd and even grids.
3. Invariants: constants map to zero; the operator is symmetric; out += accumulates.
"""
import os

import numpy as np
import pytest

import oracle as O

SHAPES = [(12, 10, 9), (16, 16, 16), (7, 5, 6), (40, 40, 40)]


@pytest.fixture(scope="module")
def ref():
    if os.path.isdir(O.REF_SRC) and not os.path.exists(O.REF):
        O.build(with_ref=True)
    r = O.ref_lib()
    if r is None:
        pytest.skip("reference harness not built (reference sources absent)")
    return r


@pytest.mark.parametrize("n", SHAPES)
def test_matches_reference_bitwise(ref, n):
    rng = np.random.default_rng(sum(n))
    ng = n[0] * n[1] * n[2]
    p = rng.standard_normal(ng)
    out0 = rng.standard_normal(ng)
    a = out0.copy()
    b = out0.copy()
    h = (0.21, 0.17, 0.29)
    O.laplacian8(n, h, -0.5, p, a)
    ref.ref_laplacian8(n[0], n[1], n[2], h[0], h[1], h[2], -0.5, O._d(p), O._d(b))
    assert np.array_equal(a, b)


def test_reference_float_instance_close(ref):
    import ctypes
    n = (16, 12, 10)
    rng = np.random.default_rng(5)
    p = rng.standard_normal(n[0] * n[1] * n[2])
    a = np.zeros_like(p)
    O.laplacian8(n, (0.2, 0.2, 0.2), -0.5, p, a)
    pf = p.astype(np.float32)
    bf = np.zeros_like(pf)
    fp = ctypes.POINTER(ctypes.c_float)
    ref.ref_laplacian8_f32(n[0], n[1], n[2], 0.2, 0.2, 0.2, -0.5, pf.ctypes.data_as(fp), bf.ctypes.data_as(fp))
    assert np.linalg.norm(bf - a) / np.linalg.norm(a) < 2e-6


def symbol_1d(theta, h):
    """Second-derivative symbol of the periodic 8th-order central difference on cos(theta i)."""
    w = [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560]
    return (w[0] + 2 * sum(w[m] * np.cos(m * theta) for m in range(1, 5))) / h ** 2


def test_weights_are_eighth_order():
    # s(k) = -k^2 (1 + O((kh)^8)): the error ratio drops by 2^8 when h halves
    k = 1.3
    errs = []
    for h in (0.4, 0.2, 0.1):
        errs.append(abs(symbol_1d(k * h, h) + k * k))
    assert errs[0] / errs[1] > 200 and errs[1] / errs[2] > 230
    assert errs[2] / (k * k) < 1e-9


@pytest.mark.parametrize("n", [(12, 10, 9), (9, 14, 8)])
def test_plane_wave_eigenvector(n):
    h = (0.3, 0.22, 0.41)
    kv = (2, 3, 1)
    iz, iy, ix = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    ph = 2 * np.pi * (kv[0] * ix / n[0] + kv[1] * iy / n[1] + kv[2] * iz / n[2])
    f = np.cos(ph).reshape(-1)
    out = np.zeros_like(f)
    O.laplacian8(n, h, 1.0, f, out)
    lam = sum(symbol_1d(2 * np.pi * kv[d] / n[d], h[d]) for d in range(3))
    assert np.abs(out - lam * f).max() < 1e-11 * abs(lam)


def test_constant_symmetric_accumulate():
    n = (10, 9, 11)
    ng = n[0] * n[1] * n[2]
    h = (0.3, 0.3, 0.3)
    c = np.full(ng, 2.5)
    out = np.zeros(ng)
    O.laplacian8(n, h, -0.5, c, out)
    assert np.abs(out).max() < 1e-10
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(ng), rng.standard_normal(ng)
    lu, lv = np.zeros(ng), np.zeros(ng)
    O.laplacian8(n, h, 1.0, u, lu)
    O.laplacian8(n, h, 1.0, v, lv)
    assert abs(u @ lv - lu @ v) < 1e-9 * abs(u @ lv)
    acc = np.ones(ng)
    O.laplacian8(n, h, 1.0, u, acc)
    assert np.allclose(acc - 1.0, lu, rtol=0, atol=1e-12)
