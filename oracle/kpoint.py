"""CPU FP64 restatement of QUMASUN's k-point (complex Bloch) Hamiltonian — TEST INFRASTRUCTURE
ONLY (the oracle; only tests/ may use it).

The cell-periodic part u of psi_k = exp(i k.r) u is acted on by
  H_k u = coef [L8 u - |k|^2 u + 2 i k . D8 u] + V u + sum_j E_j conj(P_j) (dV P_j^T u)
with coef = -1/2, L8 the 8th-order Laplacian, D8 the 8th-order gradient and
P_j = beta_j(x_u - R) exp(i k . x_u) on the projector's image blocks (no conjugation in the
projection, the conjugate in the application).

What pins it (tests/test_oracle_kpoint.py): the stencil's action on a plane wave is the
closed-form symbol coef [s2(q) - |k|^2 - 2 k . s1(q)]; at k = 0 the operator is the oracle's
pinned real Hamiltonian; the k-point projector columns equal an independent brute-force sum
over periodic images; the dense H_k is Hermitian and H_{-k} = conj(H_k) (equal spectra).
"""
import numpy as np

from .binding import hamiltonian_block  # noqa: F401  (the k = 0 reduction is checked against it)

__all__ = ["kpoint_stencil", "kpoint_hamiltonian", "kpoint_dense"]

_C2 = (8064.0, -1008.0, 128.0, -9.0)     # L8 off-centre weights x 5040
_C1 = (672.0, -168.0, 32.0, -3.0)        # D8 weights x 840


def kpoint_stencil(u, n, h, k, coef=-0.5):
    """LaplasianKpoint8th_halo_f applied to the real and the imaginary part
    (GridDifferenceKpoint8th.h:17-148): coef (L8 - |k|^2) u + 2 coef i k . D8 u.  u: (..., Ng)
    complex, x fastest.  Returns the same shape."""
    nx, ny, nz = n
    u = np.asarray(u, dtype=np.complex128)
    shp = u.shape
    p = u.reshape(shp[:-1] + (nz, ny, nx))
    k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    out = (coef * (-14350.0 / 5040.0) * sum(1.0 / (hh * hh) for hh in h) - coef * k2) * p
    for d, hh in enumerate(h):
        ax = p.ndim - 1 - d
        lap = np.zeros_like(p)
        grad = np.zeros_like(p)
        for m in range(1, 5):
            up, dn = np.roll(p, -m, axis=ax), np.roll(p, m, axis=ax)
            lap += _C2[m - 1] * (up + dn)
            grad += _C1[m - 1] * (up - dn)
        out += coef / (5040.0 * hh * hh) * lap
        if k[d] != 0.0:
            out += 2.0j * coef * k[d] / (840.0 * hh) * grad
    return out.reshape(shp)


def kpoint_hamiltonian(n, h, V, onl, k, U):
    """mHamiltonianMatrix_ddm at one k point (qumasun_kpoint.matrix.h:21-110 with ProjectionPP,
    PseudoPotOperator_kpoint_v2.h:2581-2690): the stencil, V u and the nonlocal term.
    onl: oracle.OracleNL or None; U: (nb, Ng) complex."""
    U = np.atleast_2d(np.asarray(U, dtype=np.complex128))
    dv = h[0] * h[1] * h[2]
    out = kpoint_stencil(U, n, h, k) + np.asarray(V)[None, :] * U
    if onl is not None and onl.ncoef:
        E = onl.energies()
        ng = U.shape[1]
        if onl.ncoef * ng <= 50_000_000:
            B = onl.dense_beta_k(k)
            c = dv * (U @ B.T)                 # (nb, ncoef): sum P u, no conjugate
            out += (c * E[None, :]) @ np.conj(B)
        else:                                  # one column at a time (large grids)
            for j in range(onl.ncoef):
                col = onl.dense_beta_k_column(j, k)
                c = dv * (U @ col)
                out += (E[j] * c)[:, None] * np.conj(col)[None, :]
    return out


def kpoint_dense(n, h, V, onl, k):
    """The dense Ng x Ng matrix of H_k (column i = H_k e_i), for small grids."""
    ng = n[0] * n[1] * n[2]
    return kpoint_hamiltonian(n, h, V, onl, k, np.eye(ng, dtype=np.complex128)).T
