"""CPU pins of the oracle's LOBPCG (Eigen_d_multi_mpi, lobpcg_d_multi_mpi.h:884-1572) and of
the DSYGVX wrapper (w_dsygvx.h:25-103).

1. The reference's own 8x8 (A, B) pair (test_scalapack.cpp:37-55): the oracle's DSYGVX
   call equals scipy.linalg.eigh(A, B) (eigenvalues; eigenvectors B-orthonormal).
2. A dense symmetric matrix with a known spectrum: repeated LOBPCG calls (warm start, as
   the reference's SCF loop does after a rollback) converge to numpy.linalg.eigh's lowest
   eigenvalues; Ritz vectors are orthonormal.
3. The oracle Hamiltonian of a small diamond cell is symmetric and LOBPCG converges to its
   dense eigenvalues.
"""
import json
import os

import numpy as np
import scipy.linalg

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def test_dsygvx_reference_pair():
    d = json.load(open(os.path.join(HERE, "golden", "scalapack_8x8.json")))
    A = np.array(d["A"]).reshape(8, 8)
    B = np.array(d["B"]).reshape(8, 8)
    w, Z = O.dsygvx(A, B)
    wr = scipy.linalg.eigh(A, B, eigvals_only=True)
    assert np.abs(w - wr).max() < 1e-10 * np.abs(wr).max()
    assert np.abs(Z.T @ B @ Z - np.eye(8)).max() < 1e-10
    assert np.abs(A @ Z - B @ Z * w).max() < 1e-8 * np.abs(A).max()


def _solve_restarting(A, nb, rounds=40, iters=100):
    n = A.shape[0]
    X = np.random.default_rng(1).random((nb, n))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    P = np.zeros_like(X)
    for r in range(rounds):
        e, st = O.lobpcg_dense(A, X, P, iters, tol=1e-10, scf_step=r)
        if st["rollbacks"] == 0 and 0 <= st["max_residual"] <= 1e-10:
            break
    return e, X, st


def test_lobpcg_dense_known_spectrum():
    n, nb = 300, 6
    rng = np.random.default_rng(2)
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([np.linspace(-3.0, -1.0, 10), np.linspace(0.0, 50.0, n - 10)])
    A = (Qm * lam) @ Qm.T
    e, X, st = _solve_restarting(A, nb)
    assert np.abs(e - lam[:nb]).max() < 1e-9
    assert np.abs(X @ X.T - np.eye(nb)).max() < 1e-8


def test_lobpcg_oracle_hamiltonian_small():
    import qumasun_amd as Q
    s = Q.diamond_system(1, 10)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    onl = O.OracleNL(n, h, s["positions"], np.zeros(8, np.int32), s["projectors"])
    Hd = O.hamiltonian_block(n, h, s["V"], onl, np.eye(ng)).T
    assert np.abs(Hd - Hd.T).max() < 1e-10
    w = np.linalg.eigvalsh(Hd)
    nb = next(k for k in range(4, 12) if w[k] - w[k - 1] > 1e-3)
    e, X, st = _solve_restarting(Hd, nb, rounds=20, iters=200)
    # Ritz values are upper bounds of the exact eigenvalues (Courant-Fischer); the isolated
    # lowest one converges fully, the near-degenerate ones (no preconditioner, as in the
    # reference) at least to 1e-2
    assert np.all(e >= w[:nb] - 1e-10)
    assert abs(e[0] - w[0]) < 1e-8
    assert np.abs(e - w[:nb]).max() < 1e-2
