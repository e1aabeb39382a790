"""Pins of the oracle's GramSchmidt (GramSchmidt.h:6-38): bit for bit against the reference's
own function compiled from its header (when the reference tree is present), and against
numpy's QR decomposition (the Gram-Schmidt Q is the QR factor with a positive R diagonal,
scaled to dVol-orthonormality)."""
import numpy as np
import pytest

import oracle as O


def _block(ng, nb, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (nb, ng))   # the reference's random start is U[0,1)


@pytest.mark.parametrize("ng,nb,dvol", [(1000, 7, 0.125), (4096, 16, 0.1685 ** 3), (300, 1, 2.0)])
def test_oracle_gram_schmidt_matches_qr(ng, nb, dvol):
    X = _block(ng, nb, 3)
    Y = O.gram_schmidt(X, dvol)
    q, r = np.linalg.qr(X.T)
    q = q * np.sign(np.diag(r))[None, :] / np.sqrt(dvol)
    assert np.abs(Y - q.T).max() <= 1e-10 * np.abs(q).max()
    assert np.abs(dvol * Y @ Y.T - np.eye(nb)).max() <= 1e-12


def test_oracle_gram_schmidt_matches_reference():
    ref = O.ref_lib()
    if ref is None or not hasattr(ref, "ref_gram_schmidt"):
        pytest.skip("reference harness not built here")
    X = _block(2000, 9, 5)
    Y = O.gram_schmidt(X, 0.01)
    Z = np.ascontiguousarray(X.copy())
    ref.ref_gram_schmidt(2000, 9, O._d(Z), 0.01)
    assert np.array_equal(Y, Z)
