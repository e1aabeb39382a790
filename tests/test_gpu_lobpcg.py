"""GPU tests of the Rayleigh-Ritz building blocks and the device LOBPCG.

Pins: the BF16 MFMA Gram product against numpy on the same BF16-rounded inputs; the kinetic
preconditioner against its closed form (a plane wave is an eigenvector of the periodic
8th-order kinetic operator); the LOBPCG iteration against the FP64 oracle restatement of
Eigen_d_multi_mpi on identical inputs (same number of iterations); converged eigenvalues
against scipy's dense eigh on a small cell and against the oracle fixture on the 8-atom
40^3 cell; the FP32 / mixed paths against FP64 (acceptance: 1e-6 Ha).
"""
import ctypes
import json
import os

import numpy as np
import pytest

import qumasun_amd as Q
from qumasun_amd import _native as N

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def bf16_round(a):
    """Round float32 values to bfloat16 (round to nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_split(a):
    """x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (float32 arithmetic), as float64."""
    hi = bf16_round(a)
    lo = bf16_round((np.asarray(a, np.float32) - hi.astype(np.float32)).astype(np.float32))
    return hi, lo


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("K,M,Ncol", [(10007, 40, 72), (64000, 16, 16), (4096, 128, 130)])
def test_gram_bf16(gpu, K, M, Ncol, split):
    import torch
    print(f"gram_bf16 K={K} M={M} N={Ncol} split={split}", flush=True)
    rng = np.random.default_rng(3)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((Ncol, K)).astype(np.float32)
    if split:
        ah, al = bf16_split(A)
        bh, bl = bf16_split(B)
        ref = ah @ bh.T + ah @ bl.T + al @ bh.T   # the three products the split kernel forms
        exact = A.astype(np.float64) @ B.astype(np.float64).T
        assert np.abs(ref - exact).max() / np.abs(exact).max() < 1e-4
    else:
        ref = bf16_round(A) @ bf16_round(B).T      # (M, N) exact products of the rounded inputs
    At = torch.from_numpy(A).to(gpu)
    Bt = torch.from_numpy(B).to(gpu)
    ldc = M + 3
    C = torch.zeros((Ncol, ldc), dtype=torch.float64, device=gpu)
    N.check(Q.lib().qg_gram_bf16(ctypes.c_void_p(At.data_ptr()), ctypes.c_void_p(Bt.data_ptr()), K, M, Ncol, K, K,
                                 ctypes.c_void_p(C.data_ptr()), ldc, split, None))
    got = C.cpu().numpy()[:, :M].T
    # FP32 accumulation of K products
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 2e-5, err


def _cell(ncell=1, npc=40):
    s = Q.diamond_system(ncell, npc)
    n = s["n"]
    h = s["h"]
    g = Q.Grid(n, n, n, h, h, h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    return s, g, integ, H


def test_kinetic_preconditioner_plane_wave(gpu):
    """(T + s)^-1 applied to cos(k.r) returns cos(k.r) / (T(k) + s)."""
    import torch
    print("precond plane wave", flush=True)
    n = (24, 20, 16)
    h = (0.3, 0.25, 0.35)
    g = Q.Grid(*n, *h)
    H = Q.Hamiltonian(g, np.zeros(n[0] * n[1] * n[2]))
    kv = (3, 2, 5)
    ix, iy, iz = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    ph = 2 * np.pi * (kv[0] * ix / n[0] + kv[1] * iy / n[1] + kv[2] * iz / n[2])
    f = np.cos(ph).transpose(2, 1, 0).reshape(-1)   # x fastest
    w = [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560]
    T = 0.0
    for d in range(3):
        th = 2 * np.pi * kv[d] / n[d]
        T += -0.5 * (w[0] + 2 * sum(w[m] * np.cos(m * th) for m in range(1, 5))) / h[d] ** 2
    shift = 0.7
    for dt in (torch.float64, torch.float32):
        r = torch.from_numpy(np.stack([f, 2 * f])).to(gpu).to(dt).contiguous()
        N.check(Q.lib().qg_precond_apply(H._h, N.QG_PRECOND_KINETIC, shift,
                                         N.QG_F64 if dt == torch.float64 else N.QG_F32, 2,
                                         ctypes.c_void_p(r.data_ptr()), g.size, None))
        torch.cuda.synchronize()
        got = r.double().cpu().numpy()
        ref = np.stack([f, 2 * f]) / (T + shift)
        tol = 1e-12 if dt == torch.float64 else 1e-6
        assert np.abs(got - ref).max() / np.abs(ref).max() < tol


def test_lobpcg_trajectory_matches_oracle(gpu, oracle):
    """FP64, no preconditioner, same start: after 12 iterations (2N first step, refresh at
    10, recurrences in between) the Ritz values equal the oracle restatement's."""
    import torch
    print("lobpcg trajectory", flush=True)
    s, g, integ, H = _cell(1, 24)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = g.size
    nb = 8
    X0 = Q.random_states(ng, nb, g.dvol)
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(len(s["positions"]), np.int32), s["projectors"])
    Xo = X0.copy()
    Po = np.zeros_like(Xo)
    eo, so = oracle.lobpcg_hamiltonian(n, h, s["V"], onl, Xo, Po, 12)
    X = torch.from_numpy(X0.copy()).to(gpu)
    P = torch.zeros_like(X)
    e, st = Q.Eigen_d_multi(H, X, P, 0, precision=N.QG_PREC_F64, max_iter=12)
    assert st["iterations"] == so["iterations"] == 12
    assert np.abs(e - eo).max() < 1e-9, np.abs(e - eo).max()
    # the returned vectors are the same Ritz vectors up to sign
    xg = X.cpu().numpy()
    ov = np.abs(np.sum(xg * Xo, axis=1)) * g.dvol
    assert np.all(np.abs(ov - 1) < 1e-6), ov


def _solve(H, X0, **kw):
    """Solve to kw['tol'] within kw['max_iter'] iterations in total.  A rollback ends a call
    early (as in the reference, lobpcg_d_multi_mpi.h:1322-1346); the solve then continues with
    the kept P block, as the next SCF step's call would."""
    import torch
    X = torch.from_numpy(X0.copy()).cuda()
    P = torch.zeros_like(X)
    budget = kw.pop("max_iter")
    total, rollbacks, step = 0, 0, 0
    while True:
        e, st = Q.Eigen_d_multi(H, X, P, step, max_iter=budget - total, **kw)
        total += st["iterations"]
        rollbacks += st["rollbacks"]
        step += 1
        if st["max_residual"] <= kw.get("tol", 0.0) or total >= budget or st["iterations"] == 0:
            break
    st["iterations"], st["rollbacks"] = total, rollbacks
    return e, st, X


def test_lobpcg_dense_small(gpu, oracle):
    """12^3 cell with overlapping periodic images: converged eigenvalues equal the lowest
    eigenvalues of the dense oracle Hamiltonian (numpy eigh), and the converged block spans the
    same subspace as the dense eigenvectors (equal after a rotation inside the block)."""
    print("lobpcg dense small", flush=True)
    s, g, integ, H = _cell(1, 12)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = g.size
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(len(s["positions"]), np.int32), s["projectors"])
    Hd = oracle.hamiltonian_block(n, h, s["V"], onl, np.eye(ng)).T
    assert np.abs(Hd - Hd.T).max() < 1e-10
    w, V = np.linalg.eigh(0.5 * (Hd + Hd.T))
    # a block that does not cut through a (symmetry-)degenerate multiplet
    nb = next(k for k in range(4, 12) if w[k] - w[k - 1] > 1e-3)
    e, st, X = _solve(H, Q.random_states(ng, nb, g.dvol), precision=N.QG_PREC_F64, max_iter=300, tol=1e-8,
                      precond=N.QG_PRECOND_KINETIC, precond_shift=1.0)
    assert st["max_residual"] <= 1e-8, st
    assert np.abs(e - w[:nb]).max() < 1e-9, (e, w[:nb])
    # subspace: with sqrt(dVol) X orthonormal, the projector onto its span equals V V^T of the
    # lowest nb dense eigenvectors (error ~ residual / gap)
    Xs = X.cpu().numpy() * np.sqrt(g.dvol)
    assert np.abs(Xs @ Xs.T - np.eye(nb)).max() < 1e-10
    Vb = V[:, :nb]
    # sines of the principal angles between the two subspaces: at most residual / gap
    # (Davis-Kahan), with the gap to the first eigenvalue outside the block
    sv = np.linalg.svd(Xs @ Vb, compute_uv=False)
    bound = max(1e-7, 10.0 * 1e-8 / (w[nb] - w[nb - 1]))
    assert np.sqrt(max(0.0, 1.0 - sv.min() ** 2)) < bound, (sv, bound)


def test_lobpcg_cfg0_f64_fixture(gpu):
    """8-atom 40^3 cell, 16 bands: FP64 with the kinetic preconditioner converges to the
    residual tolerance 1e-8 and to the oracle's converged eigenvalues (fixture)."""
    print("lobpcg cfg0 f64", flush=True)
    fx = json.load(open(os.path.join(HERE, "golden", "cfg0_eigenvalues.json")))
    s, g, integ, H = _cell(1, 40)
    e, st, _ = _solve(H, Q.random_states(g.size, 16, g.dvol), precision=N.QG_PREC_F64, max_iter=300, tol=1e-8,
                      precond=N.QG_PRECOND_KINETIC, precond_shift=1.0)
    assert st["max_residual"] <= 1e-8, st
    assert np.abs(e - np.array(fx["eigenvalues"])).max() < 1e-8


@pytest.mark.parametrize("prec", [N.QG_PREC_F32, N.QG_PREC_MIXED])
def test_lobpcg_cfg0_reduced_precision(gpu, prec):
    """FP32 / mixed (BF16 Gram) phases followed by FP64 refinement reach the FP64
    eigenvalues within 1e-6 Ha."""
    print(f"lobpcg cfg0 prec={prec}", flush=True)
    fx = json.load(open(os.path.join(HERE, "golden", "cfg0_eigenvalues.json")))
    s, g, integ, H = _cell(1, 40)
    e, st, _ = _solve(H, Q.random_states(g.size, 16, g.dvol), precision=prec, max_iter=300, tol=1e-8,
                      precond=N.QG_PRECOND_KINETIC, precond_shift=1.0, switch_tol=1e-3)
    assert np.abs(e - np.array(fx["eigenvalues"])).max() < 1e-6, (e, fx["eigenvalues"])


def test_lobpcg_guard_bands(gpu, oracle):
    """nconv: 16 wanted bands and 4 guard bands of the 8-atom cell (FP64, kinetic preconditioner,
    orthonormal basis, tol 1e-8): the solve stops when the lowest 16 meet the tolerance, the
    reported maximum residual is theirs (recomputed with the oracle's H), and their eigenvalues
    are the fixture's."""
    import torch
    print("lobpcg guard bands", flush=True)
    fx = json.load(open(os.path.join(HERE, "golden", "cfg0_eigenvalues.json")))
    s, g, integ, H = _cell(1, 40)
    X = torch.from_numpy(Q.random_states(g.size, 20, g.dvol)).to(gpu)
    P = torch.zeros_like(X)
    e, st = Q.Eigen_d_multi(H, X, P, 0, precision=N.QG_PREC_F64, max_iter=400, tol=1e-8, nconv=16,
                            precond=N.QG_PRECOND_KINETIC, precond_shift=1.0, basis=N.QG_BASIS_ORTHO)
    assert 0 < st["iterations"] < 400 and st["max_residual"] <= 1e-8, st
    n3, h3 = (s["n"],) * 3, (s["h"],) * 3
    nat = len(s["positions"])
    onl = oracle.OracleNL(n3, h3, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    xs = X.cpu().numpy()
    hx = oracle.hamiltonian_block(n3, h3, s["V"], onl, xs)
    res = np.sqrt(g.dvol * np.sum((hx - e[:, None] * xs) ** 2, axis=1))
    assert abs(res[:16].max() - st["max_residual"]) <= 0.05 * st["max_residual"] + 1e-12, (res, st)
    assert np.abs(e[:16] - np.array(fx["eigenvalues"])).max() < 1e-8


def test_lobpcg_foreign_operator(gpu):
    """The operator callback (the reference's OpeA) drives the same iteration: with the
    library's own H behind the callback the Ritz values equal the built-in path's."""
    import torch
    print("lobpcg foreign operator", flush=True)
    s, g, integ, H = _cell(1, 24)
    ng = g.size
    calls = []

    def op(dt, n, x, ax, stream):
        calls.append(n)
        N.check(Q.lib().qg_hamiltonian_apply(H._h, dt, n, x, ng, ax, ng, stream))

    X0 = Q.random_states(ng, 6, g.dvol)
    res = []
    for use_op in (False, True):
        X = torch.from_numpy(X0.copy()).to(gpu)
        P = torch.zeros_like(X)
        kw = dict(precision=N.QG_PREC_F64, max_iter=15)
        if use_op:
            kw["op"] = op
        e, st = Q.Eigen_d_multi(H, X, P, 0, **kw)
        res.append(e)
    assert len(calls) == 16   # X once, then R every iteration
    # same iteration; H psi sums the projector term with LDS atomics in its table form, so the
    # two runs agree to rounding, not bit for bit
    assert np.abs(res[0] - res[1]).max() <= 1e-11


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("ng,nb", [(40 ** 3, 16), (12 ** 3, 7), (5000, 1)])
def test_gram_schmidt_matches_oracle(gpu, oracle, prec, ng, nb):
    """qg_gram_schmidt (two Cholesky-QR passes) equals the oracle's band-by-band Gram-Schmidt
    (itself equal bit for bit to the reference's GramSchmidt) on the reference's U[0,1) start."""
    import torch
    print(f"gram-schmidt {prec} ng={ng} nb={nb}", flush=True)
    dvol = 0.1685 ** 3
    rng = np.random.default_rng(ng + nb)
    X = rng.uniform(0.0, 1.0, (nb, ng))
    ref = oracle.gram_schmidt(X, dvol)
    dt = torch.float64 if prec == "f64" else torch.float32
    t = torch.from_numpy(X).to(gpu).to(dt).contiguous()
    Q.GramSchmidt(t, dvol)
    torch.cuda.synchronize()
    got = t.double().cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= (1e-12 if prec == "f64" else 2e-6), err
    assert np.abs(dvol * got @ got.T - np.eye(nb)).max() <= (1e-12 if prec == "f64" else 1e-5)


def test_gram_schmidt_rejects_dependent_bands(gpu):
    import torch
    X = torch.ones(3, 1000, dtype=torch.float64, device=gpu)
    with pytest.raises(Q.QGError):
        Q.GramSchmidt(X, 0.5)
