"""GPU parity of the H·psi kernels against the FP64 oracle (same seeded inputs).

Tolerances (relative 2-norm over the whole block): FP64 <= 1e-12, FP32 <= 2e-6.
"""
import numpy as np
import pytest

import qumasun_amd as Q

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _t(x, dtype, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev).to(dtype)


def _np(t):
    return t.double().cpu().numpy()


GRIDS = [((40, 40, 40), (0.1685, 0.1685, 0.1685)),
         ((12, 10, 9), (0.2, 0.25, 0.3)),
         ((80, 16, 24), (0.17, 0.18, 0.16)),
         ((33, 7, 5), (0.3, 0.2, 0.25))]


@pytest.mark.parametrize("n,h", GRIDS)
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("accumulate", [0, 1])
def test_laplacian8(gpu, oracle, n, h, prec, accumulate):
    import torch
    print(f"laplacian8 n={n} prec={prec} acc={accumulate}", flush=True)
    rng = np.random.default_rng(1)
    ng = n[0] * n[1] * n[2]
    nb = 3
    p = rng.standard_normal((nb, ng))
    o0 = rng.standard_normal((nb, ng))
    ref = (o0.copy() if accumulate else np.zeros_like(o0))
    for b in range(nb):
        oracle.laplacian8(n, h, -0.5, p[b].copy(), ref[b])
    dt = torch.float64 if prec == "f64" else torch.float32
    g = Q.Grid(*n, *h)
    pt = _t(p, dt, gpu)
    ot = _t(o0, dt, gpu)
    Q.Laplasian8th(g, ot, pt, -0.5, accumulate=bool(accumulate))
    torch.cuda.synchronize()
    err = rel(_np(ot), ref)
    assert err <= (TOL64 if prec == "f64" else TOL32), err


def _system(ncell, npc):
    return Q.diamond_system(ncell, npc)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_nonlocal_project_apply_cfg0(gpu, oracle, prec):
    import torch
    print(f"nonlocal cfg0 prec={prec}", flush=True)
    s = _system(1, 40)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    nat = len(s["positions"])
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    rng = np.random.default_rng(2)
    nb = 5
    psi = rng.standard_normal((nb, ng))
    coef_ref = np.stack([onl.project(psi[b].copy()) for b in range(nb)])
    hp_ref = np.zeros((nb, ng))
    for b in range(nb):
        onl.apply(psi[b].copy(), hp_ref[b])
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    assert integ.num_coef == onl.ncoef == 4 * nat
    dt = torch.float64 if prec == "f64" else torch.float32
    pt = _t(psi, dt, gpu)
    coef = integ.InnerNonlocalPsi(pt).cpu().numpy()
    hp = torch.zeros_like(pt)
    integ.ProjectionPP(hp, pt)
    torch.cuda.synchronize()
    tol = TOL64 if prec == "f64" else 2e-6
    assert rel(coef, coef_ref) <= tol
    assert rel(_np(hp), hp_ref) <= tol


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hamiltonian_cfg0(gpu, oracle, prec):
    import torch
    print(f"hamiltonian cfg0 prec={prec}", flush=True)
    s = _system(1, 40)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    nat = len(s["positions"])
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    X = Q.random_states(ng, 16, h[0] ** 3)
    ref = oracle.hamiltonian_block(n, h, s["V"], onl, X)
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    xt = _t(X, dt, gpu)
    out = torch.empty_like(xt)
    H.apply(xt, out)
    torch.cuda.synchronize()
    err = rel(_np(out), ref)
    assert err <= (TOL64 if prec == "f64" else TOL32), err


def test_hamiltonian_cfg1_subset(gpu, oracle):
    """64-atom 80^3 cell: FP64 and FP32 H·psi of the 128-band block; 6 bands vs the oracle."""
    import torch
    print("hamiltonian cfg1", flush=True)
    s = _system(2, 40)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    nat = len(s["positions"])
    X = Q.random_states(ng, 128, h[0] ** 3)
    sel = [0, 1, 63, 64, 126, 127]
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    ref = oracle.hamiltonian_block(n, h, s["V"], onl, X[sel])
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    for dt, tol in ((torch.float64, TOL64), (torch.float32, TOL32)):
        xt = _t(X, dt, gpu)
        out = torch.empty_like(xt)
        H.apply(xt, out)
        torch.cuda.synchronize()
        got = out[sel].double().cpu().numpy()
        err = rel(got, ref)
        assert err <= tol, (dt, err)
