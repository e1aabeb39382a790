"""GPU tests of the batched path for independent uneven cells: the persistent multi-cell H psi
launch against per-cell launches (same arithmetic up to the order of LDS atomic additions) and
the FP64 oracle, the
autotuner's tile shape keeping oracle parity, and the batched LOBPCG (cells solved together,
H psi of all cells in one launch per round) against one-cell-at-a-time solves."""
import numpy as np
import pytest

import qumasun_amd as Q
from qumasun_amd import _native as N

pytestmark = pytest.mark.gpu

CELLS = [((1, 1, 1), 20, 3), ((2, 1, 1), 20, 5), ((1, 2, 1), 16, 8), ((1, 1, 2), 24, 2)]


def _build(dims, npc):
    c = Q.diamond_cell(dims, npc)
    g = Q.Grid(*c["n"], *c["h"])
    integ = Q.PseudoPotIntegrator(g, c["projectors"])
    integ.UpdatePosition(c["positions"])
    H = Q.Hamiltonian(g, c["V"], integ)
    return c, g, integ, H


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_batch_apply_matches_per_cell(gpu, oracle, dt):
    import torch
    print(f"batch apply {dt}", flush=True)
    tdt = torch.float64 if dt == "f64" else torch.float32
    cells = [_build(d, n) + (nb,) for d, n, nb in CELLS]
    psis, outs, refs = [], [], []
    for c, g, integ, H, nb in cells:
        X = Q.random_states(g.size, nb, g.dvol, seed=7 + nb)
        x = torch.from_numpy(X).to(gpu).to(tdt)
        psis.append(x)
        outs.append(torch.empty_like(x))
        r = torch.empty_like(x)
        H.apply(x, r)
        refs.append(r)
    Q.batch_apply([c[3] for c in cells], psis, outs)
    torch.cuda.synchronize()
    for (c, g, integ, H, nb), o, r, x in zip(cells, outs, refs, psis):
        # the projector term accumulates in LDS with atomics, so two applications agree to
        # the last bits, not bit for bit
        rel = float((o - r).abs().max() / r.abs().max())
        assert rel <= (1e-14 if dt == "f64" else 1e-6), rel
        # and against the FP64 oracle
        onl = oracle.OracleNL(c["n"], c["h"], c["positions"], np.zeros(len(c["positions"]), np.int32),
                              c["projectors"])
        ref = oracle.hamiltonian_block(c["n"], c["h"], c["V"], onl, x.double().cpu().numpy())
        err = np.linalg.norm(o.double().cpu().numpy() - ref) / np.linalg.norm(ref)
        assert err <= (1e-12 if dt == "f64" else 2e-6), err


def test_autotune_keeps_parity(gpu, oracle):
    import torch
    print("autotune", flush=True)
    c, g, integ, H = _build((2, 1, 1), 20)
    ty, ppt, ms = Q.autotune(H, "f64", nbands=4)
    assert ty >= 1 and ppt in (2, 4, 5, 8) and ms > 0
    X = Q.random_states(g.size, 4, g.dvol)
    x = torch.from_numpy(X).to(gpu)
    out = torch.empty_like(x)
    H.apply(x, out)
    torch.cuda.synchronize()
    onl = oracle.OracleNL(c["n"], c["h"], c["positions"], np.zeros(len(c["positions"]), np.int32), c["projectors"])
    ref = oracle.hamiltonian_block(c["n"], c["h"], c["V"], onl, X)
    assert np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref) <= 1e-12


def test_batch_lobpcg_matches_sequential(gpu):
    import torch
    print("batch lobpcg", flush=True)
    cells = [_build(d, n) for d, n, _ in CELLS[:3]]
    nbs = [4, 6, 8]
    opts = dict(precision=N.QG_PREC_F64, max_iter=300, tol=1e-6, precond=N.QG_PRECOND_KINETIC, precond_shift=1.0)
    res = {}
    for batched in (False, True):
        Xs = [torch.from_numpy(Q.random_states(g.size, nb, g.dvol)).to(gpu) for (c, g, i, H), nb in zip(cells, nbs)]
        Ps = [torch.zeros_like(x) for x in Xs]
        e, st = Q.batch_eigen([H for c, g, i, H in cells], Xs, Ps, 0, batched=batched, **opts)
        res[batched] = (e, st)
        for s in st:
            assert s["max_residual"] <= 1e-6, s
    for a, b in zip(res[False][0], res[True][0]):
        assert np.abs(a - b).max() < 1e-9
