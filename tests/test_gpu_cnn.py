"""GPU tests of the learned preconditioner (qumasun_amd/csrc/qg_cnn.hip).

CPU reference: tests/cnn_ref.py (torch CPU restatement, pinned by test_cnn_ref.py).  Covered: the
network and its adjoint against that reference; the BF16 inference weights; the device loss and
gradient against torch autograd on the CPU (the training gradient is the project's own
forward/backward kernels, not autograd); M = K + K N^T N K symmetric positive definite; the
device training loop lowering the loss; LOBPCG with the learned preconditioner converging to
the FP64 fixture eigenvalues, with the stall fallback exercised; and the iteration counts of
no preconditioner / kinetic / learned reported side by side.
"""
import json
import os

import numpy as np
import pytest

import cnn_ref as R
import qumasun_amd as Q
from qumasun_amd import _native as N

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _ham(nx=20, ny=16, nz=12, seed=3):
    rng = np.random.default_rng(seed)
    h = (0.3, 0.35, 0.4)
    g = Q.Grid(nx, ny, nz, *h)
    V = rng.uniform(-2.0, 1.0, nx * ny * nz)
    H = Q.Hamiltonian(g, V)
    return g, H, V, h


def _cpu_views(g, V, h):
    import torch
    n = g.n
    v = torch.from_numpy(R.vnorm(V).reshape(1, 1, n[2], n[1], n[0]))
    T = torch.from_numpy(R.kinetic_symbol(n, h))
    return v, T


def test_network_and_adjoint_match_cpu(gpu):
    import torch
    print("cnn network vs cpu", flush=True)
    g, H, V, h = _ham()
    net = Q.LearnedPreconditioner(H, 0.8, max_pairs=0, seed=11)
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(R.NPARAM) * 0.3).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    nb = 3
    x = rng.standard_normal((nb, g.size)).astype(np.float32)
    v, _ = _cpu_views(g, V, h)
    P = R.split(torch.from_numpy(p.astype(np.float64)))
    xc = torch.from_numpy(x.astype(np.float64)).reshape(nb, 1, g.n[2], g.n[1], g.n[0])
    for adjoint in (False, True):
        xd = torch.from_numpy(x).to(gpu)
        out = torch.empty_like(xd)
        net.network(xd, out, adjoint=adjoint)
        ref = (R.net_adj(P, v, xc) if adjoint else R.net(P, v, xc)).reshape(nb, -1).numpy()
        err = np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < 2e-5, (adjoint, err)
    # <N a, b> = <a, N^T b> on the device
    a = torch.from_numpy(x).to(gpu)
    b = torch.from_numpy(rng.standard_normal((nb, g.size)).astype(np.float32)).to(gpu)
    na = torch.empty_like(a)
    ntb = torch.empty_like(b)
    net.network(a, na)
    net.network(b, ntb, adjoint=True)
    lhs = float((na.double() * b.double()).sum())
    rhs = float((a.double() * ntb.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_bf16_weights(gpu):
    import torch
    print("cnn bf16 weights", flush=True)
    g, H, V, h = _ham()
    net = Q.LearnedPreconditioner(H, 0.8, max_pairs=0, seed=12)
    p = net.get_params()
    pb = torch.from_numpy(p).to(torch.bfloat16).to(torch.float64)
    v, _ = _cpu_views(g, V, h)
    x = np.random.default_rng(2).standard_normal((2, g.size)).astype(np.float32)
    xd = torch.from_numpy(x).to(gpu)
    out = torch.empty_like(xd)
    net.network(xd, out, bf16=True)
    xc = torch.from_numpy(x.astype(np.float64)).reshape(2, 1, g.n[2], g.n[1], g.n[0])
    ref = R.net(R.split(pb), v, xc).reshape(2, -1).numpy()
    assert np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max() < 2e-5


def test_mfma_network_matches_bf16_emulation(gpu):
    """The matrix-core inference kernels (BF16 weights and activations, FP32 accumulation)
    equal the CPU reference run with the same BF16 roundings."""
    import torch
    print("cnn mfma network", flush=True)
    for dims in ((20, 16, 12), (40, 36, 8), (13, 9, 7)):
        g, H, V, h = _ham(*dims)
        net = Q.LearnedPreconditioner(H, 0.8, max_pairs=0, seed=15)
        rng = np.random.default_rng(6)
        net.set_params((rng.standard_normal(R.NPARAM) * 0.3).astype(np.float32))
        pb = torch.from_numpy(net.get_params()).to(torch.bfloat16).to(torch.float64)
        v, _ = _cpu_views(g, V, h)
        nb = 2
        x = rng.standard_normal((nb, g.size)).astype(np.float32)
        xc = torch.from_numpy(x.astype(np.float64)).reshape(nb, 1, g.n[2], g.n[1], g.n[0])
        for adjoint in (False, True):
            xd = torch.from_numpy(x).to(gpu)
            out = torch.empty_like(xd)
            net.network(xd, out, adjoint=adjoint, bf16=2)
            f = R.net_adj if adjoint else R.net
            ref = f(R.split(pb), v, xc, rnd=R.bf16).reshape(nb, -1).numpy()
            err = np.linalg.norm(out.cpu().numpy() - ref) / np.linalg.norm(ref)
            assert err < 2e-3, (dims, adjoint, err)


def test_loss_gradient_matches_autograd(gpu):
    """The device gradient (two forward and two backward passes of the project's kernels)
    equals torch autograd of the CPU reference's loss."""
    import torch
    print("cnn gradient vs autograd", flush=True)
    g, H, V, h = _ham()
    shift = 0.8
    net = Q.LearnedPreconditioner(H, shift, max_pairs=0, seed=13)
    rng = np.random.default_rng(7)
    p = (rng.standard_normal(R.NPARAM) * 0.3).astype(np.float32)
    net.set_params(p)
    nb = 2
    u = rng.standard_normal((nb, g.size)).astype(np.float32)
    t = rng.standard_normal((nb, g.size)).astype(np.float32)
    znorm = 3.0 * g.size
    loss, grad = net.loss_grad(torch.from_numpy(u).to(gpu), torch.from_numpy(t).to(gpu), znorm)
    v, T = _cpu_views(g, V, h)
    pt = torch.from_numpy(p.astype(np.float64)).requires_grad_()
    sh = (nb, 1, g.n[2], g.n[1], g.n[0])
    lc = R.loss(R.split(pt), v, T, shift, torch.from_numpy(u.astype(np.float64)).reshape(sh),
                torch.from_numpy(t.astype(np.float64)).reshape(sh), znorm)
    lc.backward()
    gref = pt.grad.numpy()
    assert abs(loss - float(lc)) <= 1e-4 * abs(float(lc)), (loss, float(lc))
    rel = np.abs(grad - gref).max() / np.abs(gref).max()
    assert rel < 1e-3, rel
    # every parameter group contributes
    P = R.split(torch.from_numpy(grad.astype(np.float64)))
    for k, val in P.items():
        assert float(val.abs().max()) > 0, k


def test_preconditioner_is_spd(gpu):
    import torch
    print("cnn spd", flush=True)
    g, H, V, h = _ham()
    net = Q.LearnedPreconditioner(H, 0.8, max_pairs=0, seed=14)
    rng = np.random.default_rng(9)
    net.set_params((rng.standard_normal(R.NPARAM) * 0.5).astype(np.float32))
    a = torch.from_numpy(rng.standard_normal((4, g.size))).to(gpu)
    b = torch.from_numpy(rng.standard_normal((4, g.size))).to(gpu)
    Ma, Mb = a.clone(), b.clone()
    net.apply(Ma)
    net.apply(Mb)
    torch.cuda.synchronize()
    lhs = (Ma * b).sum(1)
    rhs = (a * Mb).sum(1)
    assert float(((lhs - rhs).abs() / lhs.abs()).max()) < 1e-4
    assert float((Ma * a).sum(1).min()) > 0


def _cell40():
    s = Q.diamond_system(1, 40)
    n = s["n"]
    g = Q.Grid(n, n, n, s["h"], s["h"], s["h"])
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    return s, g, integ, H


def _solve(H, g, nb, **kw):
    """Solve to kw['tol'] within kw['max_iter'] iterations, continuing after a rollback with
    the kept P block (as the next SCF step's call would)."""
    import torch
    X = torch.from_numpy(Q.random_states(g.size, nb, g.dvol)).cuda()
    P = torch.zeros_like(X)
    budget = kw.pop("max_iter")
    total, step, fb = 0, 0, -1
    while True:
        e, st = Q.Eigen_d_multi(H, X, P, step, max_iter=budget - total, **kw)
        if fb < 0 and st["fallback_iter"] >= 0:
            fb = total + st["fallback_iter"]
        total += st["iterations"]
        step += 1
        if st["max_residual"] <= kw.get("tol", 0.0) or total >= budget or st["iterations"] == 0:
            break
    st["iterations"], st["fallback_iter"] = total, fb
    return e, st


def test_learned_lobpcg_cfg0(gpu):
    """8-atom 40^3 cell, 16 bands: train on the solver's residuals, then LOBPCG with the
    learned preconditioner reaches the FP64 fixture eigenvalues; iteration counts of the
    three preconditioners are printed (the reference runs without one)."""
    print("cnn lobpcg cfg0", flush=True)
    fx = json.load(open(os.path.join(HERE, "golden", "cfg0_eigenvalues.json")))
    ref = np.array(fx["eigenvalues"])
    s, g, integ, H = _cell40()
    nb = 16
    sigma = -11.64      # below the lowest eigenvalue of T + V (-11.54) for this cell
    shift = 2.0         # >= max V - sigma, so (T + V - sigma)^-1 >= K
    net, info = Q.fit_learned_preconditioner(H, nb, sigma, shift, steps=400)
    print("training:", info, flush=True)
    assert info["pairs"] == nb * 5
    assert info["loss_last"] < info["loss_first"]
    net.attach()
    e_l, st_l = _solve(H, g, nb, max_iter=300, tol=1e-8, precond=N.QG_PRECOND_LEARNED, precond_shift=shift,
                       stall_window=8)
    e_k, st_k = _solve(H, g, nb, max_iter=300, tol=1e-8, precond=N.QG_PRECOND_KINETIC, precond_shift=shift)
    e_k1, st_k1 = _solve(H, g, nb, max_iter=300, tol=1e-8, precond=N.QG_PRECOND_KINETIC, precond_shift=0.5)
    print(f"iterations: learned {st_l['iterations']} (fallback {st_l['fallback_iter']}), kinetic s={shift} "
          f"{st_k['iterations']}, kinetic s=0.5 {st_k1['iterations']}", flush=True)
    for e, st in ((e_l, st_l), (e_k, st_k), (e_k1, st_k1)):
        assert st["max_residual"] <= 1e-8
        assert np.abs(e - ref).max() < 1e-8
    # the stall fallback: an FP32 run asked for a residual below FP32 resolution stalls, and
    # the kinetic preconditioner takes over after stall_window iterations without progress
    e_f, st_f = _solve(H, g, nb, max_iter=150, tol=1e-10, precision=N.QG_PREC_F32, switch_tol=1e-12,
                       precond=N.QG_PRECOND_LEARNED, precond_shift=shift, stall_window=4)
    print("fallback at iteration", st_f["fallback_iter"], "max residual", st_f["max_residual"], flush=True)
    assert st_f["fallback_iter"] >= 1
    assert np.abs(e_f - ref).max() < 1e-5
    net.detach()
