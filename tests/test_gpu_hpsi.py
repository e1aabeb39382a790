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


@pytest.mark.parametrize("analytic,split", [("0", None), ("1", None), ("2", None), ("2", "0"), ("0", "1")])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hamiltonian_projector_paths(gpu, oracle, monkeypatch, prec, analytic, split):
    """Every form of the projector term against the oracle: per-point tables or values
    evaluated in the kernel from the image centres (QG_NL_ANALYTIC 0 / 1 = FP32 projector pass /
    2 = FP64 projector pass too), in the fused pass (tables) or in a separate projector pass
    (QG_SPLIT_NL). The cell
    is orthorhombic with unequal spacings, two species with different cutoffs, atoms on the cell
    faces and a cutoff sphere wider than the cell along z (two images of one atom on the same
    rows)."""
    import torch
    print(f"projector paths prec={prec} analytic={analytic} split={split}", flush=True)
    monkeypatch.setenv("QG_NL_ANALYTIC", analytic)
    if split is not None:
        monkeypatch.setenv("QG_SPLIT_NL", split)
    n = (30, 26, 22)
    L = (6.0, 5.5, 5.0)
    h = tuple(L[d] / n[d] for d in range(3))
    pos = np.array([[0.0, 0.0, 0.0], [3.1, 2.7, 2.45], [5.9, 0.2, 4.95], [1.5, 4.0, 1.0], [4.4, 1.3, 3.3]])
    species = np.array([0, 1, 0, 1, 0], np.int32)
    projs = [[Q.gaussian_projector(0, 1.0, 0.6, 2.4), Q.gaussian_projector(1, -0.5, 0.6, 2.4)],
             [Q.gaussian_projector(0, 0.7, 0.8, 2.8)]]
    rng = np.random.default_rng(11)
    V = rng.uniform(-1.0, 0.5, n[0] * n[1] * n[2])
    ng = n[0] * n[1] * n[2]
    onl = oracle.OracleNL(n, h, pos, species, projs)
    X = Q.random_states(ng, 9, h[0] * h[1] * h[2])
    ref = oracle.hamiltonian_block(n, h, V, onl, X)
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, projs)
    integ.UpdatePosition(pos, species)
    H = Q.Hamiltonian(g, V, integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    xt = _t(X, dt, gpu)
    out = torch.empty_like(xt)
    H.apply(xt, out)
    torch.cuda.synchronize()
    assert integ.projector_form == ("tables" if analytic == "0" else "kernel")
    err = rel(_np(out), ref)
    assert err <= (TOL64 if prec == "f64" else TOL32), err
    # the projector term alone (H psi minus the local part) to the same bar
    loc = oracle.hamiltonian_block(n, h, V, None, X)
    errnl = rel(_np(out) - loc, ref - loc)
    assert errnl <= (1e-11 if prec == "f64" else 2e-5), errnl


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nb", [1, 3, 13])
def test_hamiltonian_ragged_padded(gpu, oracle, prec, nb):
    """Band counts that are not multiples of the per-workgroup band groups (1, 3, 13), on blocks
    whose leading dimension is padded (ld = Ng + 37): H psi equals the oracle and the padding
    of the output is left untouched."""
    import ctypes
    import torch
    print(f"ragged/padded prec={prec} nb={nb}", flush=True)
    s = _system(1, 40)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    ld = ng + 37
    nat = len(s["positions"])
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    X = Q.random_states(ng, nb, h[0] ** 3, seed=99 + nb)
    ref = oracle.hamiltonian_block(n, h, s["V"], onl, X)
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    xin = torch.zeros(nb, ld, dtype=dt, device=gpu)
    xin[:, :ng] = _t(X, dt, gpu)
    out = torch.full((nb, ld), 7.0, dtype=dt, device=gpu)
    code = Q.QG_F64 if prec == "f64" else Q.QG_F32
    Q.check(Q.lib().qg_hamiltonian_apply(H._h, code, nb, ctypes.c_void_p(xin.data_ptr()), ld,
                                         ctypes.c_void_p(out.data_ptr()), ld, ctypes.c_void_p(0)))
    torch.cuda.synchronize()
    got = out[:, :ng].double().cpu().numpy()
    err = rel(got, ref)
    assert err <= (TOL64 if prec == "f64" else TOL32), err
    assert bool((out[:, ng:] == 7.0).all())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_update_position_with_live_hamiltonian(gpu, oracle, prec):
    """UpdatePosition on an integrator that a Hamiltonian was built on (the reference moves the
    nuclei of its one integrator at every MD / relaxation step, qumasun_mpi.move.h:15): the
    handle is rebuilt in place, and the Hamiltonian's next application uses the new positions
    (oracle parity for the last set)."""
    import torch
    print(f"update position prec={prec}", flush=True)
    s = _system(1, 40)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    X = Q.random_states(ng, 5, h[0] ** 3)
    xt = _t(X, dt, gpu)
    out = torch.empty_like(xt)
    H.apply(xt, out)
    rng = np.random.default_rng(5)
    pos = s["positions"]
    for step in range(2):
        pos = pos + rng.uniform(-0.3, 0.3, pos.shape)
        integ.UpdatePosition(pos)
        H.apply(xt, out)
    torch.cuda.synchronize()
    onl = oracle.OracleNL(n, h, pos, np.zeros(len(pos), np.int32), s["projectors"])
    ref = oracle.hamiltonian_block(n, h, s["V"], onl, X)
    err = rel(_np(out), ref)
    assert err <= (TOL64 if prec == "f64" else TOL32), err


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_two_hamiltonians_one_integrator(gpu, oracle, prec):
    """Two Hamiltonians (different local potentials, e.g. two spin channels) built on one
    integrator, applied in one batched launch: each equals its own application and the oracle."""
    import torch
    print(f"batch shared integrator prec={prec}", flush=True)
    s = _system(1, 40)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    V2 = s["V"] + 0.3 * np.sin(np.arange(ng) * 0.01)
    Ha = Q.Hamiltonian(g, s["V"], integ)
    Hb = Q.Hamiltonian(g, V2, integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    Xa = Q.random_states(ng, 4, h[0] ** 3, seed=11)
    Xb = Q.random_states(ng, 6, h[0] ** 3, seed=12)
    xa, xb = _t(Xa, dt, gpu), _t(Xb, dt, gpu)
    ya, yb = torch.empty_like(xa), torch.empty_like(xb)
    Q.batch_apply([Ha, Hb], [xa, xb], [ya, yb])
    torch.cuda.synchronize()
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(len(s["positions"]), np.int32), s["projectors"])
    tol = TOL64 if prec == "f64" else TOL32
    assert rel(_np(ya), oracle.hamiltonian_block(n, h, s["V"], onl, Xa)) <= tol
    assert rel(_np(yb), oracle.hamiltonian_block(n, h, V2, onl, Xb)) <= tol


def _log_table(l, energy, sigma, cutoff, xi_min, dxi, npts):
    import math
    amp = math.sqrt(4.0 / (math.sqrt(math.pi) * sigma ** 3))
    xi = xi_min + dxi * np.arange(npts)
    vals = amp * np.exp(-np.exp(2 * xi) / (2 * sigma * sigma))
    return Q.table_projector(l, energy, cutoff, xi_min, dxi, vals)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_table_projectors_parity(gpu, oracle, prec):
    """Tabulated radial projectors (QG_RADIAL_TABLE, a Gaussian sampled on a log grid): H psi and
    the projector coefficients against the oracle.  The tables are short enough that sphere
    points fall in every interpolation branch of GetValueByXi (RadialGrid2.h:81-136), below
    r_min (table[0]) and beyond the last table point (zero) inside the cutoff; a second species
    has s and p channels of different cutoffs, and one nucleus sits on a grid point."""
    import torch
    print(f"table projectors prec={prec}", flush=True)
    import math
    n = (28, 24, 32)
    L = (6.0, 5.6, 6.4)
    h = tuple(L[d] / n[d] for d in range(3))
    projs = [[_log_table(0, 1.0, 0.6, 2.4, math.log(0.3), 0.25, 9), _log_table(1, -0.5, 0.6, 2.4, math.log(0.3), 0.25, 9)],
             [_log_table(1, 0.8, 0.5, 2.0, math.log(0.2), 0.3, 12), _log_table(0, -0.3, 0.7, 1.6, math.log(0.25), 0.2, 10)]]
    pos = np.array([[0.0, 0.0, 0.0], [2.9137, 3.0411, 2.8213], [10 * h[0], 12 * h[1], 9 * h[2]], [5.7, 0.3, 6.1]])
    species = np.array([0, 1, 1, 0], np.int32)
    ng = n[0] * n[1] * n[2]
    rng = np.random.default_rng(21)
    V = rng.uniform(-1.0, 0.5, ng)
    onl = oracle.OracleNL(n, h, pos, species, projs)
    X = Q.random_states(ng, 6, h[0] * h[1] * h[2], seed=77)
    ref = oracle.hamiltonian_block(n, h, V, onl, X)
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, projs)
    integ.UpdatePosition(pos, species)
    H = Q.Hamiltonian(g, V, integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    xt = _t(X, dt, gpu)
    out = torch.empty_like(xt)
    H.apply(xt, out)
    coef = integ.InnerNonlocalPsi(xt).cpu().numpy()
    torch.cuda.synchronize()
    tol = TOL64 if prec == "f64" else TOL32
    assert rel(_np(out), ref) <= tol
    coef_ref = np.stack([onl.project(X[b].copy()) for b in range(len(X))])
    assert rel(coef, coef_ref) <= (TOL64 if prec == "f64" else 2e-6)
    loc = oracle.hamiltonian_block(n, h, V, None, X)
    assert rel(_np(out) - loc, ref - loc) <= (1e-11 if prec == "f64" else 2e-5)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nb", [1, 5, 16])
def test_rows_general_projector_sets(gpu, oracle, prec, nb):
    """The row kernels' general projector form (images that do not hold exactly one s and one p
    channel): two species, one with s + p (cutoff 2.4) and one with p first and a separate
    s cutoff (2.0 / 1.6, so an atom has two images of different channel sets), an orthorhombic
    cell with x-rows of whole 16-byte vectors, atoms on cell faces and one on a grid point."""
    import torch
    print(f"rows general prec={prec} nb={nb}", flush=True)
    n = (32, 28, 24)
    L = (6.4, 5.6, 5.1)
    h = tuple(L[d] / n[d] for d in range(3))
    pos = np.array([[0.0, 0.0, 0.0], [3.1, 2.7, 2.45], [6.35, 0.2, 5.0], [10 * h[0], 7 * h[1], 3 * h[2]],
                    [4.4, 1.3, 3.3]])
    species = np.array([0, 1, 0, 1, 1], np.int32)
    projs = [[Q.gaussian_projector(0, 1.0, 0.6, 2.4), Q.gaussian_projector(1, -0.5, 0.6, 2.4)],
             [Q.gaussian_projector(1, 0.8, 0.5, 2.0), Q.gaussian_projector(0, -0.3, 0.7, 1.6)]]
    ng = n[0] * n[1] * n[2]
    rng = np.random.default_rng(31)
    V = rng.uniform(-1.0, 0.5, ng)
    onl = oracle.OracleNL(n, h, pos, species, projs)
    X = Q.random_states(ng, nb, h[0] * h[1] * h[2], seed=5 + nb)
    ref = oracle.hamiltonian_block(n, h, V, onl, X)
    g = Q.Grid(*n, *h)
    integ = Q.PseudoPotIntegrator(g, projs)
    integ.UpdatePosition(pos, species)
    H = Q.Hamiltonian(g, V, integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    xt = _t(X, dt, gpu)
    out = torch.empty_like(xt)
    H.apply(xt, out)
    coef = integ.InnerNonlocalPsi(xt).cpu().numpy()
    torch.cuda.synchronize()
    assert rel(_np(out), ref) <= (TOL64 if prec == "f64" else TOL32)
    coef_ref = np.stack([onl.project(X[b].copy()) for b in range(nb)])
    assert rel(coef, coef_ref) <= (TOL64 if prec == "f64" else 2e-6)
    loc = oracle.hamiltonian_block(n, h, V, None, X)
    assert rel(_np(out) - loc, ref - loc) <= (1e-11 if prec == "f64" else 2e-5)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_vps_projectors_parity(gpu, oracle, prec, tmp_path):
    """H psi with the projector channels of an MBK .vps file (LoadVPS, qg_vps_projector_specs,
    PseudoPotIntegrator) against the oracle given the channels the reference builds from the
    arrays that were written (PseudoPotOperator_mpi_v2.h:64-84, :108-166: energy_up, the NandL
    cutoff of each l, the mean of the two j-components, r below r[0] taking the first value):
    the 8-atom cell at 40^3 with s and p channels of different cutoffs."""
    import torch
    import vps_synth as S
    print(f"vps projectors prec={prec}", flush=True)
    d = S.synth_pp(n=160, channels=((0, 1.0, 0.9, 0.6, 1.6), (1, -0.5, -0.45, 0.55, 1.8)), scale=3.0)
    path = tmp_path / "c.vps"
    S.write_vps(str(path), d)
    pp = Q.LoadVPS(str(path))
    xi = d["xi"]
    dxi = (xi[-1] - xi[0]) / (len(xi) - 1)
    cut = {l: c for (_, _, l, c) in d["nandl"]}
    want = [[Q.table_projector(l, d["e_up"][k], cut[l], xi[0], dxi, (d["comps"][2 * k] + d["comps"][2 * k + 1]) / 2.0,
                               r_min=d["r"][0]) for k, l in enumerate(d["l"])]]
    s = Q.diamond_system(1, 40)
    n3 = (s["n"],) * 3
    h3 = (s["h"],) * 3
    ng = s["n"] ** 3
    nat = len(s["positions"])
    onl = oracle.OracleNL(n3, h3, s["positions"], np.zeros(nat, np.int32), want)
    X = Q.random_states(ng, 8, s["h"] ** 3, seed=91)
    ref = oracle.hamiltonian_block(n3, h3, s["V"], onl, X)
    g = Q.Grid(*n3, *h3)
    integ = Q.PseudoPotIntegrator(g, [pp.projector_specs()])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    dt = torch.float64 if prec == "f64" else torch.float32
    xt = _t(X, dt, gpu)
    out = torch.empty_like(xt)
    H.apply(xt, out)
    torch.cuda.synchronize()
    assert rel(_np(out), ref) <= (TOL64 if prec == "f64" else TOL32)
    # the projector part alone: its own relative error plus the rounding of the whole result
    # (FP32 rounds H psi, of which this part is a fraction)
    loc = oracle.hamiltonian_block(n3, h3, s["V"], None, X)
    nl_err = np.linalg.norm((_np(out) - loc) - (ref - loc))
    if prec == "f64":
        assert nl_err <= 1e-11 * np.linalg.norm(ref - loc)
    else:
        assert nl_err <= 2e-5 * np.linalg.norm(ref - loc) + TOL32 * np.linalg.norm(ref)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_hpsi_plans(gpu, oracle, prec):
    """Both kernel families and explicit row-kernel shapes (through qg_hamiltonian_set_plan)
    give the oracle's H psi; asking for the row kernels where they do not serve is refused."""
    import torch
    from qumasun_amd import _native as N
    print(f"hpsi plans {prec}", flush=True)
    s = Q.diamond_system(1, 24)
    n, h = s["n"], s["h"]
    g = Q.Grid(n, n, n, h, h, h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    X = Q.random_states(g.size, 6, g.dvol)
    onl = oracle.OracleNL((n,) * 3, (h,) * 3, s["positions"], np.zeros(len(s["positions"]), np.int32),
                          s["projectors"])
    ref = oracle.hamiltonian_block((n,) * 3, (h,) * 3, s["V"], onl, X)
    dt = torch.float64 if prec == "f64" else torch.float32
    x = _t(X, dt, gpu)
    tol = TOL64 if prec == "f64" else TOL32
    plans = [dict(), dict(kernels=N.QG_KERNELS_TABLES), dict(kernels=N.QG_KERNELS_ROWS, rows_ty=4, rows_tz=8),
             dict(kernels=N.QG_KERNELS_ROWS, rows_bands=1, rows_tzp=5, rows_nbp=4)]
    for plan in plans:
        H.set_plan(**plan)
        out = torch.empty_like(x)
        H.apply(x, out)
        torch.cuda.synchronize()
        assert rel(_np(out), ref) <= tol, (plan, rel(_np(out), ref))
    # Nx = 42 FP32: rows of 16-byte vectors do not fit, so the row kernels refuse
    s2 = Q.diamond_system(1, 42)
    g2 = Q.Grid(42, 42, 42, s2["h"], s2["h"], s2["h"])
    i2 = Q.PseudoPotIntegrator(g2, s2["projectors"])
    i2.UpdatePosition(s2["positions"])
    H2 = Q.Hamiltonian(g2, s2["V"], i2)
    H2.set_plan(kernels=N.QG_KERNELS_ROWS)
    x2 = torch.zeros((1, g2.size), device=gpu, dtype=torch.float32)
    with pytest.raises(Q.QGError):
        H2.apply(x2, torch.empty_like(x2))
