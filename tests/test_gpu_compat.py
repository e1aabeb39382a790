"""The reference-shaped C++ surface (include/qumasun_compat.hpp) driven as a QUMASUN call site
would drive it: tests/harness/compat_drive.cpp calls Laplasian8th, ProjectionPP(Hp, psi,
l_grid, nuclei, n), UpdatePosition, the Hamiltonian, LOBPCG::Eigen_d_multi_mpi(l_grid, dVol,
N, iter_max, ev, X, keep_P, work, keep_S_matrix, SCF_step, OpeA, watch) with the engine's
Hamiltonian and with a per-band user operator, and EigenLOBPCG_d_multi, all on host arrays.
Every result is compared with the FP64 oracle on the same seeded inputs; the watch must have
been driven with the reference's sector ids.  A second integrator takes its channels from an
MBK .vps file through Load(Z, filepath, comm).  The k-point operator runs as the reference's
OpeA(Ax, x) on SoAComplex bands and through LOBPCG::Eigen_z_multi on planar complex bands;
Gradient8th_ddm, EnergyPPnonlocal, ForceNonlocal and EnergyForceNonlocal give the nonlocal
energy and forces."""
import os
import re
import subprocess

import numpy as np
import pytest

import qumasun_amd as Q

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_compat_surface_against_oracle(gpu, oracle, tmp_path):
    from qumasun_amd import build as B
    print("compat driver", flush=True)
    exe = B.compat_driver()
    out = tmp_path / "compat.bin"
    import vps_synth as VS
    vd = VS.synth_pp(n=160)
    vps = tmp_path / "c.vps"
    VS.write_vps(str(vps), vd)
    r = subprocess.run([exe, str(out), str(vps)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-3000:]
    s = Q.diamond_system(1, 24)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    ng = n[0] ** 3
    nb = 8
    d = np.fromfile(str(out), dtype=np.float64)
    sizes = [ng, ng, 4 * ng, nb, nb * ng, nb, ng, nb, ng, ng, 4, 8 * ng, 3 * ng, 1, 1, 24, 1, 24, ng]
    assert d.size == sum(sizes), (d.size, sum(sizes))
    parts = np.split(d, np.cumsum(sizes)[:-1])
    (lap, ppv, hp, ev, X, ev_user, hp_moved, ev_single, kre, kim, evk, Zk, grad, e0, ef, f0, eb, fb,
     ppv_vps) = parts
    X0 = Q.random_states(ng, nb, h[0] ** 3)
    nat = len(s["positions"])
    onl = oracle.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    ref_lap = np.zeros(ng)
    oracle.laplacian8(n, h, -0.5, X0[0].copy(), ref_lap)
    assert rel(lap, ref_lap) <= 1e-12
    ref_pp = np.zeros(ng)
    onl.apply(X0[0].copy(), ref_pp)
    assert rel(ppv, ref_pp) <= 1e-12
    assert rel(hp.reshape(4, ng), oracle.hamiltonian_block(n, h, s["V"], onl, X0[:4])) <= 1e-12
    Xo, Po = X0.copy(), np.zeros_like(X0)
    eo, so = oracle.lobpcg_hamiltonian(n, h, s["V"], onl, Xo, Po, 12)
    assert so["iterations"] == 12
    assert np.abs(ev - eo).max() < 1e-9, np.abs(ev - eo).max()
    ov = np.abs(np.sum(X.reshape(nb, ng) * Xo, axis=1)) * h[0] ** 3
    assert np.all(np.abs(ov - 1) < 1e-6), ov
    assert np.abs(ev_user - ev).max() < 1e-10, np.abs(ev_user - ev).max()
    # UpdatePosition on the live integrator: the Hamiltonian follows the moved nuclei
    moved = s["positions"].copy()
    for i in range(8):
        moved[i] += [0.11 * (i % 3), -0.07 * (i % 2), 0.05]
    onl_m = oracle.OracleNL(n, h, moved, np.zeros(nat, np.int32), s["projectors"])
    assert rel(hp_moved, oracle.hamiltonian_block(n, h, s["V"], onl_m, X0[:1])[0]) <= 1e-12
    # the single-process form: the reference's stop rule, dVol sum r^2 < 1e-10 dVol
    Xs, Ps = X0.copy(), np.zeros_like(X0)
    es, ss = oracle.lobpcg_hamiltonian(n, h, s["V"], onl, Xs, Ps, 40, tol=np.sqrt(1e-10 * h[0] ** 3))
    assert np.abs(ev_single - es).max() < 1e-8, np.abs(ev_single - es).max()
    # the k-point operator as OpeA(Ax, x) on SoAComplex bands, and Eigen_z_multi on planar bands
    k = (0.3, -0.2, 0.25)
    refk = oracle.kpoint_hamiltonian(n, h, s["V"], onl, k, X0[0] + 1j * X0[1])[0]
    assert rel(kre + 1j * kim, refk) <= 1e-12
    Zc = Zk.reshape(4, 2, ng)
    Zc = Zc[:, 0] + 1j * Zc[:, 1]
    dv = h[0] ** 3
    assert np.abs(dv * (Zc.conj() @ Zc.T) - np.eye(4)).max() <= 1e-10
    Rk = oracle.kpoint_hamiltonian(n, h, s["V"], onl, k, Zc) - evk[:, None] * Zc
    assert np.sqrt(dv * (np.abs(Rk) ** 2).sum(1)).max() <= 1e-7
    assert np.all(np.diff(evk) >= 0)
    # Gradient8th_ddm, EnergyPPnonlocal / ForceNonlocal on one band, EnergyForceNonlocal on the block
    gref = oracle.gradient8(X0[0], n, h, -2.0 * 0.7)
    assert rel(grad, np.concatenate(gref)) <= 1e-12
    sp_ = np.zeros(nat, np.int32)
    e_one, f_one = oracle.nonlocal_energy_force(n, h, onl, X0[:1], np.array([0.7]), sp_, s["projectors"])
    assert abs(e0[0] - e_one / 0.7) <= 1e-12 * abs(e_one / 0.7)
    assert abs(ef[0] - e0[0]) <= 1e-13 * abs(e0[0])
    assert rel(f0, f_one.reshape(-1)) <= 1e-11
    occ = np.array([2.0, 2.0, 1.5, 1.0, 0.5, 0.0, 0.25, 0.0])
    e_blk, f_blk = oracle.nonlocal_energy_force(n, h, onl, X0, occ, sp_, s["projectors"])
    assert abs(eb[0] - e_blk) <= 1e-12 * abs(e_blk)
    assert rel(fb, f_blk.reshape(-1)) <= 1e-11
    # Load(Z, filepath, comm) from the .vps file: the channels the reference builds from the
    # arrays written (energy_up, NandL cutoff per l, mean of the two j-components)
    xi = vd["xi"]
    dxi = (xi[-1] - xi[0]) / (len(xi) - 1)
    cut = {l: c for (_, _, l, c) in vd["nandl"]}
    vch = [Q.table_projector(l, vd["e_up"][k], cut[l], xi[0], dxi, (vd["comps"][2 * k] + vd["comps"][2 * k + 1]) / 2.0,
                             r_min=vd["r"][0]) for k, l in enumerate(vd["l"])]
    onl_v = oracle.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), [vch])
    ref_v = np.zeros(ng)
    onl_v.apply(X0[0].copy(), ref_v)
    assert rel(ppv_vps, ref_v) <= 1e-12
    for tag in ("watch", "watch_user"):
        m = re.search(rf"^{tag} (.*)$", r.stdout, re.M)
        assert m, r.stdout
        cnt = dict(tuple(int(v) for v in kv.split(":")) for kv in m[1].split())
        for sector in (10, 11, 15, 16, 17):
            assert cnt[sector] > 0, (tag, cnt)
        assert cnt[11] >= 12 * nb, (tag, cnt)   # H psi counted in bands: 2N at entry + N per iteration
