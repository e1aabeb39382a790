#!/usr/bin/env python3
"""Benchmark of the H·psi hot path and of the time to 512 eigenpairs on one MI355X.

Default workload (the configuration the headline metric is quoted on): the 216-atom 3x3x3
diamond supercell, L = 20.22 bohr, periodic 120^3 grid, Vloc = sum -4 erf(r/0.8)/r + U(+-0.05),
Gaussian s+p KB projectors (sigma 0.6, cutoff 2.4, E = +1.0 / -0.5), a block of 512 seeded
random real bands in FP32, inputs resident in HBM.  A step is one application of H to the
whole block: the tile-centric projector-dot launch (and its per-atom reduction), then one
stencil + V + projector-apply launch that writes H psi once.

value     = algorithmic bytes of a step / step time; algorithmic bytes = 2*B*Ng*s + Ng*s (psi
            read once, H psi written once, V read once; projector values are evaluated in the
            kernels, so no table bytes).
roofline  = the same bytes over the summed duration of all H psi launches of a step (projection
            included), measured with hipEvents inside the library on the launch stream;
            traffic = HBM bytes per step from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
            summary of this workload, scaled by its copy-kernel calibration.
secondary = the same measurement on the 64-atom / 80^3 / 128-band cell.
forces    = nonlocal energy and forces of the same block, every band occupied (time per call).
kpoint    = H_k psi of the same cell at k = (pi/L)(1,1,1) on as many complex64 bands, and the
            complex LOBPCG there (32 bands + 8 guard bands to residual 1e-4, preconditioned).
cpu_baseline = the FP64 oracle (a C++ restatement of the reference's CPU loops) on all host
            cores (bands dealt over threads; one pass over a bounded sample of the same block),
            with the single-thread figure beside it.
eigen     = seconds until the 512 eigenvalues lie within 1e-6 Ha of the FP64 fixture
            (tests/golden/cfg2_eigenvalues.json): FP32 LOBPCG with the learned preconditioner
            on the kinetic one K(0.5) (training time reported and counted) followed by FP64
            sweeps; the same with K(0.5) alone in FP32 and in FP64; plain FP64 LOBPCG (no
            preconditioner, orthonormal basis) stopped at a budget of a few times the learned
            run.  --eigen cfg0 runs the 8-atom / 40^3 / 16-band comparison instead.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
WORKLOADS = {
    "cfg2": dict(ncell=3, bands=512, desc="H psi 216-atom diamond 3x3x3, L=20.22 bohr, 120^3 grid, 512 bands"),
    "cfg1": dict(ncell=2, bands=128, desc="H psi 64-atom diamond 2x2x2, L=13.48 bohr, 80^3 grid, 128 bands"),
}
PMC_SUMMARY = os.path.join(ROOT, "profiles", "hpsi_216atoms_120cube_512bands_f32_pmc.json")
FIXTURE_CFG2 = os.path.join(ROOT, "tests", "golden", "cfg2_eigenvalues.json")
FIXTURE_CFG0 = os.path.join(ROOT, "tests", "golden", "cfg0_eigenvalues.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="cfg2")
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 80^3 / 128-band secondary H psi line")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--eigen", choices=["none", "cfg0", "cfg2"], default="cfg2",
                    help="also time LOBPCG to 1e-6 Ha: FP32 + learned preconditioner vs plain FP64")
    ap.add_argument("--eigen-budget", type=float, default=0.0,
                    help="seconds after which the plain FP64 run stops (0: 3.9 x the learned run, training "
                         "included, at least 60 s and at most 200 s)")
    ap.add_argument("--eigen-compare", action="store_true",
                    help="kept for older command lines: FP32 + kinetic and FP64 + kinetic are always timed")
    ap.add_argument("--lobpcg-iters", type=int, default=0,
                    help="only run this many LOBPCG iterations per precision (FP64, FP32, mixed) on the --eigen "
                         "workload and print the per-sector times: the command the solver's GEMM profile is taken on")
    ap.add_argument("--kpoint-only", action="store_true",
                    help="only time the k-point H psi of the --workload cell (complex64 bands) and print its line")
    ap.add_argument("--forces-only", action="store_true",
                    help="only time the nonlocal energy + forces of the --workload block and print its line: the "
                         "command the force kernels' profile is taken on")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (FP32), the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters to bytes")
    return ap.parse_args()


def log(msg):
    # progress on stderr (stdout carries the one JSON line)
    print(msg, file=sys.stderr, flush=True)


def copy_calibration():
    import torch
    x = torch.ones(64 * 1024 * 1024, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": x.numel() * 4, "copy_bytes_written_per_launch": y.numel() * 4}))


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def host_threads():
    """Host cores this process may use (the box's CPU share for one GPU is 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def build_system(ncell):
    import qumasun_amd as Q
    s = Q.diamond_system(ncell, 40)
    n, h = s["n"], s["h"]
    g = Q.Grid(n, n, n, h, h, h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    return s, g, integ, Q.Hamiltonian(g, s["V"], integ)


def traffic_per_step(dtype, workload):
    """HBM bytes per step from the committed rocprofv3 counter summary, if it matches."""
    try:
        d = json.load(open(PMC_SUMMARY))
    except Exception:
        return None
    if d.get("dtype") != dtype or d.get("workload_key") != workload:
        return None
    return int(d["hbm_bytes_per_step"])


def hpsi_leg(key, dtype, steps, warmup):
    """Time H psi on workload key: (result dict, system, psi tensor)."""
    import torch
    w = WORKLOADS[key]
    s, g, integ, H = build_system(w["ncell"])
    nb = w["bands"]
    ng = g.size
    tdt = torch.float32 if dtype == "f32" else torch.float64
    esz = 4 if dtype == "f32" else 8
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20241203)
    psi = (torch.rand((nb, ng), generator=gen, device="cuda", dtype=torch.float32) - 0.5).to(tdt).contiguous()
    out = torch.empty_like(psi)
    for _ in range(warmup):
        H.apply(psi, out)
    torch.cuda.synchronize()
    # timed region: K steps, HIP events on torch's current stream (the launch stream)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        H.apply(psi, out)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / steps
    # per-launch durations of the same kernels: hipEvents around the projection and the apply
    # launch inside the library, on the launch stream
    H.set_timing(True)
    tp, ta = [], []
    for _ in range(steps):
        H.apply(psi, out)
        a, b = H.last_times()
        tp.append(a)
        ta.append(b)
    H.set_timing(False)
    t_proj, t_apply = float(np.mean(tp)), float(np.mean(ta))
    bytes_step = 2 * nb * ng * esz + ng * esz
    achieved = bytes_step / ((t_proj + t_apply) * 1e-3) / 1e9
    res = {
        "value": round(bytes_step / (ms_step * 1e-3) / 1e9, 2),
        "ms_per_step": round(ms_step, 5),
        "wall_s_timed": round(wall, 4),
        "workload": w["desc"],
        "grid": [s["n"]] * 3, "atoms": len(s["positions"]), "bands": nb,
        "roofline": {"bound": "hbm", "kernel": "rows_project_kernel + rows_coef_reduce_kernel (projection), "
                                               "rows_hpsi_kernel (stencil + V + projector apply)",
                     "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic_per_step(dtype, key),
                     "algorithmic_bytes_per_step": bytes_step,
                     "launch_ms": round(t_proj + t_apply, 5), "project_ms": round(t_proj, 5),
                     "apply_ms": round(t_apply, 5)},
    }
    del out
    return res, (s, g, integ, H), psi


def forces_leg(integ, psi, dtype, steps):
    """Nonlocal energy and forces of the whole block (every band occupied, occ = 2): the
    gradient launch and four projections per run of bands, host-synchronised per call.
    chain bytes = psi read by the projection and the gradient, the three gradient blocks written
    and read back by their projections (8 block passes); psi once is the fused minimum."""
    nb, ng = psi.shape
    esz = 4 if dtype == "f32" else 8
    occ = np.full(nb, 2.0)
    integ.EnergyForceNonlocal(psi, occ)
    k = max(1, min(steps, 5))
    t0 = time.perf_counter()
    for _ in range(k):
        e, f = integ.EnergyForceNonlocal(psi, occ)
    ms = (time.perf_counter() - t0) / k * 1e3
    chain = 8 * nb * ng * esz
    return {"workload": f"{len(f)} atoms, {nb} occupied bands, {dtype}", "ms_per_call": round(ms, 3),
            "chain_bytes": chain, "chain_GBps": round(chain / (ms * 1e-3) / 1e9, 1),
            "psi_once_GBps": round(nb * ng * esz / (ms * 1e-3) / 1e9, 1), "energy": e}


def kpoint_leg(system, nb, steps, warmup):
    """H_k psi of the same cell at k = (pi/L, pi/L, pi/L), nb complex64 bands (the Bloch phase at
    the zone boundary; the reference's k-point operator).  bytes = 2 B Ng 8 + Ng 4 (psi read
    once, H psi written once, V once)."""
    import math
    import torch
    import qumasun_amd as Q
    s, g, integ, H = system
    kk = math.pi / s["L"]
    Hk = Q.KpointHamiltonian(H, (kk, kk, kk))
    ng = g.size
    gen = torch.Generator(device="cuda")
    gen.manual_seed(20241204)
    u = torch.complex(torch.rand((nb, ng), generator=gen, device="cuda") - 0.5,
                      torch.rand((nb, ng), generator=gen, device="cuda") - 0.5).contiguous()
    out = torch.empty_like(u)
    for _ in range(warmup):
        Hk.apply(u, out)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        Hk.apply(u, out)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    bytes_step = 2 * nb * ng * 8 + ng * 4
    del u, out
    # the complex LOBPCG at the same k point: 32 wanted bands (+8 guard bands), residual 1e-4,
    # residuals preconditioned by (T_k + 0.5)^-1
    nbl, nguard = 32, 8
    rng = np.random.default_rng(20241205)
    X = torch.from_numpy(rng.standard_normal((nbl + nguard, ng)) + 1j * rng.standard_normal((nbl + nguard, ng)))
    X = X.cuda().contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eig, it, res = Hk.Eigen_z_multi(X, iter_max=200, tol=1e-4, nconv=nbl, precond_shift=0.5)
    torch.cuda.synchronize()
    t_lob = time.perf_counter() - t0
    del X
    return {"workload": f"H_k psi, {len(s['positions'])} atoms, {s['n']}^3, {nb} complex64 bands, k = (pi/L)(1,1,1)",
            "ms_per_step": round(ms, 4), "GBps": round(bytes_step / (ms * 1e-3) / 1e9, 1),
            "algorithmic_bytes_per_step": bytes_step,
            "lobpcg": {"bands": nbl, "guard_bands": nguard, "tol": 1e-4, "precond": "(T_k + 0.5)^-1",
                       "iterations": it, "seconds": round(t_lob, 3), "max_residual": res,
                       "lowest_eigenvalues": [round(float(v), 8) for v in eig[:3]]}}


def cpu_baseline(s, psi, nb_one, nb_all):
    """The FP64 oracle on one host thread (nb_one bands) and on all host cores (nb_all bands
    dealt over threads), one pass over a bounded sample of the same block."""
    import oracle as O
    lib = O.olib()
    lib.or_set_threads(1)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    nat = len(s["positions"])
    onl = O.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    ng = n[0] * n[1] * n[2]
    X = psi[:max(nb_one, nb_all)].double().cpu().numpy()

    def run(lo, hi):
        O.hamiltonian_block(n, h, s["V"], onl, X[lo:hi])

    t0 = time.perf_counter()
    run(0, nb_one)
    dt1 = time.perf_counter() - t0
    th = host_threads()
    edges = np.linspace(0, nb_all, th + 1).astype(int)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=th) as ex:
        list(ex.map(lambda k: run(edges[k], edges[k + 1]), range(th)))
    dtn = time.perf_counter() - t0
    by = lambda nb: 2 * nb * ng * 8 + ng * 8
    model = cpu_model()
    return {"value": round(by(nb_all) / dtn / 1e9, 4), "unit": "GB/s", "cores": th, "kind": "port",
            "sample": f"FP64 oracle H psi (C++ restatement of the reference's loops), {nb_all} bands of the "
                      f"{s['n']}^3 / {nat}-atom block dealt over {th} threads, one pass, {dtn:.2f} s on {model}; "
                      f"bytes counted as 2*B*Ng*8+Ng*8",
            "one_core": {"value": round(by(nb_one) / dt1 / 1e9, 4), "unit": "GB/s", "cores": 1,
                         "sample": f"{nb_one} bands on 1 thread, {dt1:.2f} s"}}


def eigen_cpu_sample(s, X0, nb=32, iters=2):
    """The FP64 oracle's LOBPCG (the C++ restatement of the reference's Eigen_d_multi loop) on a
    bounded sample of the same problem: nb bands of the seeded start, iters iterations; dense
    products through OpenBLAS on the host threads, H psi band by band on one thread."""
    import oracle as O
    lib = O.olib()
    th = host_threads()
    lib.or_set_threads(th)
    n = (s["n"],) * 3
    h = (s["h"],) * 3
    nat = len(s["positions"])
    onl = O.OracleNL(n, h, s["positions"], np.zeros(nat, np.int32), s["projectors"])
    X = np.ascontiguousarray(X0[:nb])
    P = np.zeros_like(X)
    t0 = time.perf_counter()
    _, st = O.lobpcg_hamiltonian(n, h, s["V"], onl, X, P, iters)
    dt = time.perf_counter() - t0
    lib.or_set_threads(1)
    return {"value": round(dt / max(1, st["iterations"]), 3), "unit": "s per LOBPCG iteration", "cores": th,
            "kind": "port",
            "sample": f"FP64 oracle LOBPCG, {nb} bands of the seeded start on the {s['n']}^3 / {nat}-atom cell, "
                      f"{st['iterations']} iterations in {dt:.1f} s (dense products on {th} OpenBLAS threads, "
                      f"H psi on one thread) on {cpu_model()}"}


def eigen_leg(cfg, budget, compare):
    """Seconds to eigenvalues within 1e-6 Ha of the FP64 reference."""
    import torch
    import qumasun_amd as Q
    from qumasun_amd import _native as N
    ncell, nb = (1, 16) if cfg == "cfg0" else (3, 512)
    s, g, integ, H = build_system(ncell)
    fixture = FIXTURE_CFG0 if cfg == "cfg0" else FIXTURE_CFG2
    # 512 wanted bands and 16 guard bands (the residual tests look at the lowest 512 only): the
    # block's edge then lies in no degenerate multiplet of the wanted states, whose last members
    # would otherwise converge at the rate of a zero gap
    nguard = 0 if cfg == "cfg0" else 16
    X0 = Q.random_states(g.size, nb + nguard, g.dvol)

    def progress(label, it, el, st):
        log(f"eigen {cfg} {label}: {it} iterations, {el:.1f} s, max residual {st['max_residual']:.2e}")

    def fresh():
        X = torch.from_numpy(X0.copy()).cuda()
        return X, torch.zeros_like(X)

    ref = np.array(json.load(open(fixture))["eigenvalues"])

    def until(label, ref, budget, chunk, **kw):
        """Repeated calls (a rollback or chunk end continues with the kept P) until the eigenvalues
        are within 1e-6 Ha of ref or the budget is spent."""
        X, P = fresh()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, step = 0, 0
        while True:
            e, st = Q.Eigen_d_multi(H, X, P, step, max_iter=chunk, nconv=nb, **kw)
            step += 1
            it += st["iterations"]
            el = time.perf_counter() - t0
            progress(label, it, el, st)
            if float(np.abs(e[:nb] - ref).max()) <= 1e-6:
                return el, it, True
            if el > budget or st["iterations"] == 0:
                return el, it, False

    def ladder(label, ref, tols, **kw):
        X, P = fresh()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it = 0
        st = {}
        for step, tol in enumerate(tols):
            e, st = Q.Eigen_d_multi(H, X, P, step, max_iter=100, tol=tol, nconv=nb, **kw)
            while st["max_residual"] > tol and st["iterations"] == 100:
                it += st["iterations"]
                progress(label, it, time.perf_counter() - t0, st)
                e, st = Q.Eigen_d_multi(H, X, P, 1, max_iter=100, tol=tol, nconv=nb, **kw)
            it += st["iterations"]
            el = time.perf_counter() - t0
            progress(label, it, el, st)
            if float(np.abs(e[:nb] - ref).max()) <= 1e-6:
                return el, it, True, st
            kw["precision"] = N.QG_PREC_F64   # later sweeps refine in FP64
        return el, it, False, st

    tols = (1e-5, 1e-6) if cfg == "cfg0" else (2e-3, 1.5e-3, 1e-3, 7e-4, 5e-4, 3e-4, 1e-4)
    # the kinetic shift of the best kinetic-only run; the network is trained on top of it
    shift = 0.5
    t1 = time.perf_counter()
    sigma = float(ref[0]) - 0.7
    net, info = Q.fit_learned_preconditioner(H, min(nb, 32), sigma, shift, snapshots=(2, 5, 10), steps=300)
    t_train = time.perf_counter() - t1
    log(f"eigen {cfg}: learned preconditioner trained in {t_train:.1f} s: {info}")
    net.attach()
    t_l, it_l, ok_l, st_l = ladder("fp32-learned", ref, tols, precision=N.QG_PREC_F32, precond=N.QG_PRECOND_LEARNED,
                                   precond_shift=shift, switch_tol=1e-3)
    net.detach()
    # the same solver with the kinetic preconditioner alone, in FP32 (+ FP64 sweeps) and in FP64
    t_k, it_k, ok_k, _ = ladder("fp32-kinetic", ref, tols, precision=N.QG_PREC_F32, precond=N.QG_PRECOND_KINETIC,
                                precond_shift=shift, switch_tol=1e-3)
    chunk = 10 if cfg == "cfg0" else 50
    t_d, it_d, ok_d = until("fp64-kinetic", ref, 300.0, chunk, precision=N.QG_PREC_F64,
                            precond=N.QG_PRECOND_KINETIC, precond_shift=shift)
    if budget <= 0:
        budget = min(200.0, max(60.0, 3.9 * (t_l + t_train)))
    t_p, it_p, ok_p = until("plain-fp64", ref, budget, chunk, precision=N.QG_PREC_F64, basis=N.QG_BASIS_ORTHO)
    res = {"workload": f"{len(s['positions'])} atoms, {s['n']}^3 grid, {nb} bands"
                       + (f" (+{nguard} guard bands in the block)" if nguard else ""),
           "criterion": "max |e - e_FP64| <= 1e-6 Ha (FP64 fixture " + os.path.relpath(fixture, ROOT) + ")",
           "fp32_learned": {"seconds": round(t_l, 3), "iterations": it_l, "reached": ok_l,
                            "train_seconds": round(t_train, 3), "fallback_iter": st_l.get("fallback_iter"),
                            "train": info},
           "plain_fp64": {"seconds": round(t_p, 3), "iterations": it_p, "reached": ok_p, "budget_s": round(budget, 1),
                          "basis": "orthonormal (SVQB), no preconditioner"},
           "fp32_kinetic": {"seconds": round(t_k, 3), "iterations": it_k, "reached": ok_k, "shift": shift},
           "fp64_kinetic": {"seconds": round(t_d, 3), "iterations": it_d, "reached": ok_d, "shift": shift}}

    def ratio(t_base, ok_base, t):
        if not ok_l:
            return None
        if ok_base:
            return {"value": round(t_base / t, 3), "bound": "exact"}
        return {"value": round(t_base / t, 3), "bound": "lower (plain run stopped unconverged at its budget)"}

    res["speedup_learned_vs_plain_fp64"] = ratio(t_p, ok_p, t_l)
    res["speedup_learned_with_training_vs_plain_fp64"] = ratio(t_p, ok_p, t_l + t_train)
    if ok_k and ok_d:
        res["speedup_fp32_vs_fp64_same_kinetic_preconditioner"] = round(t_d / t_k, 3)
    if ok_k and ok_l:
        res["learned_vs_kinetic_fp32"] = {"iterations": [it_l, it_k], "seconds": [round(t_l, 3), round(t_k, 3)],
                                          "note": "the network saves iterations; its apply costs more than they do"}
    if ok_k:
        res["speedup_fp32_kinetic_vs_plain_fp64"] = ratio(t_p, ok_p, t_k) if ok_l else None
    if cfg == "cfg2":
        log("eigen cfg2: CPU sample (FP64 oracle LOBPCG)")
        res["cpu_baseline"] = eigen_cpu_sample(s, X0)
        res["cpu_baseline"]["gpu_plain_fp64_s_per_iteration_528_bands"] = round(t_p / max(1, it_p), 3)
    return res


def lobpcg_iterations(cfg, iters):
    """A few orthonormal-basis LOBPCG iterations (kinetic preconditioner) per precision, from the
    seeded start, with the solver's per-sector times; the GEMM work per iteration is stated so a
    kernel trace of this run gives the Gram / update GEMM rate."""
    import torch
    import qumasun_amd as Q
    from qumasun_amd import _native as N
    ncell, nb = (1, 16) if cfg == "cfg0" else (3, 512)
    s, g, integ, H = build_system(ncell)
    X0 = Q.random_states(g.size, nb, g.dvol)
    K = g.size
    gemm = 2.0 * nb * nb * K   # flops of one N x N x Ng product
    out = {"workload": f"{len(s['positions'])} atoms, {s['n']}^3 grid, {nb} bands", "iterations": iters,
           "flops_per_NxNxNg_product": gemm}
    for name, prec in (("f64", N.QG_PREC_F64), ("f32", N.QG_PREC_F32), ("mixed", N.QG_PREC_MIXED)):
        log(f"lobpcg {name}: {iters} iterations on {out['workload']}")
        X = torch.from_numpy(X0.copy()).cuda()
        P = torch.zeros_like(X)
        # one warm-up iteration (library handles, workspaces), then the timed ones from the same start
        Q.Eigen_d_multi(H, X, P, 0, max_iter=1, precision=prec, precond=N.QG_PRECOND_KINETIC, precond_shift=0.5,
                        basis=N.QG_BASIS_ORTHO, switch_tol=0.0)
        X = torch.from_numpy(X0.copy()).cuda()
        P = torch.zeros_like(X)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e, st = Q.Eigen_d_multi(H, X, P, 0, max_iter=iters, precision=prec, precond=N.QG_PRECOND_KINETIC,
                                precond_shift=0.5, basis=N.QG_BASIS_ORTHO, switch_tol=0.0)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        it = max(1, st["iterations"])
        out[name] = {"seconds": round(el, 3), "iterations": st["iterations"],
                     "per_iteration_ms": {k[2:]: round(st[k] / it * 1e3, 2)
                                          for k in ("t_gram", "t_update", "t_eig", "t_hpsi", "t_precond")}}
        log(f"lobpcg {name}: {st['iterations']} iterations in {el:.2f} s")
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine runs the whole grid on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.kpoint_only:
        w = WORKLOADS[args.workload]
        print(json.dumps({"kpoint": kpoint_leg(build_system(w["ncell"]), w["bands"], args.steps, args.warmup)}))
        return
    if args.forces_only:
        import torch
        w = WORKLOADS[args.workload]
        s_, g, integ, _ = build_system(w["ncell"])
        gen = torch.Generator(device="cuda")
        gen.manual_seed(20241203)
        tdt = torch.float32 if args.dtype == "f32" else torch.float64
        psi = (torch.rand((w["bands"], g.size), generator=gen, device="cuda", dtype=torch.float32) - 0.5).to(tdt)
        print(json.dumps({"forces": forces_leg(integ, psi.contiguous(), args.dtype, args.steps)}))
        return
    if args.lobpcg_iters > 0:
        lobpcg_iterations(args.eigen if args.eigen != "none" else "cfg2", args.lobpcg_iters)
        return
    import torch
    log(f"H psi {args.workload} {args.dtype}: {args.warmup} warmup + {args.steps} timed steps")
    r, system, psi = hpsi_leg(args.workload, args.dtype, args.steps, args.warmup)
    res = {
        "metric": "Hpsi HBM GB/s (algorithmic bytes / step time); s to 512 eigenpairs @1e-6 Ha, FP32+learned vs "
                  "plain FP64 in 'eigen'",
        "value": r["value"],
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded diamond cell, seeded random bands)",
        "config": {"workload": r["workload"], "grid": r["grid"], "atoms": r["atoms"], "bands": r["bands"],
                   "projectors": "gaussian s+p, sigma 0.6, rc 2.4"},
        "roofline": r["roofline"],
        "wall_s_timed": r["wall_s_timed"],
    }
    if not args.no_cpu:
        log("CPU baseline (FP64 oracle)")
        nb = r["bands"]
        res["cpu_baseline"] = cpu_baseline(system[0], psi, min(nb, 32 if args.workload == "cfg2" else 64),
                                           min(nb, 256 if args.workload == "cfg2" else 128))
    log("nonlocal energy + forces")
    res["forces"] = forces_leg(system[2], psi, args.dtype, args.steps)
    del psi
    torch.cuda.empty_cache()
    log("k-point H psi")
    res["kpoint"] = kpoint_leg(system, r["bands"], args.steps, args.warmup)
    del system
    torch.cuda.empty_cache()
    if not args.no_secondary and args.workload != "cfg1":
        log("H psi cfg1 (secondary)")
        r1, sys1, psi1 = hpsi_leg("cfg1", args.dtype, args.steps, args.warmup)
        res["secondary"] = {"workload": r1["workload"], "value": r1["value"], "unit": "GB/s",
                            "ms_per_step": r1["ms_per_step"], "roofline_frac": r1["roofline"]["frac"],
                            "project_ms": r1["roofline"]["project_ms"], "apply_ms": r1["roofline"]["apply_ms"]}
        del psi1, sys1
        torch.cuda.empty_cache()
    if args.eigen != "none":
        try:
            res["eigen"] = eigen_leg(args.eigen, args.eigen_budget, args.eigen_compare)
        except Exception as exc:   # the H psi line stands on its own
            res["eigen"] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
