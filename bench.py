#!/usr/bin/env python3
"""Benchmark of the H·psi hot path (and, optionally, time to eigenpairs) on one MI355X.

H·psi workload (the configuration the H·psi bandwidth metric is quoted on): the 64-atom
2x2x2 diamond supercell, L = 13.48 bohr, periodic 80^3 grid, Vloc = sum -4 erf(r/0.8)/r +
U(+-0.05), Gaussian s+p KB projectors (sigma 0.6, cutoff 2.4, E = +1.0 / -0.5), and a block of
128 seeded random real bands in FP32.  A step is one application of H to the whole 128-band
block with inputs resident in HBM: per band chunk, the projector-dot kernel, then (FP32 default)
a stencil + V launch and a projector-term launch with in-kernel projector values, or one fused
stencil + V + projector launch reading per-point tables.

value     = algorithmic bytes of a step / step time, algorithmic bytes = 2*B*Ng*s + Ng*s (psi
            read once, H psi written once, V read once; the projector tables are reported
            separately as table_bytes).
roofline  = the H psi launches of one band chunk: its algorithmic bytes / their average
            duration (hipEvents around them, on the launch stream); traffic = HBM bytes per
            chunk from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
            workload, scaled by the copy-kernel calibration factor recorded there.
cpu_baseline = the FP64 oracle (a C++ restatement of the reference's CPU path) on 1 host
            thread, one pass over a bounded sample of the same block.
--eigen cfg0|cfg2|none (default cfg0, 8 atoms / 40^3 / 16 bands, ~15 s): also time the
            LOBPCG solve to eigenvalues within 1e-6 Ha of the FP64 reference: plain FP64 (the
            reference's algorithm, and on the orthonormal basis) vs FP32 + learned
            preconditioner.  cfg2 (216 atoms / 120^3 / 512 bands) takes ~15 minutes with
            --eigen-budget 650.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "fused_hpsi_80cube_128bands_f32_pmc.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--bands", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-bands", type=int, default=128,
                    help="bands of the CPU baseline sample (FP64 oracle, 1 thread)")
    ap.add_argument("--eigen", choices=["none", "cfg0", "cfg2"], default="cfg0",
                    help="also time LOBPCG to 1e-6 Ha: plain FP64 vs FP32 + learned preconditioner")
    ap.add_argument("--eigen-budget", type=float, default=600.0,
                    help="seconds after which the plain FP64 eigen runs stop (reported as unconverged)")
    ap.add_argument("--lobpcg-iters", type=int, default=0,
                    help="only run this many LOBPCG iterations per precision (FP64, FP32, mixed) on the --eigen "
                         "workload and print the per-sector times: the command the solver's GEMM profile is taken on")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 256 MiB (FP32), the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters to bytes")
    return ap.parse_args()


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


def cpu_baseline(sysd, X, nbands):
    """The FP64 oracle on 1 host thread, one pass over nbands bands of the 80^3 block."""
    import oracle as O
    O.olib().or_set_threads(1)
    n = (sysd["n"],) * 3
    h = (sysd["h"],) * 3
    nat = len(sysd["positions"])
    onl = O.OracleNL(n, h, sysd["positions"], np.zeros(nat, np.int32), sysd["projectors"])
    sample = np.ascontiguousarray(X[:nbands])
    t0 = time.perf_counter()
    O.hamiltonian_block(n, h, sysd["V"], onl, sample)
    dt = time.perf_counter() - t0
    ng = n[0] * n[1] * n[2]
    bytes_ = 2 * nbands * ng * 8 + ng * 8
    return dict(value=round(bytes_ / dt / 1e9, 4), unit="GB/s", cores=1, kind="port",
                sample=f"FP64 oracle H·psi (C++ restatement of the reference's loops), {nbands} bands of "
                       f"the 80^3/64-atom block, one pass, {dt:.2f} s on 1 thread of {cpu_model()}; "
                       f"bytes counted as 2*B*Ng*8+Ng*8")


def traffic_per_launch(dtype, chunk_bands):
    """HBM bytes per fused launch from the committed rocprofv3 counter summary, if it matches."""
    try:
        d = json.load(open(PMC_SUMMARY))
    except Exception:
        return None
    if d.get("dtype") != dtype or d.get("chunk_bands") != chunk_bands:
        return None
    return int(d["hbm_bytes_per_launch"])


def eigen_leg(cfg, budget):
    """Seconds to eigenvalues within 1e-6 Ha of the FP64 reference: plain FP64 LOBPCG (no
    preconditioner; the reference's [X R P] iteration, and the same on the orthonormal basis)
    vs FP32 with the learned preconditioner (and with the kinetic one). The FP32 runs are single
    solver calls with a residual tolerance; a run that has not reached 1e-6 Ha continues with
    FP64 sweeps at the next tolerance of the ladder. Plain runs stop at `budget` seconds."""
    import torch
    import qumasun_amd as Q
    from qumasun_amd import _native as N
    ncell, nb = (1, 16) if cfg == "cfg0" else (3, 512)
    s = Q.diamond_system(ncell, 40)
    n, h = s["n"], s["h"]
    g = Q.Grid(n, n, n, h, h, h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    X0 = Q.random_states(g.size, nb, g.dvol)

    def progress(label, it, el, st):
        # progress on stderr (stdout carries the one JSON line)
        print(f"eigen {cfg} {label}: {it} iterations, {el:.1f} s, max residual {st['max_residual']:.2e}",
              file=sys.stderr, flush=True)

    def solve(label, chunk, ref=None, budget=600.0, **kw):
        X = torch.from_numpy(X0.copy()).cuda()
        P = torch.zeros_like(X)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, step = 0, 0
        while True:
            e, st = Q.Eigen_d_multi(H, X, P, step, max_iter=chunk, **kw)
            step += 1
            it += st["iterations"]
            el = time.perf_counter() - t0
            progress(label, it, el, st)
            if ref is None and st["max_residual"] <= kw.get("tol", 0.0):
                return e, el, it, True
            if ref is not None and float(np.abs(e - ref).max()) <= 1e-6:
                return e, el, it, True
            if el > budget or st["iterations"] == 0:
                return e, el, it, False

    def ladder(label, ref, tols, **kw):
        X = torch.from_numpy(X0.copy()).cuda()
        P = torch.zeros_like(X)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it = 0
        for step, tol in enumerate(tols):
            # at most 100 iterations per call, so progress lines keep coming
            e, st = Q.Eigen_d_multi(H, X, P, step, max_iter=100, tol=tol, **kw)
            while st["max_residual"] > tol and st["iterations"] == 100:
                it += st["iterations"]
                progress(label, it, time.perf_counter() - t0, st)
                e, st = Q.Eigen_d_multi(H, X, P, 1, max_iter=100, tol=tol, **kw)
            it += st["iterations"]
            el = time.perf_counter() - t0
            progress(label, it, el, st)
            if float(np.abs(e - ref).max()) <= 1e-6:
                return el, it, True
            kw["precision"] = N.QG_PREC_F64   # later sweeps refine in FP64
        return el, it, False

    ref, t_ref, it_ref, _ = solve("reference", 400 if cfg == "cfg0" else 50, precision=N.QG_PREC_F64,
                                  precond=N.QG_PRECOND_KINETIC, precond_shift=0.5,
                                  tol=1e-9 if cfg == "cfg0" else 1e-6, budget=1e9)
    chunk = 10 if cfg == "cfg0" else 50
    _, t_plain, it_plain, ok_plain = solve("plain", chunk, ref=ref, budget=min(budget, 60.0),
                                           precision=N.QG_PREC_F64)
    _, t_plo, it_plo, ok_plo = solve("plain-ortho", chunk, ref=ref, budget=budget, precision=N.QG_PREC_F64,
                                     basis=N.QG_BASIS_ORTHO)
    t1 = time.perf_counter()
    sigma = float(ref[0]) - 0.7
    net, info = Q.fit_learned_preconditioner(H, min(nb, 32), sigma, 2.0, snapshots=(2, 5, 10), steps=300)
    t_train = time.perf_counter() - t1
    net.attach()
    tols = (1e-5, 1e-6) if cfg == "cfg0" else (2e-3, 1.5e-3, 1e-3, 7e-4, 5e-4, 3e-4, 1e-4)
    t_l, it_l, ok_l = ladder("fp32-learned", ref, tols, precision=N.QG_PREC_F32, precond=N.QG_PRECOND_LEARNED,
                             precond_shift=2.0, switch_tol=1e-3)
    t_k, it_k, ok_k = ladder("fp32-kinetic", ref, tols, precision=N.QG_PREC_F32, precond=N.QG_PRECOND_KINETIC,
                             precond_shift=0.5, switch_tol=1e-3)
    net.detach()

    def ratio(t_base, ok_base, base_budget, t, ok):
        if not ok:
            return None
        if ok_base:
            return {"value": round(t_base / t, 3), "bound": "exact"}
        if t_base >= base_budget:
            return {"value": round(t_base / t, 3), "bound": "lower (plain run stopped unconverged at its budget)"}
        return None   # the plain run stalled before its budget: no time to compare with

    return {"workload": f"{len(s['positions'])} atoms, {n}^3 grid, {nb} bands",
            "criterion": "max |e - e_FP64| <= 1e-6 Ha",
            "plain_fp64": {"seconds": round(t_plain, 3), "iterations": it_plain, "reached": ok_plain,
                           "basis": "reference ([X R P] generalized, rollback exit)"},
            "plain_fp64_ortho": {"seconds": round(t_plo, 3), "iterations": it_plo, "reached": ok_plo,
                                 "basis": "orthonormal (SVQB)", "budget_s": budget},
            "fp32_learned": {"seconds": round(t_l, 3), "iterations": it_l, "reached": ok_l,
                             "train_seconds": round(t_train, 3), "train": info},
            "fp32_kinetic": {"seconds": round(t_k, 3), "iterations": it_k, "reached": ok_k},
            "speedup_learned_vs_plain_ortho": ratio(t_plo, ok_plo, budget, t_l, ok_l),
            "speedup_learned_with_training_vs_plain_ortho": ratio(t_plo, ok_plo, budget, t_l + t_train, ok_l),
            "speedup_learned_vs_plain": ratio(t_plain, ok_plain, min(budget, 60.0), t_l, ok_l),
            "reference_fp64_kinetic": {"seconds": round(t_ref, 3), "iterations": it_ref}}


def lobpcg_iterations(cfg, iters):
    """A few orthonormal-basis LOBPCG iterations (kinetic preconditioner) per precision, from the
    seeded start, with the solver's per-sector times; the GEMM work per iteration is stated so a
    kernel trace of this run gives the Gram / update GEMM rate."""
    import torch
    import qumasun_amd as Q
    from qumasun_amd import _native as N
    ncell, nb = (1, 16) if cfg == "cfg0" else (3, 512)
    s = Q.diamond_system(ncell, 40)
    n, h = s["n"], s["h"]
    g = Q.Grid(n, n, n, h, h, h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    H = Q.Hamiltonian(g, s["V"], integ)
    X0 = Q.random_states(g.size, nb, g.dvol)
    K = g.size
    gemm = 2.0 * nb * nb * K   # flops of one N x N x Ng product
    out = {"workload": f"{len(s['positions'])} atoms, {n}^3 grid, {nb} bands", "iterations": iters,
           "flops_per_NxNxNg_product": gemm}
    for name, prec in (("f64", N.QG_PREC_F64), ("f32", N.QG_PREC_F32), ("mixed", N.QG_PREC_MIXED)):
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
        print(f"lobpcg {name}: {st['iterations']} iterations in {el:.2f} s", file=sys.stderr, flush=True)
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
    if args.lobpcg_iters > 0:
        lobpcg_iterations(args.eigen if args.eigen != "none" else "cfg2", args.lobpcg_iters)
        return
    import torch
    import qumasun_amd as Q

    dev = torch.device("cuda:0")
    sysd = Q.diamond_system(2, 40)
    n = sysd["n"]
    ng = n ** 3
    h = sysd["h"]
    g = Q.Grid(n, n, n, h, h, h)
    integ = Q.PseudoPotIntegrator(g, sysd["projectors"])
    integ.UpdatePosition(sysd["positions"])
    H = Q.Hamiltonian(g, sysd["V"], integ)
    X = Q.random_states(ng, args.bands, h ** 3)
    dt = torch.float32 if args.dtype == "f32" else torch.float64
    s = 4 if args.dtype == "f32" else 8
    psi = torch.from_numpy(X).to(dev).to(dt).contiguous()
    out = torch.empty_like(psi)

    for _ in range(args.warmup):
        H.apply(psi, out)
    torch.cuda.synchronize()

    # timed region: K steps, HIP events on torch's current stream (the launch stream)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        H.apply(psi, out)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_time(e1) / args.steps

    # per-launch durations of the same kernels (hipEvents around each launch inside the library)
    H.set_timing(True)
    tp, tf = [], []
    for _ in range(args.steps):
        H.apply(psi, out)
        a, b = H.last_times()
        tp.append(a)
        tf.append(b)
    nchunks = H.last_chunks()
    H.set_timing(False)
    t_proj = float(np.mean(tp))
    t_fused = float(np.mean(tf))
    chunk_bands = (args.bands + nchunks - 1) // nchunks

    bytes_step = 2 * args.bands * ng * s + ng * s
    bytes_launch = 2 * chunk_bands * ng * s + ng * s
    value = bytes_step / (ms_step * 1e-3) / 1e9
    fused_launch_ms = t_fused / nchunks
    achieved = bytes_launch / (fused_launch_ms * 1e-3) / 1e9
    table_bytes = int(round(1.51 * ng)) * 4 * s   # beta tables, 4 slots per sphere point
    split = (args.dtype == "f32" and integ.projector_form == "kernel" and os.environ.get("QG_SPLIT_NL", "") != "0") \
        or os.environ.get("QG_SPLIT_NL", "") == "1"
    kernel_desc = ("fused_hpsi_kernel: stencil+V pass + projector pass (in-kernel projector values) per band chunk"
                   if split else "fused_hpsi_kernel: one fused stencil+V+projector pass per band chunk")

    res = {
        "metric": "Hpsi HBM GB/s (algorithmic bytes / step time)",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded diamond cell, seeded random bands)",
        "config": {"workload": "Hpsi 64-atom diamond 2x2x2, L=13.48 bohr, 80^3 grid, 128 bands",
                   "grid": [n, n, n], "atoms": len(sysd["positions"]), "bands": args.bands,
                   "projectors": "gaussian s+p, sigma 0.6, rc 2.4", "band_chunks": nchunks},
        "roofline": {"bound": "hbm", "kernel": kernel_desc,
                     "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic_per_launch(args.dtype, chunk_bands),
                     "algorithmic_bytes_per_launch": bytes_launch, "launch_ms": round(fused_launch_ms, 5),
                     "project_ms_per_step": round(t_proj, 5), "fused_ms_per_step": round(t_fused, 5),
                     "table_bytes": table_bytes},
        "wall_s_timed": round(wall, 4),
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(sysd, X, min(args.cpu_bands, args.bands))
    if args.eigen != "none":
        try:
            res["eigen"] = eigen_leg(args.eigen, args.eigen_budget)
        except Exception as exc:   # the H psi line stands on its own
            res["eigen"] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
