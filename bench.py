#!/usr/bin/env python3
"""Benchmark of the H·psi hot path on one MI355X.

Workload (the configuration the headline H·psi metric is quoted on): the 64-atom 2x2x2
diamond supercell, L = 13.48 bohr, periodic 80^3 grid, Vloc = sum -4 erf(r/0.8)/r + U(+-0.05),
Gaussian s+p KB projectors (sigma 0.6, cutoff 2.4, E = +1.0 / -0.5), and a block of 128
seeded random real bands.  A step is one application of H (projector dots + the fused
stencil/V/projector kernel) to the whole 128-band block, inputs resident in HBM.

value = algorithmic bytes of the step / step time, algorithmic bytes = 2*B*Ng*s + Ng*s
(psi read once, H psi written once, V read once; analytic projectors, no tables).
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--bands", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-bands", type=int, default=128,
                    help="bands of the CPU baseline sample (FP64 oracle, 1 core)")
    return ap.parse_args()


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
    """The FP64 oracle (a C++ restatement of the reference's CPU path, 1 thread) on a bounded
    sample of the same workload: nbands bands of the 80^3 block."""
    import oracle as O
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
    return dict(value=bytes_ / dt / 1e9, unit="GB/s", cores=1, kind="port",
                sample=f"FP64 oracle H·psi, {nbands} bands of the 80^3/64-atom block, one pass, "
                       f"{dt:.2f} s on 1 thread of {cpu_model()}; bytes counted as 2*B*Ng*8+Ng*8")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(this build runs the whole grid on a single MI355X)", file=sys.stderr)
        sys.exit(2)
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

    # per-kernel durations of the same launches (sector hipEvents inside the library)
    H.set_timing(True)
    tp, tf = [], []
    for _ in range(args.steps):
        H.apply(psi, out)
        a, b = H.last_times()
        tp.append(a)
        tf.append(b)
    H.set_timing(False)
    t_proj = float(np.mean(tp))
    t_fused = float(np.mean(tf))

    bytes_step = 2 * args.bands * ng * s + ng * s
    value = bytes_step / (ms_step * 1e-3) / 1e9
    achieved = bytes_step / (t_fused * 1e-3) / 1e9

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
                   "projectors": "gaussian s+p, sigma 0.6, rc 2.4"},
        "roofline": {"bound": "hbm", "kernel": "fused_hpsi_kernel",
                     "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel_ms": round(t_fused, 5), "project_ms": round(t_proj, 5),
                     "algorithmic_bytes_per_launch": bytes_step},
        "wall_s_timed": round(wall, 4),
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(sysd, X, min(args.cpu_bands, args.bands))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
