"""CPU checks of the C ABI library: it loads, exports every entry point include/*.h
declares, rejects bad arguments with a message, and its host-side synthetic-cell helpers
reproduce their definitions (the std::mt19937 / generate_canonical stream of the
reference's random start, qumasun_mpi.initialize.h:196-214)."""
import ctypes
import math
import os
import re

import numpy as np
from scipy.special import erf

import oracle as O
import qumasun_amd as Q
from qumasun_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^(?:const\s+)?\w+\s*\*?\s*(qg_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(N.SIGNATURES) >= names, names - set(N.SIGNATURES)


def test_bad_arguments_are_rejected():
    h = ctypes.c_void_p()
    rc = Q.lib().qg_grid_create(2, 8, 8, 0.1, 0.1, 0.1, ctypes.byref(h))
    assert rc == 1
    assert b"4 points" in Q.lib().qg_last_error()
    rc = Q.lib().qg_grid_create(8, 8, 8, -0.1, 0.1, 0.1, ctypes.byref(h))
    assert rc == 1


def canonical_doubles(seed, count):
    """std::uniform_real_distribution<double>(0,1) on std::mt19937(seed): two 32-bit draws
    per value, (lo + hi * 2^32) / 2^64."""
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    raw = bg.random_raw(2 * count).astype(np.float64)
    return (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64


def test_random_states_stream():
    ng, nb, dv = 1000, 3, 0.01
    X = Q.random_states(ng, nb, dv, seed=123456789)
    u = canonical_doubles(123456789, (ng + 1) * nb)
    for b in range(nb):
        v = u[b * (ng + 1): b * (ng + 1) + ng]
        v = v / math.sqrt((v * v).sum() * dv)
        assert np.allclose(X[b], v, rtol=1e-15, atol=0)
    Xo, _ = O.random_states(ng, nb, dv, 123456789)
    assert np.array_equal(X, Xo)


def test_vloc_and_atoms():
    xyz = np.zeros(3 * 8)
    N.check(Q.lib().qg_synth_diamond_atoms(1, 6.74, N.dptr(xyz)))
    frac = xyz.reshape(8, 3) / 6.74
    assert np.allclose(frac[1], [0.25, 0.25, 0.25]) and np.allclose(frac[7], [0.25, 0.75, 0.75])
    n, L = 12, 6.74
    V = np.zeros(n ** 3)
    N.check(Q.lib().qg_synth_vloc(n, n, n, L, L, L, 8, N.dptr(xyz), 4.0, 0.8, 0.05, 99, N.dptr(V)))
    noise = 0.05 * (2 * canonical_doubles(99, n ** 3) - 1)
    iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1) * (L / n)
    ref = noise.copy()
    for R in xyz.reshape(8, 3):
        d = pts - R
        d -= L * np.round(d / L)
        r = np.sqrt((d ** 2).sum(1))
        with np.errstate(invalid="ignore", divide="ignore"):
            v = np.where(r < 1e-12, -4.0 * 2 / (math.sqrt(math.pi) * 0.8), -4.0 * erf(r / 0.8) / r)
        ref += v
    assert np.abs(V - ref).max() < 1e-12


def test_compat_header_compiles_and_links(tmp_path):
    """include/qumasun_compat.hpp (the reference-shaped C++ surface): the driver program of
    tests/harness/compat_drive.cpp, written against it alone, compiles and links against the
    library (it is run on the GPU by tests/test_gpu_compat.py)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    libdir = os.path.dirname(N.LIB_PATH)
    exe = tmp_path / "compat"
    r = subprocess.run([hipcc, "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "harness", "compat_drive.cpp"), "-o", str(exe),
                        "-L" + libdir, "-lqumasun_gpu", "-Wl,-rpath," + libdir],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    assert os.path.exists(exe)
