"""GPU test of the allocation-free H psi entry point (qg_hamiltonian_apply_noalloc): through the
C ABI its output equals qg_hamiltonian_apply's (FP64 to 1e-12, FP32 to 2e-6 relative) on two
calls in a row, the second with fewer bands; the device's free memory does not change across
its calls; a projector set that needs several band chunks of the fixed scratch gives the same
result; a grid the row kernels do not serve is rejected with QG_ERR_ARG."""
import ctypes
import os

import numpy as np
import pytest

import qumasun_amd as Q
from qumasun_amd import _native as N

pytestmark = pytest.mark.gpu


QG_ERR_ARG = 1


def last_error():
    return Q.lib().qg_last_error().decode(errors="replace")


def _cell(ncell, npc):
    s = Q.diamond_system(ncell, npc)
    n, h = s["n"], s["h"]
    g = Q.Grid(n, n, n, h, h, h)
    integ = Q.PseudoPotIntegrator(g, s["projectors"])
    integ.UpdatePosition(s["positions"])
    return s, g, integ, Q.Hamiltonian(g, s["V"], integ)


def _apply(fn, H, x, out):
    import torch
    ng = H.grid.size
    nb = x.shape[0]
    dt = N.QG_F64 if x.dtype == torch.float64 else N.QG_F32
    rc = fn(H._h, dt, nb, ctypes.c_void_p(x.data_ptr()), ng, ctypes.c_void_p(out.data_ptr()), ng, None)
    torch.cuda.synchronize()
    return rc


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_matches_apply_two_calls(gpu, prec):
    import torch
    print(f"noalloc two calls {prec}", flush=True)
    lib = Q.lib()
    s, g, integ, H = _cell(1, 40)
    dt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 2e-6
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    for nb in (24, 7):   # second call smaller than the first
        x = torch.rand((nb, g.size), generator=gen, device=gpu, dtype=torch.float64).to(dt) - 0.5
        ref = torch.empty_like(x)
        got = torch.full_like(x, float("nan"))
        assert _apply(lib.qg_hamiltonian_apply, H, x, ref) == N.QG_OK, last_error()
        free0 = torch.cuda.mem_get_info()[0]
        assert _apply(lib.qg_hamiltonian_apply_noalloc, H, x, got) == N.QG_OK, last_error()
        assert torch.cuda.mem_get_info()[0] == free0
        err = float((got - ref).norm() / ref.norm())
        assert err <= tol, (nb, err)


def test_noalloc_band_chunks(gpu, monkeypatch):
    """120^3, 216 atoms, projection slabs of 4 planes: a band needs ~37k doubles of projector
    scratch, so 512 bands run in several chunks of the fixed array."""
    import torch
    print("noalloc band chunks 120^3", flush=True)
    monkeypatch.setenv("QG_ROWS_TZP", "4")   # read when the projector geometry is prepared
    lib = Q.lib()
    s, g, integ, H = _cell(3, 40)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(12)
    x = torch.rand((512, g.size), generator=gen, device=gpu, dtype=torch.float32) - 0.5
    ref = torch.empty_like(x)
    got = torch.full_like(x, float("nan"))
    assert _apply(lib.qg_hamiltonian_apply, H, x, ref) == N.QG_OK, last_error()
    assert _apply(lib.qg_hamiltonian_apply_noalloc, H, x, got) == N.QG_OK, last_error()
    err = float((got - ref).norm() / ref.norm())
    assert err <= 2e-6, err
    del x, ref, got


def test_noalloc_rejects_unserved_grid(gpu):
    import torch
    lib = Q.lib()
    s, g, integ, H = _cell(1, 42)   # Nx = 42: FP32 rows are not whole 16-byte vectors
    x = torch.zeros((2, g.size), device=gpu, dtype=torch.float32)
    out = torch.empty_like(x)
    assert _apply(lib.qg_hamiltonian_apply_noalloc, H, x, out) == QG_ERR_ARG
    assert "row kernels" in last_error()
