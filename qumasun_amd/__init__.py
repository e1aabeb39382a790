"""qumasun_amd — MI355X engine for QUMASUN's Kohn-Sham H·psi and block-LOBPCG path.

Python mirror of the reference's C++ call surface for this path, on top of the C ABI
(include/qumasun_gpu.h, libqumasun_gpu.so).  Device arrays are torch tensors on a ROCm
device (torch is only used for device memory and streams); all arithmetic runs in the
library's HIP kernels.

Reference interfaces mirrored (QUMASUN source/):
  Laplasian8th(grid, out, p, coef, ...)           GridDifference8th.h:177-185 (accumulates)
  PseudoPotIntegrator.UpdatePosition / ProjectionPP PseudoPotOperator_mpi_v2.h:231, :465
  Hamiltonian.apply                                qumasun_mpi.matrix.h:21-43
  Eigen_d_multi(...)                               lobpcg_d_multi_mpi.h:882-1572
"""
import ctypes
import math

import numpy as np

from . import _native as N
from ._native import (QG_F32, QG_F64, QG_PREC_F32, QG_PREC_F64, QG_PREC_MIXED, QG_PRECOND_KINETIC,
                      QG_PRECOND_NONE, QGError, check, dptr, lib)

__all__ = ["Grid", "Laplasian8th", "PseudoPotIntegrator", "Hamiltonian", "Eigen_d_multi",
           "gaussian_projector", "diamond_system", "QGError"]


def _dtype_code(t):
    import torch
    if t.dtype == torch.float64:
        return QG_F64
    if t.dtype == torch.float32:
        return QG_F32
    raise QGError(f"unsupported dtype {t.dtype} (float64 or float32)")


def _block(t, ng):
    """(pointer, nbands, ld) of a band-major device block: shape (nbands, ...) or (ng,)."""
    if not t.is_cuda:
        raise QGError("device tensor required (ROCm torch tensor)")
    if not t.is_contiguous():
        raise QGError("contiguous tensor required")
    n = t.numel()
    if n % ng:
        raise QGError(f"tensor of {n} elements is not a whole number of {ng}-point bands")
    return ctypes.c_void_p(t.data_ptr()), n // ng, ng


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Grid:
    """Periodic grid (GridRange + spacings).  Index i = ix + Nx*(iy + Ny*iz)."""

    def __init__(self, nx, ny, nz, dx, dy, dz):
        self.n = (int(nx), int(ny), int(nz))
        self.h = (float(dx), float(dy), float(dz))
        h = ctypes.c_void_p()
        check(lib().qg_grid_create(nx, ny, nz, dx, dy, dz, ctypes.byref(h)))
        self._h = h

    @property
    def size(self):
        return self.n[0] * self.n[1] * self.n[2]

    @property
    def dvol(self):
        return self.h[0] * self.h[1] * self.h[2]

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().qg_grid_destroy(self._h)
            except Exception:
                pass
            self._h = None


def Laplasian8th(grid, out, p, coef, accumulate=True):
    """out += coef * Lap8 p (periodic), for every band of the block (GridDifference8th.h:177)."""
    pp, nb, ld = _block(p, grid.size)
    op, nb2, ldo = _block(out, grid.size)
    if nb != nb2 or p.dtype != out.dtype:
        raise QGError("psi/out block mismatch")
    check(lib().qg_laplacian8_apply(grid._h, _dtype_code(p), nb, pp, ld, op, ldo, float(coef),
                                    int(bool(accumulate)), _stream()))
    return out


def gaussian_projector(l, energy, sigma=0.6, cutoff=2.4):
    """Gaussian radial KB channel: p(r) = A exp(-r^2/(2 sigma^2)), int p^2 r^2 dr = 1."""
    amp = math.sqrt(4.0 / (math.sqrt(math.pi) * sigma ** 3))
    return dict(l=l, energy=energy, cutoff=cutoff, kind="gaussian", sigma=sigma, amplitude=amp)


def table_projector(l, energy, cutoff, xi_min, xi_delta, values, r_min=None):
    """Tabulated radial KB channel on the log grid xi = log r (RadialGrid2 conventions)."""
    values = np.ascontiguousarray(values, dtype=np.float64)
    if r_min is None:
        r_min = math.exp(xi_min)
    return dict(l=l, energy=energy, cutoff=cutoff, kind="table", xi_min=xi_min,
                xi_delta=xi_delta, values=values, r_min=r_min)


def _specs_array(species_projectors):
    flat = [p for sp in species_projectors for p in sp]
    arr = (N.ProjSpec * max(1, len(flat)))()
    keep = []
    for i, p in enumerate(flat):
        s = arr[i]
        s.l, s.energy, s.cutoff = p["l"], p["energy"], p["cutoff"]
        if p["kind"] == "gaussian":
            s.radial_kind = N.QG_RADIAL_GAUSSIAN
            s.sigma, s.amplitude = p["sigma"], p["amplitude"]
        else:
            s.radial_kind = N.QG_RADIAL_TABLE
            v = p["values"]
            keep.append(v)
            s.n_table, s.xi_min, s.xi_delta, s.r_min = len(v), p["xi_min"], p["xi_delta"], p["r_min"]
            s.table = dptr(v)
    return arr, keep


class PseudoPotIntegrator:
    """Nonlocal KB projectors on one GPU holding the whole grid (PseudoPotIntegrator_mpi).

    species_projectors: list (one entry per species) of lists of projector dicts.
    """

    def __init__(self, grid, species_projectors):
        self.grid = grid
        self.species_projectors = species_projectors
        self._h = None

    def UpdatePosition(self, positions, species=None):
        """positions: (natoms, 3) bohr; species: per-atom species index (default 0)."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        nat = pos.shape[0]
        sp = np.zeros(nat, dtype=np.int32) if species is None else np.ascontiguousarray(species, dtype=np.int32)
        nproj = np.array([len(s) for s in self.species_projectors], dtype=np.int32)
        specs, keep = _specs_array(self.species_projectors)
        h = ctypes.c_void_p()
        check(lib().qg_nonlocal_create(self.grid._h, nat, dptr(pos),
                                       sp.ctypes.data_as(N.P(ctypes.c_int)), len(nproj),
                                       nproj.ctypes.data_as(N.P(ctypes.c_int)), specs, ctypes.byref(h)))
        if self._h:
            lib().qg_nonlocal_destroy(self._h)
        self._h = h
        self.positions = pos
        self.species = sp

    @property
    def num_coef(self):
        return lib().qg_nonlocal_num_coef(self._h)

    def InnerNonlocalPsi(self, psi):
        """coef[b, j] = dV sum beta_j psi_b (mInnerNonlocalPsi_v2), FP64 device tensor."""
        import torch
        pp, nb, ld = _block(psi, self.grid.size)
        coef = torch.empty((nb, self.num_coef), dtype=torch.float64, device=psi.device)
        check(lib().qg_nonlocal_project(self._h, _dtype_code(psi), nb, pp, ld,
                                        ctypes.c_void_p(coef.data_ptr()), _stream()))
        return coef

    def ProjectionPP(self, Hp, psi, accumulate=True):
        """Hp += V_NL psi for every band (ProjectionPP accumulates into Hp)."""
        pp, nb, ld = _block(psi, self.grid.size)
        op, nb2, ldo = _block(Hp, self.grid.size)
        if nb != nb2:
            raise QGError("block mismatch")
        check(lib().qg_nonlocal_apply(self._h, _dtype_code(psi), nb, pp, ld, op, ldo,
                                      int(bool(accumulate)), _stream()))
        return Hp

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().qg_nonlocal_destroy(self._h)
            except Exception:
                pass
            self._h = None


class Hamiltonian:
    """H = V + V_NL - 1/2 Lap8 (mHamiltonianMatrix_ddm)."""

    def __init__(self, grid, V, integrator=None):
        self.grid = grid
        self.integrator = integrator
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        if V.size != grid.size:
            raise QGError("potential size mismatch")
        h = ctypes.c_void_p()
        check(lib().qg_hamiltonian_create(grid._h, dptr(V), integrator._h if integrator else None,
                                          ctypes.byref(h)))
        self._h = h

    def set_potential(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        check(lib().qg_hamiltonian_set_potential(self._h, dptr(V)))

    def apply(self, psi, out):
        pp, nb, ld = _block(psi, self.grid.size)
        op, nb2, ldo = _block(out, self.grid.size)
        if nb != nb2 or psi.dtype != out.dtype:
            raise QGError("block mismatch")
        check(lib().qg_hamiltonian_apply(self._h, _dtype_code(psi), nb, pp, ld, op, ldo, _stream()))
        return out

    def set_timing(self, enable=True):
        check(lib().qg_hamiltonian_set_timing(self._h, int(bool(enable))))

    def last_times(self):
        """(projector-dot ms, fused-kernel ms) of the last apply (hipEvent sector timers)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().qg_hamiltonian_last_times(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().qg_hamiltonian_destroy(self._h)
            except Exception:
                pass
            self._h = None


def lobpcg_opts(**kw):
    o = N.LobpcgOpts()
    lib().qg_lobpcg_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def Eigen_d_multi(ham, X, P, scf_step=0, **opts):
    """Block LOBPCG (Eigen_d_multi_mpi): X (nbands, Ng) float64 device, in/out; P keep block.
    Returns (eigenvalues ndarray, stats dict)."""
    import torch
    torch.cuda.synchronize()
    xp, nb, _ = _block(X, ham.grid.size)
    pp, nb2, _ = _block(P, ham.grid.size)
    if nb != nb2 or X.dtype != torch.float64 or P.dtype != torch.float64:
        raise QGError("X and P must be float64 blocks of the same size")
    o = lobpcg_opts(**opts)
    st = N.LobpcgStats()
    eig = np.zeros(nb)
    check(lib().qg_lobpcg_d_multi(ham._h, ctypes.byref(o), nb, dptr(eig), xp, pp, int(scf_step),
                                  ctypes.byref(st)))
    stats = {f: getattr(st, f) for f, _ in N.LobpcgStats._fields_}
    return eig, stats


# ---------------------------------------------------------------- synthetic cells ----
DIAMOND_A = 6.74          # bohr, cubic diamond cell edge of the workloads
SEED_STATES = 123456789   # qumasun_mpi.initialize.h:199
SEED_VLOC = 123456790     # noise stream of the local potential


def diamond_system(ncell, n_per_cell, a=DIAMOND_A, zion=4.0, rc=0.8, noise=0.05,
                   seed=SEED_VLOC, sigma=0.6, cutoff=2.4, e_s=1.0, e_p=-0.5):
    """The seeded diamond workload: ncell^3 supercell, (n_per_cell*ncell)^3 grid, local
    potential sum -zion erf(r/rc)/r + U(-noise, noise), Gaussian s + p projectors."""
    L = ncell * a
    n = n_per_cell * ncell
    nat = 8 * ncell ** 3
    xyz = np.zeros(3 * nat)
    check(lib().qg_synth_diamond_atoms(ncell, a, dptr(xyz)))
    V = np.zeros(n ** 3)
    check(lib().qg_synth_vloc(n, n, n, L, L, L, nat, dptr(xyz), zion, rc, noise, seed, dptr(V)))
    projs = [[gaussian_projector(0, e_s, sigma, cutoff), gaussian_projector(1, e_p, sigma, cutoff)]]
    return dict(n=n, L=L, h=L / n, positions=xyz.reshape(nat, 3), V=V, projectors=projs)


def random_states(ng, nbands, dvol, seed=SEED_STATES):
    X = np.zeros(ng * nbands)
    check(lib().qg_synth_random_states(ng, nbands, dvol, seed, dptr(X)))
    return X.reshape(nbands, ng)
