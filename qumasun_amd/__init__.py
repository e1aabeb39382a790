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
  LoadVPS(path) -> PseudoPot_MBK                   vps_loader.h:87-247
"""
import ctypes
import math
import os

import numpy as np

from . import _native as N
from ._native import (QG_F32, QG_F64, QG_PREC_F32, QG_PREC_F64, QG_PREC_MIXED, QG_PRECOND_KINETIC,
                      QG_PRECOND_LEARNED, QG_PRECOND_NONE, QGError, check, dptr, lib)

__all__ = ["Grid", "Laplasian8th", "PseudoPotIntegrator", "Hamiltonian", "Eigen_d_multi",
           "LearnedPreconditioner", "gaussian_projector", "diamond_system", "QGError", "LoadVPS",
           "PseudoPot_MBK", "SetOccupancy", "Density", "DensityMix", "HartreePotential", "XcLDA", "SCF",
           "Gradient8th", "KpointHamiltonian"]


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


class PseudoPot_MBK:
    """An MBK .vps pseudopotential as LoadVPS reads it (vps_loader.h:9-247): the fields of
    PseudoPot_MBK (valence_electron, num_projectors, num_radial_grids, cutoff_vlocal, cutoff_r,
    projector_quantum_l, projector_energy_up / _down, radius, V_local, projector (two
    j-components per projector, in the reference's order), pcc_charge, has_pcc_charge, xi_min,
    xi_delta), read by the engine's C++ reader (qg_vps_load)."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        check(lib().qg_vps_load(os.fsencode(path), ctypes.byref(h)))
        try:
            info = N.VpsInfo()
            check(lib().qg_vps_get_info(h, ctypes.byref(info)))
            n, npj = info.num_radial_grids, info.num_projectors
            self.valence_electron = info.valence_electron
            self.num_projectors = npj
            self.num_radial_grids = n
            self.cutoff_vlocal = info.cutoff_vlocal
            self.cutoff_r = list(info.cutoff_r)
            self.xi_min, self.xi_delta = info.xi_min, info.xi_delta
            self.has_pcc_charge = info.has_pcc_charge
            self.pcc_charge_integral = info.pcc_charge_integral
            self._total_lm, self._max_l = info.total_projector_lm, info.max_projector_l
            self.radius = np.zeros(n)
            self.V_local = np.zeros(n)
            pcc = np.zeros(n)
            check(lib().qg_vps_get_radial(h, dptr(self.radius), dptr(self.V_local), dptr(pcc)))
            self.pcc_charge = pcc if self.has_pcc_charge else None
            self.projector_quantum_l = []
            self.projector_energy_up = []
            self.projector_energy_down = []
            self.projector = []
            for k in range(npj):
                l, eu, ed = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
                up, dn = np.zeros(n), np.zeros(n)
                check(lib().qg_vps_get_projector(h, k, ctypes.byref(l), ctypes.byref(eu), ctypes.byref(ed),
                                                 dptr(up), dptr(dn)))
                self.projector_quantum_l.append(l.value)
                self.projector_energy_up.append(eu.value)
                self.projector_energy_down.append(ed.value)
                self.projector += [up, dn]
            specs = (N.ProjSpec * max(1, npj))()
            check(lib().qg_vps_projector_specs(h, specs, max(1, npj)))
            self._specs = []
            for k in range(npj):
                s = specs[k]
                vals = np.ctypeslib.as_array(s.table, shape=(s.n_table,)).copy()
                self._specs.append(table_projector(s.l, s.energy, s.cutoff, s.xi_min, s.xi_delta, vals,
                                                   r_min=s.r_min))
        finally:
            lib().qg_vps_destroy(h)

    def TotalProjectorLM(self):
        return self._total_lm

    def MaxProjectorL(self):
        return self._max_l

    def projector_specs(self):
        """The projector channels for PseudoPotIntegrator, as the reference builds them from the
        file (PseudoPotOperator_mpi_v2.h:64-84, :108-166): l, energy_up, cutoff_r[l], the mean of
        the two j-components on the file's log grid, r below radius[0] taking the first value."""
        return [dict(p) for p in self._specs]


def LoadVPS(path):
    """LoadVPS(filepath) (vps_loader.h:87): the file's PseudoPot_MBK; raises QGError where the
    reference prints an error (not MBK, missing blocks) or would read out of bounds."""
    return PseudoPot_MBK(path)


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
        if self._h:
            # move the nuclei of the live handle in place: Hamiltonians built on this
            # integrator keep a valid handle and see the new positions
            check(lib().qg_nonlocal_update(self._h, nat, dptr(pos), sp.ctypes.data_as(N.P(ctypes.c_int))))
        else:
            nproj = np.array([len(s) for s in self.species_projectors], dtype=np.int32)
            specs, keep = _specs_array(self.species_projectors)
            h = ctypes.c_void_p()
            check(lib().qg_nonlocal_create(self.grid._h, nat, dptr(pos),
                                           sp.ctypes.data_as(N.P(ctypes.c_int)), len(nproj),
                                           nproj.ctypes.data_as(N.P(ctypes.c_int)), specs, ctypes.byref(h)))
            self._h = h
        self.positions = pos
        self.species = sp

    @property
    def num_coef(self):
        return lib().qg_nonlocal_num_coef(self._h)

    @property
    def projector_form(self):
        """'kernel' when H psi evaluates the Gaussian projectors in the kernel, 'tables'
        when it reads per-point tables."""
        return "kernel" if lib().qg_nonlocal_projector_form(self._h) == 1 else "tables"

    def InnerNonlocalPsi(self, psi):
        """coef[b, j] = dV sum beta_j psi_b (mInnerNonlocalPsi_v2), FP64 device tensor."""
        import torch
        pp, nb, ld = _block(psi, self.grid.size)
        coef = torch.empty((nb, self.num_coef), dtype=torch.float64, device=psi.device)
        check(lib().qg_nonlocal_project(self._h, _dtype_code(psi), nb, pp, ld,
                                        ctypes.c_void_p(coef.data_ptr()), _stream()))
        return coef

    def EnergyForceNonlocal(self, psi, occupancy, forces=True):
        """Nonlocal energy sum_b occ_b sum E c^2 and, with forces, the (natoms, 3) forces
        (mGetEnergyPseudoNonlocal / mGetForcePseudoNonlocal with ForceNonlocal): states with
        occ > 0 of the device block psi.  Returns (energy, forces or None)."""
        pp, nb, ld = _block(psi, self.grid.size)
        occ = np.ascontiguousarray(occupancy, dtype=np.float64)
        if len(occ) != nb:
            raise QGError("one occupancy per band")
        e = ctypes.c_double()
        f = np.zeros((len(self.positions), 3)) if forces else None
        check(lib().qg_nonlocal_energy_force(self._h, _dtype_code(psi), nb, pp, ld, dptr(occ), ctypes.byref(e),
                                             dptr(f) if forces else None, _stream()))
        return e.value, f

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
        """V: host array, or a float64 device tensor of Ng values (installed without a host copy)."""
        if hasattr(V, "is_cuda") and V.is_cuda:
            vp, nb, _ = _block(V, self.grid.size)
            if nb != 1 or V.dtype.itemsize != 8:
                raise QGError("the potential is one float64 field")
            check(lib().qg_hamiltonian_set_potential_device(self._h, vp, _stream()))
            return
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1)
        check(lib().qg_hamiltonian_set_potential(self._h, dptr(V)))

    def apply(self, psi, out):
        pp, nb, ld = _block(psi, self.grid.size)
        op, nb2, ldo = _block(out, self.grid.size)
        if nb != nb2 or psi.dtype != out.dtype:
            raise QGError("block mismatch")
        check(lib().qg_hamiltonian_apply(self._h, _dtype_code(psi), nb, pp, ld, op, ldo, _stream()))
        return out

    def set_plan(self, kernels=N.QG_KERNELS_AUTO, rows_ty=0, rows_bands=0, rows_tz=0, rows_tzp=0, rows_nbp=0):
        """Kernel plan of H psi (qg_hamiltonian_set_plan): which kernel family, and the row
        kernels' tile shape (0: automatic)."""
        p = N.HpsiPlan()
        lib().qg_default_hpsi_plan(ctypes.byref(p))
        p.kernels, p.rows_ty, p.rows_bands, p.rows_tz, p.rows_tzp, p.rows_nbp = (kernels, rows_ty, rows_bands, rows_tz,
                                                                               rows_tzp, rows_nbp)
        check(lib().qg_hamiltonian_set_plan(self._h, ctypes.byref(p)))

    def set_timing(self, enable=True):
        check(lib().qg_hamiltonian_set_timing(self._h, int(bool(enable))))

    def last_times(self):
        """(projector-dot ms, fused-kernel ms) of the last apply, summed over its band chunks
        (hipEvent sector timers)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().qg_hamiltonian_last_times(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def last_chunks(self):
        """Band chunks (fused launches) of the last timed apply."""
        n = ctypes.c_int()
        check(lib().qg_hamiltonian_last_chunks(self._h, ctypes.byref(n)))
        return n.value

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
    op = opts.pop("op", None)
    o = lobpcg_opts(**opts)
    if op is not None:
        # op(dtype, nbands, x_ptr, ax_ptr, stream_ptr): the reference's OpeA on device blocks
        cb = N.OP_CALLBACK(lambda user, dt, n, x, ax, s: op(dt, n, x, ax, s))
        o.op = cb
    st = N.LobpcgStats()
    eig = np.zeros(nb)
    check(lib().qg_lobpcg_d_multi(ham._h, ctypes.byref(o), nb, dptr(eig), xp, pp, int(scf_step),
                                  ctypes.byref(st)))
    stats = {f: getattr(st, f) for f, _ in N.LobpcgStats._fields_}
    return eig, stats


# ----------------------------------------------------------------------- SCF step ----
KB_HARTREE = 3.166811563e-6    # Boltzmann constant in Hartree / K


def _field(t, ng, what):
    import torch
    if t is None:
        return None
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == ng):
        raise QGError(f"{what}: a contiguous float64 device tensor of {ng} values is needed")
    return ctypes.c_void_p(t.data_ptr())


def _kdtype(t):
    import torch
    if t.dtype == torch.complex128:
        return N.QG_Z128
    if t.dtype == torch.complex64:
        return N.QG_C64
    raise QGError(f"k-point blocks are complex128 or complex64, not {t.dtype}")


class KpointHamiltonian:
    """The Hamiltonian at one k point (QUMASUN_KPOINT's mHamiltonianMatrix_ddm,
    qumasun_kpoint.matrix.h:21-110): for the cell-periodic part u of psi = exp(i k.r) u,
    H_k u = -1/2 [L8 - |k|^2 + 2 i k . D8] u + V u + sum E conj(P) (dV P^T u) with the Bloch-phased
    projectors P = beta exp(i k . x_u).  k in bohr^-1; blocks are (nbands, Ng) complex tensors."""

    def __init__(self, ham, k):
        self.ham = ham
        self.k = tuple(float(c) for c in k)
        h = ctypes.c_void_p()
        check(lib().qg_kpoint_create(ham._h, *self.k, ctypes.byref(h)))
        self._h = h

    def apply(self, psi, out):
        """out = H_k psi (complex128 or complex64, band-major)."""
        g = self.ham.grid.size
        pp, nb, ld = _block(psi, g)
        op, nbo, ldo = _block(out, g)
        if nbo != nb or out.dtype != psi.dtype:
            raise QGError("out must match psi in bands and dtype")
        check(lib().qg_kpoint_apply(self._h, _kdtype(psi), nb, pp, ld, op, ldo, _stream()))
        return out

    def Eigen_z_multi(self, X, iter_max=300, tol=1e-6, nconv=0, precond_shift=0.0):
        """Lowest eigenpairs by complex FP64 LOBPCG (Eigen_z_multi_mpi, lobpcg_z_multi_mpi.h:988):
        X (nbands, Ng) complex128 start, overwritten with the eigenvectors; precond_shift > 0
        preconditions the residuals by (T_k + shift)^-1 (the reference runs without).  Returns
        (eigenvalues ndarray, iterations, largest residual norm)."""
        import torch
        if X.dtype != torch.complex128:
            raise QGError("the k-point LOBPCG runs in complex128")
        g = self.ham.grid.size
        pp, nb, ld = _block(X, g)
        eig = np.zeros(nb)
        it = ctypes.c_int()
        res = ctypes.c_double()
        check(lib().qg_kpoint_lobpcg(self._h, nb, pp, ld, dptr(eig), int(iter_max), float(tol), int(nconv),
                                     float(precond_shift), ctypes.byref(it), ctypes.byref(res), _stream()))
        return eig, it.value, res.value

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().qg_kpoint_destroy(self._h)
            except Exception:   # interpreter shutdown: the module's globals may be gone
                pass
            self._h = None


def Gradient8th(grid, psi, coef=1.0):
    """(d/dx, d/dy, d/dz) psi * coef per band, periodic 8th-order central differences
    (Gradient8th, GridGradient8th.h:17-160): three device blocks like psi."""
    import torch
    pp, nb, ld = _block(psi, grid.size)
    outs = [torch.empty_like(psi) for _ in range(3)]
    check(lib().qg_gradient8_apply(grid._h, _dtype_code(psi), nb, pp, ld, *[ctypes.c_void_p(o.data_ptr()) for o in outs],
                                   ld, float(coef), _stream()))
    return outs


def SetOccupancy(eigen_values, num_electrons, kbT):
    """Fermi-Dirac occupancies of the spin-unpolarised states (mSetOccupancySpinOff,
    qumasun_mpi.occupancy.h:61-108): (occupancy ndarray, mu)."""
    e = np.ascontiguousarray(eigen_values, dtype=np.float64)
    occ = np.zeros_like(e)
    mu = ctypes.c_double()
    check(lib().qg_occupancy_fermi(len(e), dptr(e), float(num_electrons), float(kbT), dptr(occ), ctypes.byref(mu)))
    return occ, mu.value


def Density(grid, X, occupancy, out=None):
    """rho = sum over states with occ > 0 of occ * psi^2 (mSetDensityOne, qumasun_mpi.density.h:10-32);
    X a band-major float64 or float32 device block; returns the float64 device field."""
    import torch
    xp, nb, ld = _block(X, grid.size)
    occ = np.ascontiguousarray(occupancy, dtype=np.float64)
    if len(occ) != nb:
        raise QGError("one occupancy per band")
    if out is None:
        out = torch.empty(grid.size, dtype=torch.float64, device=X.device)
    check(lib().qg_density(grid._h, _dtype_code(X), nb, xp, ld, dptr(occ), _field(out, grid.size, "rho"), _stream()))
    return out


def DensityMix(grid, rho, rho_prev, ratio):
    """rho <- prev + ratio (rho - prev) in place (mSetDensity, qumasun_mpi.density.h:46-64);
    returns (int |rho - prev| dV, int rho_mixed dV).  rho_prev None: no mixing."""
    d, t = ctypes.c_double(), ctypes.c_double()
    check(lib().qg_density_mix(grid._h, _field(rho, grid.size, "rho"), _field(rho_prev, grid.size, "rho_prev"),
                               float(ratio), ctypes.byref(d), ctypes.byref(t), _stream()))
    return d.value, t.value


def HartreePotential(grid, rho, out=None):
    """V_H(k) = 4 pi rho(k) / |k|^2, V_H(0) = 0, by FFT (poisson_fft.h:131-240)."""
    import torch
    if out is None:
        out = torch.empty_like(rho)
    check(lib().qg_hartree_potential(grid._h, _field(rho, grid.size, "rho"), _field(out, grid.size, "V_H"), _stream()))
    return out


def XcLDA(grid, rho, pcc=None):
    """Perdew-Zunger LDA of rho + pcc (Calc_XC_LDA, Vxc.h:46-75): (V_x, V_c, e_x, e_c)."""
    import torch
    vx, vc, ex, ec = (torch.empty_like(rho) for _ in range(4))
    f = lambda t, w: _field(t, grid.size, w)   # noqa: E731
    check(lib().qg_xc_lda(grid._h, f(rho, "rho"), f(pcc, "pcc"), f(vx, "vx"), f(vc, "vc"), f(ex, "ex"), f(ec, "ec"),
                          _stream()))
    return vx, vc, ex, ec


class SCF:
    """The SCF loop of QUMASUN_MPI::Execute (qumasun_mpi.scfloop.h:16-160), spin-unpolarised, on one
    GPU: per step the block LOBPCG (eigen_step_initial iterations on the first step,
    eigen_step_per_scf after), Fermi-Dirac occupancies, the density linearly mixed with the previous
    one, and the new potential V_ext + V_H + V_x + V_c installed in the Hamiltonian; converged when
    int |rho - rho_prev| dV / N_e < threshold (1e-7 in the reference).  Every field stays on the
    device.  The starting density is rho0 (default: the uniform density N_e / volume; the reference
    starts from its random states or from atomic orbitals, qumasun_mpi.initialize.h:321-380).
    lobpcg: options of Eigen_d_multi (precision, preconditioner, basis, tol ...)."""

    def __init__(self, ham, vext, num_electrons, nbands, kbT=300.0 * KB_HARTREE, mixing_ratio=0.5,
                 eigen_step_initial=100, eigen_step_per_scf=10, scf_step=100, threshold=1.0e-7, pcc=None,
                 **lobpcg):
        import torch
        self.H = ham
        self.grid = ham.grid
        ng = self.grid.size
        self.vext = vext if hasattr(vext, "is_cuda") else torch.from_numpy(
            np.ascontiguousarray(vext, dtype=np.float64)).cuda()
        self.pcc = pcc
        self.ne = float(num_electrons)
        self.nbands = int(nbands)
        self.kbT, self.ratio = float(kbT), float(mixing_ratio)
        self.steps = (int(eigen_step_initial), int(eigen_step_per_scf))
        self.limit, self.threshold = int(scf_step), float(threshold)
        self.lobpcg = lobpcg
        dev = self.vext.device
        self.rho = torch.empty(ng, dtype=torch.float64, device=dev)
        self.rho_prev = torch.empty_like(self.rho)
        self.vh, self.vx, self.vc, self.vtot = (torch.empty_like(self.rho) for _ in range(4))
        self.history = []

    def set_potential(self):
        """mSetPotential (qumasun_mpi.potential.h:23-170): V_tot of the current density, installed."""
        f = lambda t, w: _field(t, self.grid.size, w)   # noqa: E731
        check(lib().qg_scf_potential(self.H._h, f(self.rho, "rho"), f(self.pcc, "pcc"), f(self.vext, "vext"),
                                     f(self.vh, "vh"), f(self.vx, "vx"), f(self.vc, "vc"), f(self.vtot, "vtot"),
                                     _stream()))

    def run(self, X, P=None, rho0=None, log=None):
        """X: (nbands, Ng) float64 device states (in/out, the Kohn-Sham states at exit); P the kept
        LOBPCG block (zeros if None).  Returns (eigenvalues, occupancies, converged, steps)."""
        import torch
        vol = self.grid.size * self.grid.dvol
        if P is None:
            P = torch.zeros_like(X)
        if rho0 is None:
            self.rho.fill_(self.ne / vol)
        else:
            self.rho.copy_(rho0)
        self.set_potential()
        eig = occ = None
        converged = False
        step = 0
        for step in range(self.limit):
            eig, st = Eigen_d_multi(self.H, X, P, step,
                                    max_iter=self.steps[0] if step == 0 else self.steps[1], **self.lobpcg)
            occ, mu = SetOccupancy(eig, self.ne, self.kbT)
            self.rho_prev.copy_(self.rho)
            Density(self.grid, X, occ, out=self.rho)
            diff, total = DensityMix(self.grid, self.rho, self.rho_prev, self.ratio)
            diff /= self.ne
            self.set_potential()
            self.history.append(dict(step=step, diff_rho=diff, total_rho=total, mu=mu, eig0=float(eig[0]),
                                     lobpcg_iterations=st["iterations"], max_residual=st["max_residual"]))
            if log:
                log(self.history[-1])
            if diff < self.threshold:
                converged = True
                break
        return eig, occ, converged, step + 1


class LearnedPreconditioner:
    """The learned preconditioner M = K + K N^T N K of a Hamiltonian (qg_cnn_*): a 3-layer
    3x3x3 convolution network N (8 channels) with the local potential as gate input, trained
    on the device on r -> (T + V - sigma)^-1 r pairs built from the solver's own iterates."""

    def __init__(self, ham, shift, max_pairs=512, seed=2024):
        self.ham = ham
        self.shift = float(shift)
        h = ctypes.c_void_p()
        check(lib().qg_cnn_create(ham._h, self.shift, int(max_pairs), int(seed), ctypes.byref(h)))
        self._h = h

    @staticmethod
    def num_params():
        return lib().qg_cnn_num_params()

    def get_params(self):
        n = self.num_params()
        a = np.zeros(n, np.float32)
        check(lib().qg_cnn_get_params(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, np.float32)
        check(lib().qg_cnn_set_params(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size))

    def network(self, x, out, adjoint=False, bf16=False):
        """out = N x (or N^T x) for float32 device blocks (nbands, Ng)."""
        import torch
        torch.cuda.synchronize()
        xp, nb, _ = _block(x, self.ham.grid.size)
        op, nb2, _ = _block(out, self.ham.grid.size)
        if nb != nb2 or x.dtype != torch.float32 or out.dtype != torch.float32:
            raise QGError("float32 blocks of the same size required")
        check(lib().qg_cnn_network(self._h, int(adjoint), int(bf16), nb, xp, op, None))

    def loss_grad(self, u, t, znorm):
        import torch
        torch.cuda.synchronize()
        up, nb, _ = _block(u, self.ham.grid.size)
        tp, nb2, _ = _block(t, self.ham.grid.size)
        loss = ctypes.c_double()
        g = np.zeros(self.num_params(), np.float32)
        check(lib().qg_cnn_loss_grad(self._h, nb, up, tp, float(znorm), ctypes.byref(loss),
                                     g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return loss.value, g

    def collect(self, X, eig, sigma, tol=1e-5, maxit=200):
        """Add training pairs from FP64 iterates X (nbands, Ng) with Ritz values eig."""
        import torch
        torch.cuda.synchronize()
        xp, nb, _ = _block(X, self.ham.grid.size)
        e = np.ascontiguousarray(eig, np.float64)
        it = ctypes.c_int()
        check(lib().qg_cnn_collect(self._h, nb, xp, dptr(e), float(sigma), float(tol), int(maxit),
                                   ctypes.byref(it)))
        return it.value

    @property
    def num_pairs(self):
        return lib().qg_cnn_num_pairs(self._h)

    def train(self, steps, batch=8, lr=3e-3, seed=7):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().qg_cnn_train(self._h, int(steps), int(batch), float(lr), int(seed), ctypes.byref(a),
                                 ctypes.byref(b)))
        return a.value, b.value

    def apply(self, r):
        """r <- M r in place (float32 or float64 device block)."""
        rp, nb, ld = _block(r, self.ham.grid.size)
        check(lib().qg_cnn_apply(self._h, _dtype_code(r), nb, rp, ld, _stream()))

    def attach(self):
        check(lib().qg_hamiltonian_set_cnn(self.ham._h, self._h))

    def detach(self):
        check(lib().qg_hamiltonian_set_cnn(self.ham._h, None))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().qg_cnn_destroy(self._h)
            except Exception:
                pass
            self._h = None


def fit_learned_preconditioner(ham, nbands, sigma, shift, snapshots=(2, 5, 10, 16, 24), steps=600,
                               batch=8, lr=3e-3, seed=2024):
    """Train a LearnedPreconditioner on pairs from this Hamiltonian's own solver iterates:
    kinetic-preconditioned LOBPCG from a seeded random start, stopped after each number of
    iterations in `snapshots`; every snapshot's Ritz vectors give nbands residuals."""
    import torch
    net = LearnedPreconditioner(ham, shift, max_pairs=nbands * len(snapshots))
    X0 = random_states(ham.grid.size, nbands, ham.grid.dvol, seed=SEED_STATES + 1)
    cg = []
    for k in snapshots:
        X = torch.from_numpy(X0.copy()).cuda()
        P = torch.zeros_like(X)
        e, _ = Eigen_d_multi(ham, X, P, 0, max_iter=int(k), precond=QG_PRECOND_KINETIC, precond_shift=shift)
        cg.append(net.collect(X, e, sigma))
    first, last = net.train(steps, batch=batch, lr=lr, seed=seed)
    return net, dict(pairs=net.num_pairs, cg_iters=cg, loss_first=first, loss_last=last)


def batch_apply(hams, psis, outs):
    """H psi for independent cells in persistent batched launches (qg_batch_hamiltonian_apply)."""
    import torch
    n = len(hams)
    if not (len(psis) == len(outs) == n) or n == 0:
        raise QGError("one psi and one out block per Hamiltonian")
    dt = _dtype_code(psis[0])
    H = (ctypes.c_void_p * n)(*[h._h.value for h in hams])
    nb = (ctypes.c_int * n)()
    pi = (ctypes.c_void_p * n)()
    po = (ctypes.c_void_p * n)()
    for c, (h, x, y) in enumerate(zip(hams, psis, outs)):
        xp, b, _ = _block(x, h.grid.size)
        yp, b2, _ = _block(y, h.grid.size)
        if b != b2 or _dtype_code(x) != dt or _dtype_code(y) != dt:
            raise QGError("block mismatch in batch")
        nb[c], pi[c], po[c] = b, xp.value, yp.value
    check(lib().qg_batch_hamiltonian_apply(n, H, dt, nb, pi, po, _stream()))
    return outs


def autotune(ham, dtype="f32", nbands=32, fixed_ppt=0):
    """Pick the fused kernel's tile shape for this grid by timing candidates on the device;
    returns (rows per tile, points per thread, ms of the best application)."""
    ty, ppt, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    check(lib().qg_autotune_hamiltonian(ham._h, QG_F32 if dtype == "f32" else QG_F64, int(nbands), int(fixed_ppt),
                                        ctypes.byref(ty), ctypes.byref(ppt), ctypes.byref(ms)))
    return ty.value, ppt.value, ms.value


def batch_eigen(hams, Xs, Ps, scf_step=0, batched=True, **opts):
    """LOBPCG for independent cells: batched=True runs them together with their H psi applied
    in persistent batched launches; False runs one cell after another."""
    import torch
    torch.cuda.synchronize()
    n = len(hams)
    H = (ctypes.c_void_p * n)(*[h._h.value for h in hams])
    nb = (ctypes.c_int * n)()
    xs = (ctypes.c_void_p * n)()
    ps = (ctypes.c_void_p * n)()
    eigs = [None] * n
    ep = (ctypes.POINTER(ctypes.c_double) * n)()
    for c, (h, X, P) in enumerate(zip(hams, Xs, Ps)):
        xp, b, _ = _block(X, h.grid.size)
        pp, b2, _ = _block(P, h.grid.size)
        if b != b2 or X.dtype != torch.float64 or P.dtype != torch.float64:
            raise QGError("X and P must be float64 blocks of the same size")
        nb[c], xs[c], ps[c] = b, xp.value, pp.value
        eigs[c] = np.zeros(b)
        ep[c] = dptr(eigs[c])
    o = lobpcg_opts(**opts)
    st = (N.LobpcgStats * n)()
    check(lib().qg_batch_lobpcg(n, H, ctypes.byref(o), nb, ep, xs, ps, int(scf_step), st, int(bool(batched))))
    stats = [{f: getattr(st[c], f) for f, _ in N.LobpcgStats._fields_} for c in range(n)]
    return eigs, stats


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


def diamond_cell(dims, n_per_cell=40, a=DIAMOND_A, zion=4.0, rc=0.8, noise=0.05, seed=SEED_VLOC, sigma=0.6,
                 cutoff=2.4, e_s=1.0, e_p=-0.5):
    """A (dims[0] x dims[1] x dims[2]) diamond supercell (8 atoms per cubic cell) on an
    (n_per_cell * dims) grid, with the same potential and projectors as diamond_system."""
    base = np.zeros(24)
    check(lib().qg_synth_diamond_atoms(1, a, dptr(base)))
    base = base.reshape(8, 3)
    pos = []
    for k in range(dims[2]):
        for j in range(dims[1]):
            for i in range(dims[0]):
                pos.append(base + a * np.array([i, j, k]))
    xyz = np.ascontiguousarray(np.concatenate(pos).reshape(-1))
    L = [a * d for d in dims]
    n = [n_per_cell * d for d in dims]
    V = np.zeros(n[0] * n[1] * n[2])
    check(lib().qg_synth_vloc(n[0], n[1], n[2], L[0], L[1], L[2], len(xyz) // 3, dptr(xyz), zion, rc, noise, seed,
                              dptr(V)))
    projs = [[gaussian_projector(0, e_s, sigma, cutoff), gaussian_projector(1, e_p, sigma, cutoff)]]
    return dict(n=tuple(n), L=tuple(L), h=tuple(L[d] / n[d] for d in range(3)), positions=xyz.reshape(-1, 3), V=V,
                projectors=projs)


SEED_CELLS = 123456791   # stream of the batched uneven-cell workload


def uneven_cells(ncells=32, seed=SEED_CELLS, n_per_cell=40):
    """The batched workload: ncells independent diamond supercells with dimensions drawn from
    {1,2}^3 (8-64 atoms, grids 40^3-80^3) and band counts from {16, 32, 64, 128}, all from one
    seeded stream; each cell's potential noise uses its own seed."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(ncells):
        dims = tuple(int(x) for x in rng.integers(1, 3, size=3))
        nb = int(rng.choice([16, 32, 64, 128]))
        out.append(dict(dims=dims, nbands=nb, seed=int(SEED_VLOC + 1 + c)))
    return out


def GramSchmidt(psi, dvol, ng=None):
    """Orthonormalise the band-major device block psi (nbands x Ng) in place with the dVol
    inner product, giving the Gram-Schmidt basis (GramSchmidt.h:6-38, GramSchmidt_mpi.h:9-53)."""
    ng = psi.shape[-1] if ng is None else ng
    pp, nb, ld = _block(psi, ng)
    check(lib().qg_gram_schmidt(_dtype_code(psi), ng, nb, pp, ld, float(dvol), _stream()))
    return psi


def random_states(ng, nbands, dvol, seed=SEED_STATES):
    X = np.zeros(ng * nbands)
    check(lib().qg_synth_random_states(ng, nbands, dvol, seed, dptr(X)))
    return X.reshape(nbands, ng)
