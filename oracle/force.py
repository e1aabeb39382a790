"""CPU FP64 restatement of QUMASUN's nonlocal energy and forces — TEST INFRASTRUCTURE ONLY (the
oracle; only tests/ may use it).

What pins them (tests/test_oracle_force.py): the 8th-order gradient's symbol on a plane wave is
the closed form (i/h) (2/840) sum_m w_m sin(m k h) and it is antisymmetric; the energy is the sum of
occ E c^2 over the oracle's projections; the forces equal minus the central-difference
derivative of that energy with respect to each nucleus position, for smooth states and
projectors whose cutoff truncates nothing measurable.
"""
import numpy as np

__all__ = ["gradient8", "nonlocal_energy_force"]

_W = (672.0, -168.0, 32.0, -3.0)


def gradient8(psi, n, h, coef=1.0):
    """Gradient8th (GridGradient8th.h:17-160): coef/(840 h) sum_m w_m (p[+m] - p[-m]) per axis,
    periodic; psi (Ng,) x fastest.  Returns (gx, gy, gz)."""
    nx, ny, nz = n
    p = np.asarray(psi, dtype=np.float64).reshape(nz, ny, nx)
    out = []
    for ax, hh in ((2, h[0]), (1, h[1]), (0, h[2])):
        d = np.zeros_like(p)
        for m, w in enumerate(_W, start=1):
            d += w * (np.roll(p, -m, axis=ax) - np.roll(p, m, axis=ax))
        out.append((coef / (840.0 * hh) * d).reshape(-1))
    return tuple(out)


def nonlocal_energy_force(n, h, onl, psi, occ, species, species_projectors):
    """mGetEnergyPseudoNonlocal + mGetForcePseudoNonlocal (qumasun_mpi.energy.h:349-375,
    qumasun_mpi.force.h:226-258) over EnergyPPnonlocal / ForceNonlocal
    (PseudoPotOperator_mpi_v2.h:573-644, :905-1013): occupied bands only; the gradient is
    scaled by -2 occ before its projection.  onl: oracle.OracleNL of those atoms (coefficients in
    the order atom, projector, m).  Returns (energy, forces (natoms, 3))."""
    natoms = len(species)
    E = onl.energies()
    off = [0]
    for a in range(natoms):
        off.append(off[-1] + sum(2 * p["l"] + 1 for p in species_projectors[species[a]]))
    assert off[-1] == onl.ncoef
    forces = np.zeros((natoms, 3))
    ev = 0.0
    for b in range(psi.shape[0]):
        if occ[b] <= 0.0:
            continue
        c = onl.project(np.ascontiguousarray(psi[b]))
        g = gradient8(psi[b], n, h, -2.0 * occ[b])
        d = [onl.project(np.ascontiguousarray(gg)) for gg in g]
        e = 0.0
        for a in range(natoms):
            for j in range(off[a], off[a + 1]):
                e += c[j] * c[j] * E[j]
                for ax in range(3):
                    forces[a, ax] += d[ax][j] * c[j] * E[j]
        ev += e * occ[b]
    return ev, forces
