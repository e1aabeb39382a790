"""The library's host-side SCF step (qg_occupancy_fermi) against the oracle's restatement of
mSetOccupancySpinOff (qumasun_mpi.occupancy.h:61-108): same occupancies and chemical potential,
bit for bit (the same bisection in the same order)."""
import numpy as np
import pytest

import oracle as O
import qumasun_amd as Q


@pytest.mark.parametrize("ne,kbT,nb", [(32.0, 0.01, 30), (31.0, 1e-3, 30), (40.0, 0.05, 24), (16.0, 3e-4, 8),
                                       (7.0, 0.02, 4)])
def test_occupancy_matches_oracle(ne, kbT, nb):
    rng = np.random.default_rng(int(ne * 100 + nb))
    e = np.sort(rng.uniform(-1.0, 0.5, nb))
    occ, mu = Q.SetOccupancy(e, ne, kbT)
    occ_o, mu_o = O.occupancy_fermi(e, ne, kbT)
    assert np.array_equal(occ, occ_o) and mu == mu_o
