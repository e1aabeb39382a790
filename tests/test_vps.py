"""The MBK .vps reader (qg_vps_load, LoadVPS in Python) on synthetic files written by
tests/vps_synth.py: every value read equals the value written (text at 17 significant digits,
so equal to the last bit), comments of every form are stripped, the first occurrence of a key
wins, inconsistent files are refused, and the projector channels handed to the engine are the
reference's (PseudoPotOperator_mpi_v2.h:64-84, :108-166: l, energy_up, the largest NandL cutoff
of that l, the mean of the two j-components, r_min = r[0]).  Host code only."""
import numpy as np
import pytest

import qumasun_amd as Q
import vps_synth as S


def _load(tmp_path, d, **kw):
    p = tmp_path / "pp.vps"
    S.write_vps(str(p), d, **kw)
    return Q.LoadVPS(str(p))


def test_values_round_trip(tmp_path):
    d = S.synth_pp(extra_nandl=[(9, 3, 1, 2.1), (10, 3, 0, 1.2)])
    pp = _load(tmp_path, d)
    assert pp.num_projectors == 2 and pp.num_radial_grids == len(d["r"])
    assert pp.valence_electron == 4
    assert pp.cutoff_vlocal == d["local_cutoff"]
    # largest cutoff per l over the NandL lines (l = 1: 1.8 and 2.1)
    assert pp.cutoff_r == [1.6, 2.1, 0.0, 0.0]
    assert np.array_equal(pp.radius, d["r"]) and np.array_equal(pp.V_local, d["vloc"])
    assert pp.projector_quantum_l == d["l"]
    assert pp.projector_energy_up == d["e_up"] and pp.projector_energy_down == d["e_dn"]
    for k in range(2 * pp.num_projectors):
        assert np.array_equal(pp.projector[k], d["comps"][k])
    assert pp.xi_min == d["xi"][0]
    assert abs(pp.xi_delta - (d["xi"][-1] - d["xi"][0]) / (len(d["xi"]) - 1)) == 0.0
    assert pp.has_pcc_charge == 1 and np.array_equal(pp.pcc_charge, d["pcc"])
    # the PCC charge LoadVPS prints: 4 pi xi_delta sum pcc r^3 (vps_loader.h:210-236)
    want = 4 * np.pi * pp.xi_delta * np.sum(d["pcc"] * d["r"] ** 3)
    assert abs(pp.pcc_charge_integral - want) <= 1e-12 * abs(want)
    # and it is the integral 4 pi int pcc r^2 dr = pi^(3/2) * 0.3 for pcc = 0.3 exp(-r^2)
    assert abs(pp.pcc_charge_integral - 0.3 * np.pi ** 1.5) < 1e-3
    assert pp.TotalProjectorLM() == 1 + 3 and pp.MaxProjectorL() == 1


def test_projector_specs_are_the_references(tmp_path):
    d = S.synth_pp(extra_nandl=[(9, 3, 1, 2.1)])
    pp = _load(tmp_path, d)
    specs = pp.projector_specs()
    assert len(specs) == 2
    for k, s in enumerate(specs):
        assert s["kind"] == "table" and s["l"] == d["l"][k]
        assert s["energy"] == d["e_up"][k]
        assert s["cutoff"] == pp.cutoff_r[d["l"][k]]
        assert s["r_min"] == d["r"][0]
        assert s["xi_min"] == d["xi"][0] and s["xi_delta"] == pp.xi_delta
        mean = (d["comps"][2 * k] + d["comps"][2 * k + 1]) / 2.0
        assert np.array_equal(s["values"], mean)


def test_first_occurrence_wins_and_comment_word(tmp_path):
    d = S.synth_pp()
    pp = _load(tmp_path, d, duplicate_keys=True)
    assert pp.cutoff_vlocal == d["local_cutoff"]      # not the later 99.0
    assert pp.valence_electron == 4                   # "<VPS>" and what follows stripped


def test_no_pcc_block(tmp_path):
    d = S.synth_pp(pcc=False)
    pp = _load(tmp_path, d)
    assert pp.has_pcc_charge == 0 and pp.pcc_charge is None


@pytest.mark.parametrize("kw,msg", [(dict(vps_type="KB"), "MBK"),
                                    (dict(vps_type="mbk"), None),            # compared without case
                                    (dict(close_potentials=False), "Pseudo.Potentials"),
                                    (dict(energies_count=3), "project.energies"),
                                    (dict(pcc_lines=10), "density.PCC")])
def test_refusals(tmp_path, kw, msg):
    d = S.synth_pp()
    if msg is None:
        assert _load(tmp_path, d, **kw).num_projectors == 2
        return
    with pytest.raises(Q.QGError, match=msg):
        _load(tmp_path, d, **kw)


def test_missing_file():
    with pytest.raises(Q.QGError, match="cannot open"):
        Q.LoadVPS("/nonexistent/pp.vps")


def test_sort_by_l_as_the_reference_runs_it(tmp_path):
    """PseudoPot_MBK::SortByL (vps_loader.h:67-82) copies instead of swapping and visits
    j = 2 .. while j < n - j.  Five channels with l = 1, 1, 0, 1, 0: at j = 2, l[2] = 0 < l[1],
    so channel 2 becomes a copy of channel 1 (l, both energies, both tables); channel 4 (l = 0
    after l = 1) lies outside the visited range and stays.  A file ordered by l is unchanged."""
    ch = ((1, 0.5, 0.4, 0.5, 1.8), (1, -0.5, -0.4, 0.6, 1.8), (0, 1.0, 0.9, 0.7, 1.6), (1, 0.2, 0.1, 0.8, 1.8),
          (0, -0.2, -0.1, 0.9, 1.6))
    d = S.synth_pp(channels=ch)
    pp = _load(tmp_path, d)
    assert pp.projector_quantum_l == [1, 1, 1, 1, 0]
    assert pp.projector_energy_up == [0.5, -0.5, -0.5, 0.2, -0.2]
    assert pp.projector_energy_down == [0.4, -0.4, -0.4, 0.1, -0.1]
    want = [0, 1, 2, 3, 2, 3, 6, 7, 8, 9]
    for k in range(10):
        assert np.array_equal(pp.projector[k], d["comps"][want[k]]), k
    ordered = S.synth_pp(channels=tuple(sorted(ch, key=lambda c: c[0])))
    po = _load(tmp_path, ordered)
    assert po.projector_quantum_l == [0, 0, 1, 1, 1]
    for k in range(10):
        assert np.array_equal(po.projector[k], ordered["comps"][k])
