"""The oracle against values the reference itself holds or the physics fixes.

* Ledger energies per atom of the reference's force benchmark
  (benchmarks/adr0047.jsonl, single-thread rows; geometry of
  benchmarks/bench_force.py:152-158), for all six families: Tersoff Si,
  two-species Tersoff X-Y, SW Si, SW Ge-Sn (Tomita), Zhou EAM Ni and Cu-Ni,
  including the SW Ge-Sn and Zhou cutoffs the ledger records.
* Published values: Tersoff T3 Si a0 = 5.432 A, E_coh = 4.63 eV/atom; SW Si
  E = -2 epsilon per atom in perfect diamond (tests/test_engine_forces.py:
  1274-1338 of the reference); all forces zero in perfect diamond.
"""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd.structures import DIAMOND, cubic

from _cases import (GESN_SPECIES, LEDGER, SW_GESN, SW_SI, TERSOFF_SI, TERSOFF_XY, cuni_all,
                    gesn_all, ni, si, xy)

TOL = 1e-12


@pytest.mark.parametrize("n", [4, 8])
def test_tersoff_si_ledger(n):
    lat, pos, t = si(n)
    e, f, s = O.tersoff_flat(lat, pos, TERSOFF_SI)
    ref = LEDGER[f"tersoff_si/{len(pos)}"]
    assert abs(e / len(pos) - ref["e_per_atom"]) < TOL * abs(ref["e_per_atom"])
    assert np.abs(f.sum(axis=0)).max() < 1e-12


@pytest.mark.parametrize("n", [4, 8])
def test_tersoff_xy_ledger(n):
    lat, pos, t = xy(n)
    e, f, s = O.tersoff_multi(lat, pos, t, TERSOFF_XY, ["X", "Y"])
    ref = LEDGER[f"tersoff_xy/{len(pos)}"]
    assert abs(e / len(pos) - ref["e_per_atom"]) < TOL * abs(ref["e_per_atom"])


@pytest.mark.parametrize("n", [4, 8])
def test_sw_si_ledger(n):
    lat, pos, t = si(n)
    e, f, s = O.sw_flat(lat, pos, SW_SI)
    ref = LEDGER[f"sw_si/{len(pos)}"]
    assert abs(e / len(pos) - ref["e_per_atom"]) < TOL * abs(ref["e_per_atom"])


@pytest.mark.parametrize("n", [4, 8])
def test_sw_gesn_ledger(n):
    lat, pos, t = gesn_all()[n]
    ref = LEDGER[f"sw_gesn_tomita/{len(pos)}"]
    _, _, _, rc = O.sw_tables(SW_GESN, GESN_SPECIES)
    assert rc == ref["r_cut"]
    e, f, s = O.sw_multi(lat, pos, t, SW_GESN, GESN_SPECIES, nthreads=4)
    assert abs(e / len(pos) - ref["e_per_atom"]) < TOL * abs(ref["e_per_atom"])
    assert np.abs(f.sum(axis=0)).max() < 1e-12


@pytest.mark.parametrize("n", [5, 10])
def test_eam_ni_ledger(n):
    lat, pos, t = ni(n)
    ref = LEDGER[f"eam_ni/{len(pos)}"]
    _, _, _, _, rc, _ = O.zhou_function_table({"Ni": O.ZHOU_NI}, ["Ni"])
    assert rc == ref["r_cut"]
    e, f, s = O.eam_zhou(lat, pos, t, {"Ni": O.ZHOU_NI}, ["Ni"], nthreads=4)
    assert abs(e / len(pos) - ref["e_per_atom"]) < TOL * abs(ref["e_per_atom"])


@pytest.mark.parametrize("n", [5, 10])
def test_eam_cuni_ledger(n):
    lat, pos, t = cuni_all()[n]
    ref = LEDGER[f"eam_cuni/{len(pos)}"]
    el = {"Cu": O.ZHOU_CU, "Ni": O.ZHOU_NI}
    _, _, _, _, rc, _ = O.zhou_function_table(el, ["Cu", "Ni"])
    assert rc == ref["r_cut"]
    e, f, s = O.eam_zhou(lat, pos, t, el, ["Cu", "Ni"], nthreads=4)
    assert abs(e / len(pos) - ref["e_per_atom"]) < TOL * abs(ref["e_per_atom"])


def _ev_minimum(fn, grid):
    out = {}
    for a in grid:
        lat, pos, t = cubic(DIAMOND, a, 3, ["Si"] * 8, rattle=0.0)
        out[a] = fn(lat, pos) / len(pos)
    a_min = min(out, key=out.get)
    return a_min, out[a_min]


def test_tersoff_t3_published_lattice_constant_and_cohesive_energy():
    a_min, e_min = _ev_minimum(lambda l, p: O.tersoff_flat(l, p, TERSOFF_SI)[0],
                               np.round(np.arange(5.40, 5.46, 0.002), 3))
    assert a_min == pytest.approx(5.432, abs=0.005)
    assert abs(e_min) == pytest.approx(4.63, rel=0.01)


def test_sw_si_published_cohesive_energy():
    eps = 2.16826
    prm = dict(SW_SI, epsilon=eps)
    a_min, e_min = _ev_minimum(lambda l, p: O.sw_flat(l, p, prm)[0],
                               np.round(np.arange(5.40, 5.46, 0.002), 3))
    assert a_min == pytest.approx(5.431, abs=0.005)
    assert abs(e_min) == pytest.approx(2.0 * eps, rel=1e-3)


def test_perfect_diamond_has_zero_forces_and_no_three_body_sw():
    lat, pos, t = cubic(DIAMOND, 5.431, 3, ["Si"] * 8, rattle=0.0)
    for fn in (lambda: O.tersoff_flat(lat, pos, TERSOFF_SI),
               lambda: O.sw_flat(lat, pos, SW_SI)):
        e, f, s = fn()
        assert np.abs(f).max() < 1e-10
        assert np.allclose(s, s.T, atol=1e-12)
    # SW three-body vanishes at the tetrahedral angle: lambda has no effect
    e1 = O.sw_flat(lat, pos, SW_SI)[0]
    e2 = O.sw_flat(lat, pos, dict(SW_SI, **{"lambda": 40.0}))[0]
    assert abs(e1 - e2) < 1e-9
