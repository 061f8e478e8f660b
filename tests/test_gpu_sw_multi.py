"""Multi-species Stillinger-Weber on the HIP engine against the oracle and
the reference's Ge-Sn ledger (benchmarks/adr0047.jsonl:7-8, Tomita et al.
2018 records of benchmarks/bench_force.py:104-119).

Bar: energies, forces and stresses within 1e-10 relative of the oracle
(summation order only); energy per atom within 1e-10 of the ledger.  Row
lengths 4 and 8 (short cutoffs), 16, 32 and 64 (Verlet skins) are covered,
and so are records whose two orders (i, j, k) and (i, k, j) differ, a
three-species set with absent records, batches, the autotuner's shapes and
NVE.
"""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import NeighbourList, StillingerWeberPotential
from akitsushima_amd.md import maxwell_boltzmann, nve_run

from _cases import GESN_SPECIES, LEDGER, SW_GESN, SW_SI, gesn_all, rattled, rel_err, sheared_si

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _check(e, f, s, e0, f0, s0, label):
    assert abs(e - e0) / abs(e0) < TOL, f"{label}: E {e} vs {e0}"
    assert rel_err(f, f0) < TOL, f"{label}: F rel err {rel_err(f, f0):.3e}"
    assert rel_err(s, s0) < TOL, f"{label}: S rel err {rel_err(s, s0):.3e}"


def _cell(seed, a=5.762, shear=0.06, rattle=0.12, p=(0.7, 0.3), species=GESN_SPECIES, n=2):
    lat, pos, _ = sheared_si(n, shear=shear, rattle=rattle, seed=seed)
    lat = lat * (a / 5.431)
    pos = pos * (a / 5.431)
    rng = np.random.default_rng(seed + 100)
    return lat, pos, [str(x) for x in rng.choice(species, len(pos), p=list(p))]


def _unequal():
    recs = {k: dict(v) for k, v in SW_GESN.items()}
    recs[("Ge", "Sn", "Ge")]["costheta0"] = -0.25
    recs[("Ge", "Ge", "Sn")]["lambda"] = 35.0
    recs[("Sn", "Ge", "Sn")]["epsilon"] = 1.3
    return recs


def _small(sig):
    """Two species X, Y with SW Si shapes and a short cutoff a * sigma."""
    base = {k: SW_SI[k] for k in ("epsilon", "A", "B", "p", "q", "a", "gamma")}
    recs = {}
    for (i, j, k), s, lam in [(("X", "X", "X"), sig, 21.0), (("Y", "Y", "Y"), 1.05 * sig, 18.0),
                              (("X", "Y", "Y"), 1.02 * sig, 20.0),
                              (("Y", "X", "X"), 1.02 * sig, 20.0)]:
        recs[(i, j, k)] = dict(base, sigma=s, **{"lambda": lam, "costheta0": -1 / 3})
    for key, lam, c0 in [(("X", "X", "Y"), 19.0, -0.30), (("X", "Y", "X"), 22.0, -0.35),
                         (("Y", "X", "Y"), 17.0, -1 / 3), (("Y", "Y", "X"), 16.0, -0.31)]:
        recs[key] = {"lambda": lam, "costheta0": c0, "epsilon": SW_SI["epsilon"]}
    return recs


@pytest.mark.parametrize("n", [4, 8])
def test_gesn_matches_ledger_and_oracle(n):
    lat, pos, t = gesn_all()[n]
    pot = StillingerWeberPotential(SW_GESN, species=GESN_SPECIES)
    assert pot.r_cut == LEDGER[f"sw_gesn_tomita/{len(pos)}"]["r_cut"]
    e, f, s = pot.compute(lat, pos, t)
    ref = LEDGER[f"sw_gesn_tomita/{len(pos)}"]["e_per_atom"]
    assert abs(e / len(pos) - ref) / abs(ref) < TOL
    e0, f0, s0 = O.sw_multi(lat, pos, t, SW_GESN, GESN_SPECIES, nthreads=4)
    _check(e, f, s, e0, f0, s0, f"gesn {n}")


def test_gesn_rows_of_sixteen():
    """No skin: the 512-atom ledger cell has at most 16 neighbours."""
    lat, pos, t = gesn_all()[4]
    pot = StillingerWeberPotential(SW_GESN, species=GESN_SPECIES)
    nl = NeighbourList(pot.r_cut, skin_width=0.0)
    nl.build(lat, pos)
    e, f, s = pot.compute(lat, pos, t, neigh_list=nl)
    e0, f0, s0 = O.sw_multi(lat, pos, t, SW_GESN, GESN_SPECIES, nthreads=4)
    _check(e, f, s, e0, f0, s0, "gesn rows of 16")


@pytest.mark.parametrize("skin", [0.3, 1.0, 1.8])
@pytest.mark.parametrize("recs", ["tomita", "unequal"])
def test_gesn_sheared_long_rows(recs, skin):
    """A passed list with a wide skin gives rows of 32 or 64 slots."""
    params = SW_GESN if recs == "tomita" else _unequal()
    lat, pos, t = _cell(21, n=3)
    pot = StillingerWeberPotential(params, species=GESN_SPECIES)
    nl = NeighbourList(pot.r_cut, skin_width=skin)
    nl.build(lat, pos)
    e, f, s = pot.compute(lat, pos, t, neigh_list=nl)
    e0, f0, s0 = O.sw_multi(lat, pos, t, params, GESN_SPECIES, nthreads=4)
    _check(e, f, s, e0, f0, s0, f"{recs} skin {skin}")


@pytest.mark.parametrize("sig,rattle", [(1.40, 0.02), (1.70, 0.12)])
def test_short_rows(sig, rattle):
    """Cutoffs inside the first shell and just past it: rows of 4 and 8."""
    recs = _small(sig)
    lat, pos, t = _cell(31, a=5.431, shear=0.03, rattle=rattle, species=["X", "Y"], n=3)
    pot = StillingerWeberPotential(recs, species=["X", "Y"])
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.sw_multi(lat, pos, t, recs, ["X", "Y"], nthreads=4)
    _check(e, f, s, e0, f0, s0, f"short rows sigma {sig}")


def test_three_species_with_absent_records():
    recs = {k: dict(v) for k, v in _unequal().items()}
    recs[("Si", "Si", "Si")] = dict(SW_SI, epsilon=1.0, A=15.0)
    recs[("Si", "Ge", "Ge")] = {"sigma": 2.14, "a": 1.80, "gamma": 1.2, "lambda": 25.0,
                                "A": 12.0, "B": 0.64, "p": 4.0, "q": 0.0, "epsilon": 1.0}
    recs[("Ge", "Si", "Si")] = {"sigma": 2.14, "a": 1.80, "gamma": 1.1, "lambda": 23.0,
                                "A": 12.0, "B": 0.64, "p": 4.0, "q": 0.0, "epsilon": 1.0}
    recs[("Si", "Ge", "Si")] = {"lambda": 24.0, "costheta0": -0.3}
    recs[("Ge", "Ge", "Si")] = {"lambda": 26.0}
    # no Si-Sn records at all: those pairs and angles do not interact
    sp = ["Ge", "Sn", "Si"]
    lat, pos, t = _cell(41, a=5.6, p=(0.5, 0.25, 0.25), species=sp, n=3)
    pot = StillingerWeberPotential(recs, species=sp)
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.sw_multi(lat, pos, t, recs, sp, nthreads=4)
    _check(e, f, s, e0, f0, s0, "three species")


def test_one_species_records_equal_flat():
    lat, pos, t = sheared_si(3, shear=0.05, rattle=0.1, seed=17)
    e1, f1, s1 = StillingerWeberPotential(SW_SI, species=["Si"]).compute(lat, pos, t)
    recs = O.sw_flat_as_records(SW_SI)
    e2, f2, s2 = StillingerWeberPotential(recs, species=["Si"]).compute(lat, pos, t)
    assert abs(e1 - e2) < 1e-12 * abs(e1)
    assert rel_err(f2, f1) < 1e-11 and rel_err(s2, s1) < 1e-11


def test_relabelling_atoms_permutes_forces():
    lat, pos, t = _cell(51, n=3)
    perm = np.random.default_rng(2).permutation(len(pos))
    pot = StillingerWeberPotential(SW_GESN, species=GESN_SPECIES)
    e1, f1, s1 = pot.compute(lat, pos, t)
    e2, f2, s2 = pot.compute(lat, pos[perm], [t[k] for k in perm])
    assert abs(e1 - e2) < 1e-12 * abs(e1)
    assert rel_err(f2, f1[perm]) < 1e-11 and rel_err(s2, s1) < 1e-11


def test_batched_and_autotuned():
    structs = []
    for k, n in enumerate([2, 3, 2, 4, 3, 2]):
        lat, pos, t = _cell(60 + k, n=n, shear=0.02 * k, rattle=0.05 + 0.02 * k)
        structs.append((lat, pos, t))
    pot = StillingerWeberPotential(_unequal(), species=GESN_SPECIES)
    ref = [O.sw_multi(l, p, t, _unequal(), GESN_SPECIES, nthreads=4) for l, p, t in structs]
    E, F, S = pot.compute_batched(structs)
    for k in range(len(structs)):
        _check(E[k], F[k], S[k], *ref[k], f"batch {k}")
    res = pot.autotune(structs, reps=2)
    assert len(res["tried"]) >= 3
    for shape in res["tried"]:
        pot.set_launch(shape["bond_tile"], shape["block_threads"], shape["grid_cap"])
        E, F, S = pot.compute_batched(structs)
        for k in range(len(structs)):
            _check(E[k], F[k], S[k], *ref[k], f"shape {shape} batch {k}")


def test_gesn_nve_matches_oracle_trajectory():
    lat, pos, t = gesn_all()[4]
    masses = np.array([72.63 if x == "Ge" else 118.71 for x in t])
    vel = maxwell_boltzmann(800.0, masses, 5)
    pot = StillingerWeberPotential(SW_GESN, species=GESN_SPECIES)
    n_steps = 60
    res = nve_run(pot, lat, pos, vel, t, dt=1.0, n_steps=n_steps, skin_width=0.3,
                  masses=masses, use_graph=True)
    force = lambda l, p, pr: O.sw_multi(l, p, t, SW_GESN, GESN_SPECIES, pairs=pr, nthreads=4)
    ep, ek, p1, v1, nreb = O.nve_run(force, lat, pos, vel, masses, 1.0, n_steps, pot.r_cut, 0.3)
    assert np.abs(res["e_pot"] - ep).max() < 1e-9 * np.abs(ep).max()
    assert np.abs(res["e_kin"] - ek).max() < 1e-9 * np.abs(ek).max()
    assert res["n_rebuilds"] == nreb
    d = res["positions"] - p1
    d -= np.round(d @ np.linalg.inv(lat)) @ lat
    assert np.abs(d).max() < 1e-8
