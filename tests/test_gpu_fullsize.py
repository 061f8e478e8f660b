"""The measured workloads at their full sizes against the oracle (1e-10
relative in E, F, sigma) and against size-independent properties: zero net
force, symmetric stress, and, for the batch, equality with single calls.

Sizes: Tersoff Si 16^3 diamond cells (32768 atoms); SW Si 20^3 (64000);
Zhou Cu-Ni fcc 40^3 (256000); 512 Tersoff structures of 64-2048 atoms; the
benchmark's multi-species SW Ge-Sn 20^3 cell (64000), setfl Cu-Ni at the
EAM size, and the Green-Kubo flux of the 64000-atom SW cell."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential, TersoffPotential, zhou_eam
from akitsushima_amd.published import SW_SI, TERSOFF_SI_T3, ZHOU_2004
from akitsushima_amd.structures import DIAMOND, FCC, cubic, diamond_batch, random_species

from _cases import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _check(e, f, s, e0, f0, s0, label):
    assert abs(e - e0) / abs(e0) < TOL, f"{label}: E {e} vs {e0}"
    assert rel_err(f, f0) < TOL, f"{label}: F rel err {rel_err(f, f0):.3e}"
    assert rel_err(s, s0) < TOL, f"{label}: S rel err {rel_err(s, s0):.3e}"
    # size-independent: net force ~ 0 (translation invariance), symmetric stress
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * np.abs(f).max() * len(f)
    assert np.abs(s - s.T).max() < 1e-12 * max(1.0, np.abs(s).max())


@pytest.mark.parametrize("tile", [1, 2])
def test_tersoff_32768_atoms(tile):
    lat, pos, t = cubic(DIAMOND, 5.431, 16, ["Si"] * 8, rattle=0.05, seed=0)
    pot = TersoffPotential(dict(TERSOFF_SI_T3), species=["Si"])
    pot.set_bond_tile(tile)
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.tersoff_flat(lat, pos, dict(TERSOFF_SI_T3), nthreads=8)
    _check(e, f, s, e0, f0, s0, "tersoff 32768")


def test_sw_64000_atoms():
    lat, pos, t = cubic(DIAMOND, 5.431, 20, ["Si"] * 8, rattle=0.05, seed=2)
    pot = StillingerWeberPotential(dict(SW_SI), species=["Si"])
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.sw_flat(lat, pos, dict(SW_SI), nthreads=8)
    _check(e, f, s, e0, f0, s0, "sw 64000")


def test_eam_cuni_256000_atoms():
    lat, pos, _ = cubic(FCC, 3.56, 40, ["X"] * 4, rattle=0.03, seed=1)
    types = random_species(np.random.default_rng(0), ["Cu", "Ni"], len(pos))
    pot = zhou_eam({"Cu": dict(ZHOU_2004["Cu"]), "Ni": dict(ZHOU_2004["Ni"])}, ["Cu", "Ni"])
    e, f, s = pot.compute(lat, pos, types)
    e0, f0, s0 = O.eam_zhou(lat, pos, types, {"Cu": O.ZHOU_CU, "Ni": O.ZHOU_NI}, ["Cu", "Ni"],
                            nthreads=8)
    _check(e, f, s, e0, f0, s0, "eam 256000")


def test_batched_512_structures():
    structs = diamond_batch(512, 64, 2048, 0.02, 0.02, seed=5, rep_range=(2, 8))
    pot = TersoffPotential(dict(TERSOFF_SI_T3), species=["Si"])
    pot.autotune(structs, reps=2)
    E, F, S = pot.compute_batched(structs)
    single = TersoffPotential(dict(TERSOFF_SI_T3), species=["Si"])
    for k in range(0, 512, 37):
        lat, p, t = structs[k]
        e0, f0, s0 = O.tersoff_flat(lat, p, dict(TERSOFF_SI_T3), nthreads=8)
        _check(E[k], F[k], S[k], e0, f0, s0, f"batch {k}")
        e1, f1, s1 = single.compute(lat, p, t)
        assert abs(E[k] - e1) <= 1e-12 * abs(e1) and rel_err(F[k], f1) < 1e-12


def test_sw_gesn_64000_atoms():
    from _cases import GESN_SPECIES, SW_GESN

    lat, pos, _ = cubic(DIAMOND, 5.762, 20, ["Ge"] * 8)
    rng = np.random.default_rng(1)
    t = [str(x) for x in rng.choice(GESN_SPECIES, len(pos), p=[0.875, 0.125])]
    pot = StillingerWeberPotential(SW_GESN, species=GESN_SPECIES)
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.sw_multi(lat, pos, t, SW_GESN, GESN_SPECIES, nthreads=8)
    _check(e, f, s, e0, f0, s0, "gesn 64000")


def test_setfl_cuni_256000_atoms(tmp_path):
    from akitsushima_amd import read_setfl
    from akitsushima_amd.published import ATOMIC_MASS

    el = {"Cu": dict(ZHOU_2004["Cu"]), "Ni": dict(ZHOU_2004["Ni"])}
    path = str(tmp_path / "CuNi.eam.alloy")
    O.write_setfl_zhou(path, el, ["Cu", "Ni"], ATOMIC_MASS)
    lat, pos, _ = cubic(FCC, 3.56, 40, ["X"] * 4, rattle=0.03, seed=1)
    types = random_species(np.random.default_rng(0), ["Cu", "Ni"], len(pos))
    pot = read_setfl(path)
    e, f, s = pot.compute(lat, pos, types)
    tab, names = O.read_setfl(path)
    e0, f0, s0 = O.eam_table(lat, pos, types, tab, names, nthreads=8)
    _check(e, f, s, e0, f0, s0, "setfl 256000")


def test_heat_flux_64000_atoms():
    import oracle.site_terms as ST
    from akitsushima_amd.md import maxwell_boltzmann

    lat, pos, t = cubic(DIAMOND, 5.431, 20, ["Si"] * 8, rattle=0.05, seed=2)
    pot = StillingerWeberPotential(dict(SW_SI), species=["Si"])
    m = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(700.0, m, 5)
    J, e = pot.heat_flux(lat, pos, vel, t, m)
    ref = ST.sw(lat, pos, t, O.sw_flat_as_records(dict(SW_SI)), ["Si"])
    J0 = ref.heat_flux(vel, m, abs(np.linalg.det(lat)))
    assert rel_err(J, J0) < 1e-10 and rel_err(e, ref.e_site) < 1e-10
