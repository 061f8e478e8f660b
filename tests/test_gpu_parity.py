"""HIP engine against the oracle and the reference's golden energies.

Bar: neighbour pair sets identical (bitwise, same order); energies,
forces and stresses within 1e-10 relative of the oracle (summation order
only), energies per atom within 1e-10 relative of the reference's ledger.
"""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import (NeighbourList, StillingerWeberPotential, TersoffPotential,
                             zhou_eam)
from akitsushima_amd.published import ZHOU_2004
from akitsushima_amd.structures import FCC, cubic

from _cases import (LEDGER, SW_SI, TERSOFF_SI, TERSOFF_XY, cuni_all, ni, rattled, rel_err,
                    sheared_si, si, xy)

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _check(e, f, s, e0, f0, s0, n, label):
    assert abs(e - e0) / abs(e0) < TOL, f"{label}: E {e} vs {e0}"
    assert rel_err(f, f0) < TOL, f"{label}: F rel err {rel_err(f, f0):.3e}"
    assert rel_err(s, s0) < TOL, f"{label}: S rel err {rel_err(s, s0):.3e}"
    assert np.abs(f.sum(axis=0)).max() < 1e-9 * max(1.0, np.abs(f).max()) * np.sqrt(n)


# --------------------------------------------------------------------------
#  Neighbour list: identical candidate pairs in canonical order
# --------------------------------------------------------------------------

NL_CASES = [
    ("si4_rl3.3", lambda: si(4), 3.0, 0.3),
    ("si8_rl4.07", lambda: si(8), 3.77118, 0.3),
    ("si2_nc2", lambda: si(2), 3.77118, 0.3),
    ("sheared_si2", lambda: sheared_si(2), 3.0, 0.3),
    ("sheared_si3", lambda: sheared_si(3, shear=0.12), 3.77118, 0.5),
    ("ni5", lambda: ni(5), 5.748, 0.3),
]


@pytest.mark.parametrize("name,builder,rc,skin", NL_CASES, ids=[c[0] for c in NL_CASES])
def test_neighbour_pairs_bitwise(name, builder, rc, skin):
    lat, pos, _ = builder()
    nl = NeighbourList(rc, skin_width=skin)
    nl.build(lat, pos)
    i, j = nl.get_candidate_pairs()
    oi, oj, _, _ = O.nl_build(lat, pos, rc + skin)
    assert len(i) == len(oi)
    assert np.array_equal(i, oi) and np.array_equal(j, oj)
    bi, bj = O.brute_force_pairs(lat, pos, rc + skin)
    assert np.array_equal(i, bi) and np.array_equal(j, bj)


def test_neighbour_list_too_thin_cell_raises():
    lat, pos, _ = si(1)
    nl = NeighbourList(3.0, skin_width=0.3)
    with pytest.raises(ValueError, match="half the smallest perpendicular"):
        nl.build(lat, pos)


# --------------------------------------------------------------------------
#  Tersoff
# --------------------------------------------------------------------------

TILES = [1, 2]   # site kernel: lane per bond / lane per (bond, partner)


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("n", [4, 8])
def test_tersoff_si_matches_ledger_and_oracle(n, tile):
    lat, pos, t = si(n)
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    pot.set_bond_tile(tile)
    e, f, s = pot.compute(lat, pos, t)
    N = len(pos)
    ref = LEDGER[f"tersoff_si/{N}"]["e_per_atom"]
    assert abs(e / N - ref) / abs(ref) < TOL
    e0, f0, s0 = O.tersoff_flat(lat, pos, TERSOFF_SI)
    _check(e, f, s, e0, f0, s0, N, f"tersoff si{n}")


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("n", [4, 8])
def test_tersoff_two_species_matches_ledger_and_oracle(n, tile):
    lat, pos, t = xy(n)
    pot = TersoffPotential(TERSOFF_XY, species=["X", "Y"])
    pot.set_bond_tile(tile)
    e, f, s = pot.compute(lat, pos, t)
    N = len(pos)
    ref = LEDGER[f"tersoff_xy/{N}"]["e_per_atom"]
    assert abs(e / N - ref) / abs(ref) < TOL
    e0, f0, s0 = O.tersoff_multi(lat, pos, t, TERSOFF_XY, ["X", "Y"])
    _check(e, f, s, e0, f0, s0, N, f"tersoff xy{n}")


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("m,gamma,l3", [(3.0, 1.0, 1.7322), (3.0, 1.07, 1.3), (1.0, 0.93, 0.9)])
def test_tersoff_fourteen_parameter_form(m, gamma, l3, tile):
    """Explicit m, gamma, lambda3 (the multi-species record form) on a
    liquid-like rattled, sheared cell."""
    lat, pos, t = sheared_si(3, shear=0.05, rattle=0.15, seed=11)
    rec = dict(TERSOFF_SI, lambda3=l3, m=m, gamma=gamma)
    recs = {("Si", "Si", "Si"): rec}
    pot = TersoffPotential(recs, species=["Si"])
    pot.set_bond_tile(tile)
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.tersoff_multi(lat, pos, t, recs, ["Si"])
    _check(e, f, s, e0, f0, s0, len(pos), f"tersoff m={m}")


# --------------------------------------------------------------------------
#  Stillinger-Weber
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [4, 8])
def test_sw_si_matches_ledger_and_oracle(n):
    lat, pos, t = si(n)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    e, f, s = pot.compute(lat, pos, t)
    N = len(pos)
    ref = LEDGER[f"sw_si/{N}"]["e_per_atom"]
    assert abs(e / N - ref) / abs(ref) < TOL
    e0, f0, s0 = O.sw_flat(lat, pos, SW_SI)
    _check(e, f, s, e0, f0, s0, N, f"sw si{n}")


def test_sw_liquid_like_sheared():
    lat, pos, t = sheared_si(3, shear=0.1, rattle=0.2, seed=5)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.sw_flat(lat, pos, SW_SI)
    _check(e, f, s, e0, f0, s0, len(pos), "sw sheared")


# --------------------------------------------------------------------------
#  Zhou EAM
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [5, 10])
def test_eam_ni_matches_ledger_and_oracle(n):
    lat, pos, t = ni(n)
    pot = zhou_eam({"Ni": ZHOU_2004["Ni"]}, ["Ni"])
    assert pot.r_cut == LEDGER[f"eam_ni/{len(pos)}"]["r_cut"]
    e, f, s = pot.compute(lat, pos, t)
    N = len(pos)
    ref = LEDGER[f"eam_ni/{N}"]["e_per_atom"]
    assert abs(e / N - ref) / abs(ref) < TOL
    e0, f0, s0 = O.eam_zhou(lat, pos, t, {"Ni": O.ZHOU_NI}, ["Ni"], nthreads=8)
    _check(e, f, s, e0, f0, s0, N, f"eam ni{n}")


@pytest.mark.parametrize("n", [5, 10])
def test_eam_cuni_matches_ledger_and_oracle(n):
    lat, pos, t = cuni_all()[n]
    el = {"Cu": ZHOU_2004["Cu"], "Ni": ZHOU_2004["Ni"]}
    pot = zhou_eam(el, ["Cu", "Ni"])
    assert pot.r_cut == LEDGER[f"eam_cuni/{len(pos)}"]["r_cut"]
    e, f, s = pot.compute(lat, pos, t)
    N = len(pos)
    ref = LEDGER[f"eam_cuni/{N}"]["e_per_atom"]
    assert abs(e / N - ref) / abs(ref) < TOL
    e0, f0, s0 = O.eam_zhou(lat, pos, t, {"Cu": O.ZHOU_CU, "Ni": O.ZHOU_NI}, ["Cu", "Ni"],
                            nthreads=8)
    _check(e, f, s, e0, f0, s0, N, f"eam cuni{n}")


def test_eam_cuni_on_a_sheared_cell_matches_the_oracle():
    # a triclinic cell: the Zhou kernels' general minimum image (a cubic
    # cell takes the diagonal one)
    lat, pos, _ = cubic(FCC, 3.58, 5, ["Cu"] * 4, rattle=0.0)
    T = np.eye(3)
    T[0, 1], T[2, 0], T[1, 2] = 0.06, -0.03, 0.02
    frac = pos @ np.linalg.inv(lat)
    lat = lat @ T
    pos = frac @ lat + np.random.default_rng(11).normal(0.0, 0.04, pos.shape)
    t = [str(x) for x in np.random.default_rng(12).choice(["Cu", "Ni"], len(pos))]
    pot = zhou_eam({"Cu": ZHOU_2004["Cu"], "Ni": ZHOU_2004["Ni"]}, ["Cu", "Ni"])
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.eam_zhou(lat, pos, t, {"Cu": O.ZHOU_CU, "Ni": O.ZHOU_NI}, ["Cu", "Ni"],
                            nthreads=8)
    _check(e, f, s, e0, f0, s0, len(pos), "eam cuni sheared")


# --------------------------------------------------------------------------
#  Lists, determinism, batching
# --------------------------------------------------------------------------

def test_owned_list_rebuilds_when_atoms_move():
    lat, pos, t = si(4)
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    pot.compute(lat, pos, t)
    for k, shift in enumerate((0.05, 0.2, 0.5)):
        p2 = rattled(lat, pos, shift, seed=20 + k)
        e, f, s = pot.compute(lat, p2, t)
        e0, f0, s0 = O.tersoff_flat(lat, p2, TERSOFF_SI)
        _check(e, f, s, e0, f0, s0, len(p2), f"moved {shift}")


def test_owned_list_follows_a_strained_cell():
    lat, pos, t = si(4)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    pot.compute(lat, pos, t)
    for eps in (0.001, 0.02, -0.03):
        T = np.eye(3) + np.array([[eps, 0.3 * eps, 0], [0.3 * eps, -eps, 0], [0, 0, 0.5 * eps]])
        lat2 = lat @ T
        pos2 = pos @ np.linalg.inv(lat) @ lat2
        e, f, s = pot.compute(lat2, pos2, t)
        e0, f0, s0 = O.sw_flat(lat2, pos2, SW_SI)
        _check(e, f, s, e0, f0, s0, len(pos2), f"strain {eps}")


def test_passed_device_list_is_used():
    lat, pos, t = si(4)
    nl = NeighbourList(3.0, skin_width=0.5)
    nl.build(lat, pos)
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    p2 = rattled(lat, pos, 0.05, seed=7)
    assert not nl.check_and_rebuild(lat, p2)
    e, f, s = pot.compute(lat, p2, t, neigh_list=nl)
    e0, f0, s0 = O.tersoff_flat(lat, p2, TERSOFF_SI)
    _check(e, f, s, e0, f0, s0, len(p2), "passed list")


def test_results_are_bitwise_reproducible():
    lat, pos, t = cuni_all()[5]
    el = {"Cu": ZHOU_2004["Cu"], "Ni": ZHOU_2004["Ni"]}
    pot = zhou_eam(el, ["Cu", "Ni"])
    a = pot.compute(lat, pos, t)
    b = pot.compute(lat, pos, t)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    lat, pos, t = si(4)
    pt = TersoffPotential(TERSOFF_SI, species=["Si"])
    a = pt.compute(lat, pos, t)
    b = pt.compute(lat, pos, t)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_batched_equals_single_calls():
    from akitsushima_amd.structures import diamond_batch

    structs = diamond_batch(12, 64, 512, 0.02, 0.05, seed=4, rep_range=(2, 4))
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    E, F, S = pot.compute_batched(structs)
    for k, (lat, pos, t) in enumerate(structs):
        e0, f0, s0 = O.tersoff_flat(lat, pos, TERSOFF_SI)
        _check(E[k], F[k], S[k], e0, f0, s0, len(pos), f"batch {k}")


# --------------------------------------------------------------------------
#  Tersoff (1989) mixing records (engine/tersoff.py:1968-2034)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("tile", TILES)
def test_tersoff1989_sic_matches_oracle_and_table_ii(tile):
    from scipy.optimize import minimize_scalar

    from akitsushima_amd import tersoff1989_to_triplets
    from akitsushima_amd.published import TERSOFF_1989
    from akitsushima_amd.structures import DIAMOND, cubic

    recs = tersoff1989_to_triplets(TERSOFF_1989["elements"], TERSOFF_1989["chi"], ["Si", "C"])
    pot = TersoffPotential(recs, species=["Si", "C"])
    pot.set_bond_tile(tile)
    lat, pos, t = cubic(DIAMOND, 4.32, 3, ["Si"] * 4 + ["C"] * 4, rattle=0.06, seed=3)
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.tersoff_multi(lat, pos, t, recs, ["Si", "C"])
    _check(e, f, s, e0, f0, s0, len(pos), "SiC 1989")
    efn = lambda a: pot.compute(*cubic(DIAMOND, a, 2, ["Si"] * 4 + ["C"] * 4, rattle=0.0))[0] / 64
    r = minimize_scalar(efn, bounds=(4.2, 4.45), method="bounded", options={"xatol": 1e-8})
    assert abs(r.x - 4.32) < 0.005 and abs(-r.fun - 6.165) < 0.0058
