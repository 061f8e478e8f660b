"""Edge cases of the force call that the reference defines: no pair inside
the cutoff gives zero forces and stress (and zero energy for the bond-order
potentials; engine/tersoff.py:1293-1294, engine/stillinger_weber.py), an
empty structure gives zeros, a ragged batch with an isolated atom and a
no-pair structure equals the single calls, and bad input raises the
reference's exception types (KeyError for an unknown species, ValueError
for non-finite positions or a degenerate lattice; engine/neighborlist.py:
259-263)."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential, TersoffPotential, zhou_eam
from akitsushima_amd.published import ZHOU_2004

from _cases import SW_SI, TERSOFF_SI, TERSOFF_XY, sheared_si

pytestmark = pytest.mark.gpu

CUNI = {s: dict(ZHOU_2004[s]) for s in ("Cu", "Ni")}
SPARSE = np.array([[1.0, 1.0, 1.0], [12.0, 1.0, 1.0], [1.0, 12.0, 1.0], [1.0, 1.0, 12.0],
                   [20.0, 20.0, 20.0], [30.0, 5.0, 17.0]])


def _pots():
    return {"tersoff": (TersoffPotential(TERSOFF_SI, species=["Si"]), ["Si"] * 6),
            "tersoff_xy": (TersoffPotential(TERSOFF_XY, species=["X", "Y"]),
                           ["X", "Y", "X", "Y", "Y", "X"]),
            "sw": (StillingerWeberPotential(SW_SI, species=["Si"]), ["Si"] * 6),
            "eam": (zhou_eam(CUNI, ["Cu", "Ni"]), ["Cu", "Ni", "Ni", "Cu", "Cu", "Ni"])}


@pytest.mark.parametrize("name", ["tersoff", "tersoff_xy", "sw", "eam"])
def test_no_pairs_inside_the_cutoff(name):
    pot, t = _pots()[name]
    lat = np.eye(3) * 40.0
    e, f, s = pot.compute(lat, SPARSE, t)
    assert not f.any() and not s.any()
    # no bond inside r_c: zeros, for EAM without the isolated atoms'
    # embedding energy F(0) (engine/eam.py:236-237), as the oracle
    assert e == 0.0
    if name == "eam":
        e0, f0, _ = O.eam_zhou(lat, SPARSE, t, CUNI, ["Cu", "Ni"])
        assert e0 == 0.0 and not f0.any()
    e1, f1, s1 = pot.compute(lat, SPARSE[:1], t[:1])
    assert e1 == 0.0 and not f1.any() and not s1.any()


def test_eam_isolated_atom_beside_a_bond_keeps_its_embedding_energy():
    pot = zhou_eam(CUNI, ["Cu", "Ni"])
    lat = np.eye(3) * 40.0
    pos = np.array([[5.0, 5.0, 5.0], [7.5, 5.0, 5.0], [25.0, 25.0, 25.0]])
    t = ["Cu", "Ni", "Ni"]
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.eam_zhou(lat, pos, t, CUNI, ["Cu", "Ni"])
    assert abs(e - e0) <= 1e-12 * abs(e0) and e0 != 0.0
    assert np.abs(f - f0).max() <= 1e-12 * np.abs(f0).max() and not f[2].any()
    # the third atom's F(0) is part of the energy: the dimer alone differs
    e2, _, _ = pot.compute(lat, pos[:2], t[:2])
    assert e2 != e
    # and a call after a no-pair call is unaffected by it
    pot.compute(lat, SPARSE, ["Cu"] * 6)
    assert pot.compute(lat, pos, t)[0] == e


@pytest.mark.parametrize("name", ["tersoff", "sw", "eam"])
def test_empty_structure_gives_zeros(name):
    pot, _ = _pots()[name]
    e, f, s = pot.compute(np.eye(3) * 40.0, np.zeros((0, 3)), [])
    assert e == 0.0 and f.shape == (0, 3) and not s.any()


def test_ragged_batch_with_isolated_structures_equals_single_calls():
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    lat, pos, _ = sheared_si(2, shear=0.05, rattle=0.1, seed=3)
    cells = [(lat, pos, ["Si"] * len(pos)), (np.eye(3) * 40.0, SPARSE[:1], ["Si"]),
             (np.eye(3) * 40.0, SPARSE, ["Si"] * 6), (lat, pos[::-1].copy(), ["Si"] * len(pos))]
    E, F, S = pot.compute_batched(cells)
    for k, (l, p, t) in enumerate(cells):
        e, f, s = pot.compute(l, p, t)
        assert abs(E[k] - e) <= 1e-12 * max(1.0, abs(e))
        assert np.abs(F[k] - f).max() <= 1e-12 * max(1.0, np.abs(f).max())
        assert np.abs(S[k] - s).max() <= 1e-12 * max(1e-3, np.abs(s).max())
    assert E[1] == 0.0 and E[2] == 0.0 and not F[2].any()


def test_bad_input_raises_the_reference_exception_types():
    pot, t = _pots()["tersoff_xy"]
    lat = np.eye(3) * 40.0
    with pytest.raises(KeyError):
        pot.compute(lat, SPARSE, ["X", "Y", "Z", "X", "Y", "X"])
    bad = SPARSE.copy()
    bad[2, 1] = np.nan
    with pytest.raises(ValueError):
        pot.compute(lat, bad, t)
    flat = np.array([[5.0, 0.0, 0.0], [0.0, 5.0, 0.0], [5.0, 5.0, 0.0]])
    with pytest.raises(ValueError):
        pot.compute(flat, SPARSE, t)


def _compressed(n=5, factor=0.85, rattle=0.03, seed=11):
    from akitsushima_amd.structures import DIAMOND, cubic

    return cubic(DIAMOND, 5.431 * factor, n, ["Si"] * 8, rattle=rattle, seed=seed)


@pytest.mark.parametrize("name", ["tersoff", "sw"])
@pytest.mark.parametrize("skin", [1.0, 3.0])
def test_compressed_cell_with_a_wide_skin_matches_the_oracle(name, skin):
    """15 % compressed diamond Si with a 1.0 A skin (rows of up to 64 slots)
    and a 3.0 A skin (rows of 128 slots, longer than a wavefront, which the
    bond kernels evaluate on a list of the pairs inside r_cut)."""
    from akitsushima_amd import NeighbourList

    from _cases import rel_err

    lat, pos, t = _compressed()
    pot = (TersoffPotential(TERSOFF_SI, species=["Si"]) if name == "tersoff"
           else StillingerWeberPotential(SW_SI, species=["Si"]))
    nl = NeighbourList(pot.r_cut, skin_width=skin)
    nl.build(lat, pos)
    print(f"compressed {name} skin {skin}: {len(pos)} atoms", flush=True)
    e, f, s = pot.compute(lat, pos, t, neigh_list=nl)
    e0, f0, s0 = (O.tersoff_flat(lat, pos, TERSOFF_SI) if name == "tersoff"
                  else O.sw_flat(lat, pos, SW_SI))
    assert abs(e - e0) <= 1e-10 * abs(e0)
    assert rel_err(f, f0) < 1e-10 and rel_err(s, s0) < 1e-10
    if skin > 2.0:
        with pytest.raises(Exception, match="64 neighbours"):
            pot.site_terms(lat, pos, t, neigh_list=nl)


def test_compressed_cell_md_with_a_wide_skin_matches_the_oracle():
    from akitsushima_amd.md import maxwell_boltzmann, nve_run

    lat, pos, t = _compressed()
    m = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(600.0, m, 2)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    res = nve_run(pot, lat, pos, vel, t, dt=1.0, n_steps=20, skin_width=1.0, masses=m)
    force = lambda l, p, pr: O.sw_flat(l, p, SW_SI, pairs=pr)
    ep, ek, p1, v1, nreb = O.nve_run(force, lat, pos, vel, m, 1.0, 20, pot.r_cut, 1.0)
    assert np.abs(res["e_pot"] - ep).max() < 1e-9 * np.abs(ep).max()
    assert np.abs(res["e_kin"] - ek).max() < 1e-9 * np.abs(ek).max()
    assert res["n_rebuilds"] == nreb
