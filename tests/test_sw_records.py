"""Host logic of the multi-species SW records (no GPU): the two-body
symmetry rule of engine/stillinger_weber.py:713-762, the cutoff of
:860-871, the dense tables of :1255-1315 and the config forms of
fitting/registry.py:572-592."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential, force_fn_from_config
from akitsushima_amd.potentials import sw_tables

from _cases import GESN_SPECIES, LEDGER, SW_GESN


def test_tomita_records_are_accepted_with_the_ledger_cutoff():
    pot = StillingerWeberPotential(SW_GESN, species=GESN_SPECIES)
    assert pot._multi
    assert pot.r_cut == LEDGER["sw_gesn_tomita/512"]["r_cut"]


def test_tables_follow_the_reference_resolution():
    tb, leg, tri = sw_tables(SW_GESN, GESN_SPECIES)
    tb0, leg0, tri0, rc0 = O.sw_tables(SW_GESN, GESN_SPECIES)
    assert np.array_equal(tb, tb0) and np.array_equal(leg, leg0) and np.array_equal(tri, tri0)
    # every Tomita triple has both leg records
    assert tri[..., 0].all()
    # (Ge, Ge, Sn) and (Ge, Sn, Ge) are separate records
    assert tri[0, 0, 1, 2] == SW_GESN[("Ge", "Ge", "Sn")]["lambda"]


def test_absent_leg_record_switches_the_triple_off():
    recs = {k: dict(v) for k, v in SW_GESN.items()}
    del recs[("Sn", "Ge", "Ge")]
    del recs[("Ge", "Sn", "Sn")]
    tb, leg, tri = sw_tables(recs, GESN_SPECIES)
    assert tb[0, 1, 0] == 0.0 and tb[1, 0, 0] == 0.0
    assert leg[0, 1, 1] == 0.0          # a = 0: no Ge-Sn leg
    assert tri[0, 0, 1, 0] == 0.0 and tri[0, 1, 0, 0] == 0.0
    assert tri[0, 0, 0, 0] == 1.0


def test_asymmetric_two_body_records_are_refused():
    recs = {k: dict(v) for k, v in SW_GESN.items()}
    recs[("Ge", "Sn", "Sn")]["A"] *= 1.2
    with pytest.raises(ValueError, match="disagree"):
        StillingerWeberPotential(recs, species=GESN_SPECIES)


def test_one_sided_two_body_record_is_refused():
    recs = {k: dict(v) for k, v in SW_GESN.items()}
    recs[("Sn", "Ge", "Ge")]["A"] = 0.0
    with pytest.raises(ValueError, match="does not"):
        StillingerWeberPotential(recs, species=GESN_SPECIES)


def test_three_body_only_differences_are_allowed():
    recs = {k: dict(v) for k, v in SW_GESN.items()}
    recs[("Ge", "Sn", "Sn")]["gamma"] = 1.3
    recs[("Ge", "Sn", "Sn")]["lambda"] = 20.0
    StillingerWeberPotential(recs, species=GESN_SPECIES)


def test_records_need_species():
    with pytest.raises(ValueError, match="species"):
        StillingerWeberPotential(SW_GESN)


def test_config_with_record_list():
    rows = [dict(v, i=i, j=j, k=k) for (i, j, k), v in SW_GESN.items()]
    fn = force_fn_from_config({"type": "SW", "species": GESN_SPECIES, "params": rows})
    assert fn.r_cut == LEDGER["sw_gesn_tomita/512"]["r_cut"]
