"""The oracle's Nose-Hoover NVT and isotropic NPT integrators against the
reference's own recorded runs, no GPU.

tests/golden/nhc_legacy_si64.npz holds what the reference recorded for two
300-step runs of 64-atom Si with SW Si at 600 K (its
tests/data/nhc_legacy_0.3.0a7, used by tests/test_nhc_chain.py:136-170): the
per-step energies, temperatures, pressures and volumes of steps 1..150, the
chain (and barostat) state after step 150 and the positions, velocities and
cell after step 300.  The start is rebuilt here from the same seeds
(structure: default_rng(1); Maxwell-Boltzmann velocities: default_rng(7)),
so the oracle runs the whole trajectory and must land where the reference
did.  The reference compares step-300 states to 1e-9 relative (its forces
moved by ~1e-9 eV/A since the recording); a wrong chain mass, a dropped
factor in the barostat force or a swapped half-step is off by orders of
magnitude more after 300 steps of a chaotic trajectory.

Those runs wrote their thermo rows after steps 1, 11, ..., 141 (the row
labelled c held the state after c + 1 steps in that version).
"""
import os

import numpy as np
import pytest

import oracle as O

from _cases import SW_SI, si64_recorded_md

HERE = os.path.dirname(os.path.abspath(__file__))
REC = np.load(os.path.join(HERE, "golden", "nhc_legacy_si64.npz"), allow_pickle=False)
R_CUT = 1.80 * 2.0951
N_DOF = 3 * 64 - 3


def _start():
    lat, pos, _ = si64_recorded_md()
    m = np.full(len(pos), 28.086)
    return lat, pos, O.mb_velocities(600.0, m, 7), m


def _force(l, p, pr):
    return O.sw_flat(l, p, SW_SI, pairs=pr)


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def test_start_is_the_recorded_one():
    lat, pos, vel, m = _start()
    ek = 0.5 * np.sum(m[:, None] * vel ** 2) / O.EV_ANG_TO_AMU_ANG_FS2
    assert abs(2 * ek / (N_DOF * O.K_B) - 600.0) < 1e-9
    assert np.abs((m[:, None] * vel).sum(axis=0)).max() < 1e-12


@pytest.fixture(scope="module")
def nvt():
    lat, pos, vel, m = _start()
    st = {}
    half = O.md_run(_force, lat, pos, vel, m, 1.0, 150, R_CUT, 0.5, nhc=(600.0, 100.0, 2),
                    state_out=st)
    full = O.md_run(_force, lat, pos, vel, m, 1.0, 300, R_CUT, 0.5, nhc=(600.0, 100.0, 2))
    return half, st, full


def test_nvt_energies_of_every_step(nvt):
    _, _, (ep, ek, ec, *_) = nvt
    assert _rel(ep[1:151], REC["nvt_e_pot"]) < 1e-11
    assert _rel(ek[1:151], REC["nvt_e_kin"]) < 1e-9
    assert _rel(2 * ek[1:151] / (N_DOF * O.K_B), REC["nvt_T"]) < 1e-9
    rows = REC["nvt_row_step"] + 1
    assert np.abs(ec[rows] - REC["nvt_row_H"]).max() < 1e-9


def test_nvt_chain_state_after_150_steps(nvt):
    _, st, _ = nvt
    xi1, xi2, v1, v2 = REC["nvt_chain150"]
    assert _rel(st["xi"], [xi1, xi2]) < 1e-9
    assert _rel(st["vxi"], [v1, v2]) < 1e-8


def test_nvt_state_after_300_steps(nvt):
    *_, (_, _, _, pos, vel, _, _) = nvt
    assert _rel(pos, REC["nvt_pos300"]) < 1e-9
    assert _rel(vel, REC["nvt_vel300"]) < 1e-9


@pytest.fixture(scope="module")
def npt():
    lat, pos, vel, m = _start()
    args = (600.0, 100.0, 2, 0.0, 500.0)
    return (O.npt_run(_force, lat, pos, vel, m, 1.0, 150, R_CUT, 0.5, *args),
            O.npt_run(_force, lat, pos, vel, m, 1.0, 300, R_CUT, 0.5, *args))


def test_npt_pressure_volume_and_temperature_of_every_step(npt):
    _, r = npt
    assert _rel(r["v"][1:151], REC["npt_V"]) < 1e-11
    assert np.abs(r["p"][1:151] - REC["npt_P"]).max() < 1e-9
    assert _rel(2 * r["ek"][1:151] / (N_DOF * O.K_B), REC["npt_T"]) < 1e-9
    assert abs(r["h"][1] - float(REC["npt_H_first_row"])) < 1e-9


def test_npt_chain_and_barostat_state_after_150_steps(npt):
    r, _ = npt
    s = REC["npt_state150"]
    assert _rel(r["xi"], s[0:2]) < 1e-9
    assert _rel(r["vxi"], s[2:4]) < 1e-8
    assert abs(r["p_eps"] - s[4]) < 1e-8 * abs(s[4])
    assert _rel(r["xi_b"], s[5:7]) < 1e-9
    assert _rel(r["vxi_b"], s[7:9]) < 1e-8


def test_npt_state_after_300_steps(npt):
    _, r = npt
    assert _rel(r["positions"], REC["npt_pos300"]) < 1e-9
    assert _rel(r["velocities"], REC["npt_vel300"]) < 1e-9
    assert _rel(r["lattice"], REC["npt_lat300"]) < 1e-12


# The reference's own restart test (tests/test_nhc_chain.py:143-170): the
# step-150 checkpoint -- positions, velocities, cell, chain (and barostat)
# scalars -- carried on for 150 steps lands on the uninterrupted run.

def test_nvt_continued_from_the_checkpoint():
    m = np.full(64, 28.086)
    c = REC["nvt_chain150"]
    *_, pos, vel, _, _ = O.md_run(_force, REC["nvt_lat150"], REC["nvt_pos150"],
                                  REC["nvt_vel150"], m, 1.0, 150, R_CUT, 0.5,
                                  nhc=(600.0, 100.0, 2), nhc_state=(c[0:2], c[2:4]))
    assert _rel(pos, REC["nvt_pos300"]) < 1e-9
    assert _rel(vel, REC["nvt_vel300"]) < 1e-9


def test_npt_continued_from_the_checkpoint():
    m = np.full(64, 28.086)
    s = REC["npt_state150"]
    start = {"xi": s[0:2], "vxi": s[2:4], "p_eps": s[4], "xi_b": s[5:7], "vxi_b": s[7:9]}
    r = O.npt_run(_force, REC["npt_lat150"], REC["npt_pos150"], REC["npt_vel150"], m, 1.0, 150,
                  R_CUT, 0.5, 600.0, 100.0, 2, 0.0, 500.0, start=start)
    assert _rel(r["positions"], REC["npt_pos300"]) < 1e-9
    assert _rel(r["velocities"], REC["npt_vel300"]) < 1e-9
    assert _rel(r["lattice"], REC["npt_lat300"]) < 1e-12
    # a restart that forgets the barostat momentum lands elsewhere
    bad = O.npt_run(_force, REC["npt_lat150"], REC["npt_pos150"], REC["npt_vel150"], m, 1.0,
                    150, R_CUT, 0.5, 600.0, 100.0, 2, 0.0, 500.0, start=dict(start, p_eps=0.0))
    assert _rel(bad["lattice"], REC["npt_lat300"]) > 1e-6
