"""The device's Nose-Hoover NVT and isotropic NPT runs against the
reference's own recorded runs.

tests/golden/nhc_legacy_si64.npz holds what the reference recorded for two
300-step runs of 64-atom Si with SW Si at 600 K, a two-link chain with
tau = 100 fs and, for NPT, P = 0 with tau_baro = 500 fs (its
tests/data/nhc_legacy_0.3.0a7; tests/test_nhc_chain.py:136-170).  Started
from the same seeds, the device runs the 300 steps in one captured-graph
loop each and must land where the reference did, to the reference's own
1e-9 relative bound on the step-300 state; energies, pressures and volumes
of steps 1..150 follow the recorded ones step by step.  Continued from the
recorded step-150 checkpoint (positions, velocities, cell and the chain and
barostat scalars, handed to the device as its thermostat state), the
device lands on the same step-300 state: the reference's own restart test
(tests/test_nhc_chain.py:143-170).  A run split in two with the state handed
over equals the uninterrupted run for every thermostat.
"""
import os

import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential
from akitsushima_amd.md import (csvr_run, langevin_run, npt_run, nvt_run, state_fields,
                                thermostat_state)

from _cases import SW_SI, si64_recorded_md

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REC = np.load(os.path.join(HERE, "golden", "nhc_legacy_si64.npz"), allow_pickle=False)
N_DOF = 3 * 64 - 3


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def _mic_rel(pos, want, lat):
    d = pos - want
    d -= np.round(d @ np.linalg.inv(lat)) @ lat
    return float(np.max(np.abs(d)) / np.max(np.abs(want)))


def _start():
    lat, pos, t = si64_recorded_md()
    m = np.full(len(pos), 28.086)
    return lat, pos, t, O.mb_velocities(600.0, m, 7), m


@pytest.mark.parametrize("use_graph", [True, False])
def test_device_nvt_lands_on_the_recorded_run(use_graph):
    lat, pos, t, vel, m = _start()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    r = nvt_run(pot, lat, pos, vel, t, 600.0, dt=1.0, n_steps=300, tau_nhc=100.0, nhc_chain=2,
                skin_width=0.5, masses=m, use_graph=use_graph)
    assert _rel(r["e_pot"][1:151], REC["nvt_e_pot"]) < 1e-11
    assert _rel(r["T"][1:151], REC["nvt_T"]) < 1e-9
    assert np.abs(r["e_cons"][REC["nvt_row_step"] + 1] - REC["nvt_row_H"]).max() < 1e-9
    assert _mic_rel(r["positions"], REC["nvt_pos300"], lat) < 1e-9
    assert _rel(r["velocities"], REC["nvt_vel300"]) < 1e-9


def test_device_npt_lands_on_the_recorded_run():
    lat, pos, t, vel, m = _start()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    r = npt_run(pot, lat, pos, vel, t, 600.0, pressure=0.0, dt=1.0, n_steps=300,
                tau_nhc=100.0, nhc_chain=2, tau_baro=500.0, skin_width=0.5, masses=m)
    assert _rel(r["volume"][1:151], REC["npt_V"]) < 1e-11
    assert np.abs(r["pressure_gpa"][1:151] - REC["npt_P"]).max() < 1e-9
    assert _rel(r["T"][1:151], REC["npt_T"]) < 1e-9
    assert abs(r["e_cons"][1] - float(REC["npt_H_first_row"])) < 1e-9
    assert _rel(r["lattice"], REC["npt_lat300"]) < 1e-12
    assert _mic_rel(r["positions"], REC["npt_pos300"], r["lattice"]) < 1e-9
    assert _rel(r["velocities"], REC["npt_vel300"]) < 1e-9


def test_device_nvt_continues_the_recorded_checkpoint():
    m = np.full(64, 28.086)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    c = REC["nvt_chain150"]
    st = thermostat_state(xi=c[0:2], vxi=c[2:4])
    r = nvt_run(pot, REC["nvt_lat150"], REC["nvt_pos150"], REC["nvt_vel150"], ["Si"] * 64, 600.0,
                dt=1.0, n_steps=150, tau_nhc=100.0, nhc_chain=2, masses=m, state=st)
    assert _mic_rel(r["positions"], REC["nvt_pos300"], REC["nvt_lat300"]) < 1e-9
    assert _rel(r["velocities"], REC["nvt_vel300"]) < 1e-9


def test_device_npt_continues_the_recorded_checkpoint():
    m = np.full(64, 28.086)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    s = REC["npt_state150"]
    st = thermostat_state(xi=s[0:2], vxi=s[2:4], p_eps=s[4], xi_b=s[5:7], vxi_b=s[7:9])
    r = npt_run(pot, REC["npt_lat150"], REC["npt_pos150"], REC["npt_vel150"], ["Si"] * 64, 600.0,
                pressure=0.0, dt=1.0, n_steps=150, tau_nhc=100.0, nhc_chain=2, tau_baro=500.0,
                masses=m, state=st)
    assert _rel(r["lattice"], REC["npt_lat300"]) < 1e-12
    assert _mic_rel(r["positions"], REC["npt_pos300"], REC["npt_lat300"]) < 1e-9
    assert _rel(r["velocities"], REC["npt_vel300"]) < 1e-9


RUNS = {
    "nhc": lambda *a, **k: nvt_run(*a, tau_nhc=50.0, nhc_chain=3, **k),
    "langevin": lambda *a, **k: langevin_run(*a, damping=80.0, seed=11, **k),
    "csvr": lambda *a, **k: csvr_run(*a, damping=80.0, seed=12, **k),
    "npt": lambda *a, **k: npt_run(*a, pressure=0.3, tau_nhc=50.0, nhc_chain=3, tau_baro=200.0,
                                   **k),
}


@pytest.mark.parametrize("kind", list(RUNS))
def test_a_split_run_equals_the_uninterrupted_one(kind):
    lat, pos, t, vel, m = _start()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    run = RUNS[kind]
    full = run(pot, lat, pos, vel, t, 600.0, dt=1.0, n_steps=200, masses=m)
    a = run(pot, lat, pos, vel, t, 600.0, dt=1.0, n_steps=100, masses=m)
    b = run(pot, a.get("lattice", lat), a["positions"], a["velocities"], t, 600.0, dt=1.0,
            n_steps=100, masses=m, state=a["state"])
    lat_end = full.get("lattice", lat)
    assert _mic_rel(b["positions"], full["positions"], lat_end) < 1e-9
    assert _rel(b["velocities"], full["velocities"]) < 1e-9
    assert _rel(b["e_cons"], full["e_cons"][100:]) < 1e-10
    assert _rel(b["e_pot"], full["e_pot"][100:]) < 1e-10
    fb, ff = state_fields(b["state"]), state_fields(full["state"])
    for key in ("xi", "vxi", "xi_b", "vxi_b"):
        assert np.allclose(fb[key], ff[key], rtol=1e-8, atol=1e-14), key
    assert abs(fb["p_eps"] - ff["p_eps"]) <= 1e-8 * max(abs(ff["p_eps"]), 1e-300)
    assert abs(fb["heat"] - ff["heat"]) <= 1e-9 * max(abs(ff["heat"]), 1e-12)
    assert fb["rng_draws"] == ff["rng_draws"]
    if kind in ("langevin", "csvr"):
        assert ff["rng_draws"] == (200 if kind == "langevin" else 400)
        assert abs(ff["heat"]) > 0.0
    if kind == "npt":
        assert _rel(b["lattice"], full["lattice"]) < 1e-12
        assert abs(full["volume"][-1] - full["volume"][0]) > 1e-4 * full["volume"][0]
