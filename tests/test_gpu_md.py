"""Device-resident NVE (velocity Verlet + on-device list rebuild test,
hipGraph replay) against the oracle's velocity Verlet with the reference's
rebuild rule."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential, TersoffPotential
from akitsushima_amd.md import maxwell_boltzmann, nve_run

from _cases import SW_SI, TERSOFF_SI, si

pytestmark = pytest.mark.gpu


def _setup(n=4, T=1000.0, seed=3):
    lat, pos, t = si(n)
    masses = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(T, masses, seed)
    return lat, pos, t, masses, vel


def test_maxwell_boltzmann_matches_oracle_draw():
    lat, pos, t, masses, vel = _setup()
    assert np.array_equal(vel, O.mb_velocities(1000.0, masses, 3))


@pytest.mark.parametrize("use_graph", [True, False])
def test_sw_nve_trajectory_matches_oracle(use_graph):
    lat, pos, t, masses, vel = _setup()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    n_steps = 100
    res = nve_run(pot, lat, pos, vel, t, dt=1.0, n_steps=n_steps, skin_width=0.3,
                  masses=masses, use_graph=use_graph)
    force = lambda l, p, pr: O.sw_flat(l, p, SW_SI, pairs=pr)
    ep, ek, p1, v1, nreb = O.nve_run(force, lat, pos, vel, masses, 1.0, n_steps, 3.77118, 0.3)
    assert np.abs(res["e_pot"] - ep).max() < 1e-9 * np.abs(ep).max()
    assert np.abs(res["e_kin"] - ek).max() < 1e-9 * np.abs(ek).max()
    assert res["n_rebuilds"] == nreb
    d = res["positions"] - p1
    d -= np.round(d @ np.linalg.inv(lat)) @ lat
    assert np.abs(d).max() < 1e-8
    assert np.abs(res["drift"]).max() / len(pos) < 2e-4


def test_graph_and_stream_runs_are_bitwise_equal():
    lat, pos, t, masses, vel = _setup(seed=8)
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    a = nve_run(pot, lat, pos, vel, t, n_steps=50, masses=masses, use_graph=True)
    b = nve_run(pot, lat, pos, vel, t, n_steps=50, masses=masses, use_graph=False)
    assert np.array_equal(a["positions"], b["positions"])
    assert np.array_equal(a["e_pot"], b["e_pot"])
