"""Flexible-cell NPT and the Langevin barostat on the device against the
oracle (pinned in tests/test_oracle_npt_cell.py and, for the isotropic
barostat, by the reference's recorded runs).

Every cell mode of NPTSimulation (cell_dof aniso, tri, xy, xy_iso, z;
solvers/molecular_dynamics.py:2113-2138, 2566-2692) with the Nose-Hoover
barostat chain, and the Langevin barostat (:2596-2617) with the isotropic,
orthorhombic and triclinic cell, on strained and sheared cells: E_pot, E_kin,
H_NPT, V and P_inst of every step, the final cell, positions and velocities
and the rebuild count.  The Langevin noise is the Philox stream the oracle
regenerates.  Also for Tersoff X-Y and Zhou Cu-Ni (their off-diagonal stress
drives the shear), the frozen-barostat limit, and a split run with the
strain momenta, heat and random draws handed over.
"""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential, TersoffPotential, zhou_eam
from akitsushima_amd.md import maxwell_boltzmann, npt_run, nvt_run, state_fields
from akitsushima_amd.published import ZHOU_2004
from akitsushima_amd.structures import FCC, cubic

from _cases import SW_SI, TERSOFF_XY, sheared_si

pytestmark = pytest.mark.gpu

CUNI = {s: dict(ZHOU_2004[s]) for s in ("Cu", "Ni")}


def _strained(lat, pos, strain):
    T = np.eye(3) + np.asarray(strain, float)
    frac = pos @ np.linalg.inv(lat)
    return lat @ T, frac @ (lat @ T)


STRAIN = [[-0.015, 0.012, 0.0], [0.012, 0.01, 0.0], [0.0, 0.0, 0.005]]


def _case(name):
    if name == "sw":
        lat, pos, t = sheared_si(3, shear=0.0, rattle=0.05, seed=1)
        pot = StillingerWeberPotential(SW_SI, species=["Si"])
        force = lambda l, p, pr: O.sw_flat(l, p, SW_SI, pairs=pr)
        m = np.full(len(pos), 28.086)
    elif name == "tersoff_xy":
        lat, pos, _ = sheared_si(3, shear=0.0, rattle=0.05, seed=4)
        t = (["X"] * 4 + ["Y"] * 4) * (len(pos) // 8)
        pot = TersoffPotential(TERSOFF_XY, species=["X", "Y"])
        force = lambda l, p, pr: O.tersoff_multi(l, p, t, TERSOFF_XY, ["X", "Y"], pairs=pr)
        m = np.full(len(pos), 28.086)
    else:
        lat, pos, _ = cubic(FCC, 3.6, 4, ["Cu"] * 4, rattle=0.05, seed=4)
        t = [str(x) for x in np.random.default_rng(4).choice(["Cu", "Ni"], len(pos))]
        pot = zhou_eam(CUNI, ["Cu", "Ni"])
        force = lambda l, p, pr: O.eam_zhou(l, p, t, CUNI, ["Cu", "Ni"], pairs=pr)
        m = np.array([63.546 if x == "Cu" else 58.693 for x in t])
    if name == "sw":
        t = ["Si"] * len(pos)
    lat, pos = _strained(lat, pos, STRAIN)
    return pot, force, lat, pos, t, m


def _compare(res, r, n_steps):
    for a, b in (("e_pot", "ep"), ("e_kin", "ek"), ("e_cons", "h"), ("volume", "v"),
                 ("pressure_gpa", "p")):
        assert np.abs(res[a] - r[b]).max() < 1e-9 * np.abs(r[b]).max(), a
    assert res["n_rebuilds"] == r["rebuilds"]
    assert np.abs(res["lattice"] - r["lattice"]).max() < 1e-10 * np.abs(r["lattice"]).max()
    d = res["positions"] - r["positions"]
    d -= np.round(d @ np.linalg.inv(res["lattice"])) @ res["lattice"]
    assert np.abs(d).max() < 1e-8
    assert np.abs(res["velocities"] - r["velocities"]).max() < \
        1e-9 * np.abs(r["velocities"]).max()


@pytest.mark.parametrize("mode,baro,name", [
    ("aniso", "nhc", "sw"), ("tri", "nhc", "sw"), ("xy", "nhc", "sw"), ("xy_iso", "nhc", "sw"),
    ("z", "nhc", "sw"), ("iso", "langevin", "sw"), ("aniso", "langevin", "sw"),
    ("tri", "langevin", "sw"), ("tri", "nhc", "tersoff_xy"), ("tri", "langevin", "eam_cuni")])
def test_npt_cell_modes_match_the_oracle(mode, baro, name):
    pot, force, lat, pos, t, m = _case(name)
    vel = maxwell_boltzmann(500.0, m, 4)
    n_steps, skin = 60, 0.5
    res = npt_run(pot, lat, pos, vel, t, 500.0, pressure=0.5, dt=1.0, n_steps=n_steps,
                  tau_nhc=40.0, nhc_chain=3, tau_baro=150.0, skin_width=skin, masses=m,
                  cell_dof=mode, baro_thermostat=baro, seed=17)
    r = O.npt_run(force, lat, pos, vel, m, 1.0, n_steps, pot.r_cut, skin, 500.0, 40.0, 3, 0.5,
                  150.0, cell_dof=mode, baro=baro, seed=17)
    _compare(res, r, n_steps)
    assert np.abs(res["lattices"] - r["lat"]).max() < 1e-10 * np.abs(r["lat"]).max()
    if mode != "iso":
        idx = ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1))
        want = np.stack([r["ptensor"][:, i, j] for i, j in idx], axis=1)
        assert np.abs(res["pressure_tensor_gpa"] - want).max() < 1e-8 * np.abs(want).max()
    L0, L = lat, res["lattice"]
    assert abs(res["volume"][-1] - res["volume"][0]) > 1e-4 * res["volume"][0]
    if mode == "z":
        assert np.array_equal(L[:, :2], L0[:, :2])
    if mode in ("xy", "xy_iso"):
        assert np.array_equal(L[:, 2], L0[:, 2])
    if mode == "tri":
        assert np.abs((L - L0)[0, 1]) > 1e-6
    f = state_fields(res["state"])
    if baro == "langevin":
        assert f["rng_draws"] == 2 * n_steps and f["heat"] == pytest.approx(r["heat"],
                                                                            rel=1e-8, abs=1e-12)
    if mode != "iso":
        assert np.allclose(f["p_g"][:len(r["p_g"])], r["p_g"], rtol=1e-7, atol=1e-9)


def test_frozen_flexible_barostat_is_the_device_nvt():
    pot, _, lat, pos, t, m = _case("sw")
    vel = maxwell_boltzmann(600.0, m, 2)
    b = nvt_run(pot, lat, pos, vel, t, 600.0, dt=1.0, n_steps=200, tau_nhc=50.0)
    for mode in ("aniso", "tri"):
        a = npt_run(pot, lat, pos, vel, t, 600.0, dt=1.0, n_steps=200, tau_nhc=50.0,
                    tau_baro=1e12, cell_dof=mode)
        assert np.abs(a["e_pot"] - b["e_pot"]).max() < 1e-9 * np.abs(b["e_pot"]).max()
        assert np.abs(a["positions"] - b["positions"]).max() < 1e-8
        if mode == "aniso":
            assert np.array_equal(a["lattice"], lat)


@pytest.mark.parametrize("mode,baro", [("tri", "langevin"), ("aniso", "nhc")])
def test_split_flexible_run_equals_the_uninterrupted_one(mode, baro):
    pot, _, lat, pos, t, m = _case("sw")
    vel = maxwell_boltzmann(500.0, m, 4)
    kw = dict(pressure=0.5, dt=1.0, tau_nhc=40.0, tau_baro=150.0, masses=m, cell_dof=mode,
              baro_thermostat=baro, seed=9)
    full = npt_run(pot, lat, pos, vel, t, 500.0, n_steps=120, **kw)
    a = npt_run(pot, lat, pos, vel, t, 500.0, n_steps=60, **kw)
    b = npt_run(pot, a["lattice"], a["positions"], a["velocities"], t, 500.0, n_steps=60,
                state=a["state"], **kw)
    assert np.abs(b["lattice"] - full["lattice"]).max() < 1e-10 * np.abs(lat).max()
    assert np.abs(b["e_cons"] - full["e_cons"][60:]).max() < 1e-9 * np.abs(full["e_cons"]).max()
    assert np.abs(b["velocities"] - full["velocities"]).max() < \
        1e-9 * np.abs(full["velocities"]).max()
    fa, fb = state_fields(b["state"]), state_fields(full["state"])
    assert np.allclose(fa["p_g"], fb["p_g"], rtol=1e-8, atol=1e-10)
    assert fa["rng_draws"] == fb["rng_draws"]
