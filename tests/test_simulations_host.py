"""Host side of the simulation classes (akitsushima_amd.simulations), no
GPU: the reference's argument checks (tests/test_nhc_chain.py:90-104,
152-163), the Maxwell-Boltzmann start with frozen components, reading a
checkpoint in the reference's npz layout (the 0.3.0a7 scalar chain and the
current arrays), and the POSCAR / XDATCAR writers."""
import json

import numpy as np
import pytest

from akitsushima_amd import (CSVRSimulation, LangevinSimulation, NPTSimulation, NVESimulation,
                             NVTSimulation, StillingerWeberPotential, make_force_fn)
from akitsushima_amd.md import maxwell_boltzmann, state_fields
from akitsushima_amd.potentials import PotentialEngine
from akitsushima_amd.simulations import write_poscar

from _cases import SW_SI, si64_recorded_md


def _fn():
    return make_force_fn(PotentialEngine().add(StillingerWeberPotential(SW_SI,
                                                                        species=["Si"])))


def _kw(tmp_path, **over):
    kw = dict(force_fn=_fn(), temperature=600.0, timestep=1.0, thermo_every=10, traj_every=50,
              seed=7, verbose=False, save_plots=False, output_dir=str(tmp_path))
    kw.update(over)
    return kw


@pytest.mark.parametrize("bad", [1, 0, 2.5, True, -3])
def test_bad_chain_length_is_refused(tmp_path, bad):
    lat, pos, t = si64_recorded_md()
    for cls in (NVTSimulation, NPTSimulation):
        with pytest.raises(ValueError, match="nhc_chain"):
            cls(lat, pos, t, n_steps=1, nhc_chain=bad, **_kw(tmp_path))


def test_default_chain_and_argument_checks(tmp_path):
    lat, pos, t = si64_recorded_md()
    assert NVTSimulation(lat, pos, t, n_steps=1, **_kw(tmp_path)).nhc_chain == 3
    assert NPTSimulation(lat, pos, t, n_steps=1, **_kw(tmp_path)).nhc_chain == 3
    with pytest.raises(ValueError, match="velocity_rescale_steps"):
        LangevinSimulation(lat, pos, t, n_steps=1, velocity_rescale_steps=5, **_kw(tmp_path))
    with pytest.raises(ValueError, match="damping"):
        CSVRSimulation(lat, pos, t, n_steps=1, damping=0.0, **_kw(tmp_path))
    with pytest.raises(ValueError, match="cell_dof"):
        NPTSimulation(lat, pos, t, n_steps=1, cell_dof="xyz", **_kw(tmp_path))
    with pytest.raises(ValueError, match="baro_thermostat"):
        NPTSimulation(lat, pos, t, n_steps=1, baro_thermostat="berendsen", **_kw(tmp_path))
    with pytest.raises(ValueError, match="fix_atoms"):
        NVTSimulation(lat, pos, t, n_steps=1, fix_atoms=np.zeros((3, 3), bool), **_kw(tmp_path))
    with pytest.raises(ValueError, match="timestep"):
        NVESimulation(lat, pos, t, n_steps=1, **_kw(tmp_path, timestep=0.0))
    with pytest.raises(ValueError, match="Buckingham"):
        NVTSimulation(lat, pos, t, n_steps=1, buck_params={("Si", "Si"): (1, 1, 1)},
                      **_kw(tmp_path))


def test_maxwell_boltzmann_start_with_frozen_components(tmp_path):
    lat, pos, t = si64_recorded_md()
    fix = np.zeros((64, 3), bool)
    fix[:4] = True
    sim = NVTSimulation(lat, pos, t, n_steps=1, fix_atoms=fix, **_kw(tmp_path))
    v = maxwell_boltzmann(600.0, np.full(64, 28.086), 7)
    assert np.all(sim.velocities[:4] == 0.0)
    assert np.array_equal(sim.velocities[4:], v[4:])
    assert sim.n_dof == 3 * 64 - 12


def _legacy_checkpoint(path, ensemble, chain=2):
    lat, pos, _ = si64_recorded_md()
    rng = np.random.default_rng(0)
    sc = {"xi1": 0.1, "xi2": -0.2, "vxi1": 1e-3, "vxi2": -2e-3}
    settings = {"ensemble": ensemble, "n_atoms": 64, "atom_types": ["Si"] * 64,
                "timestep": 1.0, "temperature": 600.0, "tau_nhc": 100.0}
    if ensemble == "npt":
        sc.update(p_eps=258.6, xi1_b=-0.03, xi2_b=-0.04, vxi1_b=-3e-4, vxi2_b=-6e-4)
    if chain != 2:
        settings["nhc_chain"] = chain
    meta = {"step": 150, "settings": settings, "scalars": sc, "extra": {}, "frame": 4}
    np.savez(path, meta=np.array(json.dumps(meta)), positions=pos, lattice=lat,
             velocities=rng.normal(size=(64, 3)) * 1e-3)
    return sc


@pytest.mark.parametrize("ens,cls", [("nvt", NVTSimulation), ("npt", NPTSimulation)])
def test_a_legacy_checkpoint_is_read_into_the_device_state(tmp_path, ens, cls):
    ck = str(tmp_path / "md_checkpoint.npz")
    sc = _legacy_checkpoint(ck, ens)
    lat, pos, t = si64_recorded_md()
    sim = cls(lat, pos, t, n_steps=300, nhc_chain=2, **_kw(tmp_path))
    step, *_ = sim._ck_load(ck)
    f = state_fields(sim._state, 2)
    assert step == 150 and f["steps"] == 150
    assert np.array_equal(f["xi"], [sc["xi1"], sc["xi2"]])
    assert np.array_equal(f["vxi"], [sc["vxi1"], sc["vxi2"]])
    if ens == "npt":
        assert f["p_eps"] == sc["p_eps"]
        assert np.array_equal(f["xi_b"], [sc["xi1_b"], sc["xi2_b"]])
        assert np.array_equal(f["vxi_b"], [sc["vxi1_b"], sc["vxi2_b"]])


def test_a_checkpoint_of_another_chain_length_is_refused(tmp_path):
    ck = str(tmp_path / "md_checkpoint.npz")
    _legacy_checkpoint(ck, "nvt")
    lat, pos, t = si64_recorded_md()
    sim = NVTSimulation(lat, pos, t, n_steps=300, nhc_chain=3, **_kw(tmp_path))
    with pytest.raises(ValueError, match="nhc_chain"):
        sim._ck_load(ck)


def test_poscar_groups_species_and_writes_direct_coordinates(tmp_path):
    lat = np.diag([4.0, 5.0, 6.0])
    pos = np.array([[1.0, 1.0, 1.0], [2.0, 2.5, 3.0], [0.0, 0.0, 0.0]])
    write_poscar(str(tmp_path / "P"), lat, pos, ["Ge", "Sn", "Ge"], title="t")
    lines = open(tmp_path / "P").read().splitlines()
    assert lines[0] == "t" and lines[1] == "1.0"
    assert lines[5].split() == ["Ge", "Sn"] and lines[6].split() == ["2", "1"]
    assert lines[7] == "Direct"
    rows = np.array([[float(x) for x in line.split()] for line in lines[8:]])
    assert np.allclose(rows, [[0.25, 0.2, 1 / 6], [0, 0, 0], [0.5, 0.5, 0.5]])
