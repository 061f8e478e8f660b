"""The simulation classes on the device (akitsushima_amd.simulations).

* The reference's restart test (tests/test_nhc_chain.py:136-170) through the
  class API: a 0.3.0a7 checkpoint at step 150 (written here in that layout
  from the recorded state) continued to step 300 by NVTSimulation and by
  NPTSimulation lands on the recorded step-300 state to 1e-9;
* thermo.log rows, the XDATCAR frames and the run summary equal the device
  run's own records (a fresh NVT, NVE and NPT run);
* a run interrupted after a checkpoint and started again continues exactly
  like the uninterrupted run (NVT, Langevin, flexible-cell NPT).
"""
import json
import os

import numpy as np
import pytest

from akitsushima_amd import (LangevinSimulation, NPTSimulation, NVESimulation, NVTSimulation,
                             StillingerWeberPotential, make_force_fn)
from akitsushima_amd.md import md_run
from akitsushima_amd.potentials import PotentialEngine

from _cases import SW_SI, si64_recorded_md

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REC = np.load(os.path.join(HERE, "golden", "nhc_legacy_si64.npz"), allow_pickle=False)


def _fn():
    return make_force_fn(PotentialEngine().add(StillingerWeberPotential(SW_SI,
                                                                        species=["Si"])))


def _kw(out, **over):
    kw = dict(force_fn=_fn(), temperature=600.0, timestep=1.0, thermo_every=10, traj_every=50,
              seed=7, verbose=False, save_plots=False, output_dir=str(out))
    kw.update(over)
    return kw


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def _mic_rel(pos, want, lat):
    d = pos - want
    d -= np.round(d @ np.linalg.inv(lat)) @ lat
    return float(np.max(np.abs(d)) / np.max(np.abs(want)))


def _write_legacy(path, name):
    """The step-150 checkpoint of the recorded run in the 0.3.0a7 layout."""
    s = REC[f"{name}_state150"] if name == "npt" else REC["nvt_chain150"]
    sc = dict(zip(("xi1", "xi2", "vxi1", "vxi2"), map(float, s[:4])))
    settings = {"ensemble": name, "n_atoms": 64, "atom_types": ["Si"] * 64, "timestep": 1.0,
                "equil_steps": 0, "thermo_every": 10, "traj_every": 50,
                "velocity_rescale_steps": 0, "temperature": 600.0, "tau_nhc": 100.0}
    if name == "npt":
        sc.update(zip(("p_eps", "xi1_b", "xi2_b", "vxi1_b", "vxi2_b"), map(float, s[4:9])))
        settings.update(tau_baro=500.0, pressure_gpa=0.0)
    meta = {"step": 150, "settings": settings, "scalars": sc, "extra": {}, "frame": 4}
    np.savez(path, meta=np.array(json.dumps(meta)), positions=REC[f"{name}_pos150"],
             velocities=REC[f"{name}_vel150"], lattice=REC[f"{name}_lat150"])


@pytest.mark.parametrize("name,cls,extra", [
    ("nvt", NVTSimulation, {}),
    ("npt", NPTSimulation, dict(pressure=0.0, tau_baro=500.0))])
def test_legacy_checkpoint_continues_like_the_uninterrupted_run(tmp_path, name, cls, extra):
    ck = str(tmp_path / "md_checkpoint.npz")
    _write_legacy(ck, name)
    lat, pos, t = si64_recorded_md()
    sim = cls(lat, pos, t, n_steps=300, nhc_chain=2, **_kw(tmp_path), **extra)
    sim.run(checkpoint=ck)
    assert _mic_rel(sim.positions, REC[f"{name}_pos300"], REC[f"{name}_lat300"]) < 1e-9
    assert _rel(sim.velocities, REC[f"{name}_vel300"]) < 1e-9
    assert _rel(sim.lattice, REC[f"{name}_lat300"]) < 1e-12
    log = open(tmp_path / "thermo.log").read()
    assert "ADR-0048" in log and log.strip().splitlines()[-1].split()[0] == "300"


def _rows(path):
    return [line.split() for line in open(path).read().splitlines()[1:]
            if line.strip() and not line.startswith("#")]


def test_fresh_nvt_run_writes_the_device_records(tmp_path):
    lat, pos, t = si64_recorded_md()
    sim = NVTSimulation(lat, pos, t, n_steps=100, equil_steps=20, tau_nhc=50.0,
                        **_kw(tmp_path, traj_every=20))
    v0 = sim.velocities.copy()
    out = sim.run()
    r = md_run(sim.potential, lat, pos, v0, t, dt=1.0, n_steps=120, skin_width=0.5,
               thermostat=("nhc", 600.0, 50.0, 3))
    rows = _rows(tmp_path / "thermo.log")
    assert [int(x[0]) for x in rows] == list(range(0, 121, 10))
    for row in rows:
        c = int(row[0])
        assert abs(float(row[2]) - r["e_pot"][c]) < 2e-6
        assert abs(float(row[6]) - r["e_cons"][c]) < 2e-6
    T = 2 * r["e_kin"] / (189 * 8.617333262e-5)
    assert abs(out["T_mean"] - T[21:].mean()) < 1e-6 * T[21:].mean()
    assert out["thermo_log"][-1]["step"] == 120
    xd = open(tmp_path / "XDATCAR_nvt").read().splitlines()
    assert sum(1 for line in xd if line.startswith("Direct configuration")) == 6
    last = np.array([[float(v) for v in line.split()] for line in xd[-64:]])
    frac = r["positions"] @ np.linalg.inv(lat)
    assert np.abs(last - frac).max() < 2e-8
    assert _rel(out["positions"], r["positions"]) < 1e-12
    assert os.path.isfile(tmp_path / "POSCAR_final_MD")


def test_nve_and_npt_runs_write_their_columns(tmp_path):
    lat, pos, t = si64_recorded_md()
    nve = NVESimulation(lat, pos, t, n_steps=60, **_kw(tmp_path / "nve"))
    out = nve.run()
    rows = _rows(tmp_path / "nve" / "thermo.log")
    assert len(rows) == 7 and abs(float(rows[0][7])) == 0.0
    assert abs(out["e_tot_drift"]) < 2e-3 and "thermo_log" not in out
    npt = NPTSimulation(lat, pos, t, n_steps=60, pressure=1.0, tau_baro=200.0, tau_nhc=50.0,
                        cell_dof="aniso", **_kw(tmp_path / "npt", traj_every=20))
    out = npt.run()
    rows = _rows(tmp_path / "npt" / "thermo.log")
    hdr = open(tmp_path / "npt" / "thermo.log").readline().split()
    assert hdr[-6:] == ["Pxx", "Pyy", "Pzz", "a", "b", "c"] and len(rows[-1]) == len(hdr)
    assert len(rows) == 7 and float(rows[-1][6]) == pytest.approx(out["V_mean"], rel=0.02)
    assert float(rows[-1][-3]) == pytest.approx(np.linalg.norm(out["lattice"][0]), abs=1e-4)
    assert out["cell_dof"] == "aniso" and out["P_tensor_mean_gpa"].shape == (3, 3)
    assert np.allclose(out["abc_mean"], np.diag(out["lattice_mean"]), rtol=1e-12)
    xd = open(tmp_path / "npt" / "XDATCAR_npt").read()
    assert xd.count("Direct configuration") == 4 and xd.count("akitsushima.py NPT-MD") == 4
    assert not np.array_equal(out["lattice"], lat)
    assert set(out) >= {"P_mean_gpa", "V_mean", "H_npt_drift", "T_mean"}


@pytest.mark.parametrize("cls,extra", [
    (NVTSimulation, dict(tau_nhc=50.0, velocity_rescale_steps=30)),
    (LangevinSimulation, dict(damping=50.0)),
    (NPTSimulation, dict(pressure=0.5, tau_baro=200.0, cell_dof="tri",
                         baro_thermostat="langevin"))])
def test_interrupted_run_resumes_like_the_uninterrupted_one(tmp_path, cls, extra):
    lat, pos, t = si64_recorded_md()
    full = cls(lat, pos, t, n_steps=120, **_kw(tmp_path / "full"), **extra)
    ref = full.run()
    ck = str(tmp_path / "part" / "ck.npz")
    os.makedirs(tmp_path / "part")
    a = cls(lat, pos, t, n_steps=50, **_kw(tmp_path / "part"), **extra)
    a.run(checkpoint=ck, checkpoint_every=25)
    b = cls(lat, pos, t, n_steps=120, **_kw(tmp_path / "part"), **extra)
    out = b.run(checkpoint=ck, checkpoint_every=25)
    assert _rel(out["velocities"], ref["velocities"]) < 1e-9
    assert _rel(out["lattice"], ref["lattice"]) < 1e-12
    full_rows = _rows(tmp_path / "full" / "thermo.log")
    part_rows = _rows(tmp_path / "part" / "thermo.log")
    assert [r[0] for r in part_rows] == [r[0] for r in full_rows]
    # every column but NL_builds: a resumed run rebuilds its list at the
    # restart, so its rebuild count may differ from the uninterrupted one
    hdr = open(tmp_path / "full" / "thermo.log").readline().split()
    cols = [i for i, h in enumerate(hdr) if i > 0 and h != "NL_builds"]
    for x, y in zip(part_rows, full_rows):
        assert np.allclose([float(x[i]) for i in cols], [float(y[i]) for i in cols], rtol=0,
                           atol=2e-6)
    assert out["T_mean"] == pytest.approx(ref["T_mean"], rel=1e-9)
