"""NEMD with two heat baths on the device (ak_nemd_run, NEMDSimulation;
solvers/nemd.py) against the oracle (tests/test_oracle_nemd.py pins it):
energies, bath heats and bath kinetic energies of every step, the final
state, and the per-bin samples (heat current, kinetic energy, atom count)
against the oracle's per-atom heat currents; a run split in two continues
the random stream; NEMDSimulation moves heat from the hot to the cold bath,
keeps E - Q, writes the thermo log, and a run killed after a checkpoint and
re-run ends where the uninterrupted run does, bit for bit (the reference's
test_restart_continues_bit_for_bit, tests/test_nemd.py:124-150)."""
import numpy as np
import pytest

import oracle as O
import oracle.site_terms as ST
from akitsushima_amd import NEMDSimulation, StillingerWeberPotential
from akitsushima_amd import nemd as nemd_mod
from akitsushima_amd.md import nemd_run

from _cases import SW_SI
from test_oracle_nemd import _force, _slab

pytestmark = pytest.mark.gpu


def test_device_nemd_matches_the_oracle():
    lat, pos, vel, m, bath = _slab()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    t = ["Si"] * len(pos)
    nb, se, ph, skin = 8, 10, 1, 0.5
    res = nemd_run(pot, lat, pos, vel, t, bath, 900.0, 300.0, 20.0, dt=1.0, n_steps=60, seed=4,
                   skin_width=skin, masses=m, sample_every=se, sample_phase=ph, axis=2,
                   n_bins=nb)
    V = abs(np.linalg.det(lat))
    recs = O.sw_flat_as_records(SW_SI)

    def sample(p, v):
        Ji = ST.sw(lat, p, t, recs, ["Si"]).atom_heat_current(v, m)
        f = (p @ np.linalg.inv(lat))[:, 2]
        f -= np.floor(f)
        b = np.minimum((f * nb).astype(int), nb - 1)
        out = np.zeros((nb, 5))
        np.add.at(out[:, :3], b, Ji / V)
        np.add.at(out[:, 3], b, 0.5 * m * np.sum(v * v, axis=1) / O.EV_ANG_TO_AMU_ANG_FS2)
        np.add.at(out[:, 4], b, 1.0)
        return out

    r = O.nemd_run(_force, lat, pos, vel, m, 1.0, 60, pot.r_cut, skin, bath, 900.0, 300.0, 20.0,
                   seed=4, sample=sample, every=se, phase=ph)
    for a, b in (("e_pot", "ep"), ("e_kin", "ek"), ("e_cons", "econs")):
        assert np.abs(res[a] - r[b]).max() < 1e-10 * np.abs(r[b]).max(), a
    assert np.abs(res["bath"] - r["rows"]).max() < 1e-10 * np.abs(r["rows"]).max()
    assert res["n_rebuilds"] == r["rebuilds"]
    d = res["positions"] - r["positions"]
    d -= np.round(d @ np.linalg.inv(lat)) @ lat
    assert np.abs(d).max() < 1e-9
    assert np.abs(res["velocities"] - r["velocities"]).max() < \
        1e-9 * np.abs(r["velocities"]).max()
    want = np.array(r["samples"])
    got = res["bins"]
    assert got.shape == want.shape == (6, nb, 5)
    assert np.array_equal(got[:, :, 4], want[:, :, 4])
    assert np.abs(got[:, :, 3] - want[:, :, 3]).max() < 1e-10 * np.abs(want[:, :, 3]).max()
    assert np.abs(got[:, :, :3] - want[:, :, :3]).max() < 1e-8 * np.abs(want[:, :, :3]).max()
    assert r["rows"][1:, 0].sum() > 0 > r["rows"][1:, 1].sum()


def test_split_nemd_run_continues_the_random_stream():
    lat, pos, vel, m, bath = _slab()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    t = ["Si"] * len(pos)
    kw = dict(dt=1.0, seed=11, masses=m)
    full = nemd_run(pot, lat, pos, vel, t, bath, 900.0, 300.0, 20.0, n_steps=80, **kw)
    a = nemd_run(pot, lat, pos, vel, t, bath, 900.0, 300.0, 20.0, n_steps=40, **kw)
    b = nemd_run(pot, lat, a["positions"], a["velocities"], t, bath, 900.0, 300.0, 20.0,
                 n_steps=40, state=a["state"], **kw)
    assert np.abs(b["velocities"] - full["velocities"]).max() < \
        1e-9 * np.abs(full["velocities"]).max()
    assert np.abs(b["e_cons"] - full["e_cons"][40:]).max() < 1e-9 * np.abs(full["e_cons"]).max()
    assert b["state"][44] == full["state"][44] == 80


def _sim():
    lat, pos, _, m, _ = _slab()
    return NEMDSimulation(lat, pos, ["Si"] * len(pos), StillingerWeberPotential(SW_SI,
                                                                                species=["Si"]),
                          masses=m, T_hot=900.0, T_cold=300.0, axis="c", hot=(0.0, 0.15),
                          cold=(0.5, 0.65), damping=20.0, timestep=1.0, n_bins=16, seed=7)


def test_nemd_simulation_moves_heat_and_keeps_e_minus_q(tmp_path):
    sim = _sim()
    st = sim.run(50, 400, sample_every=5, thermo_path=str(tmp_path / "thermo.log"),
                 verbose=False)
    assert st["Q_hot_prod"] > 0 > st["Q_cold_prod"]
    Q = abs(st["Q_hot"]) + abs(st["Q_cold"])
    assert abs(st["E_minus_Q_drift_eV"]) < 5e-3 * Q
    assert st["n_samples"] == 80 and st["n_prod"] == 400
    res = sim.results()
    T = res["T_K"]
    assert T.shape == (16,) and T[0] > T[8]
    assert np.isclose(res["atoms_per_bin"].sum(), len(sim.types))
    assert len(res["segments"]) == 2 and np.all(np.isfinite(res["local_flux_eV_A2_fs"]))
    assert all(len(s["bins"]) >= 2 for s in res["segments"]) and np.isfinite(res["kappa_W_mK"])
    rows = [ln.split() for ln in open(tmp_path / "thermo.log") if not ln.startswith("#")]
    assert [int(r[0]) for r in rows] == list(range(0, 451, 50))
    assert np.isclose(float(rows[-1][3]), st["Q_hot"], rtol=1e-7)


def test_restart_continues_bit_for_bit(tmp_path, monkeypatch):
    ref = _sim()
    ref.run(40, 160, sample_every=5, checkpoint=str(tmp_path / "a.npz"), checkpoint_every=50,
            verbose=False)
    ck = str(tmp_path / "b.npz")
    calls = {"n": 0}
    real = nemd_mod.nemd_run

    def dying(*a, **k):
        calls["n"] += 1
        if calls["n"] == 4:
            raise RuntimeError("killed")
        return real(*a, **k)

    monkeypatch.setattr(nemd_mod, "nemd_run", dying)
    with pytest.raises(RuntimeError):
        _sim().run(40, 160, sample_every=5, checkpoint=ck, checkpoint_every=50, verbose=False)
    monkeypatch.setattr(nemd_mod, "nemd_run", real)
    sim = _sim()
    st = sim.run(40, 160, sample_every=5, checkpoint=ck, checkpoint_every=50, verbose=False)
    for k in ("pos", "vel", "ke_bin", "n_bin", "j_bin", "q_block"):
        assert np.array_equal(st[k], ref.state[k]), k
    for k in ("Q_hot", "Q_cold", "E0", "n_samples", "E_minus_Q_drift_eV"):
        assert st[k] == ref.state[k], k


def test_zero_step_run_with_sampling_records_the_start_only():
    # the two set-up steps record at slots 0 and 1 of the history buffers;
    # with n_steps = 0 the run still returns just the start state
    lat, pos, vel, m, bath = _slab()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    t = ["Si"] * len(pos)
    print("nemd_run n_steps=0 sample_every=5", flush=True)
    res = nemd_run(pot, lat, pos, vel, t, bath, 900.0, 300.0, 20.0, dt=1.0, n_steps=0, seed=3,
                   masses=m, sample_every=5, axis=2, n_bins=8)
    assert res["e_pot"].shape == (1,) and res["bath"].shape == (1, 4)
    assert np.array_equal(res["velocities"], vel)
    assert len(res["bins"]) == 0


def test_finished_checkpoint_returns_without_a_device_call(tmp_path, monkeypatch):
    ck = str(tmp_path / "done.npz")
    first = _sim().run(20, 40, sample_every=5, checkpoint=ck, checkpoint_every=30,
                       verbose=False)

    def forbidden(*a, **k):
        raise AssertionError("a finished run called the device")

    monkeypatch.setattr(nemd_mod, "nemd_run", forbidden)
    again = _sim().run(20, 40, sample_every=5, checkpoint=ck, checkpoint_every=30,
                       verbose=False)
    for k in ("pos", "vel", "ke_bin", "n_bin", "j_bin", "q_block"):
        assert np.array_equal(again[k], first[k]), k
    for k in ("Q_hot", "Q_cold", "E0", "n_samples", "n_prod", "E_minus_Q_drift_eV"):
        assert again[k] == first[k], k
