"""Host logic of NEMDSimulation (akitsushima_amd/nemd.py; solvers/nemd.py):
bath slabs (wrap-around, the 1e-9 shift of a layer at 0, overlap and empty
refusals), the cross-section geometry of a triclinic cell, and results() on
hand-built accumulators: linear temperature segments give their gradient,
kappa = J / |grad T| and the sign checks; a temperature jump at an interface
gives delta_T and R = delta_T / J.  No device call is made."""
import numpy as np
import pytest

from akitsushima_amd import NEMDSimulation, StillingerWeberPotential
from akitsushima_amd.md import K_B
from akitsushima_amd.nemd import EV_ANG2_FS_TO_W_M2, EV_ANG_K_FS_TO_W_MK

from _cases import SW_SI

POT = StillingerWeberPotential(SW_SI, species=["Si"])
LAT = np.array([[6.0, 0.0, 0.0], [1.0, 7.0, 0.0], [0.5, 0.3, 40.0]])


def _sim(zs, **kw):
    frac = np.stack([np.full(len(zs), 0.3), np.full(len(zs), 0.6), zs], axis=1)
    args = dict(hot=(0.0, 0.1), cold=(0.5, 0.6))
    args.update(kw)
    return NEMDSimulation(LAT, frac @ LAT, ["Si"] * len(zs), POT, n_bins=20, **args)


def test_bath_slabs_and_cross_section():
    sim = _sim(np.array([-1e-12, 0.02, 0.5, 0.52, 0.9, 0.97, 0.3]), hot=(0.95, 0.05),
               cold=(0.5, 0.55))
    assert sim.hot.tolist() == [0, 1, 5] and sim.cold.tolist() == [2, 3]
    assert sim.bath.tolist() == [1, 1, 2, 2, 0, 1, 0]
    cross = np.cross(LAT[0], LAT[1])
    assert np.isclose(sim.area, np.linalg.norm(cross))
    assert np.isclose(sim.length * sim.area, abs(np.linalg.det(LAT)))
    assert np.allclose(sim.normal, cross / np.linalg.norm(cross)) and sim.normal @ LAT[2] > 0
    with pytest.raises(ValueError, match="overlap"):
        _sim(np.array([0.02, 0.5]), hot=(0.0, 0.6), cold=(0.4, 0.55))
    with pytest.raises(ValueError, match="no atoms"):
        _sim(np.array([0.02, 0.3]))


def _state(T, sim, n_prod=1000, q=2.0):
    nb = sim.n_bins
    n_bin = np.full((4, nb), 10.0)
    st = {"ke_bin": 1.5 * K_B * T[None, :] * n_bin, "n_bin": n_bin,
          "j_bin": np.ones((4, nb)), "n_jsamp": np.full(4, 5.0), "n_samples": 20,
          "Q_hot_prod": q, "Q_cold_prod": -q, "n_prod": n_prod,
          "E_minus_Q_drift_eV": 0.0}
    sim.state = st
    return st


def test_results_from_linear_segments_and_an_interface_jump():
    sim = _sim(np.array([0.0, 0.05, 0.099, 0.5, 0.55, 0.599, 0.2, 0.3, 0.7, 0.8]))
    L = sim.length
    z = (np.arange(20) + 0.5) / 20
    T = np.where(z < 0.5, 800.0 - 0.5 * z * L, 400.0 + 0.5 * z * L)
    _state(T, sim)
    res = sim.results()
    J = 2.0 / 1000.0 / (2.0 * sim.area)
    seg = {s["segment"]: s for s in res["segments"]}
    a, b = seg["hot->cold (+axis)"], seg["cold->hot (+axis)"]
    assert a["bins"] == [4, 5, 6, 7] and b["bins"] == [14, 15, 16, 17]
    assert np.isclose(a["gradient_K_per_A"], -0.5) and np.isclose(b["gradient_K_per_A"], 0.5)
    assert a["gradient_sign_ok"] and b["gradient_sign_ok"]
    kap = J / 0.5 * EV_ANG_K_FS_TO_W_MK
    assert np.isclose(a["kappa_W_mK"], kap) and np.isclose(res["kappa_W_mK"], kap)
    assert a["kappa_err_W_mK"] == pytest.approx(0.0, abs=1e-9 * kap)
    assert np.isclose(res["flux_W_m2"], J * EV_ANG2_FS_TO_W_M2)
    assert res["bath_balance"] == 0.0
    assert np.allclose(res["T_K"], T) and np.allclose(res["local_flux_eV_A2_fs"], 4.0 / 20.0)
    # a jump of 25 K at z = 0.3 inside a linear profile
    _state(800.0 - 0.5 * z * L - 25.0 * (z > 0.3), sim)
    k = sim.results(interfaces=[0.3])["kapitza"][0]
    assert np.isclose(k["delta_T_K"], 25.0)
    assert np.isclose(k["R_m2K_per_W"], 25.0 / (J * EV_ANG2_FS_TO_W_M2))


def test_bath_interval_rule():
    from akitsushima_amd.nemd import _bath_members

    f = np.array([0.0, 0.3, 0.55, 0.999999, -1e-12, 0.7])
    # [lo, hi), a wrapping interval, and lo == hi (a full period, as (0, 1))
    assert _bath_members(f, (0.3, 0.55)).tolist() == [1]
    assert _bath_members(f, (0.6, 0.31)).tolist() == [0, 1, 3, 4, 5]
    assert _bath_members(f, (0.0, 1.0)).tolist() == list(range(6))
    # a layer a rounding error below 0 belongs to a slab starting at 0
    assert 4 in _bath_members(f, (0.0, 0.05)).tolist()


def test_production_sums_round_trip_through_a_checkpoint(tmp_path):
    from akitsushima_amd.nemd import _Production

    p = _Production(6)
    for k in range(-3, 9):
        p.add_step(0.1 * k, -0.05 * k, k, 8, 1.0)
    p.add_sample(2, np.arange(6.0), np.ones(6), np.full(6, 0.5))
    p.add_sample(3, np.arange(6.0), np.ones(6))
    p.step, p.E0, p.E_last = 11, -5.0, -4.5
    assert p.n_prod == 8 and p.blocks[:, 2].tolist() == [2.0, 2.0, 2.0, 2.0]
    assert np.isclose(p.blocks[0, 0], 0.1 + 0.2) and p.n_jsamp.tolist() == [0, 0, 1, 0]
    np.savez(tmp_path / "p.npz", **p.arrays())
    q = _Production.from_arrays(np.load(tmp_path / "p.npz"))
    a, b = p.as_dict(), q.as_dict()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k
    assert _Production.from_arrays(np.load(tmp_path / "p.npz")).E0 == -5.0
