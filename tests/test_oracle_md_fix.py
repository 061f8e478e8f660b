"""Pins of fix_atoms and velocity_rescale_steps in the oracle integrators
(solvers/molecular_dynamics.py:248-267, 341-361, 1154-1167, 1552-1565,
2907-2920), no GPU.

* an all-False mask is no mask, bit for bit;
* frozen components keep zero velocity and frozen atoms stay where they
  are, under every thermostat;
* with a quarter of the atoms frozen, the chain holds the target mean
  temperature counted with n_dof = 3N - (frozen components) -- the 3N - 3
  count would be 25 % off -- and H_NHC is conserved;
* the Langevin O step with frozen atoms keeps the centre-of-mass velocity
  instead of removing it, and the mean temperature is the target;
* velocity_rescale_steps puts T_inst exactly on the target after each of the
  first k steps, and only those.
"""
import numpy as np
import pytest

import oracle as O

from _cases import SW_SI, si64_recorded_md

R_CUT = 1.80 * 2.0951


def _force(l, p, pr):
    return O.sw_flat(l, p, SW_SI, pairs=pr)


def _start(T=600.0, seed=7):
    lat, pos, _ = si64_recorded_md()
    m = np.full(len(pos), 28.086)
    return lat, pos, O.mb_velocities(T, m, seed), m


def _fix(n=64, frozen=range(16)):
    f = np.zeros((n, 3), bool)
    f[list(frozen)] = True
    return f


RUNS = {
    "nve": {},
    "nhc": {"nhc": (600.0, 50.0, 3)},
    "langevin": {"langevin": (600.0, 50.0, 3)},
    "csvr": {"csvr": (600.0, 50.0, 4)},
}


@pytest.mark.parametrize("kind", list(RUNS))
def test_an_empty_mask_is_no_mask(kind):
    lat, pos, vel, m = _start()
    a = O.md_run(_force, lat, pos, vel, m, 1.0, 60, R_CUT, 0.5, **RUNS[kind])
    b = O.md_run(_force, lat, pos, vel, m, 1.0, 60, R_CUT, 0.5, fix=np.zeros((64, 3), bool),
                 **RUNS[kind])
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])


@pytest.mark.parametrize("kind", list(RUNS))
def test_frozen_atoms_do_not_move(kind):
    lat, pos, vel, m = _start()
    fix = _fix()
    fix[20, 2] = True                      # one frozen component of a free atom
    *_, p, v, _, _ = O.md_run(_force, lat, pos, vel, m, 1.0, 100, R_CUT, 0.5, fix=fix,
                              **RUNS[kind])
    assert np.all(v[fix] == 0.0)
    d = p - pos
    d -= np.round(d @ np.linalg.inv(lat)) @ lat         # the run wraps into the cell
    assert np.abs(d[:16]).max() < 1e-12
    assert abs(d[20, 2]) < 1e-12 and abs(d[20, 0]) > 1e-4
    assert np.abs(d[16:]).max() > 1e-3


def test_frozen_dof_leave_the_chain_temperature_and_h_nhc():
    lat, pos, vel, m = _start()
    fix = _fix()
    n_dof = 3 * 64 - 48
    ep, ek, ec, *_ = O.md_run(_force, lat, pos, vel, m, 1.0, 4000, R_CUT, 0.5,
                              nhc=(900.0, 50.0, 4), fix=fix)
    T = 2 * ek / (n_dof * O.K_B)
    assert abs(T[1000:].mean() - 900.0) < 0.06 * 900.0
    assert abs(2 * ek[1000:].mean() / (189 * O.K_B) - 900.0) > 0.15 * 900.0
    et = ep + ek
    assert ec.max() - ec.min() < 2e-3 * (et.max() - et.min())


def test_langevin_with_frozen_atoms_keeps_the_com_velocity():
    lat, pos, vel, m = _start()
    fix = _fix()
    *_, v_free, _, _ = O.md_run(_force, lat, pos, vel, m, 1.0, 50, R_CUT, 0.5,
                                langevin=(600.0, 50.0, 3))
    assert np.abs((m[:, None] * v_free).sum(axis=0)).max() < 1e-12
    ep, ek, _, _, v, _, _ = O.md_run(_force, lat, pos, vel, m, 1.0, 3000, R_CUT, 0.5,
                                     langevin=(600.0, 50.0, 3), fix=fix)
    assert np.abs((m[:, None] * v).sum(axis=0)).max() > 1e-3
    T = 2 * ek / ((3 * 64 - 48) * O.K_B)
    assert abs(T[500:].mean() - 600.0) < 0.06 * 600.0


@pytest.mark.parametrize("ens", ["nvt", "npt"])
def test_velocity_rescale_steps(ens):
    lat, pos, vel, m = _start(T=300.0)
    k = 20
    if ens == "nvt":
        ep, ek, *_ = O.md_run(_force, lat, pos, vel, m, 1.0, 60, R_CUT, 0.5,
                              nhc=(900.0, 100.0, 3), rescale_steps=k)
    else:
        r = O.npt_run(_force, lat, pos, vel, m, 1.0, 60, R_CUT, 0.5, 900.0, 100.0, 3, 0.0,
                      300.0, rescale_steps=k)
        ek = r["ek"]
    T = 2 * ek / (189 * O.K_B)
    assert np.abs(T[1:k + 1] - 900.0).max() < 1e-9
    assert np.abs(T[k + 2:] - 900.0).min() > 1e-3
    assert abs(T[0] - 300.0) < 1e-9
