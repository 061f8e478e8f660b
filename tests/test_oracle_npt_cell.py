"""Pins of the oracle's flexible-cell NPT (cell_dof aniso, tri, xy, xy_iso,
z; solvers/molecular_dynamics.py:2113-2138, 2566-2692) and of the Langevin
barostat (:2596-2617), no GPU.

* H_NPT is conserved up to the integrator's O(dt^2) error in every mode and
  with either barostat thermostat: its range falls about fourfold when dt
  halves and stays far below the range of E_tot.  A dropped term of the
  MTTK equations (the Tr(E_k) 2 E_kin / n_dof barostat force, the Tr(G)/n_dof
  velocity coupling, the P V or p^2 / 2W terms) breaks it;
* with a frozen barostat (tau_baro -> infinity) a diagonal mode is the NVT
  chain run bit for bit, the shear mode to rounding;
* cell rows move only along the generators: z leaves x and y untouched
  bit for bit, xy leaves z, xy_iso moves x and y by one factor;
* a cell strained along one axis relaxes to the target pressure in every
  diagonal component (aniso) and, sheared, to zero shear stress (tri);
* the single generator E = I is the isotropic barostat up to the one
  difference the reference has between them (the kinetic pressure taken
  before or after the particle chain's half-step).
"""
import numpy as np
import pytest

import oracle as O

from _cases import SW_SI, si64_recorded_md

R_CUT = 1.80 * 2.0951


def _force(l, p, pr):
    return O.sw_flat(l, p, SW_SI, pairs=pr)


def _start(strain=None, seed=7):
    lat, pos, _ = si64_recorded_md()
    if strain is not None:
        T = np.eye(3) + np.asarray(strain, float)
        frac = pos @ np.linalg.inv(lat)
        lat = lat @ T
        pos = frac @ lat
    m = np.full(len(pos), 28.086)
    return lat, pos, O.mb_velocities(600.0, m, seed), m


def _run(mode, dt, n, baro="nhc", tau_b=300.0, p=0.5, strain=None, **kw):
    lat, pos, vel, m = _start(strain)
    return O.npt_run(_force, lat, pos, vel, m, dt, n, R_CUT, 0.5, 600.0, 50.0, 3, p, tau_b,
                     cell_dof=mode, baro=baro, seed=3, **kw)


@pytest.mark.parametrize("mode,baro", [("aniso", "nhc"), ("tri", "nhc"), ("xy", "nhc"),
                                       ("xy_iso", "nhc"), ("z", "nhc"), ("iso", "langevin"),
                                       ("tri", "langevin")])
def test_h_npt_is_conserved_to_second_order(mode, baro):
    strain = [[-0.015, 0.01, 0.0], [0.01, 0.01, 0.0], [0.0, 0.0, 0.0]]
    ranges = []
    for dt in (1.0, 0.5):
        r = _run(mode, dt, int(200 / dt), baro=baro, strain=strain)
        h, et = r["h"], r["ep"] + r["ek"]
        ranges.append(h.max() - h.min())
        assert ranges[-1] < 2e-2 * (et.max() - et.min())
        assert abs(r["v"][-1] - r["v"][0]) > 1e-4 * r["v"][0]
    assert 2.5 < ranges[0] / ranges[1] < 5.5
    if baro == "langevin":
        assert r["heat"] != 0.0 and r["draws"] == 2 * 400


def test_frozen_barostat_is_the_nvt_run():
    lat, pos, vel, m = _start()
    ref = O.md_run(_force, lat, pos, vel, m, 1.0, 150, R_CUT, 0.5, nhc=(600.0, 50.0, 3))
    for mode in ("aniso", "z", "tri"):
        r = O.npt_run(_force, lat, pos, vel, m, 1.0, 150, R_CUT, 0.5, 600.0, 50.0, 3, 0.0, 1e12,
                      cell_dof=mode)
        if mode == "tri":
            assert np.abs(r["positions"] - ref[3]).max() < 1e-10
            assert np.abs(r["ep"] - ref[0]).max() < 1e-10
        else:
            assert np.array_equal(r["positions"], ref[3])
            assert np.array_equal(r["velocities"], ref[4])
            assert np.array_equal(r["lattice"], lat)


def test_axes_without_a_generator_do_not_move():
    lat0 = si64_recorded_md()[0]
    z = _run("z", 1.0, 100)["lattice"]
    assert np.array_equal(z[:, :2], lat0[:, :2]) and z[2, 2] != lat0[2, 2]
    xy = _run("xy", 1.0, 100)["lattice"]
    assert np.array_equal(xy[:, 2], lat0[:, 2]) and xy[0, 0] != xy[1, 1]
    xyi = _run("xy_iso", 1.0, 100)["lattice"]
    assert xyi[0, 0] == xyi[1, 1] and xyi[2, 2] == lat0[2, 2] and xyi[0, 0] != lat0[0, 0]
    a = _run("aniso", 1.0, 100)["lattice"]
    assert np.count_nonzero(a - np.diag(np.diag(a))) == 0


@pytest.mark.parametrize("baro", ["nhc", "langevin"])
def test_aniso_relaxes_each_axis_to_the_target_pressure(baro):
    lat, pos, vel, m = _start(np.diag([-0.02, 0.015, 0.0]))
    r = O.npt_run(_force, lat, pos, vel, m, 1.0, 3000, R_CUT, 0.5, 600.0, 50.0, 3, 0.0, 200.0,
                  cell_dof="aniso", baro=baro, seed=5)
    pt0 = r["ptensor"][:50].mean(axis=0).diagonal()
    assert pt0[0] > 0.3 and pt0[1] < -0.5          # x compressed, y stretched
    L = r["lat"][1500:].mean(axis=0).diagonal()
    # the unstrained SW Si cell is cubic: the axes equalise
    assert np.ptp(L) < 0.1 * abs(lat[0, 0] - lat[1, 1])
    assert np.all(np.abs(r["ptensor"][1500:].mean(axis=0).diagonal()) < 0.2)


@pytest.mark.parametrize("baro", ["nhc", "langevin"])
def test_tri_relaxes_a_shear(baro):
    lat, pos, vel, m = _start([[0.0, 0.03, 0.0], [0.03, 0.0, 0.0], [0.0, 0.0, 0.0]])
    r = O.npt_run(_force, lat, pos, vel, m, 1.0, 3000, R_CUT, 0.5, 600.0, 50.0, 3, 0.0, 200.0,
                  cell_dof="tri", baro=baro, seed=6)
    assert r["ptensor"][:50].mean(axis=0)[0, 1] < -0.2
    L = r["lat"][1500:].mean(axis=0)
    assert abs(L[0, 1]) < 0.1 * lat[0, 1] and abs(L[1, 0]) < 0.1 * lat[1, 0]
    assert abs(r["ptensor"][1500:].mean(axis=0)[0, 1]) < 0.05


def test_the_identity_generator_is_the_isotropic_barostat():
    a = _run("iso", 1.0, 100)
    b = _run("aniso", 1.0, 100, generators=[np.eye(3)])
    assert np.abs(a["v"] - b["v"]).max() < 1e-4 * a["v"][0]
    assert np.abs(a["lattice"] - b["lattice"]).max() < 1e-5
    assert np.abs(a["v"] - b["v"]).max() > 0.0
