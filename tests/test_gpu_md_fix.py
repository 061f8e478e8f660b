"""fix_atoms and velocity_rescale_steps on the device against the oracle
(pinned in tests/test_oracle_md_fix.py): every thermostat and the NPT
barostat with a quarter of the atoms frozen plus one frozen component of a
free atom, the chain and NPT also with velocity rescaling over the first 20
steps -- energies and conserved quantities of every step and the final
state to 1e-9; frozen components exactly zero; a run split inside the
rescaling window equals the uninterrupted one.
"""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential
from akitsushima_amd.md import csvr_run, langevin_run, npt_run, nve_run, nvt_run

from _cases import SW_SI, si64_recorded_md

pytestmark = pytest.mark.gpu

R_CUT = 1.80 * 2.0951


def _force(l, p, pr):
    return O.sw_flat(l, p, SW_SI, pairs=pr)


def _start():
    lat, pos, t = si64_recorded_md()
    m = np.full(len(pos), 28.086)
    fix = np.zeros((64, 3), bool)
    fix[:16] = True
    fix[20, 2] = True
    return lat, pos, t, O.mb_velocities(450.0, m, 7), m, fix


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("kind", ["nve", "nhc", "nhc_rescale", "langevin", "csvr", "npt_iso",
                                  "npt_tri_rescale"])
def test_fixed_atoms_match_the_oracle(kind):
    lat, pos, t, vel, m, fix = _start()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    n, skin = 80, 0.5
    kw = dict(dt=1.0, n_steps=n, skin_width=skin, masses=m, fix_atoms=fix)
    ok = dict(fix=fix)
    if kind == "nve":
        res = nve_run(pot, lat, pos, vel, t, **kw)
        r = O.md_run(_force, lat, pos, vel, m, 1.0, n, R_CUT, skin, **ok)
    elif kind.startswith("nhc"):
        k = 20 if kind == "nhc_rescale" else 0
        res = nvt_run(pot, lat, pos, vel, t, 600.0, tau_nhc=50.0, nhc_chain=3,
                      velocity_rescale_steps=k, **kw)
        r = O.md_run(_force, lat, pos, vel, m, 1.0, n, R_CUT, skin, nhc=(600.0, 50.0, 3),
                     rescale_steps=k, **ok)
    elif kind == "langevin":
        res = langevin_run(pot, lat, pos, vel, t, 600.0, damping=50.0, seed=3, **kw)
        r = O.md_run(_force, lat, pos, vel, m, 1.0, n, R_CUT, skin, langevin=(600.0, 50.0, 3),
                     **ok)
    elif kind == "csvr":
        res = csvr_run(pot, lat, pos, vel, t, 600.0, damping=50.0, seed=4, **kw)
        r = O.md_run(_force, lat, pos, vel, m, 1.0, n, R_CUT, skin, csvr=(600.0, 50.0, 4), **ok)
    else:
        mode = "iso" if kind == "npt_iso" else "tri"
        k = 20 if kind.endswith("rescale") else 0
        res = npt_run(pot, lat, pos, vel, t, 600.0, pressure=0.5, tau_nhc=50.0, nhc_chain=3,
                      tau_baro=200.0, cell_dof=mode, velocity_rescale_steps=k, **kw)
        q = O.npt_run(_force, lat, pos, vel, m, 1.0, n, R_CUT, skin, 600.0, 50.0, 3, 0.5, 200.0,
                      cell_dof=mode, rescale_steps=k, **ok)
        r = (q["ep"], q["ek"], q["h"], q["positions"], q["velocities"], q["rebuilds"])
        assert _rel(res["volume"], q["v"]) < 1e-10
        assert _rel(res["lattice"], q["lattice"]) < 1e-10
    ep, ek, ec, p, v, rb = r[:6]
    assert _rel(res["e_pot"], ep) < 1e-10
    assert _rel(res["e_kin"], ek) < 1e-9
    assert _rel(res["e_cons"], ec) < 1e-10
    L = res.get("lattice", lat)
    d = res["positions"] - p
    d -= np.round(d @ np.linalg.inv(L)) @ L
    assert np.abs(d).max() < 1e-8
    assert _rel(res["velocities"], v) < 1e-9
    assert np.all(res["velocities"][fix] == 0.0)
    assert res["n_rebuilds"] == rb
    if "rescale" in kind:
        assert np.abs(res["T"][1:21] - 600.0).max() < 1e-9


def test_split_inside_the_rescale_window():
    lat, pos, t, vel, m, fix = _start()
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    kw = dict(dt=1.0, tau_nhc=50.0, masses=m, fix_atoms=fix, velocity_rescale_steps=30)
    full = nvt_run(pot, lat, pos, vel, t, 600.0, n_steps=80, **kw)
    a = nvt_run(pot, lat, pos, vel, t, 600.0, n_steps=20, **kw)
    b = nvt_run(pot, lat, a["positions"], a["velocities"], t, 600.0, n_steps=60,
                state=a["state"], **kw)
    assert _rel(b["e_cons"], full["e_cons"][20:]) < 1e-10
    assert _rel(b["velocities"], full["velocities"]) < 1e-9
    assert np.abs(full["T"][1:31] - 600.0).max() < 1e-9
    assert np.abs(b["T"][1:11] - 600.0).max() < 1e-9 and abs(b["T"][12] - 600.0) > 1e-6
