"""Pins of the oracle's NEMD step (oracle.nemd_run; NEMDSimulation.run,
solvers/nemd.py:113-117, 184-193):

* E_tot - Q is conserved up to the Verlet error, which falls about fourfold
  when dt halves (the reference's own check, tests/test_nemd.py:113-121:
  the energy changes only by the bath heat);
* with the baths off (damping -> infinity, c = 1) the run is the NVE run of
  md_run bit for bit;
* free bath atoms relax to the exact Ornstein-Uhlenbeck variance k_B T / m
  per component (an update with sqrt(1 - c) in place of sqrt(1 - c^2), or
  the wrong bath's temperature, fails).
"""
import numpy as np

import oracle as O

from _cases import SW_SI, sheared_si


def _slab():
    lat, pos, _ = sheared_si(2, shear=0.0, rattle=0.05, seed=3)
    lat = lat.copy()
    pos = np.concatenate([pos, pos + lat[2]])
    lat[2] *= 2
    f = (pos @ np.linalg.inv(lat))[:, 2] % 1.0
    bath = np.where(f < 0.15, 1, np.where((f >= 0.5) & (f < 0.65), 2, 0))
    m = np.full(len(pos), 28.0855)
    vel = O.mb_velocities(600.0, m, 5)
    return lat, pos, vel, m, bath


def _force(lat, p, pr):
    return O.sw_flat(lat, p, SW_SI, pairs=pr)


def test_energy_changes_only_by_the_bath_heat():
    lat, pos, vel, m, bath = _slab()
    drift = {}
    for dt, n in ((1.0, 100), (0.5, 200)):
        r = O.nemd_run(_force, lat, pos, vel, m, dt, n, 3.77118, 0.5, bath, 900.0, 300.0,
                       20.0, seed=4)
        Q = np.abs(r["rows"][:, 0]).sum() + np.abs(r["rows"][:, 1]).sum()
        assert r["rows"][1:, 0].sum() > 0 > r["rows"][1:, 1].sum()
        assert Q > 0.5
        drift[dt] = np.abs(r["econs"] - r["econs"][0]).max()
        assert drift[dt] < 2e-3 * Q
    assert drift[1.0] / drift[0.5] > 2.5


def test_baths_off_is_the_nve_run():
    lat, pos, vel, m, bath = _slab()
    r = O.nemd_run(_force, lat, pos, vel, m, 1.0, 40, 3.77118, 0.5, bath, 900.0, 300.0,
                   1e300, seed=4)
    ep, ek, ec, p, v, rb, _ = O.md_run(_force, lat, pos, vel, m, 1.0, 40, 3.77118, 0.5)
    assert np.array_equal(r["positions"], p) and np.array_equal(r["velocities"], v)
    assert np.array_equal(r["ep"], ep) and np.array_equal(r["ek"], ek)
    assert r["rebuilds"] == rb and not r["rows"][:, :2].any()


def test_free_bath_atoms_reach_the_ou_variance():
    n = 4000
    g = np.arange(n)
    pos = np.stack([g % 20, (g // 20) % 20, g // 400], axis=1) * 3.0 + 0.5
    lat = np.diag([60.0, 60.0, 30.0])
    m = np.random.default_rng(2).uniform(10.0, 100.0, n)
    bath = np.where(g < n // 2, 1, 2)
    zero = lambda l, p, pr: (0.0, np.zeros_like(p), np.zeros((3, 3)))
    samples = []
    r = O.nemd_run(zero, lat, pos, np.zeros_like(pos), m, 1.0, 40, 1.0, 0.3, bath, 800.0,
                   200.0, 1.0, seed=9, sample=lambda p, v: v.copy(), every=1)
    for v in r["samples"][10:]:
        for idx, T in ((g < n // 2, 800.0), (g >= n // 2, 200.0)):
            samples.append((np.mean(m[idx, None] * v[idx] ** 2) / O.EV_ANG_TO_AMU_ANG_FS2,
                            O.K_B * T))
    ratio = np.array([a / b for a, b in samples]).reshape(-1, 2).mean(axis=0)
    assert np.allclose(ratio, 1.0, atol=0.02), ratio
    # the heat is the kinetic energy the baths hold (no forces)
    assert np.isclose(r["rows"][1:, 0].sum(), r["rows"][-1, 2], rtol=1e-10)
    assert np.isclose(r["rows"][1:, 1].sum(), r["rows"][-1, 3], rtol=1e-10)
