"""Pins of the oracle's Nose-Hoover chain integrator (oracle.md_run with
nhc) and of the Green-Kubo post-processing (akitsushima_amd.thermal), no GPU.

* H_NHC (solvers/molecular_dynamics.py:947-969) is conserved up to the
  integrator's O(dt^2) error: its range falls fourfold when dt halves, and
  stays three orders below the range of E_tot, which the thermostat moves;
* the chain holds the canonical mean temperature;
* autocorr_fft equals the direct lag sum sum_t a[t+k] b[t] / (n - k)
  (the reference's lag convention, solvers/thermal_conductivity.py:1107-1170);
* compute_kappa on an AR(1) series recovers the closed-form integral of its
  autocorrelation, and gives zero for a constant flux.
"""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd.thermal import EV_ANG_K_FS_TO_W_MK, ThermalConductivityCalculator, autocorr_fft

from _cases import SW_SI, si


def _run(dt, n, T=900.0, tau=50.0, M=3):
    lat, pos, t = si(2)
    m = np.full(len(pos), 28.086)
    vel = O.mb_velocities(300.0, m, 3)
    force = lambda l, p, pr: O.sw_flat(l, p, SW_SI, pairs=pr)
    return O.md_run(force, lat, pos, vel, m, dt, n, 3.77118, 0.3, nhc=(T, tau, M)), len(pos)


def test_nhc_conserves_h_to_second_order():
    ranges = []
    for dt in (0.5, 0.25):
        (ep, ek, ec, *_), _ = _run(dt, int(600 / dt))
        et = ep + ek
        ranges.append(ec.max() - ec.min())
        assert ranges[-1] < 1e-3 * (et.max() - et.min())
    assert 3.0 < ranges[0] / ranges[1] < 5.0


def test_nhc_holds_the_target_temperature():
    (ep, ek, ec, *_), n = _run(1.0, 3000, M=4)
    T = 2 * ek / ((3 * n - 3) * O.K_B)
    assert abs(T[1500:].mean() - 900.0) < 0.05 * 900.0


def test_autocorr_is_the_direct_lag_sum():
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=257), rng.normal(size=257)
    for sub in (True, False):
        c = autocorr_fft(a, b, 40, subtract_mean=sub)
        x = a - a.mean() if sub else a
        y = b - b.mean() if sub else b
        d = [np.dot(x[k:], y[:len(y) - k]) / (len(x) - k) for k in range(40)]
        assert np.allclose(c, d, rtol=1e-12, atol=1e-14)


def test_kappa_of_an_ar1_flux_matches_its_closed_form():
    rng = np.random.default_rng(5)
    n, phi, sig, dt = 400_000, 0.9, 1e-6, 2.0
    from scipy.signal import lfilter

    J = lfilter([1.0], [1.0, -phi], rng.normal(scale=sig, size=(n, 3)), axis=0)
    lat = np.eye(3) * 10.0
    tc = ThermalConductivityCalculator.__new__(ThermalConductivityCalculator)
    tc.lattice, tc.temperature, tc._J_series, tc._dt_hf = lat, 500.0, None, None
    tau_max = 0.3                                   # ps: 150 lags, phi^150 ~ 1e-7
    kappa = tc.compute_kappa(J, dt_hf=dt, tau_max=tau_max, n_blocks=5)
    n_lag = int(tau_max * 1e3 / dt)
    C0 = sig ** 2 / (1 - phi ** 2)
    k = np.arange(n_lag)
    integral = dt * (C0 * np.sum(phi ** k) - 0.5 * C0 - 0.5 * C0 * phi ** (n_lag - 1))
    ref = 1000.0 / (O.K_B * 500.0 ** 2) * integral * EV_ANG_K_FS_TO_W_MK
    # statistical scatter of one component ~4 % at this length
    d = np.diag(kappa)
    assert abs(d.mean() - ref) < 0.05 * ref
    assert np.all(np.abs(d - ref) < 0.12 * ref)
    assert np.all(np.abs(kappa - np.diag(d)) < 0.1 * ref)
    assert np.all(np.isfinite(tc._kappa_std))
    kc = tc.compute_kappa(np.ones((1000, 3)) * 3e-5, dt_hf=dt, tau_max=0.2, n_blocks=2)
    assert np.abs(kc).max() < 1e-20


# --------------------------------------------------------------------------
#  stochastic thermostats: the Philox stream and its distributions, then the
#  integrators' E - Q conservation and canonical statistics
# --------------------------------------------------------------------------

def test_philox_known_answers():
    """Known-answer vectors of Philox4x32-10 (Random123's kat_vectors)."""
    for c, k, out in [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
                      ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
                       (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
                      ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
                       (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]:
        got = O.philox4x32(list(c), *k)
        assert [int(x) for x in got] == list(out)


def test_normals_and_gamma_have_their_moments():
    n = 400_000
    a = np.arange(n, dtype=np.uint64)
    z0, z1 = O.normal2(O.philox4x32([a, np.zeros(n), np.ones(n), np.zeros(n)], 11, 0))
    z = np.concatenate([z0, z1])
    assert abs(z.mean()) < 5e-3 and abs(z.var() - 1.0) < 1e-2
    assert abs(np.mean(z ** 4) - 3.0) < 0.05
    for shape in (1.5, 40.0):
        g = np.array([O.gamma_mt(shape, h, 5, 9) for h in range(4000)])
        assert abs(g.mean() - shape) < 4 * np.sqrt(shape / 4000)
        assert abs(g.var() / shape - 1.0) < 0.15


def _stoch(kind, dt, n):
    lat, pos, t = si(2)
    m = np.full(len(pos), 28.086)
    vel = O.mb_velocities(300.0, m, 3)
    force = lambda l, p, pr: O.sw_flat(l, p, SW_SI, pairs=pr)
    out = O.md_run(force, lat, pos, vel, m, dt, n, 3.77118, 0.3, **{kind: (600.0, 50.0, 7)})
    return out, len(pos)


@pytest.mark.parametrize("kind", ["langevin", "csvr"])
def test_stochastic_thermostats_conserve_e_minus_q_and_sample_600k(kind):
    ranges = []
    for dt in (1.0, 0.5):
        (ep, ek, ec, *_), n = _stoch(kind, dt, int(2000 / dt))
        et = ep + ek
        ranges.append(ec.max() - ec.min())
        assert ranges[-1] < 2e-3 * (et.max() - et.min())
    assert 3.0 < ranges[0] / ranges[1] < 5.5
    (ep, ek, ec, *_), n = _stoch(kind, 1.0, 4000)
    nf = 3 * n - 3
    T = 2 * ek / (nf * O.K_B)
    assert abs(T[1000:].mean() - 600.0) < 0.05 * 600.0
    if kind == "csvr":   # canonical kinetic-energy variance (n_f / 2) (kT)^2
        r = ek[1000:].var() / (0.5 * nf * (O.K_B * 600.0) ** 2)
        assert 0.6 < r < 1.6


# --------------------------------------------------------------------------
#  NPT (isotropic barostat)
# --------------------------------------------------------------------------

def _npt_cell():
    from akitsushima_amd.structures import DIAMOND, cubic
    lat, pos, t = cubic(DIAMOND, 5.40, 2, ["Si"] * 8, rattle=0.03, seed=1)
    m = np.full(len(pos), 28.086)
    return lat, pos, m, O.mb_velocities(600.0, m, 3)


def _sw(l, p, pr):
    return O.sw_flat(l, p, SW_SI, pairs=pr)


def test_npt_with_a_frozen_barostat_is_nvt():
    """tau_baro -> infinity: W -> infinity, the cell stays, and the step is
    the Nose-Hoover NVT step operation for operation."""
    lat, pos, m, vel = _npt_cell()
    a = O.npt_run(_sw, lat, pos, vel, m, 1.0, 200, 3.77118, 0.5, 600.0, 50.0, 3, 0.0, 1e12)
    ep, ek, ec, p1, v1, rb, _ = O.md_run(_sw, lat, pos, vel, m, 1.0, 200, 3.77118, 0.5,
                                         nhc=(600.0, 50.0, 3))
    assert np.array_equal(a["ep"], ep) and np.array_equal(a["positions"], p1)
    assert a["rebuilds"] == rb


def test_npt_conserves_h_and_relaxes_to_the_target_pressure():
    lat, pos, m, vel = _npt_cell()
    ranges = []
    for dt in (1.0, 0.5):
        r = O.npt_run(_sw, lat, pos, vel, m, dt, int(1500 / dt), 3.77118, 0.5, 600.0, 50.0, 3,
                      0.0, 200.0)
        et = r["ep"] + r["ek"]
        ranges.append(r["h"].max() - r["h"].min())
        assert ranges[-1] < 2e-3 * (et.max() - et.min())
    assert 2.5 < ranges[0] / ranges[1] < 5.0
    # compressed start (P_inst 2.3 GPa): the cell expands and P averages to 0
    assert r["p"][0] > 2.0 and r["v"][-1] > 1.02 * r["v"][0]
    assert abs(r["p"][len(r["p"]) // 2:].mean()) < 0.2
