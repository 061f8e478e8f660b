"""Tersoff (1989) multicomponent mixing (engine/tersoff.py:1968-2034) and the
oracle's multi-species Tersoff path on its records, checked against what the
paper fixes: eq. (1e) by hand, Table II's 3C-SiC a = 4.32 A and bulk modulus
2.2 Mbar, the quoted cohesive energies of C (7.37 eV/atom at 3.566 A), Si
(4.63 at 5.432) and SiC (6.165, within the +-0.0058 eV the rounding of
Table I propagates to), and forces / stress as energy derivatives on a
rattled, strained SiC cell (tests/test_multispecies_fit.py:70-156 of the
reference makes the same checks)."""
import math

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

import oracle as O
from akitsushima_amd import force_fn_from_config, tersoff1989_to_triplets
from akitsushima_amd.published import TERSOFF_1989
from akitsushima_amd.structures import DIAMOND, cubic

EL, CHI = TERSOFF_1989["elements"], TERSOFF_1989["chi"]


def zincblende(a, x="Si", y="C", n=2):
    return cubic(DIAMOND, a, n, [x] * 4 + [y] * 4, rattle=0.0)


def rattled_sic(seed=3, amp=0.06, shear=0.02):
    lat, pos, t = zincblende(4.32)
    rng = np.random.default_rng(seed)
    e = rng.uniform(-shear, shear, (3, 3))
    T = np.eye(3) + 0.5 * (e + e.T)
    return lat @ T, pos @ T + rng.normal(0, amp, pos.shape), t


def test_expansion_follows_eq_1e():
    t = tersoff1989_to_triplets(EL, CHI, ["Si", "C"])
    assert len(t) == 8
    sic = t[("Si", "C", "C")]
    assert sic["A"] == pytest.approx(math.sqrt(1830.8 * 1393.6), rel=1e-14)
    assert sic["B"] == pytest.approx(0.9776 * math.sqrt(471.18 * 346.7), rel=1e-14)
    assert sic["lambda1"] == pytest.approx((2.4799 + 3.4879) / 2, rel=1e-14)
    assert sic["lambda2"] == pytest.approx((1.7322 + 2.2119) / 2, rel=1e-14)
    r, s = math.sqrt(2.7 * 1.8), math.sqrt(3.0 * 2.1)
    assert sic["R"] == pytest.approx((r + s) / 2, rel=1e-14)
    assert sic["D"] == pytest.approx((s - r) / 2, rel=1e-14)
    assert (sic["beta"], sic["n"]) == (EL["Si"]["beta"], EL["Si"]["n"])
    for trip in (("Si", "C", "C"), ("Si", "C", "Si"), ("Si", "Si", "C")):
        assert (t[trip]["c"], t[trip]["d"], t[trip]["h"]) == (EL["Si"]["c"], EL["Si"]["d"],
                                                              EL["Si"]["h"])
        assert t[trip]["lambda3"] == 0.0 and t[trip]["gamma"] == 1.0 and t[trip]["m"] == 3.0
    cs = t[("C", "Si", "Si")]
    assert (cs["beta"], cs["c"]) == (EL["C"]["beta"], EL["C"]["c"])
    assert cs["A"] == sic["A"] and cs["B"] == sic["B"]
    assert t[("Si", "Si", "Si")]["B"] == 471.18
    assert "A" not in t[("Si", "C", "Si")]
    om = tersoff1989_to_triplets(EL, CHI, ["Si", "C"], omega={"Si-C": 0.9})
    assert om[("Si", "Si", "C")]["gamma"] == 0.9 and om[("C", "C", "Si")]["gamma"] == 0.9
    assert om[("Si", "C", "Si")]["gamma"] == 1.0


def test_missing_chi_is_an_error():
    with pytest.raises(ValueError, match="chi for the"):
        tersoff1989_to_triplets(EL, {}, ["C", "Ge"])
    with pytest.raises(ValueError, match="lacks"):
        tersoff1989_to_triplets({"Si": {"A": 1.0}}, {}, ["Si"])


def _emin(recs, species, builder, bounds):
    e = lambda a: O.tersoff_multi(*builder(a), recs, species)[0] / 64
    r = minimize_scalar(e, bounds=bounds, method="bounded", options={"xatol": 1e-8})
    return r.x, -r.fun, e


@pytest.mark.parametrize("el,a0,e_coh", [("C", 3.566, 7.37), ("Si", 5.432, 4.63)])
def test_oracle_elements_reproduce_quoted_cohesive_energies(el, a0, e_coh):
    recs = tersoff1989_to_triplets(EL, {}, [el])
    a, e, _ = _emin(recs, [el], lambda a: cubic(DIAMOND, a, 2, [el] * 8, rattle=0.0),
                    (0.95 * a0, 1.05 * a0))
    assert abs(e - e_coh) < 0.005
    assert abs(a - a0) < 0.001


def test_oracle_sic_reproduces_table_ii():
    recs = tersoff1989_to_triplets(EL, CHI, ["Si", "C"])
    a, e, efn = _emin(recs, ["Si", "C"], zincblende, (4.2, 4.45))
    assert abs(a - 4.32) < 0.005
    assert abs(e - 6.165) < 0.0058
    h = 1e-3
    ee = [efn(a * (1 + k * h)) for k in (-1, 0, 1)]
    B = a * a * (ee[0] - 2 * ee[1] + ee[2]) / (a * h) ** 2 / (9 * a ** 3 / 8) * 160.21766208
    assert abs(B - 220.0) <= 5.0


def _fd4(fun, h):
    return (fun(-2 * h) - 8 * fun(-h) + 8 * fun(h) - fun(2 * h)) / (12 * h)


def test_oracle_sic_forces_and_stress_are_energy_derivatives():
    recs = tersoff1989_to_triplets(EL, CHI, ["Si", "C"])
    lat, pos, t = rattled_sic()
    sp = ["Si", "C"]
    _, f, s = O.tersoff_multi(lat, pos, t, recs, sp)
    for i in (0, 33, 45):
        for d in range(3):
            def e(x):
                p = pos.copy()
                p[i, d] += x
                return O.tersoff_multi(lat, p, t, recs, sp)[0]
            assert abs(f[i, d] + _fd4(e, 5e-4)) < 5e-9
    frac, vol = pos @ np.linalg.inv(lat), abs(np.linalg.det(lat))
    for a_, b_ in ((0, 0), (1, 2), (2, 1)):
        def e(x):
            T = np.eye(3)
            T[a_, b_] += x
            return O.tersoff_multi(lat @ T, frac @ (lat @ T), t, recs, sp)[0]
        assert abs(s[a_, b_] + _fd4(e, 1e-4) / vol) < 1e-9


def test_config_mixing_form():
    fn = force_fn_from_config({"type": "Tersoff", "species": ["Si", "C"],
                               "params": {"mixing": "tersoff1989", "elements": EL,
                                          "chi": CHI}})
    rc = max(v["R"] + v["D"] for (i, j, k), v in
             tersoff1989_to_triplets(EL, CHI, ["Si", "C"]).items() if j == k)
    assert fn.r_cut == rc
