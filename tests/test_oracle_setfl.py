"""The oracle's tabulated (setfl) EAM against what fixes it independently:

* the spline evaluation (engine/_eam_kernels.py:128-145) equals scipy's own
  PPoly evaluation of the same CubicSpline, inside the knots, at the knots
  and beyond both ends (extrapolation by the end polynomials), and the
  s(r) / r form its quotient rule;
* a Zhou Al-Ni potential written to a setfl file and read back gives the
  analytic potential's energy, forces and stress to the interpolation error
  the reference's own round-trip test allows (tests/test_multispecies_fit.py:
  295-311: 1e-10 eV/atom, 1e-5 eV/A, 1e-7 eV/A^3) -- a format slip (wrong
  order, phi for r*phi, a shifted grid) is of order 0.1;
* forces and stress of the spline potential are its energy derivatives.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

import oracle as O
from akitsushima_amd.published import ATOMIC_MASS, ZHOU_2004
from akitsushima_amd.structures import FCC, cubic

EL = {s: dict(ZHOU_2004[s]) for s in ("Al", "Ni")}


def ni3al(a=3.646, n=4, seed=6, amp=0.08, shear=0.02):
    lat, pos, t = cubic(FCC, a, n, ["Al", "Ni", "Ni", "Ni"], rattle=0.0)
    rng = np.random.default_rng(seed)
    e = rng.uniform(-shear, shear, (3, 3))
    T = np.eye(3) + 0.5 * (e + e.T)
    return lat @ T, pos @ T + rng.normal(0, amp, pos.shape), t


@pytest.fixture(scope="module")
def setfl_path(tmp_path_factory):
    path = tmp_path_factory.mktemp("setfl") / "AlNi.eam.alloy"
    O.write_setfl_zhou(str(path), EL, ["Al", "Ni"], ATOMIC_MASS)
    return str(path)


@pytest.mark.parametrize("uniform", [True, False])
def test_spline_evaluation_is_scipys(uniform):
    rng = np.random.default_rng(1)
    x = np.linspace(0.5, 6.0, 41) if uniform else np.sort(rng.uniform(0.5, 6.0, 41))
    y = np.sin(2 * x) * np.exp(-x) + 0.1 * rng.normal(size=x.size)
    cs = CubicSpline(x, y)
    rows = np.zeros((2, O.ZW))
    rows[0, 0], rows[1, 0] = 1, 2
    tab = O.EamTable(rows, [0], [0], [[1]], 6.0, {0: (cs.x, cs.c), 1: (cs.x, cs.c)})
    r = np.concatenate([rng.uniform(0.0, 7.0, 2000), x, [x[0] - 0.3, x[-1] + 0.4]])
    v, dv = tab.feval(0, r)
    assert np.allclose(v, cs(r), rtol=1e-13, atol=1e-13)
    assert np.allclose(dv, cs(r, 1), rtol=1e-12, atol=1e-12)
    w, dw = tab.feval(1, r)
    assert np.allclose(w, cs(r) / r, rtol=1e-13, atol=1e-13)
    assert np.allclose(dw, (cs(r, 1) * r - cs(r)) / r ** 2, rtol=1e-12, atol=1e-11)


def test_setfl_round_trip_of_zhou_al_ni(setfl_path):
    tab, names = O.read_setfl(setfl_path)
    assert names == ["Al", "Ni"]
    ztab = O.zhou_table(EL, ["Al", "Ni"])
    assert tab.r_cut == pytest.approx(ztab.r_cut, rel=1e-15)
    lat, pos, t = ni3al()
    a = O.eam_table(lat, pos, t, ztab, ["Al", "Ni"], nthreads=4)
    b = O.eam_table(lat, pos, t, tab, ["Al", "Ni"], nthreads=4)
    assert abs(a[0] - b[0]) / len(t) < 1e-10
    assert np.abs(a[1] - b[1]).max() < 1e-5
    assert np.abs(a[2] - b[2]).max() < 1e-7


def test_setfl_forces_and_stress_are_energy_derivatives(setfl_path):
    tab, _ = O.read_setfl(setfl_path)
    lat, pos, t = ni3al(seed=2, amp=0.06)
    sp = ["Al", "Ni"]
    _, f, s = O.eam_table(lat, pos, t, tab, sp)

    def fd4(fun, h):
        return (fun(-2 * h) - 8 * fun(-h) + 8 * fun(h) - fun(2 * h)) / (12 * h)

    for i in (0, 17, 130):
        for d in range(3):
            def e(x):
                p = pos.copy()
                p[i, d] += x
                return O.eam_table(lat, p, t, tab, sp)[0]
            assert abs(f[i, d] + fd4(e, 1e-4)) < 2e-8
    frac, vol = pos @ np.linalg.inv(lat), abs(np.linalg.det(lat))
    for a_, b_ in ((0, 0), (1, 2), (2, 1)):
        def e(x):
            T = np.eye(3)
            T[a_, b_] += x
            return O.eam_table(lat @ T, frac @ (lat @ T), t, tab, sp)[0]
        assert abs(s[a_, b_] + fd4(e, 1e-5) / vol) < 1e-9
