"""Host side of the setfl reader (no GPU): the tables the engine receives
are the oracle's (scipy CubicSpline coefficients, knot layout, index map)."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import read_setfl
from akitsushima_amd.published import ATOMIC_MASS, ZHOU_2004


@pytest.fixture(scope="module")
def setfl_path(tmp_path_factory):
    path = tmp_path_factory.mktemp("setfl") / "CuNi.eam.alloy"
    el = {s: dict(ZHOU_2004[s]) for s in ("Cu", "Ni")}
    O.write_setfl_zhou(str(path), el, ["Cu", "Ni"], ATOMIC_MASS, n_rho=800, n_r=900)
    return str(path)


def test_tables_are_the_oracles(setfl_path):
    pot = read_setfl(setfl_path)
    tab, names = O.read_setfl(setfl_path)
    assert pot.species == names == ["Cu", "Ni"]
    ns = 2
    m = pot._map
    for s in range(ns):
        for f_eng, f_orc in ((m[s], tab.rho_f[s]), (m[ns + s], tab.emb_f[s])):
            row = pot._fn[f_eng]
            o, n = int(row[1]), int(row[2])
            assert row[0] == tab.rows[f_orc][0] == 1
            xo, nx = tab.xo[f_orc], tab.nx[f_orc]
            assert n == nx
            assert np.array_equal(pot._knots[o:o + n], tab.X[xo:xo + nx])
            assert np.array_equal(pot._coeffs[o:o + n].T, tab.C[:, xo:xo + nx])
    for a in range(ns):
        for b in range(ns):
            assert pot._fn[m[2 * ns + a * ns + b]][0] == 2
            assert m[2 * ns + a * ns + b] == m[2 * ns + b * ns + a]


def test_unknown_species_is_refused(setfl_path):
    with pytest.raises(KeyError):
        read_setfl(setfl_path, species=["Cu", "Fe"])


class _StubEAM:
    """Analytic functions in place of the device tabulation (host only)."""

    r_cut = 5.0
    zhou_params = None

    def function_index(self, kind, a, b=None):
        return (kind, a, b)

    def evaluate(self, f, x):
        kind, a, b = f
        k = 1.0 + ["Cu", "Ni"].index(a) + (0.5 * ["Cu", "Ni"].index(b) if b else 0.0)
        if kind == "F":
            return -k * np.sqrt(x), None
        if kind == "rho":
            return k * np.exp(-x), None
        return k * np.exp(-2.0 * x) / (1.0 + x), None


def test_writer_layout_reads_back_through_the_reference_reader(tmp_path):
    from akitsushima_amd import write_setfl

    path = str(tmp_path / "stub.eam.alloy")
    pot = _StubEAM()
    write_setfl(path, pot, ["Cu", "Ni"], n_rho=23, rho_max=4.0, n_r=17,
                comments=("a", "b\nc", ""), lattice={"Cu": (3.615, "fcc")})
    lines = open(path).read().splitlines()
    assert lines[:5] == ["a", "b c", "", "2 Cu Ni",
                         f"23 {4.0 / 22:.16e} 17 {5.0 / 16:.16e} {5.0:.16e}"]
    assert lines[5] == f"29 {ATOMIC_MASS['Cu']:.6f} 3.615000 fcc"
    # 23 values: 4 full lines and one of 3; 17 values: 3 lines and one of 2
    assert [len(l.split()) for l in lines[6:11]] == [5, 5, 5, 5, 3]
    assert [len(l.split()) for l in lines[11:15]] == [5, 5, 5, 2]
    assert lines[15] == f"28 {ATOMIC_MASS['Ni']:.6f} 0.000000 fcc"
    assert len(lines) == 5 + 2 * (1 + 5 + 4) + 3 * 4
    assert lines[11].startswith(" 1.0000000000000000e+00  ")  # % .16e: a sign column
    tab, names = O.read_setfl(path)
    assert names == ["Cu", "Ni"]
    rho = np.arange(23) * (4.0 / 22)
    r = np.arange(17) * (5.0 / 16)
    for k, s in enumerate(names):
        f, g = tab.emb_f[k], tab.rho_f[k]
        assert np.array_equal(tab.X[tab.xo[f]:tab.xo[f] + tab.nx[f]], rho)
        assert np.array_equal(tab.X[tab.xo[g]:tab.xo[g] + tab.nx[g]], r)
    # the written numbers are the functions to the 17 digits of the format
    vals = [float(v) for v in lines[6].split()]
    assert np.allclose(vals, pot.evaluate(("F", "Cu", None), rho[:5])[0], rtol=1e-16, atol=0)
    pair = [float(v) for l in lines[-4:] for v in l.split()]
    assert np.allclose(pair, r * pot.evaluate(("phi", "Ni", "Ni"), r)[0], rtol=1e-16, atol=0)
