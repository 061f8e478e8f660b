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
