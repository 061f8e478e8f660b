"""Batched phonon and elastic drivers (the callers of the batched force
call): against the reference's published SW Si elastic constants (its own
test, tests/test_sw_three_body_counting.py:214-230), against the same
drivers run over oracle forces and stresses, and against the invariants of
force constants (symmetry, acoustic sum rule, three zero modes at Gamma)."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential, TersoffPotential
from akitsushima_amd.elastic import elastic_constants
from akitsushima_amd.phonon import force_constants, gamma_frequencies
from akitsushima_amd.published import SW_SI, TERSOFF_SI_T3
from akitsushima_amd.structures import DIAMOND, supercell

pytestmark = pytest.mark.gpu

A_SI = 5.430950


class OracleBatch:
    """The oracle behind the compute_batched interface (checker only)."""

    def __init__(self, fn):
        self.fn = fn

    def compute_batched(self, cells):
        out = [self.fn(l, p) for l, p, _ in cells]
        return (np.array([o[0] for o in out]), [o[1] for o in out],
                np.array([o[2] for o in out]))


def _cell(n=2):
    lat, pos, t = supercell(DIAMOND, A_SI, (n, n, n), ["Si"] * 8)
    return lat, pos, t


def test_sw_si_elastic_constants_are_the_published_ones():
    lat, pos, t = _cell()
    pot = StillingerWeberPotential(dict(SW_SI), species=["Si"])
    res = elastic_constants(pot, lat, pos, t, relax_ions=True)
    C = res["C"]
    assert C[0, 0] == pytest.approx(151.4, abs=0.5)
    assert C[0, 1] == pytest.approx(76.4, abs=0.5)
    assert C[3, 3] == pytest.approx(56.4, abs=0.5)
    assert res["bulk_modulus"] == pytest.approx(101.4, abs=0.3)
    # cubic symmetry and the internal strain only under shear
    assert C[1, 1] == pytest.approx(C[0, 0], abs=1e-6)
    assert C[4, 4] == pytest.approx(C[3, 3], abs=1e-6)
    assert res["C_clamped"][3, 3] > C[3, 3] + 30.0


def test_clamped_tensor_equals_the_oracle_driven_one():
    lat, pos, t = _cell()
    rng = np.random.default_rng(3)
    pos = pos + rng.normal(0.0, 0.02, pos.shape)
    pot = TersoffPotential(dict(TERSOFF_SI_T3), species=["Si"])
    g = elastic_constants(pot, lat, pos, t)["C_clamped"]
    ref = OracleBatch(lambda l, p: O.tersoff_flat(l, p, dict(TERSOFF_SI_T3)))
    c = elastic_constants(ref, lat, pos, t)["C_clamped"]
    assert np.abs(g - c).max() < 1e-7 * np.abs(c).max()


@pytest.mark.parametrize("kind", ["sw", "tersoff"])
def test_force_constants_equal_the_oracle_driven_ones(kind):
    lat, pos, t = _cell()
    pos = pos + np.random.default_rng(5).normal(0.0, 0.03, pos.shape)
    if kind == "sw":
        pot = StillingerWeberPotential(dict(SW_SI), species=["Si"])
        ref = OracleBatch(lambda l, p: O.sw_flat(l, p, dict(SW_SI)))
    else:
        pot = TersoffPotential(dict(TERSOFF_SI_T3), species=["Si"])
        ref = OracleBatch(lambda l, p: O.tersoff_flat(l, p, dict(TERSOFF_SI_T3)))
    phi = force_constants(pot, lat, pos, t, delta=0.01)
    phi0 = force_constants(ref, lat, pos, t, delta=0.01)
    assert np.abs(phi - phi0).max() < 1e-7 * np.abs(phi0).max()


def test_force_constant_invariants_and_gamma_acoustic_modes():
    lat, pos, t = _cell()
    pot = StillingerWeberPotential(dict(SW_SI), species=["Si"])
    phi = force_constants(pot, lat, pos, t, delta=0.01)
    n = len(pos)
    assert np.abs(phi - phi.T).max() < 1e-12 * np.abs(phi).max()
    rows = phi.reshape(n, 3, n, 3).sum(axis=2)
    assert np.abs(rows).max() < 1e-10 * np.abs(phi).max()
    f = gamma_frequencies(phi, np.full(n, 28.086))
    assert np.sum(np.abs(f) < 1e-3) == 3
    assert f.min() > -1e-3          # perfect diamond is dynamically stable
