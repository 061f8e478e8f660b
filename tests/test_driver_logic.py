"""Host logic of the phonon and elastic drivers (no GPU): strain points and
matrices in the reference's order and convention, the acoustic-sum-rule
finish of the force constants, and the fits on synthetic data with known
answers."""
import numpy as np

from akitsushima_amd.elastic import (VOIGT_PAIRS, elastic_constants, strain_matrix,
                                     strain_points)
from akitsushima_amd.phonon import displaced_cells, finish_ifc, gamma_frequencies


def test_strain_points_follow_the_reference_order():
    pts = strain_points((0.01, 0.0025))
    assert pts[0] == (None, 0.0)
    assert pts[1:5] == [(0, 0.0025), (0, -0.0025), (0, 0.01), (0, -0.01)]
    assert len(pts) == 1 + 6 * 2 * 2


def test_shear_strain_is_engineering():
    e = strain_matrix(3, 0.02)
    assert e[1, 2] == e[2, 1] == 0.01 and e.trace() == 0.0
    assert VOIGT_PAIRS[5] == (0, 1)


def test_finish_ifc_symmetrises_and_imposes_the_sum_rule():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(12, 12))
    f = finish_ifc(a.copy())
    assert np.allclose(f.reshape(4, 3, 4, 3).sum(axis=2), 0.0, atol=1e-12)
    off = ~np.kron(np.eye(4, dtype=bool), np.ones((3, 3), dtype=bool))
    assert np.allclose(f[off], (0.5 * (a + a.T))[off])


def test_displaced_cells_reach_minus_through_plus():
    pos = np.array([[0.1, 0.2, 0.3], [1.0, 1.1, 1.2]])
    cells, tasks = displaced_cells(np.eye(3) * 5, pos, ["X", "X"], 0.01)
    assert len(cells) == 12 and tasks[1] == (0, 1)
    assert cells[1][1][0, 0] == (0.1 + 0.01) - 0.02


class Harmonic:
    """E = 1/2 V eta C eta for a cell with fixed fractional coordinates, and
    harmonic springs for the force constants: a known answer for the fits."""

    def __init__(self, C, lat0, k):
        self.C, self.lat0, self.k = C, lat0, k

    def compute_batched(self, cells):
        E, F, S = [], [], []
        V0 = abs(np.linalg.det(self.lat0))
        for lat, pos, _ in cells:
            eps = np.linalg.solve(self.lat0, lat) - np.eye(3)
            eta = np.array([eps[a, b] * (1 if a == b else 2) for a, b in VOIGT_PAIRS])
            sig_std = self.C @ eta / 160.2176621
            sig = np.zeros((3, 3))
            for q, (a, b) in enumerate(VOIGT_PAIRS):
                sig[a, b] = sig[b, a] = -sig_std[q]
            E.append(0.5 * V0 * eta @ sig_std)
            F.append(-self.k * (pos - pos.mean(axis=0)))
            S.append(sig)
        return np.array(E), F, np.array(S)


def test_clamped_fit_recovers_a_known_tensor():
    C = np.array([[150, 75, 75, 0, 0, 0], [75, 150, 75, 0, 0, 0], [75, 75, 150, 0, 0, 0],
                  [0, 0, 0, 60, 0, 0], [0, 0, 0, 0, 60, 0], [0, 0, 0, 0, 0, 60.0]])
    lat = np.eye(3) * 10.0
    pos = np.array([[1.0, 1.0, 1.0], [6.0, 6.0, 6.0]])
    res = elastic_constants(Harmonic(C, lat, 1.0), lat, pos, ["X", "X"])
    assert np.allclose(res["C_clamped"], C, atol=1e-9)


def test_gamma_frequencies_of_two_springs():
    k, m = 2.0, 3.0
    phi = np.zeros((6, 6))
    for a in range(3):
        phi[a, a] = phi[3 + a, 3 + a] = k
        phi[a, 3 + a] = phi[3 + a, a] = -k
    f = np.sort(gamma_frequencies(phi, [m, m]))
    w = np.sqrt(2 * k / m) * np.sqrt(0.009648533212331001) * 1e15 / (2 * np.pi) / 1e12
    assert np.allclose(f[:3], 0.0, atol=1e-9) and np.allclose(f[3:], w, rtol=1e-9)
