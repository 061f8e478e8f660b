"""Workloads of the reference's force benchmark (benchmarks/bench_force.py:
52-61, 64-131, 152-158, 223-234) rebuilt from the same seeds, plus the
parameter sets the tests use."""
import json
import os

import numpy as np

from akitsushima_amd.structures import DIAMOND, FCC, cubic, supercell

HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(HERE, "golden", "ledger_energies.json")) as fh:
    LEDGER = json.load(fh)

TERSOFF_SI = {"A": 1830.8, "B": 471.18, "lambda1": 2.4799, "lambda2": 1.7322,
              "lambda3": 0.0, "beta": 1.1e-6, "n": 0.78734, "c": 100390.0,
              "d": 16.217, "h": -0.59825, "R": 2.85, "D": 0.15}
SW_SI = {"epsilon": 2.1683, "sigma": 2.0951, "a": 1.80, "gamma": 1.20, "lambda": 21.0,
         "A": 7.049556277, "B": 0.6022245584, "p": 4.0, "q": 0.0, "costheta0": -1 / 3}
TERSOFF_XY = {
    ("X", "Y", "Y"): {"A": 1830.8, "B": 471.18, "lambda1": 2.4799, "lambda2": 1.7322,
                      "R": 2.85, "D": 0.15, "beta": 1.1e-6, "n": 0.78734, "c": 100390.0,
                      "d": 16.217, "h": -0.59825, "lambda3": 1.7322},
    ("Y", "X", "X"): {"A": 1500.0, "B": 400.0, "lambda1": 2.2, "lambda2": 1.5, "R": 2.85,
                      "D": 0.15, "beta": 0.9e-6, "n": 0.7, "c": 90000.0, "d": 14.0,
                      "h": -0.55, "lambda3": 1.5},
}

# Tomita et al., Jpn. J. Appl. Phys. 57, 04FB04 (2018), Table I, as
# benchmarks/bench_force.py:104-119 gives it (epsilon 1, costheta0 -0.333333).
_TOMITA_KEYS = ("sigma", "a", "lambda", "gamma", "A", "B", "p", "q")
_TOMITA_ROWS = [
    ("Ge", "Ge", "Ge", 2.181, 1.8110, 32.101, 1.1718, 10.093, 0.6862, 4.6973, 0.6588),
    ("Sn", "Sn", "Sn", 2.503, 1.7825, 28.747, 1.1610, 9.4315, 0.7012, 4.9827, 0.6252),
    ("Ge", "Ge", "Sn", 2.337, 1.7988, 29.858, 1.8136, 0.0, 0.0, 0.0, 0.0),
    ("Ge", "Sn", "Ge", 2.337, 1.7988, 29.858, 0.7232, 0.0, 0.0, 0.0, 0.0),
    ("Ge", "Sn", "Sn", 2.337, 1.7988, 24.460, 1.1484, 9.8864, 0.7021, 4.8537, 0.6482),
    ("Sn", "Ge", "Ge", 2.337, 1.7988, 24.460, 1.1484, 9.8864, 0.7021, 4.8537, 0.6482),
    ("Sn", "Ge", "Sn", 2.337, 1.7988, 26.074, 1.3564, 0.0, 0.0, 0.0, 0.0),
    ("Sn", "Sn", "Ge", 2.337, 1.7988, 26.074, 1.0924, 0.0, 0.0, 0.0, 0.0),
]
SW_GESN = {tuple(r[:3]): dict(zip(_TOMITA_KEYS, r[3:]), epsilon=1.0, costheta0=-0.333333)
           for r in _TOMITA_ROWS}
GESN_SPECIES = ["Ge", "Sn"]


def si(n):
    return cubic(DIAMOND, 5.431, n, ["Si"] * 8)


def ni(n):
    return cubic(FCC, 3.52, n, ["Ni"] * 4)


def xy(n):
    lat, pos, _ = si(n)
    return lat, pos, (["X"] * 4 + ["Y"] * 4) * (len(pos) // 8)


def gesn_all(cells_diamond=(4, 8)):
    """The Ge0.875Sn0.125 cells of bench_force.build_cases (:223-229)."""
    rng = np.random.default_rng(1)
    out = {}
    for n in cells_diamond:
        lat, pos, types = cubic(DIAMOND, 5.762, n, ["Ge"] * 8)
        out[n] = (lat, pos, [str(x) for x in rng.choice(["Ge", "Sn"], len(types),
                                                         p=[0.875, 0.125])])
    return out


def cuni_all(cells_diamond=(4, 8), cells_fcc=(5, 10)):
    """The Cu-Ni cells of bench_force.build_cases: species drawn from
    default_rng(1) after the Ge-Sn draws."""
    rng = np.random.default_rng(1)
    for n in cells_diamond:
        lat, pos, types = cubic(DIAMOND, 5.762, n, ["Ge"] * 8)
        rng.choice(["Ge", "Sn"], len(types), p=[0.875, 0.125])
    out = {}
    for n in cells_fcc:
        lat, pos, types = cubic(FCC, 3.58, n, ["Cu"] * 4)
        out[n] = (lat, pos, [str(x) for x in rng.choice(["Cu", "Ni"], len(types))])
    return out


def rattled(lat, pos, sigma, seed):
    rng = np.random.default_rng(seed)
    return pos + rng.normal(0.0, sigma, pos.shape)


def sheared_si(n=2, shear=0.07, rattle=0.1, seed=3):
    """Diamond Si with an off-diagonal cell strain (triclinic cell)."""
    lat, pos, t = cubic(DIAMOND, 5.431, n, ["Si"] * 8, rattle=0.0)
    T = np.eye(3)
    T[0, 1] = shear
    T[2, 0] = -0.5 * shear
    frac = pos @ np.linalg.inv(lat)
    lat2 = lat @ T
    pos2 = frac @ lat2
    return lat2, rattled(lat2, pos2, rattle, seed), t


def si64_recorded_md():
    """The 64-atom Si of the reference's recorded thermostat runs
    (_perturbed_si_supercell(1.0, n=2, sigma=0.05, seed=1),
    tests/test_solvers.py:223-241): 2x2x2 conventional diamond at
    a = 5.430950 A, displaced by N(0, 0.05) draws of default_rng(1) with
    their mean removed."""
    lat, pos, t = supercell(DIAMOND, 5.430950, (2, 2, 2), ["Si"] * 8)
    d = np.random.default_rng(1).normal(0.0, 0.05, pos.shape)
    return lat, pos + (d - d.mean(axis=0)), t


def rel_err(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)
