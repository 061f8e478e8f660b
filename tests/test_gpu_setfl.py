"""Tabulated (setfl) EAM on the HIP engine against the oracle.

The oracle's setfl path is pinned in tests/test_oracle_setfl.py (scipy's own
spline evaluation, the Zhou round trip, finite differences).  Here: the
engine reading the same file gives the oracle's energy, forces and stress to
1e-10 relative; the engine's device tabulation (write_setfl) writes the
oracle's tables; the engine's Zhou and setfl potentials agree to the
interpolation error of the reference's round-trip test; the setfl config
form, batches and every EAM lane count.
"""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import force_fn_from_config, read_setfl, write_setfl, zhou_eam
from akitsushima_amd.published import ATOMIC_MASS, ZHOU_2004
from akitsushima_amd.structures import FCC, cubic

from _cases import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
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


def _check(a, b, label):
    e, f, s = a
    e0, f0, s0 = b
    assert abs(e - e0) / abs(e0) < TOL, f"{label}: E {e} vs {e0}"
    assert rel_err(f, f0) < TOL, f"{label}: F {rel_err(f, f0):.3e}"
    assert rel_err(s, s0) < TOL, f"{label}: S {rel_err(s, s0):.3e}"


@pytest.mark.parametrize("lanes", [0, 4, 8, 16, 32, 64])
def test_setfl_matches_oracle(setfl_path, lanes):
    pot = read_setfl(setfl_path)
    if lanes:
        pot.set_lanes(lanes)
    tab, names = O.read_setfl(setfl_path)
    assert pot.species == names and pot.r_cut == tab.r_cut
    lat, pos, t = ni3al()
    _check(pot.compute(lat, pos, t), O.eam_table(lat, pos, t, tab, names, nthreads=4),
           f"setfl lanes {lanes}")


@pytest.mark.parametrize("block,cap", [(64, 3), (256, 5)])
def test_setfl_persistent_force_grid(setfl_path, block, cap):
    # general tables run the EAM force pass on at most `cap` blocks, each
    # looping over chunks of atoms
    pot = read_setfl(setfl_path)
    pot.set_lanes(8)
    pot.set_launch(0, block, cap)
    tab, names = O.read_setfl(setfl_path)
    lat, pos, t = ni3al(seed=9)
    _check(pot.compute(lat, pos, t), O.eam_table(lat, pos, t, tab, names, nthreads=4),
           f"setfl block {block} cap {cap}")


def test_setfl_species_order_is_the_callers(setfl_path):
    pot = read_setfl(setfl_path, species=["Ni", "Al"])
    tab, names = O.read_setfl(setfl_path)
    lat, pos, t = ni3al(seed=8)
    _check(pot.compute(lat, pos, t), O.eam_table(lat, pos, t, tab, names, nthreads=4),
           "species order")


def test_round_trip_of_the_engines_zhou_potential(tmp_path):
    """Device tabulation -> file -> splines: the reference's round-trip bar."""
    zp = zhou_eam(EL, ["Al", "Ni"])
    path = str(tmp_path / "dev.eam.alloy")
    write_setfl(path, zp, ["Al", "Ni"])
    tab = read_setfl(path)
    assert tab.species == ["Al", "Ni"] and abs(tab.r_cut - zp.r_cut) < 1e-12
    lat, pos, t = ni3al()
    a, b = zp.compute(lat, pos, t), tab.compute(lat, pos, t)
    assert abs(a[0] - b[0]) / len(t) < 1e-10
    assert np.abs(a[1] - b[1]).max() < 1e-5
    assert np.abs(a[2] - b[2]).max() < 1e-7
    # the device's tables are the oracle's to the last written digits
    ref = str(tmp_path / "oracle.eam.alloy")
    O.write_setfl_zhou(ref, EL, ["Al", "Ni"], ATOMIC_MASS)
    rd = lambda p: np.array([float(v) for line in open(p).read().splitlines()[5:]
                             for v in line.split() if len(line.split()) != 4 or "fcc" not in line])
    x, y = rd(path), rd(ref)
    assert x.shape == y.shape
    assert np.abs(x - y).max() <= 1e-13 * np.abs(y).max()


def test_device_function_values_match_the_oracle():
    zp = zhou_eam(EL, ["Al", "Ni"])
    ztab = O.zhou_table(EL, ["Al", "Ni"])
    r = np.linspace(0.5, zp.r_cut, 997)
    rho = np.linspace(0.0, 60.0, 1001)
    for kind, s, b, x in [("rho", "Al", None, r), ("F", "Ni", None, rho), ("phi", "Al", "Ni", r),
                          ("phi", "Ni", "Ni", r)]:
        f = zp.function_index(kind, s, b)
        v, dv = zp.evaluate(f, x)
        v0, dv0 = ztab.feval(f, x)
        assert rel_err(v, v0) < 1e-13 and rel_err(dv, dv0) < 1e-12, (kind, s, b)


def test_config_form_and_batch(setfl_path):
    fn = force_fn_from_config({"type": "EAM", "setfl": setfl_path, "species": ["Al", "Ni"]})
    tab, names = O.read_setfl(setfl_path)
    structs = [ni3al(seed=k, amp=0.03 + 0.02 * k) for k in range(4)]
    for lat, pos, t in structs[:1]:
        e, f, s = fn(lat, pos, t)
        _check((e, f, s), O.eam_table(lat, pos, t, tab, names, nthreads=4), "config")
    pot = read_setfl(setfl_path)
    E, F, S = pot.compute_batched(structs)
    for k, (lat, pos, t) in enumerate(structs):
        _check((E[k], F[k], S[k]), O.eam_table(lat, pos, t, tab, names, nthreads=4),
               f"batch {k}")
