"""Re-entrancy: the reference's callers fan force calls out to a thread pool
(solvers/phonon.py:474-480, solvers/elastic.py:289-291) with one owned list
per thread (engine/neighborlist.py:1098-1100).  Worker threads calling one
potential object and one force function of make_force_fn on phonon-style
displaced cells get the results of the same calls made serially (each
thread holds its own device handle, stream and owned list): bitwise for
Tersoff and SW, whose lists here hold no candidates in the skin; EAM rows
hold skin candidates, and which ones depends on the cell a thread's list was
built from, which moves the per-lane summation order: 1e-14 relative."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from akitsushima_amd import (PotentialEngine, StillingerWeberPotential, TersoffPotential,
                             make_force_fn, zhou_eam)
from akitsushima_amd.published import ZHOU_2004
from akitsushima_amd.structures import FCC, cubic

from _cases import SW_SI, TERSOFF_SI, si

pytestmark = pytest.mark.gpu


def _displaced(lat, pos, h=0.01, atoms=(0, 1, 5)):
    cells = []
    for a in atoms:
        for ax in range(3):
            for sgn in (1.0, -1.0):
                p = pos.copy()
                p[a, ax] += sgn * h
                cells.append(p)
    return cells


def _same(a, b, exact):
    (e0, f0, s0), (e1, f1, s1) = a, b
    if exact:
        return e0 == e1 and np.array_equal(f0, f1) and np.array_equal(s0, s1)
    return (abs(e0 - e1) <= 1e-14 * abs(e0) and np.abs(f0 - f1).max() <= 1e-14 * np.abs(f0).max()
            and np.abs(s0 - s1).max() <= 1e-14 * np.abs(s0).max())


def _run_both(fn, cells, workers):
    serial = [fn(c) for c in cells]
    with ThreadPoolExecutor(max_workers=workers) as ex:
        threaded = list(ex.map(fn, cells))
    return serial, threaded


@pytest.mark.parametrize("kind", ["tersoff", "sw", "eam"])
@pytest.mark.parametrize("workers", [4, 8])
def test_threaded_calls_equal_serial_calls(kind, workers):
    if kind == "eam":
        lat, pos, _ = cubic(FCC, 3.56, 4, ["X"] * 4, rattle=0.02, seed=4)
        types = ["Cu" if k % 3 else "Ni" for k in range(len(pos))]
        pot = zhou_eam({"Cu": dict(ZHOU_2004["Cu"]), "Ni": dict(ZHOU_2004["Ni"])},
                       ["Cu", "Ni"])
    else:
        lat, pos, types = si(3)
        pot = (TersoffPotential(TERSOFF_SI, species=["Si"]) if kind == "tersoff"
               else StillingerWeberPotential(SW_SI, species=["Si"]))
    cells = _displaced(lat, pos)
    # the potential object itself, from every thread
    call = lambda p: pot.compute(lat, p, types)
    exact = kind != "eam"
    serial, threaded = _run_both(call, cells, workers)
    assert all(_same(a, b, exact) for a, b in zip(serial, threaded))
    # a force function of the engine, as the phonon and elastic solvers call it
    fn = make_force_fn(PotentialEngine().add(pot))
    call = lambda p: fn(lat, p, types)
    serial, threaded = _run_both(call, cells, workers)
    assert all(_same(a, b, exact) for a, b in zip(serial, threaded))
    # and the displaced cells really differ from one another
    assert len({float(e) for e, _, _ in serial}) > 1
