"""Launch shapes and the autotuner: every shape the autotuner may pick gives
the oracle's energies, forces and stresses (1e-10 relative) on a batch of
uneven structures, including the persistent form where blocks loop over
more chunks than the grid has blocks."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import StillingerWeberPotential, TersoffPotential, zhou_eam
from akitsushima_amd.published import ZHOU_2004
from akitsushima_amd.structures import DIAMOND, cubic, diamond_batch

from _cases import SW_SI, TERSOFF_SI, cuni_all, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10

# block 512: the four-slot kernel; with the bond-pair form (tile 2) the plan
# falls back to 256
SHAPES = [(1, 256, 0), (1, 64, 0), (1, 128, 3), (1, 512, 0), (2, 256, 0), (2, 64, 5),
          (2, 128, 1), (2, 512, 0)]


def _batch():
    return diamond_batch(9, 64, 512, 0.02, 0.08, seed=21, rep_range=(2, 4))


def _compare(E, F, S, refs):
    for k, (e0, f0, s0) in enumerate(refs):
        assert abs(E[k] - e0) / abs(e0) < TOL, (k, E[k], e0)
        assert rel_err(F[k], f0) < TOL, (k, rel_err(F[k], f0))
        assert rel_err(S[k], s0) < TOL, (k, rel_err(S[k], s0))


@pytest.mark.parametrize("shape", SHAPES, ids=[f"tile{t}-b{b}-cap{c}" for t, b, c in SHAPES])
def test_tersoff_batch_every_launch_shape(shape):
    structs = _batch()
    refs = [O.tersoff_flat(l, p, TERSOFF_SI) for l, p, _ in structs]
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    pot.set_launch(*shape)
    E, F, S = pot.compute_batched(structs)
    _compare(E, F, S, refs)


@pytest.mark.parametrize("block", [256, 512])
@pytest.mark.parametrize("cells", [4, 5])
def test_single_structure_kernels_equal_the_batched_ones_bit_for_bit(cells, block):
    """One structure on rows of four runs the single-structure forms of the
    four-slot site kernel and its gather (fewer kernel arguments, chunks
    from the block index); the same structure twice in a batch runs the
    batched forms over the same chunks (the last one partial at 1000
    atoms). Energy, forces and stress agree bit for bit."""
    lat, pos, types = cubic(DIAMOND, 5.431, cells, ["Si"] * 8, rattle=0.05, seed=3)
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    pot.set_launch(1, block, 0)
    e, f, s = pot.compute(lat, pos, types)
    E, F, S = pot.compute_batched([(lat, pos, types), (lat, pos, types)])
    for k in range(2):
        assert E[k] == e, (k, E[k], e)
        assert np.array_equal(F[k], f), (k, np.abs(F[k] - f).max())
        assert np.array_equal(S[k], s), (k, np.abs(S[k] - s).max())


@pytest.mark.parametrize("shape", [(0, 64, 0), (0, 128, 2), (0, 256, 7), (0, 512, 0)])
def test_sw_batch_launch_shapes(shape):
    structs = _batch()
    refs = [O.sw_flat(l, p, SW_SI) for l, p, _ in structs]
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    pot.set_launch(*shape)
    E, F, S = pot.compute_batched(structs)
    _compare(E, F, S, refs)


@pytest.mark.parametrize("lanes,block,cap", [(8, 64, 0), (32, 128, 3), (64, 256, 0)])
def test_eam_launch_shapes(lanes, block, cap):
    lat, pos, t = cuni_all()[5]
    el = {"Cu": ZHOU_2004["Cu"], "Ni": ZHOU_2004["Ni"]}
    pot = zhou_eam(el, ["Cu", "Ni"])
    pot.set_lanes(lanes)
    pot.set_launch(0, block, cap)
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.eam_zhou(lat, pos, t, {"Cu": O.ZHOU_CU, "Ni": O.ZHOU_NI}, ["Cu", "Ni"],
                            nthreads=8)
    assert abs(e - e0) / abs(e0) < TOL
    assert rel_err(f, f0) < TOL and rel_err(s, s0) < TOL


def test_autotuner_picks_a_measured_shape_and_stays_exact():
    structs = _batch()
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    res = pot.autotune(structs, reps=3)
    tried = res["tried"]
    assert len(tried) == 7  # two forms x three block sizes, and 512 for the four-slot kernel
    assert all(r["ms"] > 0.0 for r in tried)
    assert res["best_ms"] == min(r["ms"] for r in tried)
    # both Tersoff forms; the bond kernels run one block per chunk, so no grid cap
    assert any(r["bond_tile"] == 2 for r in tried) and all(r["grid_cap"] == 0 for r in tried)
    refs = [O.tersoff_flat(l, p, TERSOFF_SI) for l, p, _ in structs]
    E, F, S = pot.compute_batched(structs)
    _compare(E, F, S, refs)


def test_autotuner_eam_lanes():
    lat, pos, t = cuni_all()[5]
    el = {"Cu": ZHOU_2004["Cu"], "Ni": ZHOU_2004["Ni"]}
    pot = zhou_eam(el, ["Cu", "Ni"])
    res = pot.autotune([(lat, pos, t)], reps=2)
    assert {r["eam_lanes"] for r in res["tried"]} == {8, 16, 32, 64}
    e, f, s = pot.compute(lat, pos, t)
    e0, f0, s0 = O.eam_zhou(lat, pos, t, {"Cu": O.ZHOU_CU, "Ni": O.ZHOU_NI}, ["Cu", "Ni"],
                            nthreads=8)
    assert abs(e - e0) / abs(e0) < TOL and rel_err(f, f0) < TOL
