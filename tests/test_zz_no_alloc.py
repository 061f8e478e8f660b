"""The allocation-free force call (ak_compute_device_noalloc) against
ak_compute_device through the C ABI: the same kernels on the same inputs, so
results are bitwise equal; two calls in a row, the second on a smaller cell,
share the library's scratch array.  One case is also held to the oracle at
1e-10, and an input beyond the scratch array is refused before any launch."""
import ctypes

import numpy as np
import pytest

import oracle as O
from akitsushima_amd import NeighbourList, StillingerWeberPotential, TersoffPotential, _lib
from akitsushima_amd import zhou_eam
from akitsushima_amd.published import ZHOU_2004

from _cases import SW_SI, TERSOFF_SI, rattled, rel_err, si

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers through the HIP runtime the engine library is linked
    against (already loaded with it)."""

    def __init__(self):
        _lib.lib()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self, *ps):
        for p in ps:
            assert self.rt.hipFree(p) == 0


def _device_call(entry, pot, nl, lat, pos, types):
    hip = _Hip()
    n = len(pos)
    F0 = np.full((n, 3), np.nan)
    es0 = np.full(10, np.nan)
    d_pos = hip.put(np.asarray(pos, float))
    d_spc = hip.put(np.asarray(pot.species_ids(types), np.int32))
    d_F = hip.put(F0)
    d_es = hip.put(es0)
    latc = np.ascontiguousarray(lat, dtype=float)
    invc = np.ascontiguousarray(np.linalg.inv(latc))
    try:
        rc = getattr(_lib.lib(), entry)(
            pot.handle(), nl.handle(), n, latc.ctypes.data_as(_lib._pf64),
            invc.ctypes.data_as(_lib._pf64), d_pos, d_spc, d_F, d_es, None)
        _lib.check(rc)
        return hip.get(d_F, F0), hip.get(d_es, es0)
    finally:
        hip.rt.hipDeviceSynchronize()
        hip.free(d_pos, d_spc, d_F, d_es)


def _both(pot, lat, pos, types, skin):
    nl = NeighbourList(pot.r_cut, skin_width=skin)
    nl.build(lat, pos)
    a = _device_call("ak_compute_device_noalloc", pot, nl, lat, pos, types)
    b = _device_call("ak_compute_device", pot, nl, lat, pos, types)
    return a, b


def _cells(kind):
    if kind == "tersoff":
        out = []
        for n, seed in ((4, 1), (3, 2)):
            lat, pos, t = si(n)
            out.append((lat, rattled(lat, pos, 0.05, seed), t))
        return TersoffPotential(TERSOFF_SI, species=["Si"]), out, 0.3
    if kind == "sw_long_rows":
        out = []
        for n, seed in ((4, 3), (3, 4)):
            lat, pos, t = si(n)
            out.append((lat, rattled(lat, pos, 0.1, seed), t))
        return StillingerWeberPotential(SW_SI, species=["Si"]), out, 1.0
    from akitsushima_amd.structures import FCC, cubic

    out = []
    for n, seed in ((5, 5), (4, 6)):
        lat, pos, _ = cubic(FCC, 3.56, n, ["X"] * 4, rattle=0.03, seed=seed)
        types = [str(x) for x in np.random.default_rng(seed).choice(["Cu", "Ni"], len(pos))]
        out.append((lat, pos, types))
    pot = zhou_eam({"Cu": dict(ZHOU_2004["Cu"]), "Ni": dict(ZHOU_2004["Ni"])}, ["Cu", "Ni"])
    return pot, out, 0.3


@pytest.mark.parametrize("kind", ["tersoff", "sw_long_rows", "eam_cuni"])
def test_noalloc_equals_device_call_twice_second_smaller(kind):
    print(f"no-alloc force call: {kind}", flush=True)
    pot, cells, skin = _cells(kind)
    assert len(cells[1][1]) < len(cells[0][1])
    for lat, pos, types in cells:
        (fa, ea), (fb, eb) = _both(pot, lat, pos, types, skin)
        assert np.all(np.isfinite(fa)) and np.all(np.isfinite(ea))
        assert np.array_equal(fa, fb), kind
        assert np.array_equal(ea, eb), kind


def test_noalloc_tersoff_matches_oracle():
    lat, pos, t = si(4)
    pos = rattled(lat, pos, 0.05, 7)
    pot = TersoffPotential(TERSOFF_SI, species=["Si"])
    nl = NeighbourList(pot.r_cut, skin_width=0.3)
    nl.build(lat, pos)
    F, es = _device_call("ak_compute_device_noalloc", pot, nl, lat, pos, t)
    e0, f0, s0 = O.tersoff_flat(lat, pos, TERSOFF_SI)
    assert abs(es[0] - e0) <= 1e-10 * abs(e0)
    assert rel_err(F, f0) < 1e-10
    assert rel_err(es[1:].reshape(3, 3), s0) < 1e-10


def test_noalloc_refuses_inputs_beyond_its_scratch():
    # 64000 atoms with a 2 A skin: 34 neighbours within 5.77 A, rows of 64
    # slots, about 130 MB of bond gradients
    lat, pos, t = si(20)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    nl = NeighbourList(pot.r_cut, skin_width=2.0)
    nl.build(lat, pos)
    with pytest.raises(Exception, match="scratch"):
        _device_call("ak_compute_device_noalloc", pot, nl, lat, pos, t)
