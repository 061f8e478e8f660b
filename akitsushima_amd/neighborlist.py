"""Device neighbour list with the reference's NeighbourList interface
(engine/neighborlist.py:390-714): ``build``, ``check_and_rebuild``,
``get_candidate_pairs``, ``rebuild_count`` / ``step_count``.

The list lives in GPU memory as full rows (both directions) sorted by
neighbour index; ``get_candidate_pairs`` copies out the canonical half list
(i < j, sorted by (i, j)).  The pair predicate is the reference's
``r2 < r_list**2 and r2 > 1e-20`` on minimum-image displacements computed
with numpy's inverse lattice, so the pair set is the reference's bit for bit.
"""
from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import check, f64, lib


def _destroy(ptr):
    try:
        lib().ak_nlist_destroy(ptr)
    except Exception:
        pass


class NeighbourList:
    def __init__(self, r_cut, skin_width=0.5, backend="auto", use_numba=None,
                 canonical_order=True):
        if r_cut <= 0.0:
            raise ValueError(f"NeighbourList: r_cut must be positive, got {r_cut}.")
        self.r_cut = float(r_cut)
        self.skin_width = float(skin_width)
        self.r_list = self.r_cut + self.skin_width
        self.canonical_order = True
        h = ctypes.c_void_p()
        check(lib().ak_nlist_create(self.r_cut, self.skin_width, ctypes.byref(h)))
        self._h = h
        self._fin = weakref.finalize(self, _destroy, h)
        self._n_steps = 0
        self._dirty = False

    def handle(self):
        return self._h

    def build(self, lattice, positions):
        lat, plat = f64(lattice)
        pos, ppos = f64(positions)
        inv, pinv = f64(np.linalg.inv(lat))
        check(lib().ak_nlist_build(self._h, len(pos), plat, pinv, ppos))

    def mark_dirty(self):
        self._dirty = True

    def check_and_rebuild(self, lattice, positions):
        self._n_steps += 1
        if self._dirty:
            self._dirty = False
            self.build(lattice, positions)
            return True
        lat, plat = f64(lattice)
        pos, ppos = f64(positions)
        inv, pinv = f64(np.linalg.inv(lat))
        rb = ctypes.c_int32(0)
        check(lib().ak_nlist_check_and_rebuild(self._h, len(pos), plat, pinv, ppos,
                                               ctypes.byref(rb)))
        return bool(rb.value)

    def get_candidate_pairs(self):
        n = ctypes.c_int64(0)
        check(lib().ak_nlist_candidate_pairs(self._h, ctypes.byref(n), None, None, 0))
        i = np.empty(n.value, np.int32)
        j = np.empty(n.value, np.int32)
        check(lib().ak_nlist_candidate_pairs(self._h, ctypes.byref(n),
                                             i.ctypes.data_as(_lib._pi32),
                                             j.ctypes.data_as(_lib._pi32), len(i)))
        return i, j

    @property
    def rebuild_count(self):
        b = ctypes.c_int64(0)
        check(lib().ak_nlist_counters(self._h, ctypes.byref(b), None))
        return int(b.value)

    @property
    def step_count(self):
        return self._n_steps
