"""Site terms and the heat flux from the device (engine/site_terms.py).

For E = sum_i U_i with U_i a function of the bond vectors v_m = r_nbr - r_c
of its entries, ``SiteTerms`` holds e_site (N) = U_i and, per entry,
centre, nbr (M), vec (M, 3) and dUdv (M, 3) = dU_centre / dv_m; the engine's
bond kernels compute exactly these (their per-bond gradients) along with the
forces.  From them follow the forces, the stress and the heat flux

    J V = sum_i E_i v_i - sum_m v_m (dUdv_m . v_nbr),   E_i = K_i + U_i

(Fan et al., Phys. Rev. B 92, 094301 (2015), eq. 21), which ``heat_flux``
evaluates on the device in one pass.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, f64, i32, lib

EV_ANG_TO_AMU_ANG_FS2 = (1.60217663e-19 / 1.0e-10) / (1.66053906660e-27) * 1.0e-20


@dataclass
class SiteTerms:
    """Per-atom site energies and the per-entry data of the reference's
    SiteTerms (engine/site_terms.py:34-70), with the quantities derived from
    them: total energy, forces F_c += dUdv, F_nbr -= dUdv, the symmetrised
    stress -(1/V) sum_m dUdv_m (x) v_m, and the virial part of the flux."""

    e_site: np.ndarray
    centre: np.ndarray
    nbr: np.ndarray
    vec: np.ndarray
    dUdv: np.ndarray

    @property
    def energy(self) -> float:
        return float(np.sum(self.e_site))

    def forces(self) -> np.ndarray:
        n = len(self.e_site)
        return np.stack([np.bincount(self.centre, self.dUdv[:, a], minlength=n)
                         - np.bincount(self.nbr, self.dUdv[:, a], minlength=n)
                         for a in range(3)], axis=1)

    def stress(self, volume: float) -> np.ndarray:
        W = np.tensordot(self.dUdv, self.vec, axes=(0, 0))
        return -(W + W.T) / (2.0 * volume)

    def virial_flux(self, velocities: np.ndarray) -> np.ndarray:
        """-sum_m v_m (dUdv_m . vel_nbr) [eV A/fs], not divided by V."""
        proj = np.sum(self.dUdv * np.asarray(velocities, float)[self.nbr], axis=1)
        return -(self.vec.T @ proj)

    @staticmethod
    def combine(parts, n_atoms):
        """Site terms of a sum of potentials: energies added per atom, the
        entries of every part listed one after another."""
        present = [p for p in parts if p is not None]
        if not present:
            none = np.zeros(0, dtype=np.int64)
            return SiteTerms(np.zeros(n_atoms), none, none, np.zeros((0, 3)), np.zeros((0, 3)))
        joined = {f: np.concatenate([getattr(p, f) for p in present])
                  for f in ("centre", "nbr", "vec", "dUdv")}
        e = np.zeros(n_atoms)
        for p in present:
            e = e + p.e_site
        return SiteTerms(e, **joined)


def _inputs(potential, lattice, positions, atom_types):
    lat, plat = f64(lattice)
    pos, ppos = f64(positions)
    n = len(pos)
    if lat.shape != (3, 3) or pos.ndim != 2 or pos.shape[1] != 3 or len(atom_types) != n:
        raise ValueError("lattice (3, 3), positions (N, 3) and N atom types expected")
    inv, pinv = f64(np.linalg.inv(lat))
    spc, pspc = i32(potential.species_ids(atom_types))
    return n, (lat, inv, pos, spc), plat, pinv, ppos, pspc


def site_terms(potential, lattice, positions, atom_types, neigh_list=None) -> SiteTerms:
    """The site terms of one structure (potential.site_terms,
    engine/site_terms.py:179-475)."""
    from .neighborlist import NeighbourList

    n, keep, plat, pinv, ppos, pspc = _inputs(potential, lattice, positions, atom_types)
    nl = neigh_list.handle() if isinstance(neigh_list, NeighbourList) else None
    m = ctypes.c_int64(0)
    e = np.zeros(n)
    L = lib()
    check(L.ak_site_terms(potential.handle(), nl, n, plat, pinv, ppos, pspc, None, 0, None,
                          None, None, None, ctypes.byref(m)))
    M = int(m.value)
    centre = np.zeros(M, np.int32)
    nbr = np.zeros(M, np.int32)
    vec = np.zeros((M, 3))
    g = np.zeros((M, 3))
    check(L.ak_site_terms(potential.handle(), nl, n, plat, pinv, ppos, pspc,
                          e.ctypes.data_as(_lib._pf64), M,
                          centre.ctypes.data_as(_lib._pi32), nbr.ctypes.data_as(_lib._pi32),
                          vec.ctypes.data_as(_lib._pf64), g.ctypes.data_as(_lib._pf64),
                          ctypes.byref(m)))
    return SiteTerms(e, centre.astype(np.int64), nbr.astype(np.int64), vec, g)


def heat_flux(potential, lattice, positions, velocities, atom_types, masses, neigh_list=None):
    """(J (3,) [eV/(A^2 fs)], e_site (N,)) of one state, evaluated on the
    device (ThermalConductivityCalculator._compute_heat_flux with the site-term
    virial flux, solvers/thermal_conductivity.py:168-303)."""
    from .neighborlist import NeighbourList

    n, keep, plat, pinv, ppos, pspc = _inputs(potential, lattice, positions, atom_types)
    vel, pvel = f64(velocities)
    m, pm = f64(masses)
    if vel.shape != (n, 3) or m.shape != (n,):
        raise ValueError("velocities (N, 3) and masses (N,) expected")
    nl = neigh_list.handle() if isinstance(neigh_list, NeighbourList) else None
    J = np.zeros(3)
    e = np.zeros(n)
    check(lib().ak_heat_flux(potential.handle(), nl, n, plat, pinv, ppos, pvel, pm, pspc,
                             J.ctypes.data_as(_lib._pf64), e.ctypes.data_as(_lib._pf64)))
    return J, e
