"""Device-resident NVE molecular dynamics over the HIP engine.

``nve_run`` integrates velocity Verlet entirely on the GPU
(NVESimulation.run, solvers/molecular_dynamics.py:1969-1996): half kick,
drift, wrap into the cell, Verlet-skin list test and rebuild
(engine/neighborlist.py:653-673), forces, half kick; one step is captured in
a hipGraph and replayed.  E_pot and E_kin of every step are recorded on the
device and returned at the end.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .published import ATOMIC_MASS

K_B = 1.380649e-23 * 6.241509074e18                                        # eV/K
EV_ANG_TO_AMU_ANG_FS2 = (1.60217663e-19 / 1.0e-10) / (1.66053906660e-27) * 1.0e-20


def maxwell_boltzmann(T, masses, seed):
    """Velocities [A/fs] at temperature T as _init_velocities_mb
    (solvers/molecular_dynamics.py:195-221): normal draws scaled by
    sqrt(k_B T conv / m), centre-of-mass momentum removed, rescaled to T with
    3N - 3 degrees of freedom."""
    masses = np.asarray(masses, float)
    rng = np.random.default_rng(seed)
    N = len(masses)
    sigma = np.sqrt(K_B * T * EV_ANG_TO_AMU_ANG_FS2 / masses)
    v = rng.standard_normal((N, 3)) * sigma[:, None]
    v -= np.sum(masses[:, None] * v, axis=0) / masses.sum()
    e_kin = 0.5 * float(np.sum(masses[:, None] * v ** 2)) / EV_ANG_TO_AMU_ANG_FS2
    T_act = 2.0 * e_kin / (max(3 * N - 3, 1) * K_B)
    if T_act > 1e-12:
        v *= np.sqrt(T / T_act)
    return v


def masses_of(atom_types):
    return np.array([ATOMIC_MASS.get(t, 1.0) for t in atom_types], float)


def nve_run(potential, lattice, positions, velocities, atom_types, dt=1.0, n_steps=1000,
            skin_width=0.3, masses=None, use_graph=True):
    """Run n_steps of NVE on the device.  Returns a dict with the final
    positions and velocities, e_pot, e_kin and e_tot per step (n_steps + 1,
    index 0 the initial state), drift = e_tot - e_tot[0], the number of list
    rebuilds and the wall time of the stepping loop in seconds."""
    lat = np.ascontiguousarray(lattice, float)
    pos = np.ascontiguousarray(positions, float).copy()
    vel = np.ascontiguousarray(velocities, float).copy()
    n = len(pos)
    m = np.ascontiguousarray(masses if masses is not None else masses_of(atom_types), float)
    inv = np.ascontiguousarray(np.linalg.inv(lat))
    spc = np.ascontiguousarray(potential.species_ids(atom_types), np.int32)
    ep = np.zeros(n_steps + 1)
    ek = np.zeros(n_steps + 1)
    nreb = ctypes.c_int32(0)
    secs = ctypes.c_double(0.0)
    P = _lib._pf64
    check(lib().ak_md_nve(potential.handle(), n, lat.ctypes.data_as(P), inv.ctypes.data_as(P),
                          pos.ctypes.data_as(P), vel.ctypes.data_as(P), m.ctypes.data_as(P),
                          spc.ctypes.data_as(_lib._pi32), float(dt), int(n_steps),
                          float(skin_width), 1 if use_graph else 0, ep.ctypes.data_as(P),
                          ek.ctypes.data_as(P), ctypes.byref(nreb), ctypes.byref(secs)))
    et = ep + ek
    return {"positions": pos, "velocities": vel, "e_pot": ep, "e_kin": ek, "e_tot": et,
            "drift": et - et[0], "n_rebuilds": int(nreb.value), "seconds": float(secs.value)}
