"""Device-resident molecular dynamics over the HIP engine.

``nve_run`` integrates velocity Verlet entirely on the GPU
(NVESimulation.run, solvers/molecular_dynamics.py:1969-1996): half kick,
drift, wrap into the cell, Verlet-skin list test and rebuild
(engine/neighborlist.py:653-673), forces, half kick; one step is captured in
a hipGraph and replayed.  ``nvt_run`` adds the two half-steps of a
Nose-Hoover chain (NVTSimulation.run, :1113-1160).  E_pot, E_kin and the
conserved quantity of every step, and optionally the heat flux every k
steps, are recorded on the device and returned at the end.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .published import ATOMIC_MASS

K_B = 1.380649e-23 * 6.241509074e18                                        # eV/K
EV_ANG3_TO_GPA = 160.2176621
EV_ANG_TO_AMU_ANG_FS2 = (1.60217663e-19 / 1.0e-10) / (1.66053906660e-27) * 1.0e-20

# cell degrees of freedom of NPT (CELL_DOF_MODES, solvers/molecular_dynamics.py:2113)
CELL_DOF_MODES = ("iso", "aniso", "tri", "xy", "xy_iso", "z")

# thermostat state carried between device runs (AK_MD_STATE_LEN doubles)
STATE_LEN = 48
_ST = {"xi": (0, 16), "vxi": (16, 32), "p_eps": (32, 33), "xi_b": (33, 35), "vxi_b": (35, 37),
       "p_g": (37, 43), "heat": (43, 44), "rng_draws": (44, 45), "steps": (45, 46)}


def thermostat_state(xi=(), vxi=(), p_eps=0.0, xi_b=(0.0, 0.0), vxi_b=(0.0, 0.0), p_g=(),
                     heat=0.0, rng_draws=0):
    """A state array for ``md_run(state=...)`` from the quantities a
    reference checkpoint carries (NVTSimulation xi_nhc / vxi_nhc,
    NPTSimulation p_eps, xi1_b, xi2_b, vxi1_b, vxi2_b, p_g, baro_heat;
    solvers/molecular_dynamics.py:823-830, 2228, 2557-2564): chain positions
    and velocities (up to 16 links), the barostat momentum (iso) or the
    strain momenta of a flexible cell and the barostat chain, the heat a
    stochastic thermostat or barostat added and the random draws used."""
    st = np.zeros(STATE_LEN)
    for key, val in (("xi", xi), ("vxi", vxi), ("p_eps", p_eps), ("xi_b", xi_b),
                     ("vxi_b", vxi_b), ("p_g", p_g), ("heat", heat), ("rng_draws", rng_draws)):
        a, b = _ST[key]
        v = np.atleast_1d(np.asarray(val, float))
        if len(v) > b - a:
            raise ValueError(f"thermostat_state: {key} has more than {b - a} entries")
        st[a:a + len(v)] = v
    return st


def state_fields(state, chain=None):
    """The named quantities of a state array (chain entries cut to
    ``chain`` links when given)."""
    st = np.asarray(state, float)
    out = {k: st[a:b].copy() for k, (a, b) in _ST.items()}
    if chain is not None:
        out["xi"], out["vxi"] = out["xi"][:chain], out["vxi"][:chain]
    out["p_eps"], out["heat"] = float(out["p_eps"][0]), float(out["heat"][0])
    out["rng_draws"] = int(out["rng_draws"][0])
    out["steps"] = int(out["steps"][0])
    return out


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


def md_run(potential, lattice, positions, velocities, atom_types, dt=1.0, n_steps=1000,
           skin_width=0.3, masses=None, use_graph=True, thermostat=None, flux_every=0,
           state=None, fix_atoms=None, velocity_rescale_steps=0, frame_every=0,
           flux_groups=None, atom_flux=False):
    """n_steps of velocity Verlet on the device (``ak_md_run``).  thermostat:
    None (NVE), ("nhc", temperature K, tau fs, chain links), ("langevin",
    temperature, damping fs, seed), ("csvr", temperature, damping, seed) or
    ("npt", temperature, tau_nhc, chain links, pressure GPa, tau_baro fs[,
    cell_dof, baro_thermostat, seed]).  Returns a
    dict with the final positions and velocities, e_pot, e_kin, e_cons
    (E_tot, H_NHC, E_tot - Q or H_NPT) and pressure_gpa (P_inst = 2 E_kin /
    3V + Tr(sigma) / 3) per step (n_steps + 1, index 0 the initial state), the number of list rebuilds, the loop's wall time and,
    with flux_every > 0, the heat flux J (frames, 3) [eV/(A^2 fs)] of the
    state after every step s with s % flux_every == 0.  ``state`` (the
    "state" array a previous run returned, one of ``thermostat_state``, or
    None to start from rest) continues the thermostat, barostat and random
    stream where that run ended; the final one is returned as "state".
    fix_atoms: (N, 3) bool, True = frozen Cartesian component (every MD
    class's fix_atoms, solvers/molecular_dynamics.py:248-267, 341-361);
    velocity_rescale_steps: direct rescaling to the target temperature after
    each of the first that many steps (chain and NPT only).  frame_every > 0
    returns "frames" (n_steps // frame_every, N, 3) and "frame_lattices", the
    positions and cell after every frame_every steps, recorded on the
    device.  flux_groups = (labels (N,) or None, n_labels, (axis, bounds) or
    None) also records "flux_groups" (frames, n_labels, n_regions, 3), the
    heat flux of each group of atoms (per-atom heat currents; labels times
    regions along a cell axis between fractional bounds).  atom_flux=True
    also returns "atom_flux" (frames, N, 3) float32, the per-atom heat
    currents J_i / V of every frame."""
    lat = np.ascontiguousarray(lattice, float)
    pos = np.ascontiguousarray(positions, float).copy()
    vel = np.ascontiguousarray(velocities, float).copy()
    n = len(pos)
    m = np.ascontiguousarray(masses if masses is not None else masses_of(atom_types), float)
    inv = np.ascontiguousarray(np.linalg.inv(lat))
    spc = np.ascontiguousarray(potential.species_ids(atom_types), np.int32)
    ep = np.zeros(n_steps + 1)
    ek = np.zeros(n_steps + 1)
    ec = np.zeros(n_steps + 1)
    every = int(flux_every)
    frames = (n_steps + every - 1) // every if every > 0 else 0
    J = np.zeros((max(frames, 1), 3))
    th = None
    if thermostat is not None:
        kind = thermostat[0]
        if kind == "nhc":
            _, T, tau, chain = thermostat
            th = np.array([1.0, float(T), float(tau), float(chain), 0.0, 0.0, 0.0, 0.0, 0.0,
                           float(int(velocity_rescale_steps))])
        elif kind in ("langevin", "csvr"):
            _, T, damping, seed = thermostat
            th = np.array([2.0 if kind == "langevin" else 3.0, float(T), float(damping), 0.0,
                           float(int(seed)), 0.0, 0.0, 0.0, 0.0,
                           float(int(velocity_rescale_steps))])
        elif kind == "npt":
            _, T, tau, chain, p_gpa, tau_b, *rest = thermostat
            cell, baro, seed = (list(rest) + ["iso", "nhc", 0][len(rest):])[:3]
            if cell not in CELL_DOF_MODES:
                raise ValueError(f"NPT: cell_dof must be one of {' | '.join(CELL_DOF_MODES)}, "
                                 f"got {cell!r}.")
            if baro not in ("nhc", "langevin"):
                raise ValueError(f"NPT: baro_thermostat must be nhc | langevin, got {baro!r}.")
            th = np.array([4.0, float(T), float(tau), float(chain), float(int(seed)),
                           float(p_gpa), float(tau_b), float(CELL_DOF_MODES.index(cell)),
                           1.0 if baro == "langevin" else 0.0,
                           float(int(velocity_rescale_steps))])
        else:
            raise ValueError(f"thermostat: unknown kind {kind!r} (nhc, langevin, csvr, npt)")
    lat_out = np.zeros((3, 3))
    vol = np.zeros(n_steps + 1)
    prs = np.zeros(n_steps + 1)
    st = np.zeros(STATE_LEN) if state is None else np.array(state, float).reshape(STATE_LEN)
    if th is None and velocity_rescale_steps:
        raise ValueError("velocity_rescale_steps needs a thermostat (nhc or npt)")
    fix = None
    if fix_atoms is not None:
        fa = np.asarray(fix_atoms, dtype=bool)
        if fa.shape != (n, 3):
            raise ValueError(f"fix_atoms must have shape ({n}, 3), got {fa.shape}.")
        if fa.any():
            fix = np.ascontiguousarray(fa, dtype=np.uint8)
    fe = int(frame_every)
    if fe < 0:
        raise ValueError("frame_every must be >= 0")
    n_fr = n_steps // fe if fe else 0
    fr = np.zeros((max(n_fr, 1), n, 3))
    frl = np.zeros((max(n_fr, 1), 3, 3))
    glab, n_lab, g_axis, gb, n_reg, gJ = None, 1, -1, None, 1, None
    if flux_groups is not None:
        if every <= 0:
            raise ValueError("flux_groups needs flux_every > 0")
        labels, n_lab, region = flux_groups
        n_lab = int(n_lab)
        if labels is not None:
            glab = np.ascontiguousarray(labels, np.int32)
            if glab.shape != (n,):
                raise ValueError("flux_groups labels must have one entry per atom")
        if region is not None:
            g_axis = int(region[0])
            gb = np.ascontiguousarray(region[1], float)
            n_reg = len(gb) - 1
        gJ = np.zeros((max(frames, 1), n_lab * n_reg, 3))
    aJ = None
    if atom_flux:
        if every <= 0:
            raise ValueError("atom_flux needs flux_every > 0")
        aJ = np.zeros((max(frames, 1), n, 3), np.float32)
    nbc = np.zeros(n_steps + 1, np.int32)
    lath = np.zeros((n_steps + 1, 3, 3))
    pth = np.zeros((n_steps + 1, 6))
    nreb = ctypes.c_int32(0)
    secs = ctypes.c_double(0.0)
    P = _lib._pf64
    check(lib().ak_md_run(potential.handle(), n, lat.ctypes.data_as(P), inv.ctypes.data_as(P),
                          pos.ctypes.data_as(P), vel.ctypes.data_as(P), m.ctypes.data_as(P),
                          spc.ctypes.data_as(_lib._pi32), float(dt), int(n_steps),
                          float(skin_width), 1 if use_graph else 0,
                          th.ctypes.data_as(P) if th is not None else None,
                          st.ctypes.data_as(P),
                          fix.ctypes.data_as(_lib._pu8) if fix is not None else None, every,
                          fe, fr.ctypes.data_as(P) if fe else None,
                          frl.ctypes.data_as(P) if fe else None,
                          ep.ctypes.data_as(P), ek.ctypes.data_as(P), ec.ctypes.data_as(P),
                          J.ctypes.data_as(P), lat_out.ctypes.data_as(P), vol.ctypes.data_as(P),
                          prs.ctypes.data_as(P), nbc.ctypes.data_as(_lib._pi32),
                          lath.ctypes.data_as(P), pth.ctypes.data_as(P),
                          glab.ctypes.data_as(_lib._pi32) if glab is not None else None,
                          n_lab, g_axis, gb.ctypes.data_as(P) if gb is not None else None, n_reg,
                          gJ.ctypes.data_as(P) if gJ is not None else None,
                          aJ.ctypes.data_as(_lib._pf32) if aJ is not None else None,
                          ctypes.byref(nreb), ctypes.byref(secs)))
    et = ep + ek
    out = {"positions": pos, "velocities": vel, "e_pot": ep, "e_kin": ek, "e_tot": et,
           "e_cons": ec, "drift": et - et[0], "n_rebuilds": int(nreb.value),
           "seconds": float(secs.value), "state": st, "rebuild_counts": nbc}
    if every > 0:
        out["flux"] = J[:frames]
        if gJ is not None:
            out["flux_groups"] = gJ[:frames].reshape(frames, n_lab, n_reg, 3)
        if aJ is not None:
            out["atom_flux"] = aJ[:frames]
    if fe:
        out["frames"] = fr[:n_fr]
        out["frame_lattices"] = frl[:n_fr]
    out["pressure_gpa"] = prs * EV_ANG3_TO_GPA
    if th is not None and th[0] == 4.0:
        out["lattice"] = lat_out
        out["volume"] = vol
        out["lattices"] = lath
        if th[7] != 0.0:
            out["pressure_tensor_gpa"] = pth * EV_ANG3_TO_GPA
    return out


def n_dof_of(n_atoms, fix_atoms=None):
    """Degrees of freedom of the temperature (_compute_n_dof,
    solvers/molecular_dynamics.py:341-361): 3N - 3, or 3N minus the frozen
    components when any is frozen."""
    n_fixed = int(np.sum(np.asarray(fix_atoms, dtype=bool))) if fix_atoms is not None else 0
    return max(3 * n_atoms - 3, 1) if n_fixed == 0 else max(3 * n_atoms - n_fixed, 1)


def nve_run(potential, lattice, positions, velocities, atom_types, dt=1.0, n_steps=1000,
            skin_width=0.3, masses=None, use_graph=True, flux_every=0, fix_atoms=None):
    """Run n_steps of NVE on the device (NVESimulation.run,
    solvers/molecular_dynamics.py:1969-1996).  Returns md_run's dict;
    drift = e_tot - e_tot[0]."""
    return md_run(potential, lattice, positions, velocities, atom_types, dt, n_steps,
                  skin_width, masses, use_graph, None, flux_every, fix_atoms=fix_atoms)


def nvt_run(potential, lattice, positions, velocities, atom_types, temperature, dt=1.0,
            n_steps=1000, tau_nhc=100.0, nhc_chain=3, skin_width=0.5, masses=None,
            use_graph=True, flux_every=0, state=None, fix_atoms=None, velocity_rescale_steps=0):
    """NVT with an M-link Nose-Hoover chain on the device (NVTSimulation.run,
    solvers/molecular_dynamics.py:1113-1167, _nhc_step :635-727; defaults of
    NVTSimulation: tau_nhc 100 fs, 3 links, skin 0.5 A).  e_cons is H_NHC
    (:947-969); T_inst = 2 E_kin / (n_dof k_B) with ``n_dof_of``."""
    out = md_run(potential, lattice, positions, velocities, atom_types, dt, n_steps,
                 skin_width, masses, use_graph, ("nhc", temperature, tau_nhc, nhc_chain),
                 flux_every, state, fix_atoms, velocity_rescale_steps)
    out["T"] = 2.0 * out["e_kin"] / (n_dof_of(len(positions), fix_atoms) * K_B)
    return out


def langevin_run(potential, lattice, positions, velocities, atom_types, temperature, dt=1.0,
                 n_steps=1000, damping=100.0, seed=42, skin_width=0.5, masses=None,
                 use_graph=True, flux_every=0, state=None, fix_atoms=None):
    """NVT with a Langevin thermostat integrated by BAOAB on the device
    (LangevinSimulation, solvers/molecular_dynamics.py:1511-1587): e_cons is
    E_tot - Q, Q the heat the O steps and the centre-of-mass removal added.
    The noise is Philox4x32-10 keyed by ``seed`` (the reference draws from
    numpy's generator: same process, other realisations)."""
    out = md_run(potential, lattice, positions, velocities, atom_types, dt, n_steps,
                 skin_width, masses, use_graph, ("langevin", temperature, damping, seed),
                 flux_every, state, fix_atoms)
    out["T"] = 2.0 * out["e_kin"] / (n_dof_of(len(positions), fix_atoms) * K_B)
    return out


def csvr_run(potential, lattice, positions, velocities, atom_types, temperature, dt=1.0,
             n_steps=1000, damping=100.0, seed=42, skin_width=0.5, masses=None,
             use_graph=True, flux_every=0, state=None, fix_atoms=None):
    """NVT with the Bussi-Donadio-Parrinello velocity rescaling on the device
    (CSVRSimulation, solvers/molecular_dynamics.py:1589-1687): R B A F B R,
    e_cons = E_tot - Q; random numbers as in :func:`langevin_run`."""
    out = md_run(potential, lattice, positions, velocities, atom_types, dt, n_steps,
                 skin_width, masses, use_graph, ("csvr", temperature, damping, seed), flux_every,
                 state, fix_atoms)
    out["T"] = 2.0 * out["e_kin"] / (n_dof_of(len(positions), fix_atoms) * K_B)
    return out


def npt_run(potential, lattice, positions, velocities, atom_types, temperature, pressure=0.0,
            dt=1.0, n_steps=1000, tau_nhc=100.0, nhc_chain=3, tau_baro=1000.0, skin_width=0.5,
            masses=None, use_graph=True, state=None, cell_dof="iso", baro_thermostat="nhc",
            seed=42, fix_atoms=None, velocity_rescale_steps=0):
    """NPT with the MTTK barostat and the particle Nose-Hoover chain, on the
    device (NPTSimulation, solvers/molecular_dynamics.py:2141-2692; defaults
    tau_nhc 100 fs, 3 links, tau_baro 1000 fs, cell_dof "iso",
    baro_thermostat "nhc", skin 0.5 A).  cell_dof: "iso" (volume only),
    "aniso", "tri", "xy", "xy_iso" or "z" (the strain generators of
    ADR-0036); baro_thermostat: "nhc" (a 2-link chain on the barostat
    momenta) or "langevin" (an Ornstein-Uhlenbeck step with friction time
    tau_baro, noise from the Philox stream keyed by ``seed``; its heat is
    subtracted from H_NPT).  pressure in GPa.  Returns md_run's dict with
    the final "lattice", "volume" [A^3] and "pressure_gpa" (P_inst) per
    step, e_cons = H_NPT and T.  A perpendicular cell width may shrink by up
    to 5 % in one call."""
    out = md_run(potential, lattice, positions, velocities, atom_types, dt, n_steps,
                 skin_width, masses, use_graph,
                 ("npt", temperature, tau_nhc, nhc_chain, pressure, tau_baro, cell_dof,
                  baro_thermostat, seed), 0, state, fix_atoms, velocity_rescale_steps)
    out["T"] = 2.0 * out["e_kin"] / (n_dof_of(len(positions), fix_atoms) * K_B)
    return out


def n_samples(n_steps, every, phase=0):
    """Records of the steps s < n_steps with (s + phase) % every == 0."""
    if every <= 0 or n_steps <= 0:
        return 0
    return (n_steps - 1 + phase) // every - (1 if phase > 0 else 0) + 1


def nemd_run(potential, lattice, positions, velocities, atom_types, bath, T_hot, T_cold,
             damping, dt=1.0, n_steps=1000, seed=42, skin_width=0.5, masses=None,
             use_graph=True, state=None, sample_every=0, sample_phase=0, axis=2, n_bins=20):
    """n_steps of NEMD on the device (``ak_nemd_run``; NEMDSimulation.run's
    step, solvers/nemd.py:184-193): velocity Verlet, then the exact
    Ornstein-Uhlenbeck update of the bath atoms (bath (N,): 0 none, 1 hot at
    T_hot, 2 cold at T_cold; c = exp(-dt / damping)), the noise Philox4x32-10
    keyed by ``seed``.  Returns the final positions and velocities, e_pot,
    e_kin, e_cons (E_tot - Q), "bath" (n_steps + 1, 4: dQ_hot, dQ_cold of the
    step, K_hot, K_cold after it), the rebuild count, the wall time, the
    final "state" and, with sample_every > 0, "bins" (samples, n_bins, 5):
    per equal bin of the fractional coordinate along ``axis``, sum_i J_i / V
    (3), the kinetic energy and the atom count after the steps s with
    (s + sample_phase) % sample_every == 0."""
    lat = np.ascontiguousarray(lattice, float)
    pos = np.ascontiguousarray(positions, float).copy()
    vel = np.ascontiguousarray(velocities, float).copy()
    n = len(pos)
    m = np.ascontiguousarray(masses if masses is not None else masses_of(atom_types), float)
    inv = np.ascontiguousarray(np.linalg.inv(lat))
    spc = np.ascontiguousarray(potential.species_ids(atom_types), np.int32)
    b = np.ascontiguousarray(bath, dtype=np.uint8)
    if b.shape != (n,):
        raise ValueError(f"bath must have one entry per atom ({n}), got shape {b.shape}")
    se, ph = int(sample_every), int(sample_phase)
    if se < 0 or (se > 0 and not 0 <= ph < se):
        raise ValueError("sample_every must be >= 0 and sample_phase in [0, sample_every)")
    ns = n_samples(n_steps, se, ph)
    bins = np.zeros((max(ns, 1), int(n_bins), 5))
    ep = np.zeros(n_steps + 1)
    ek = np.zeros(n_steps + 1)
    ec = np.zeros(n_steps + 1)
    rows = np.zeros((n_steps + 1, 4))
    st = np.zeros(STATE_LEN) if state is None else np.array(state, float).reshape(STATE_LEN)
    nreb = ctypes.c_int32(0)
    secs = ctypes.c_double(0.0)
    P = _lib._pf64
    check(lib().ak_nemd_run(potential.handle(), n, lat.ctypes.data_as(P), inv.ctypes.data_as(P),
                            pos.ctypes.data_as(P), vel.ctypes.data_as(P), m.ctypes.data_as(P),
                            spc.ctypes.data_as(_lib._pi32), float(dt), int(n_steps),
                            float(skin_width), 1 if use_graph else 0,
                            b.ctypes.data_as(_lib._pu8), float(T_hot), float(T_cold),
                            float(damping), float(int(seed)), st.ctypes.data_as(P), se, ph,
                            int(axis), int(n_bins), ep.ctypes.data_as(P), ek.ctypes.data_as(P),
                            ec.ctypes.data_as(P), rows.ctypes.data_as(P),
                            bins.ctypes.data_as(P) if se > 0 else None, ctypes.byref(nreb),
                            ctypes.byref(secs)))
    out = {"positions": pos, "velocities": vel, "e_pot": ep, "e_kin": ek, "e_cons": ec,
           "bath": rows, "n_rebuilds": int(nreb.value), "seconds": float(secs.value),
           "state": st}
    if se > 0:
        out["bins"] = bins[:ns]
    return out
