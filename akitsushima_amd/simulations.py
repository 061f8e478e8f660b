"""The reference's molecular-dynamics classes over the device integrators.

``NVESimulation``, ``NVTSimulation``, ``LangevinSimulation``,
``CSVRSimulation`` and ``NPTSimulation`` take the constructor arguments of
solvers/molecular_dynamics.py (:829-930, 1320-1333, 1737-1830, 2233-2400)
and produce what their ``run`` produces: ``thermo.log`` with the same columns
and number formats (rows labelled by completed steps, every
``thermo_every``), the XDATCAR trajectory (frames every ``traj_every`` steps
from ``equil_steps`` on), the final POSCAR, and the same result dictionary.
The steps themselves run on the device in one captured-graph loop per call
(``md.md_run``): energies, pressures and rebuild counts of every step and
the trajectory frames are recorded there and read back once.

Checkpoints (``run(checkpoint=path, checkpoint_every=k)``) use the
reference's npz layout (positions, velocities, lattice, a JSON ``meta`` with
step, settings and thermostat scalars, ``state_*`` arrays, ``list_*``
production series, ``tlog_*`` thermo rows; ADR-0031), so a checkpoint the
reference wrote -- also one of 0.3.0a7 with the chain as scalars xi1 .. vxi2
-- is continued here.  The device's own thermostat state rides along in
``meta["extra"]["device_state"]``.  The stochastic thermostats draw from the
device's Philox stream, not numpy's generator: a reference checkpoint of a
Langevin or CSVR run continues with other (equally valid) random numbers.

Not carried over: the Buckingham/Coulomb keyword path (use a force function
of the device potentials), trajectory formats other than XDATCAR, the
ensemble-uncertainty column, plots and timing tables.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from .md import (CELL_DOF_MODES, K_B, STATE_LEN, masses_of, maxwell_boltzmann, md_run,
                 n_dof_of, state_fields, thermostat_state)

EV_ANG3_TO_GPA = 160.2176621
_LEGACY = ("buck_params", "charges", "ewald_alpha")


def _potential_of(force_fn):
    pots = getattr(getattr(force_fn, "engine", None), "_potentials", None)
    if pots is not None:
        if len(pots) != 1:
            raise ValueError("the device MD runs one potential; the engine holds "
                             f"{len(pots)}")
        return pots[0]
    if force_fn is not None and hasattr(force_fn, "handle"):
        return force_fn
    raise ValueError("force_fn must be a force function of make_force_fn or a device "
                     "potential (the Buckingham/Coulomb keyword path is not available)")


def _species_order(atom_types):
    seen, order = set(), []
    for t in atom_types:
        if t not in seen:
            order.append(t)
            seen.add(t)
    return order


def _write_cell_header(fh, lattice, atom_types, comment):
    order = _species_order(atom_types)
    fh.write(f"{comment}\n1.0\n")
    for vec in lattice:
        fh.write(f"  {vec[0]:20.14f}  {vec[1]:20.14f}  {vec[2]:20.14f}\n")
    fh.write("  " + "  ".join(order) + "\n")
    fh.write("  " + "  ".join(str(list(atom_types).count(s)) for s in order) + "\n")


def _write_frac(fh, positions, lattice, atom_types):
    frac = positions @ np.linalg.inv(lattice)
    for sp in _species_order(atom_types):
        for i, t in enumerate(atom_types):
            if t == sp:
                p = frac[i]
                fh.write(f"  {p[0]:12.8f}  {p[1]:12.8f}  {p[2]:12.8f}\n")


def write_poscar(path, lattice, positions, atom_types, title="Written by akitsushima_amd"):
    """VASP 5 POSCAR in direct coordinates, grouped by species
    (io/poscar.py:101-175)."""
    with open(path, "w") as fh:
        _write_cell_header(fh, lattice, atom_types, title)
        fh.write("Direct\n")
        _write_frac(fh, positions, lattice, atom_types)


class _DeviceMD:
    _TASK = "nvt"
    _ENSEMBLE = "nvt"
    _POSCAR = "POSCAR_final_MD"
    _POSCAR_TITLE = "MD final frame"
    _TRAJ_COMMENT = "akitsushima.py MD"
    _SCALARS = ()

    def __init__(self, lattice, positions, atom_types, force_fn=None, r_cut=None, *,
                 temperature=300.0, timestep=1.0, n_steps=10000, equil_steps=0, traj_every=10,
                 traj_format="xdatcar", thermo_every=100, output_dir=".", skin_width=0.5,
                 seed=42, verbose=True, masses=None, fix_atoms=None, timing=False,
                 velocity_rescale_steps=0, flush_every_step=True, fsync_every_step=False,
                 initial_velocities=None, save_plots=True, nhc_chain=3, tau_nhc=100.0,
                 **legacy):
        cls = type(self).__name__
        for k, v in legacy.items():
            if k in _LEGACY or k in ("coulomb_method", "pppm_accuracy", "pppm_mesh",
                                     "pppm_order"):
                if k in _LEGACY and v is not None:
                    raise ValueError(f"{cls}: {k} (the Buckingham/Coulomb path) is not "
                                     "available on the device engine; pass force_fn")
            else:
                raise TypeError(f"{cls}: unexpected keyword argument {k!r}")
        if not timestep > 0.0:
            raise ValueError(f"{cls}: timestep must be > 0, got {timestep}.")
        if n_steps < 0 or equil_steps < 0:
            raise ValueError(f"{cls}: n_steps and equil_steps must be >= 0.")
        if traj_every < 1 or thermo_every < 1:
            raise ValueError(f"{cls}: traj_every and thermo_every must be >= 1.")
        if traj_format != "xdatcar":
            raise ValueError(f"{cls}: traj_format {traj_format!r} is not available here "
                             "(xdatcar only)")
        self.potential = _potential_of(force_fn)
        self.r_cut = float(self.potential.r_cut if r_cut is None else r_cut)
        self.lattice = np.array(lattice, float)
        self.positions = np.array(positions, float)
        self.atom_types = list(atom_types)
        self.N = len(self.positions)
        self.temperature = float(temperature)
        self.dt = float(timestep)
        self.n_steps = int(n_steps)
        self.equil_steps = int(equil_steps)
        self.traj_every = int(traj_every)
        self.thermo_every = int(thermo_every)
        self.output_dir = output_dir
        self.skin_width = float(skin_width)
        self.seed = int(seed)
        self.verbose = bool(verbose)
        self.velocity_rescale_steps = int(velocity_rescale_steps)
        self.nhc_chain = int(nhc_chain)
        self.tau_nhc = float(tau_nhc)
        self.masses = (np.asarray(masses, float) if masses is not None
                       else masses_of(self.atom_types))
        if fix_atoms is not None:
            fa = np.asarray(fix_atoms, dtype=bool)
            if fa.shape != (self.N, 3):
                raise ValueError(f"{cls}: fix_atoms must have shape ({self.N}, 3), "
                                 f"got {fa.shape}.")
            self._fix = fa
        else:
            self._fix = np.zeros((self.N, 3), bool)
        self.n_dof = n_dof_of(self.N, self._fix)
        if initial_velocities is not None:
            v = np.array(initial_velocities, float)
            if v.shape != (self.N, 3):
                raise ValueError(f"{cls}: initial_velocities must have shape ({self.N}, 3)")
        else:
            v = maxwell_boltzmann(self.temperature, self.masses, self.seed)
        v[self._fix] = 0.0
        self.velocities = v
        self._state = np.zeros(STATE_LEN)
        self._thermo_log = []

    # ---- per-class pieces -------------------------------------------------
    def _thermostat(self):
        return None

    def _header(self):
        return (f"{'Step':>8}  {'Time(ps)':>10}  {'E_pot(eV)':>14}  {'E_kin(eV)':>12}  "
                f"{'E_tot(eV)':>12}  {'T_inst(K)':>10}  {'H_NHC(eV)':>14}  {'NL_builds':>10}")

    def _row(self, c, r, k, nb):
        t_ps = c * self.dt * 1e-3
        ep, ek = r["e_pot"][k], r["e_kin"][k]
        T = 2.0 * ek / (self.n_dof * K_B)
        text = (f"{c:8d}  {t_ps:10.4f}  {ep:14.6f}  {ek:12.6f}  {ep + ek:12.6f}  {T:10.2f}  "
                f"{r['e_cons'][k]:14.6f}  {nb:10d}")
        return text, {"step": c, "time_ps": t_ps, "e_pot": ep, "e_kin": ek, "e_tot": ep + ek,
                      "T_inst": T, "H_nhc": r["e_cons"][k], "nl_builds": nb}

    def _settings(self):
        out = {"ensemble": self._ENSEMBLE, "n_atoms": self.N, "atom_types": self.atom_types,
               "timestep": self.dt, "equil_steps": self.equil_steps,
               "thermo_every": self.thermo_every, "traj_every": self.traj_every,
               "velocity_rescale_steps": self.velocity_rescale_steps}
        if self.nhc_chain != 2 and self._ENSEMBLE in ("nvt", "npt"):
            out["nhc_chain"] = self.nhc_chain
        return out

    # ---- checkpoints (the reference's npz layout, ADR-0031) -----------------
    def _ck_load(self, path):
        z = np.load(path, allow_pickle=False)
        meta = json.loads(str(z["meta"]))
        cls = type(self).__name__
        st = meta.get("settings", {})
        if st.get("ensemble") not in (None, self._ENSEMBLE):
            raise ValueError(f"{cls}: the checkpoint {path} is of ensemble "
                             f"{st.get('ensemble')!r}")
        if st.get("n_atoms", self.N) != self.N:
            raise ValueError(f"{cls}: the checkpoint {path} holds {st.get('n_atoms')} atoms")
        if self._ENSEMBLE in ("nvt", "npt") and st.get("nhc_chain", 2) != self.nhc_chain:
            raise ValueError(f"{cls}: the checkpoint {path} was written with nhc_chain = "
                             f"{st.get('nhc_chain', 2)}, this run has nhc_chain = "
                             f"{self.nhc_chain}")
        step = int(meta["step"])
        if step > self.equil_steps + self.n_steps:
            raise ValueError(f"{cls}: the checkpoint {path} is at step {step}, beyond "
                             f"equil_steps + n_steps = {self.equil_steps + self.n_steps}")
        self.positions = np.array(z["positions"], float)
        self.velocities = np.array(z["velocities"], float)
        self.lattice = np.array(z["lattice"], float)
        extra = meta.get("extra", {}) or {}
        if "device_state" in extra:
            self._state = np.array(extra["device_state"], float)
        else:
            sc = meta.get("scalars", {})
            if "state_xi_nhc" in z.files:
                xi, vxi = z["state_xi_nhc"], z["state_vxi_nhc"]
            elif "xi1" in sc:
                xi, vxi = [sc["xi1"], sc["xi2"]], [sc["vxi1"], sc["vxi2"]]
            else:
                xi, vxi = [], []
            kw = {}
            if "p_eps" in sc:
                kw = dict(p_eps=sc["p_eps"], xi_b=(sc["xi1_b"], sc["xi2_b"]),
                          vxi_b=(sc["vxi1_b"], sc["vxi2_b"]))
            if "state_p_g" in z.files:
                kw["p_g"] = z["state_p_g"]
            heat = float(extra.get("Q", 0.0))
            if "state_baro_heat" in z.files:
                heat = float(np.asarray(z["state_baro_heat"]).ravel()[0])
            self._state = thermostat_state(xi=xi, vxi=vxi, heat=heat, **kw)
        self._state[45] = step
        lists = {k[5:]: list(np.asarray(z[k], float)) for k in z.files if k.startswith("list_")}
        if "thermo_log_keys" in meta:
            cols = [np.asarray(z["tlog_" + k]).tolist() for k in meta["thermo_log_keys"]]
            self._thermo_log = [dict(zip(meta["thermo_log_keys"], row)) for row in zip(*cols)]
        return (step, lists, extra, int(meta.get("frame", 1)), meta.get("labels"),
                meta.get("sizes", {}))

    def _ck_save(self, path, step, lists, extra, frame, sizes):
        meta = {"step": int(step), "settings": self._settings(),
                "scalars": self._ck_scalars(), "extra": dict(extra),
                "frame": int(frame), "labels": "completed", "sizes": sizes}
        meta["extra"]["device_state"] = [float(x) for x in self._state]
        arrays = {"positions": self.positions, "velocities": self.velocities,
                  "lattice": self.lattice}
        f = state_fields(self._state, self.nhc_chain)
        if self._ENSEMBLE in ("nvt", "npt"):
            arrays["state_xi_nhc"] = f["xi"]
            arrays["state_vxi_nhc"] = f["vxi"]
        for k, v in lists.items():
            arrays["list_" + k] = np.asarray(v, float)
        if self._thermo_log:
            keys = list(self._thermo_log[0])
            meta["thermo_log_keys"] = keys
            for k in keys:
                arrays["tlog_" + k] = np.array([row[k] for row in self._thermo_log])
        tmp = path + ".tmp.npz"
        np.savez(tmp, meta=np.array(json.dumps(meta)), **arrays)
        os.replace(tmp, path)

    def _ck_scalars(self):
        return {}

    def _ck_extra(self):
        return {}

    def _ck_restore_extra(self, extra):
        pass

    # ---- the run ------------------------------------------------------------
    def _chunk(self, n, frame_every):
        return md_run(self.potential, self.lattice, self.positions, self.velocities,
                      self.atom_types, dt=self.dt, n_steps=n, skin_width=self.skin_width,
                      masses=self.masses, thermostat=self._thermostat(), state=self._state,
                      fix_atoms=self._fix if self._fix.any() else None,
                      velocity_rescale_steps=self._rescale_steps(), frame_every=frame_every)

    def _rescale_steps(self):
        return 0

    def _variable_cell(self):
        return False

    def run(self, checkpoint=None, checkpoint_every=0, checkpoint_meta=None):
        t_start = time.perf_counter()
        os.makedirs(self.output_dir, exist_ok=True)
        total = self.equil_steps + self.n_steps
        traj_path = os.path.join(self.output_dir, f"XDATCAR_{self._TASK}")
        thermo_path = os.path.join(self.output_dir, "thermo.log")
        start, lists, extra, frame = 0, {}, {}, 1
        resumed = bool(checkpoint) and os.path.isfile(checkpoint)
        if resumed:
            start, lists, extra, frame, labels, sizes = self._ck_load(checkpoint)
            self._ck_restore_extra(extra)
            # output written after the checkpoint is cut off (truncate_outputs)
            for key, path in (("thermo", thermo_path), ("traj", traj_path)):
                if key in sizes and os.path.isfile(path) and \
                        os.path.getsize(path) > int(sizes[key]):
                    with open(path, "r+b") as fh:
                        fh.truncate(int(sizes[key]))
            if labels != "completed":
                # rows written before ADR-0048 were labelled by completed steps - 1
                with open(thermo_path, "a") as tfh:
                    tfh.write("# ADR-0048: rows below are labelled by completed steps (row 0 = "
                              "initial state); rows above, written before 0.4.0a4, by "
                              "completed steps - 1\n")
        for k in self._list_keys():
            lists.setdefault(k, [])
        n_builds = 1 + int(extra.get("nl_rebuilds", 0))
        var_cell = self._variable_cell()
        with open(thermo_path, "a" if resumed else "w") as tfh, \
                open(traj_path, "a" if resumed else "w") as xfh:
            if not resumed:
                tfh.write(self._header() + "\n")
                if not var_cell:
                    _write_cell_header(xfh, self.lattice, self.atom_types,
                                       f"{self._TRAJ_COMMENT}  T={self.temperature}K")
            c = start
            first = not resumed
            while c < total or first:
                stop = total
                if checkpoint and checkpoint_every:
                    stop = min(stop, (c // checkpoint_every + 1) * checkpoint_every)
                aligned = c % self.traj_every == 0
                if not aligned:
                    stop = min(stop, (c // self.traj_every + 1) * self.traj_every)
                n = stop - c
                r = self._chunk(n, self.traj_every if aligned else 0)
                nb = r["rebuild_counts"]
                k0 = 0 if first else 1
                for k in range(k0, n + 1):
                    cc = c + k
                    if k >= 1 and cc - 1 >= self.equil_steps:
                        self._collect(lists, r, k)
                    if cc % self.thermo_every == 0:
                        text, row = self._row(cc, r, k, n_builds + int(nb[k]))
                        tfh.write(text + "\n")
                        self._thermo_log.append(row)
                        if self.verbose:
                            print("  " + text)
                    if cc >= self.equil_steps and cc % self.traj_every == 0:
                        if k == 0:
                            pos, lat = self.positions, self.lattice
                        elif k == n:
                            pos, lat = r["positions"], r.get("lattice", self.lattice)
                        else:
                            f = k // self.traj_every - 1
                            pos, lat = r["frames"][f], r["frame_lattices"][f]
                        if var_cell:
                            _write_cell_header(xfh, lat, self.atom_types,
                                               f"{self._TRAJ_COMMENT}  step={cc}")
                        xfh.write(f"Direct configuration={frame:6d}\n")
                        _write_frac(xfh, pos, lat, self.atom_types)
                        frame += 1
                self.positions = r["positions"]
                self.velocities = r["velocities"]
                if "lattice" in r:
                    self.lattice = r["lattice"]
                self._state = r["state"]
                n_builds += int(nb[n])
                self._last = r
                c = stop
                first = False
                tfh.flush()
                xfh.flush()
                if checkpoint and (checkpoint_every and c % checkpoint_every == 0 or c == total):
                    extra = dict(extra, nl_rebuilds=n_builds - 1, **self._ck_extra())
                    self._ck_save(checkpoint, c, lists, extra, frame,
                                  {"thermo": os.path.getsize(thermo_path),
                                   "traj": os.path.getsize(traj_path)})
        write_poscar(os.path.join(self.output_dir, self._POSCAR), self.lattice, self.positions,
                     self.atom_types, title=self._POSCAR_TITLE)
        out = self._summary(lists)
        out["timing"] = {"enabled": False, "wall_s": time.perf_counter() - t_start}
        return out

    def _list_keys(self):
        return ("T", "e_pot", "e_kin")

    def _collect(self, lists, r, k):
        lists["T"].append(2.0 * r["e_kin"][k] / (self.n_dof * K_B))
        lists["e_pot"].append(r["e_pot"][k])
        lists["e_kin"].append(r["e_kin"][k])

    def _summary(self, lists):
        T = np.array(lists["T"])
        ep = np.array(lists["e_pot"])
        return {"positions": self.positions.copy(), "velocities": self.velocities.copy(),
                "lattice": self.lattice.copy(),
                "T_mean": float(T.mean()) if len(T) else 0.0,
                "T_std": float(T.std()) if len(T) else 0.0,
                "e_pot_mean": float(ep.mean()) if len(ep) else 0.0,
                "thermo_log": self._thermo_log}


def _check_chain(name, kw):
    """_init_nhc_chain (:575-588): an integer of at least two links (at most
    16 on the device); temperature and tau_nhc positive."""
    chain = kw.get("nhc_chain", 3)
    if isinstance(chain, bool) or not float(chain).is_integer() or not 2 <= int(chain) <= 16:
        raise ValueError(f"{name}: nhc_chain must be an integer from 2 to 16, got {chain!r}")
    if not kw.get("temperature", 300.0) > 0.0 or not kw.get("tau_nhc", 100.0) > 0.0:
        raise ValueError(f"{name}: temperature and tau_nhc must be > 0")


class NVTSimulation(_DeviceMD):
    """NVT with an M-link Nose-Hoover chain (solvers/molecular_dynamics.py:
    774-1297) on the device."""
    _TASK = "nvt"
    _ENSEMBLE = "nvt"
    _TRAJ_COMMENT = "akitsushima.py NVT-MD"

    def __init__(self, lattice, positions, atom_types, force_fn=None, r_cut=None, **kw):
        _check_chain("NVTSimulation", kw)
        super().__init__(lattice, positions, atom_types, force_fn, r_cut, **kw)

    def _thermostat(self):
        return ("nhc", self.temperature, self.tau_nhc, self.nhc_chain)

    def _rescale_steps(self):
        return self.velocity_rescale_steps

    def _settings(self):
        return dict(super()._settings(), temperature=self.temperature, tau_nhc=self.tau_nhc)


class _StochasticNVT(_DeviceMD):
    _KIND = ""

    def __init__(self, lattice, positions, atom_types, force_fn=None, r_cut=None, *,
                 damping=100.0, velocity_rescale_steps=0, **kw):
        name = type(self).__name__
        if int(velocity_rescale_steps) != 0:
            raise ValueError(f"{name}: velocity_rescale_steps must be 0, got "
                             f"{velocity_rescale_steps}.")
        if not damping > 0.0:
            raise ValueError(f"{name}: damping must be > 0 fs, got {damping}.")
        super().__init__(lattice, positions, atom_types, force_fn, r_cut, **kw)
        self.damping = float(damping)

    def _thermostat(self):
        return (self._KIND, self.temperature, self.damping, self.seed)

    def _header(self):
        return (f"{'Step':>8}  {'Time(ps)':>10}  {'E_pot(eV)':>14}  {'E_kin(eV)':>12}  "
                f"{'E_tot(eV)':>12}  {'T_inst(K)':>10}  {'E-Q(eV)':>14}  {'NL_builds':>10}")

    def _row(self, c, r, k, nb):
        text, row = super()._row(c, r, k, nb)
        row.pop("H_nhc")
        row["e_minus_q"] = r["e_cons"][k]
        return text, row

    def _summary(self, lists):
        out = super()._summary(lists)
        out["heat_added"] = float(state_fields(self._state)["heat"])
        return out

    def _settings(self):
        return dict(super()._settings(), temperature=self.temperature, damping=self.damping)


class LangevinSimulation(_StochasticNVT):
    """NVT with the Langevin thermostat, BAOAB (:1511-1587), on the device."""
    _KIND = "langevin"
    _ENSEMBLE = "langevin"
    _TRAJ_COMMENT = "akitsushima.py Langevin NVT-MD"
    _POSCAR_TITLE = "Langevin MD final frame"


class CSVRSimulation(_StochasticNVT):
    """NVT with the Bussi-Donadio-Parrinello thermostat (:1589-1687), on the
    device."""
    _KIND = "csvr"
    _ENSEMBLE = "csvr"
    _TRAJ_COMMENT = "akitsushima.py CSVR NVT-MD"
    _POSCAR_TITLE = "CSVR MD final frame"


class NVESimulation(_DeviceMD):
    """Velocity Verlet NVE (:1695-2070) on the device; ``velocities`` as the
    reference names the initial velocities (else Maxwell-Boltzmann at
    ``temperature``)."""
    _TASK = "nve"
    _ENSEMBLE = "nve"
    _POSCAR = "POSCAR_final_NVE"
    _POSCAR_TITLE = "NVE final frame"
    _TRAJ_COMMENT = "akitsushima.py NVE-MD"

    def __init__(self, lattice, positions, atom_types, force_fn=None, r_cut=None, *,
                 velocities=None, **kw):
        kw.pop("velocity_rescale_steps", None)
        if velocities is not None:
            kw["initial_velocities"] = velocities
        super().__init__(lattice, positions, atom_types, force_fn, r_cut, **kw)
        self._e_ref = None

    def _header(self):
        return (f"{'Step':>8}  {'Time(ps)':>10}  {'E_pot(eV)':>14}  {'E_kin(eV)':>12}  "
                f"{'E_tot(eV)':>14}  {'T_inst(K)':>10}  {'P_inst(GPa)':>12}  "
                f"{'dE_drift(eV)':>14}")

    def _row(self, c, r, k, nb):
        t_ps = c * self.dt * 1e-3
        ep, ek = r["e_pot"][k], r["e_kin"][k]
        if self._e_ref is None:
            self._e_ref = ep + ek
        T = 2.0 * ek / (self.n_dof * K_B)
        P = r["pressure_gpa"][k]
        d = ep + ek - self._e_ref
        text = (f"{c:8d}  {t_ps:10.4f}  {ep:14.6f}  {ek:12.6f}  {ep + ek:14.6f}  {T:10.2f}  "
                f"{P:12.4f}  {d:14.6f}")
        return text, {"step": c, "time_ps": t_ps, "e_pot": ep, "e_kin": ek, "e_tot": ep + ek,
                      "T_inst": T, "P_gpa": P, "drift": d}

    def run(self, checkpoint=None, checkpoint_every=0, checkpoint_meta=None):
        """E_tot drift against the initial state (or the reference energy a
        checkpoint carries)."""
        self._e_ref = None
        out = super().run(checkpoint, checkpoint_every, checkpoint_meta)
        r = self._last
        et = r["e_pot"] + r["e_kin"]
        ref = self._e_ref if self._e_ref is not None else et[0]
        out["e_tot_drift"] = float(et[-1] - ref)
        return out

    def _ck_extra(self):
        return {"e_tot_ref": self._e_ref} if self._e_ref is not None else {}

    def _ck_restore_extra(self, extra):
        if "e_tot_ref" in extra:
            self._e_ref = float(extra["e_tot_ref"])

    def _summary(self, lists):
        out = super()._summary(lists)
        out.pop("thermo_log")
        return out


class NPTSimulation(_DeviceMD):
    """NPT with the MTTK barostat (:2141-3060) on the device: every cell_dof
    and baro_thermostat of the reference, with the per-axis thermo columns and
    summary of the flexible cells."""
    _TASK = "npt"
    _ENSEMBLE = "npt"
    _POSCAR = "POSCAR_final_NPT"
    _POSCAR_TITLE = "NPT final frame"
    _TRAJ_COMMENT = "akitsushima.py NPT-MD"

    def __init__(self, lattice, positions, atom_types, force_fn=None, r_cut=None, *,
                 pressure=0.0, tau_baro=1000.0, cell_dof="iso", baro_thermostat="nhc", **kw):
        cell_dof = str(cell_dof).lower()
        if cell_dof not in CELL_DOF_MODES:
            raise ValueError(f"NPTSimulation: cell_dof must be one of "
                             f"{' | '.join(CELL_DOF_MODES)}, got {cell_dof!r}.")
        baro = str(baro_thermostat).lower()
        if baro not in ("nhc", "langevin"):
            raise ValueError("NPTSimulation: baro_thermostat must be nhc | langevin, got "
                             f"{baro_thermostat!r}.")
        if not tau_baro > 0.0:
            raise ValueError("NPTSimulation: tau_baro must be > 0")
        _check_chain("NPTSimulation", kw)
        super().__init__(lattice, positions, atom_types, force_fn, r_cut, **kw)
        self.pressure_gpa = float(pressure)
        self.P_ext = self.pressure_gpa / EV_ANG3_TO_GPA
        self.tau_baro = float(tau_baro)
        self.cell_dof = cell_dof
        self.baro_thermostat = baro

    def _thermostat(self):
        return ("npt", self.temperature, self.tau_nhc, self.nhc_chain, self.pressure_gpa,
                self.tau_baro, self.cell_dof, self.baro_thermostat, self.seed)

    def _rescale_steps(self):
        return self.velocity_rescale_steps

    def _variable_cell(self):
        return True

    # per-axis output of the flexible cells (_extra_cols, _extra_values,
    # _extra_summary :3015-3054)
    _P_KEYS = ("Pxx", "Pyy", "Pzz", "Pyz", "Pxz", "Pxy")

    def _extra_cols(self):
        if self.cell_dof == "tri":
            return (*self._P_KEYS, "a", "b", "c", "alpha", "beta", "gamma")
        return ("Pxx", "Pyy", "Pzz", "a", "b", "c")

    def _extra_values(self, r, k):
        L = r["lattices"][k]
        out = dict(zip(self._P_KEYS, map(float, r["pressure_tensor_gpa"][k])))
        for key, x in zip("abc", np.linalg.norm(L, axis=1)):
            out[key] = float(x)

        def ang(u, w):
            return float(np.degrees(np.arccos(np.clip(
                u @ w / (np.linalg.norm(u) * np.linalg.norm(w)), -1.0, 1.0))))
        out.update(alpha=ang(L[1], L[2]), beta=ang(L[0], L[2]), gamma=ang(L[0], L[1]))
        for i, x in enumerate(L.ravel()):
            out[f"L{i}"] = float(x)
        return out

    def _header(self):
        h = (f"{'Step':>8}  {'Time(ps)':>10}  {'E_pot(eV)':>14}  {'E_kin(eV)':>12}  "
             f"{'T_inst(K)':>10}  {'P_inst(GPa)':>12}  {'V(Ang3)':>12}  {'H_NPT(eV)':>14}")
        if self.cell_dof != "iso":
            h += "".join(f"  {c:>10}" for c in self._extra_cols())
        return h

    def _row(self, c, r, k, nb):
        t_ps = c * self.dt * 1e-3
        ep, ek = r["e_pot"][k], r["e_kin"][k]
        T = 2.0 * ek / (self.n_dof * K_B)
        P, V, H = r["pressure_gpa"][k], r["volume"][k], r["e_cons"][k]
        text = (f"{c:8d}  {t_ps:10.4f}  {ep:14.6f}  {ek:12.6f}  {T:10.2f}  {P:12.4f}  "
                f"{V:12.4f}  {H:14.6f}")
        if self.cell_dof != "iso":
            xv = self._extra_values(r, k)
            text += "".join(f"  {xv[key]:10.4f}" for key in self._extra_cols())
        return text, {"step": c, "time_ps": t_ps, "e_pot": ep, "e_kin": ek, "T_inst": T,
                      "P_gpa": P, "V": V, "H_npt": H}

    def _list_keys(self):
        keys = ("T", "P", "V", "e_pot", "e_kin", "H")
        if self.cell_dof != "iso":
            keys += (*self._P_KEYS, "a", "b", "c", *(f"L{i}" for i in range(9)))
        return keys

    def _collect(self, lists, r, k):
        super()._collect(lists, r, k)
        lists["P"].append(r["pressure_gpa"][k])
        lists["V"].append(r["volume"][k])
        lists["H"].append(r["e_cons"][k])
        if self.cell_dof != "iso":
            xv = self._extra_values(r, k)
            for key in (*self._P_KEYS, "a", "b", "c", *(f"L{i}" for i in range(9))):
                lists[key].append(xv[key])

    def _ck_scalars(self):
        f = state_fields(self._state)
        return {"p_eps": f["p_eps"], "xi1_b": float(f["xi_b"][0]), "xi2_b": float(f["xi_b"][1]),
                "vxi1_b": float(f["vxi_b"][0]), "vxi2_b": float(f["vxi_b"][1])}

    def _settings(self):
        out = dict(super()._settings(), temperature=self.temperature, tau_nhc=self.tau_nhc,
                   tau_baro=self.tau_baro, pressure_gpa=self.pressure_gpa)
        if self.cell_dof != "iso":
            out["cell_dof"] = self.cell_dof
        if self.baro_thermostat != "nhc":
            out["baro_thermostat"] = self.baro_thermostat
        return out

    def _summary(self, lists):
        T, P, V, H = (np.array(lists[k]) for k in ("T", "P", "V", "H"))
        r = self._last
        h_ref = H[0] if len(H) else r["e_cons"][0]
        out = {"positions": self.positions.copy(), "velocities": self.velocities.copy(),
               "lattice": self.lattice.copy(),
               "T_mean": float(T.mean()) if len(T) else 0.0,
               "T_std": float(T.std()) if len(T) else 0.0,
               "P_mean_gpa": float(P.mean()) if len(P) else 0.0,
               "P_std_gpa": float(P.std()) if len(P) else 0.0,
               "V_mean": float(V.mean()) if len(V) else 0.0,
               "H_npt_drift": float(r["e_cons"][-1] - h_ref)}
        if self.cell_dof != "iso" and len(lists["a"]):
            Pm = np.zeros((3, 3))
            for key, (i, j) in zip(self._P_KEYS, ((0, 0), (1, 1), (2, 2), (1, 2), (0, 2),
                                                   (0, 1))):
                Pm[i, j] = Pm[j, i] = float(np.mean(lists[key]))
            abc = np.array([lists["a"], lists["b"], lists["c"]])
            out.update({"cell_dof": self.cell_dof,
                        "lattice_mean": np.array([np.mean(lists[f"L{i}"])
                                                  for i in range(9)]).reshape(3, 3),
                        "P_tensor_mean_gpa": Pm, "abc_mean": abc.mean(axis=1),
                        "abc_std": abc.std(axis=1)})
        return out
