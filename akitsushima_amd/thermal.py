"""Green-Kubo thermal conductivity with the MD and the heat flux on the device.

``ThermalConductivityCalculator`` keeps the reference's interface
(solvers/thermal_conductivity.py:50-130, run_md :414-677, compute_kappa
:679-831): ``run_md`` equilibrates with a Nose-Hoover chain and records the
exact many-body heat flux of an NVE production run every
``heat_flux_every`` steps -- both runs device-resident, one hipGraph step
each, the flux formed on the device from the bond kernels' site terms --
and ``compute_kappa`` integrates the heat-flux autocorrelation on the host:

    kappa_ab = V / (k_B T^2) int_0^tau_max <J_a J_b>(t) dt   [W/(m K)],

with the reference's FFT correlation, trapezoid running integral and block
error estimate.
"""
from __future__ import annotations

import os
import warnings

import numpy as np

from .md import EV_ANG3_TO_GPA, K_B, maxwell_boltzmann, masses_of, md_run, nvt_run

EV_ANG_K_FS_TO_W_MK = (1.60217663e-19) / (1e-10 * 1e-15)


def autocorr_fft(a, b, n_lag, subtract_mean=True):
    """c[k] = sum_t a[t + k] b[t] / (n - k), k < n_lag, by zero-padded FFTs
    (the lag convention of the reference's _autocorr_fft,
    solvers/thermal_conductivity.py:1107-1170); series mean-subtracted by
    default."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    n = len(a)
    if subtract_mean:
        a = a - a.mean()
        b = b - b.mean()
    nfft = 1
    while nfft < 2 * n:
        nfft *= 2
    c = np.fft.irfft(np.fft.rfft(a, n=nfft) * np.conj(np.fft.rfft(b, n=nfft)), n=nfft)[:n]
    return (c / (n - np.arange(n)))[:n_lag]


def _running_integral(h, dt):
    run = np.zeros_like(h)
    for k in range(1, len(h)):
        run[k] = run[k - 1] + 0.5 * dt * (h[k - 1] + h[k])
    return run


def _cross_cf(A, B, n_lag):
    """(n_lag, 3, 3) correlation tensor c[k, a, b] = <A_a(t + k) B_b(t)> of
    two (n, 3) series."""
    h = np.zeros((n_lag, 3, 3))
    for a in range(3):
        for b in range(3):
            h[:, a, b] = autocorr_fft(A[:, a], B[:, b], n_lag)
    return h


def _hcacf(J, n_lag):
    return _cross_cf(J, J, n_lag)


class ThermalConductivityCalculator:
    """Green-Kubo kappa of one cell with a GPU potential (any of the
    engine's potential classes, or a force function of ``make_force_fn``
    whose engine holds one of them)."""

    def __init__(self, lattice, positions, atom_types, potential, masses=None,
                 temperature=300.0):
        self.lattice = np.asarray(lattice, float).copy()
        self.positions = np.asarray(positions, float).copy()
        self.atom_types = list(atom_types)
        pots = getattr(getattr(potential, "engine", None), "_potentials", None)
        if pots is not None:
            if len(pots) != 1:
                raise ValueError("ThermalConductivityCalculator: the engine must hold one "
                                 "potential (its site terms give the heat flux)")
            potential = pots[0]
        if not hasattr(potential, "site_terms"):
            raise ValueError("ThermalConductivityCalculator: the potential has no site-energy "
                             "decomposition for the heat flux")
        self.potential = potential
        self.temperature = float(temperature)
        self.N = len(self.atom_types)
        self.masses = (np.asarray(masses, float) if masses is not None
                       else masses_of(self.atom_types))
        self._J_series = None
        self._dt_hf = None
        self._hcacf = self._tau_axis = self._kappa = self._kappa_run = self._kappa_std = None

    def run_md(self, n_steps=100_000, timestep=0.5, equil_steps=10_000, velocities=None,
               heat_flux_every=5, traj_every=100, thermo_every=1000, output_dir=None, seed=42,
               verbose=False, decompose=None, *, tau_nhc=100.0, nhc_chain=3, skin_width=0.5):
        """NVT (Nose-Hoover chain) equilibration for equil_steps, then NVE
        for n_steps recording J after every step s with s %
        heat_flux_every == 0.  Without ``velocities`` the start is a
        Maxwell-Boltzmann draw at the temperature from ``seed``.
        ``decompose`` = {"species": bool, "regions": {"axis": "a"|"b"|"c",
        "bounds": [fractional edges]}, "save_atom_flux": bool} also records
        the heat flux of each group of atoms (per-atom heat currents,
        regions by each frame's positions) for compute_kappa_decomposition;
        save_atom_flux writes the per-atom currents J_i / V of every frame,
        float32, to output_dir/atom_heat_flux.npy (frames, N, 3).  With
        output_dir, the NVE run's thermo_nve.log (every thermo_every steps:
        T = 2 E_kin / (3 N k_B), E_pot, E_tot, P = Tr(sigma) / 3),
        trajectory_tc.xyz (every traj_every steps, frames recorded on the
        device) and heat_flux.dat are written in the reference's formats
        (:560-620).  Returns {"J_series" (frames, 3)
        [eV/(A^2 fs)], "dt_hf" fs, "positions", "velocities", "e_tot"
        (NVE), "equil" (the NVT run's record or None)}."""
        decompose = dict(decompose or {})
        save_atoms = bool(decompose.get("save_atom_flux"))
        if save_atoms and not output_dir:
            raise ValueError("run_md: save_atom_flux writes atom_heat_flux.npy and needs "
                             "output_dir")
        pos = self.positions.copy()
        vel = (np.asarray(velocities, float).copy() if velocities is not None
               else maxwell_boltzmann(self.temperature, self.masses, seed))
        eq = None
        if equil_steps > 0:
            eq = nvt_run(self.potential, self.lattice, pos, vel, self.atom_types,
                         self.temperature, dt=timestep, n_steps=equil_steps, tau_nhc=tau_nhc,
                         nhc_chain=nhc_chain, skin_width=skin_width, masses=self.masses)
            pos, vel = eq["positions"], eq["velocities"]
        groups = None
        species_names = region_names = None
        if decompose.get("species") or decompose.get("regions"):
            labels, n_lab, region = None, 1, None
            if decompose.get("species"):
                species_names = sorted(set(self.atom_types), key=self.atom_types.index)
                labels = np.array([species_names.index(t) for t in self.atom_types])
                n_lab = len(species_names)
            reg = decompose.get("regions")
            if reg:
                ax = {"a": 0, "b": 1, "c": 2}.get(str(reg.get("axis", "c")), reg.get("axis"))
                bounds = np.asarray(reg["bounds"], dtype=float)
                region = (int(ax), bounds)
                region_names = [f"[{bounds[k]:.3f},{bounds[k + 1]:.3f})"
                                for k in range(len(bounds) - 1)]
            groups = (labels, n_lab, region)
        traj = int(traj_every) if output_dir and 0 < int(traj_every) <= n_steps else 0
        pos0 = pos.copy()
        prod = md_run(self.potential, self.lattice, pos, vel, self.atom_types, dt=timestep,
                      n_steps=n_steps, skin_width=skin_width, masses=self.masses,
                      flux_every=int(heat_flux_every), flux_groups=groups,
                      atom_flux=save_atoms, frame_every=traj)
        J = prod["flux"]
        self._J_series = J
        self._J_groups = {}
        if groups is not None:
            Jg = prod["flux_groups"]                  # (frames, labels, regions, 3)
            if species_names is not None:
                self._J_groups["species"] = (species_names, Jg.sum(axis=2))
            if region_names is not None:
                self._J_groups["regions"] = (region_names, Jg.sum(axis=1))
        self._dt_hf = float(timestep * heat_flux_every)
        if output_dir:
            os.makedirs(output_dir, exist_ok=True)
            with open(os.path.join(output_dir, "heat_flux.dat"), "w") as fh:
                fh.write("step  Jx_eV_Ang2_fs  Jy  Jz\n")
                for k, j in enumerate(J):
                    fh.write(f"{k * heat_flux_every + 1}  {j[0]:.8e}  {j[1]:.8e}  {j[2]:.8e}\n")
            if save_atoms:
                np.save(os.path.join(output_dir, "atom_heat_flux.npy"), prod["atom_flux"])
            self._write_nve_logs(output_dir, prod, pos0, n_steps, int(thermo_every), traj)
        if verbose:
            print(f"  NVE production: {n_steps} steps, {len(J)} heat-flux frames, "
                  f"{prod['seconds']:.2f} s on the device")
        return {"J_series": J, "dt_hf": self._dt_hf, "positions": prod["positions"],
                "velocities": prod["velocities"], "e_tot": prod["e_tot"], "equil": eq}

    def _write_nve_logs(self, output_dir, prod, pos0, n_steps, thermo_every, traj_every):
        """thermo_nve.log and trajectory_tc.xyz of the NVE run from the
        device records (row c = the state after c steps, c = 0 the start)."""
        V = abs(np.linalg.det(self.lattice))
        N = self.N
        ek, ep = prod["e_kin"], prod["e_pot"]
        # the log's P is the virial part Tr(sigma) / 3 of the recorded P_inst
        p_vir = prod["pressure_gpa"] - 2.0 * ek / (3.0 * V) * EV_ANG3_TO_GPA
        with open(os.path.join(output_dir, "thermo_nve.log"), "w") as fh:
            fh.write(f"{'Step':>8}  {'T(K)':>9}  {'E_pot(eV)':>14}  {'E_tot(eV)':>14}  "
                     f"{'P(GPa)':>9}\n")
            for c in range(0, n_steps + 1, max(thermo_every, 1)):
                T = ek[c] * 2.0 / (3.0 * N * K_B)
                fh.write(f"{c:8d}  {T:9.2f}  {ep[c]:14.6f}  {ep[c] + ek[c]:14.6f}  "
                         f"{p_vir[c]:9.4f}\n")
        if traj_every <= 0:
            return
        a, b, cc = self.lattice
        head = (f"Lattice=\"{a[0]:.6f} {a[1]:.6f} {a[2]:.6f} {b[0]:.6f} {b[1]:.6f} "
                f"{b[2]:.6f} {cc[0]:.6f} {cc[1]:.6f} {cc[2]:.6f}\" "
                f"Properties=species:S:1:pos:R:3 ")
        frames = [(0, pos0)] + [((k + 1) * traj_every, f)
                                for k, f in enumerate(prod["frames"])]
        with open(os.path.join(output_dir, "trajectory_tc.xyz"), "w") as fh:
            for c, p in frames:
                fh.write(f"{N}\n{head}step={c}\n")
                for sp, x in zip(self.atom_types, p):
                    fh.write(f"{sp}  {x[0]:.8f}  {x[1]:.8f}  {x[2]:.8f}\n")

    def save_results(self, output_dir=".", verbose=True):
        """hcacf.csv, kappa_running.csv and kappa_tensor.csv of the last
        compute_kappa in the reference's layouts (save_results,
        solvers/thermal_conductivity.py:1017-1069; no plots)."""
        os.makedirs(output_dir, exist_ok=True)
        if self._hcacf is not None and self._tau_axis is not None:
            c = self._hcacf
            path = os.path.join(output_dir, "hcacf.csv")
            with open(path, "w") as fh:
                fh.write("tau_ps,Cxx,Cyy,Czz,Cxy,Cxz,Cyz\n")
                for i, tau in enumerate(self._tau_axis):
                    fh.write(f"{tau:.6f},{c[i, 0, 0]:.8e},{c[i, 1, 1]:.8e},{c[i, 2, 2]:.8e},"
                             f"{c[i, 0, 1]:.8e},{c[i, 0, 2]:.8e},{c[i, 1, 2]:.8e}\n")
            if verbose:
                print(f"  HCACF      -> {path}")
        if self._kappa_run is not None:
            path = os.path.join(output_dir, "kappa_running.csv")
            with open(path, "w") as fh:
                fh.write("tau_ps,kappa_xx,kappa_yy,kappa_zz,kappa_avg  # W/(m.K)\n")
                for i, tau in enumerate(self._tau_axis):
                    k = self._kappa_run[i]
                    avg = (k[0, 0] + k[1, 1] + k[2, 2]) / 3
                    fh.write(f"{tau:.6f},{k[0, 0]:.4f},{k[1, 1]:.4f},{k[2, 2]:.4f},"
                             f"{avg:.4f}\n")
            if verbose:
                print(f"  Running kappa -> {path}")
        if self._kappa is not None:
            path = os.path.join(output_dir, "kappa_tensor.csv")
            std = self._kappa_std if self._kappa_std is not None else np.zeros((3, 3))
            with open(path, "w") as fh:
                fh.write("# Thermal conductivity tensor [W/(m.K)]\n")
                fh.write("alpha,beta,kappa_WmK,kappa_std_WmK\n")
                for a in range(3):
                    for b in range(3):
                        fh.write(f"{a},{b},{self._kappa[a, b]:.4f},{std[a, b]:.4f}\n")
            if verbose:
                print(f"  kappa tensor -> {path}")

    def compute_kappa(self, J_series=None, dt_hf=None, tau_max=20.0, n_blocks=5, volume=None,
                      verbose=False):
        """kappa (3, 3) [W/(m K)] from the heat-flux ACF integrated to tau_max
        ps (compute_kappa, solvers/thermal_conductivity.py:679-831); block
        standard error in ``self._kappa_std``."""
        if J_series is None:
            if self._J_series is None:
                raise RuntimeError("compute_kappa: no J_series available. Call run_md() first "
                                   "or supply J_series=.")
            J_series = self._J_series
        if dt_hf is None:
            dt_hf = self._dt_hf if self._dt_hf else 1.0
        J = np.asarray(J_series, float)
        n = len(J)
        dt = float(dt_hf)
        T = self.temperature
        V = float(volume) if volume is not None else abs(np.linalg.det(self.lattice))
        n_lag = min(int(tau_max * 1e3 / dt), n // 2)
        pref = V / (K_B * T ** 2)
        h = _hcacf(J, n_lag)
        run = _running_integral(h, dt) * pref
        kappa = run[-1] * EV_ANG_K_FS_TO_W_MK
        block = max(n // n_blocks, n_lag)
        nb = n // block if block > 0 else 0
        if nb < 2:
            warnings.warn("compute_kappa: trajectory too short relative to tau_max and n_blocks "
                          "for a block-averaged uncertainty; kappa_std is NaN", UserWarning,
                          stacklevel=2)
            std = np.full((3, 3), np.nan)
        else:
            kb = np.array([_running_integral(_hcacf(J[i * block:(i + 1) * block], n_lag), dt)[-1]
                           * pref * EV_ANG_K_FS_TO_W_MK for i in range(nb)])
            std = np.std(kb, axis=0) / np.sqrt(nb)
        self._hcacf = h
        self._tau_axis = np.arange(n_lag) * dt * 1e-3
        self._kappa = kappa
        self._kappa_run = run * EV_ANG_K_FS_TO_W_MK
        self._kappa_std = std
        if verbose:
            d = np.diag(kappa)
            print(f"  Green-Kubo (T={T:.0f} K, tau_max={tau_max:.1f} ps): k = {d} W/(m K), "
                  f"mean {d.mean():.3f}")
        return kappa

    def compute_kappa_decomposition(self, tau_max=20.0, verbose=True):
        """Group contributions to the Green-Kubo tensor
        (compute_kappa_decomposition, solvers/thermal_conductivity.py:
        832-883): kappa_g,ab = V / (k_B T^2) int <J_g,a(0) J_b(tau)> dt with
        J_g the group's flux; linear in J_g, so the groups of one
        decomposition add up to the total.  Needs run_md(decompose=...).
        Returns {decomposition: {group: (3, 3) W/(m K)}, "total": (3, 3),
        "sum_rule_residual": {decomposition: max |sum_g kappa_g - kappa|}}."""
        groups = getattr(self, "_J_groups", None)
        if not groups:
            raise RuntimeError("compute_kappa_decomposition: run run_md(decompose=...) first.")
        J = self._J_series
        dt = float(self._dt_hf)
        n_lag = min(int(tau_max * 1e3 / dt), len(J) // 2)
        scale = (abs(np.linalg.det(self.lattice)) / (K_B * self.temperature ** 2)
                 * EV_ANG_K_FS_TO_W_MK)
        # the Green-Kubo integral up to the last lag, trapezoid rule over the
        # lags as in compute_kappa, of <flux(t + tau) J(t)>
        gk = lambda flux: _running_integral(_cross_cf(flux, J, n_lag), dt)[-1] * scale
        total = gk(J)
        out = {}
        residual = {}
        for decomposition, (labels, series) in groups.items():
            parts = {label: gk(series[:, k]) for k, label in enumerate(labels)}
            out[decomposition] = parts
            residual[decomposition] = float(np.abs(np.sum(list(parts.values()), axis=0)
                                                   - total).max())
        if verbose:
            for decomposition, parts in out.items():
                print(f"\n  Green-Kubo kappa by {decomposition} [W/(m K)], diagonal:")
                width = max(len(str(g)) for g in parts)
                for label, k in parts.items():
                    print("    " + str(label).rjust(width) + "  "
                          + "  ".join(f"{v:10.4f}" for v in np.diag(k)))
        out["sum_rule_residual"] = residual
        out["total"] = total
        return out
