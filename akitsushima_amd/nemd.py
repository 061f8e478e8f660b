"""Non-equilibrium MD with two Langevin heat baths on the device
(NEMDSimulation, solvers/nemd.py:1-337): temperature profile, local heat
flux from the per-atom heat currents, thermal conductivity from the
gradient and the Kapitza resistance of given interfaces.

The velocity Verlet step, the Ornstein-Uhlenbeck update of the bath atoms
and the per-bin sums of the samples (kinetic energy, atom count, per-atom
heat currents) run in one captured hipGraph step (``ak_nemd_run``); the host
adds the per-step bath heats and the per-sample bin sums into the reference's
accumulators (four production blocks) and evaluates ``results`` from them.
The initial velocities are the reference's draw from numpy's generator (the
same numbers for the same seed); the bath noise is the Philox4x32-10 stream
keyed by the seed (the reference continues numpy's generator: same process,
other realisations).  A checkpoint holds positions, velocities, accumulators
and the device state (random draws used, heat), so a run continued from it
ends where one written with the same checkpoint spacing does, bit for bit.
"""
from __future__ import annotations

import json
import os

import numpy as np

from .md import EV_ANG_TO_AMU_ANG_FS2, K_B, STATE_LEN, masses_of, n_samples, nemd_run
from .simulations import _potential_of

__all__ = ["NEMDSimulation"]

# the unit factors of solvers/nemd.py:26-27
EV_ANG_K_FS_TO_W_MK = 1.602176634e-19 / (1e-10 * 1e-15)
EV_ANG2_FS_TO_W_M2 = 1.602176634e-19 / (1e-20 * 1e-15)
_AX = {"a": 0, "b": 1, "c": 2, 0: 0, 1: 1, 2: 2}


class NEMDSimulation:
    """The reference's constructor arguments (solvers/nemd.py:52-100): the
    system, ``force_fn`` a device potential or a force function of
    ``make_force_fn`` holding one; bath temperatures [K]; transport ``axis``
    'a' | 'b' | 'c'; the baths as fractional slabs [lo, hi) of the initial
    positions along it (or ``hot_atoms`` / ``cold_atoms``); ``damping`` [fs];
    ``timestep`` [fs]; ``n_bins`` along the axis; ``local_flux`` (per-bin
    heat flux from the per-atom currents); ``seed``.  Keywords:
    ``skin_width`` of the device neighbour list, ``use_graph``."""

    def __init__(self, lattice, positions, atom_types, force_fn, masses=None, T_hot=330.0,
                 T_cold=270.0, axis="c", hot=(0.0, 0.05), cold=(0.5, 0.55), hot_atoms=None,
                 cold_atoms=None, damping=100.0, timestep=1.0, n_bins=20, local_flux=True,
                 seed=42, *, skin_width=0.5, use_graph=True):
        self.potential = _potential_of(force_fn)
        self.lattice = np.asarray(lattice, float).copy()
        self.pos = np.asarray(positions, float).copy()
        self.types = list(atom_types)
        self.N = len(self.types)
        self.masses = (np.asarray(masses, float) if masses is not None
                       else masses_of(self.types))
        self.T_hot, self.T_cold = float(T_hot), float(T_cold)
        self.axis = _AX[axis]
        self.damping, self.dt = float(damping), float(timestep)
        self.n_bins = int(n_bins)
        self.local_flux = bool(local_flux)
        self.seed = int(seed)
        self.skin_width = float(skin_width)
        self.use_graph = bool(use_graph)
        if self.local_flux and not hasattr(self.potential, "site_terms"):
            raise ValueError("NEMDSimulation: local_flux needs a potential with a site-energy "
                             "decomposition")
        frac = self._frac(self.pos)
        ax = frac[:, self.axis]

        def slab(iv):
            lo, hi = float(iv[0]) % 1.0, float(iv[1]) % 1.0
            # a layer at 0 may sit at -1e-17 and wrap to just below 1
            f = (ax + 1e-9) % 1.0
            inside = (f >= lo) & (f < hi) if lo < hi else (f >= lo) | (f < hi)
            return np.flatnonzero(inside)

        self.hot = np.asarray(hot_atoms, int) if hot_atoms is not None else slab(hot)
        self.cold = np.asarray(cold_atoms, int) if cold_atoms is not None else slab(cold)
        if len(self.hot) == 0 or len(self.cold) == 0:
            raise ValueError("NEMDSimulation: a heat bath has no atoms.")
        if np.intersect1d(self.hot, self.cold).size:
            raise ValueError("NEMDSimulation: the hot and cold baths overlap.")
        self.bath = np.zeros(self.N, np.uint8)
        self.bath[self.hot] = 1
        self.bath[self.cold] = 2
        self.hot_iv = (ax[self.hot].min(), ax[self.hot].max())
        self.cold_iv = (ax[self.cold].min(), ax[self.cold].max())
        i, j = [k for k in range(3) if k != self.axis]
        cross = np.cross(self.lattice[i], self.lattice[j])
        self.volume = abs(np.linalg.det(self.lattice))
        self.area = float(np.linalg.norm(cross))
        self.length = self.volume / self.area          # width perpendicular to the section
        # unit normal of the cross-section along +axis
        self.normal = cross / self.area * np.sign(cross @ self.lattice[self.axis])
        self.state = None

    # ------------------------------------------------------------------
    def _frac(self, pos):
        f = pos @ np.linalg.inv(self.lattice)
        return f - np.floor(f)

    def _temp_of(self, k_bath, n):
        return 2.0 * k_bath / (3.0 * n * K_B)

    def _fresh_state(self):
        self.rng = np.random.default_rng(self.seed)
        T0 = 0.5 * (self.T_hot + self.T_cold)
        sig = np.sqrt(K_B * T0 * EV_ANG_TO_AMU_ANG_FS2 / self.masses)
        vel = self.rng.standard_normal((self.N, 3)) * sig[:, None]
        vel -= (self.masses[:, None] * vel).sum(0) / self.masses.sum()
        nb = self.n_bins
        return {"step": 0, "pos": self.pos.copy(), "vel": vel,
                "Q_hot": 0.0, "Q_cold": 0.0, "Q_hot_prod": 0.0, "Q_cold_prod": 0.0,
                "E0": None, "n_prod": 0, "n_samples": 0,
                "ke_bin": np.zeros((4, nb)), "n_bin": np.zeros((4, nb)),
                "j_bin": np.zeros((4, nb)), "n_jsamp": np.zeros(4),
                "q_block": np.zeros((4, 2)), "t_block": np.zeros(4),
                "device_state": np.zeros(STATE_LEN)}

    def _save(self, path, st):
        tmp = path + ".tmp.npz"
        arrays = {k: v for k, v in st.items() if isinstance(v, np.ndarray)}
        scal = {k: v for k, v in st.items() if not isinstance(v, np.ndarray)}
        np.savez(tmp, **arrays, scalars=json.dumps(scal),
                 rng=json.dumps(self.rng.bit_generator.state))
        os.replace(tmp, path)

    def _load(self, path):
        d = np.load(path, allow_pickle=False)
        st = json.loads(str(d["scalars"]))
        for k in d.files:
            if k not in ("scalars", "rng"):
                st[k] = d[k].copy()
        self.rng = np.random.default_rng()
        self.rng.bit_generator.state = json.loads(str(d["rng"]))
        if "device_state" not in st:
            raise ValueError(f"NEMDSimulation: {path} holds no device state (written by "
                             "another implementation); start a fresh run")
        return st

    # ------------------------------------------------------------------
    def run(self, equil_steps, n_steps, sample_every=10, checkpoint=None, checkpoint_every=0,
            thermo_path=None, verbose=True):
        """``equil_steps`` with the baths on (not sampled), then ``n_steps``
        of production sampled every ``sample_every`` steps into four
        consecutive blocks (NEMDSimulation.run, solvers/nemd.py:150-232).
        With ``checkpoint``, the state is written there every
        ``checkpoint_every`` steps (atomically) and a run finding the file
        continues from it.  Returns the accumulator dictionary (also kept
        in ``self.state``) with "E_minus_Q_drift_eV"."""
        st = (self._load(checkpoint) if checkpoint and os.path.isfile(checkpoint)
              else self._fresh_state())
        if verbose and st["step"]:
            print(f"  NEMD: restarting from {checkpoint} at step {st['step']}")
        equil_steps, n_steps = int(equil_steps), int(n_steps)
        se = int(sample_every)
        total = equil_steps + n_steps
        ce = int(checkpoint_every) if checkpoint else 0
        row_every = max(se * 10, 1)
        th = open(thermo_path, "a") if thermo_path else None
        header = th is not None and st["step"] == 0
        n_hot, n_cold = len(self.hot), len(self.cold)
        last = None
        try:
            while last is None or st["step"] < total:
                g0 = int(st["step"])
                end = total
                if ce > 0:
                    end = min(end, (g0 // ce + 1) * ce)
                if g0 < equil_steps:
                    end = min(end, equil_steps)
                n = max(end - g0, 0)
                prod = g0 >= equil_steps and se > 0
                k0 = g0 - equil_steps
                phase = (k0 + 1) % se if prod else 0
                out = nemd_run(self.potential, self.lattice, st["pos"], st["vel"], self.types,
                               self.bath, self.T_hot, self.T_cold, self.damping, dt=self.dt,
                               n_steps=n, seed=self.seed, skin_width=self.skin_width,
                               masses=self.masses, use_graph=self.use_graph,
                               state=st["device_state"], sample_every=se if prod else 0,
                               sample_phase=phase, axis=self.axis, n_bins=self.n_bins)
                last = out
                ep, ek, rows = out["e_pot"], out["e_kin"], out["bath"]
                if st["E0"] is None:
                    st["E0"] = float(ep[0] + ek[0])
                if header:
                    th.write("# step T_hot_K T_cold_K Q_hot_eV Q_cold_eV E_minus_Q_drift_eV\n")
                    drift = float(ep[0] + ek[0]) - st["E0"] - st["Q_hot"] - st["Q_cold"]
                    th.write(f"0 {self._temp_of(rows[0, 2], n_hot):.3f} "
                             f"{self._temp_of(rows[0, 3], n_cold):.3f} {st['Q_hot']:.8e} "
                             f"{st['Q_cold']:.8e} {drift:.6e}\n")
                    header = False
                for s in range(1, n + 1):
                    dh, dc = float(rows[s, 0]), float(rows[s, 1])
                    st["Q_hot"] += dh
                    st["Q_cold"] += dc
                    step = g0 + s
                    k = step - equil_steps
                    if k > 0:
                        blk = min(4 * (k - 1) // max(n_steps, 1), 3)
                        st["Q_hot_prod"] += dh
                        st["Q_cold_prod"] += dc
                        st["n_prod"] += 1
                        st["q_block"][blk] += (dh, dc)
                        st["t_block"][blk] += self.dt
                    if th is not None and step % row_every == 0:
                        drift = (float(ep[s] + ek[s]) - st["E0"] - st["Q_hot"]
                                 - st["Q_cold"])
                        th.write(f"{step} {self._temp_of(rows[s, 2], n_hot):.3f} "
                                 f"{self._temp_of(rows[s, 3], n_cold):.3f} "
                                 f"{st['Q_hot']:.8e} {st['Q_cold']:.8e} {drift:.6e}\n")
                if prod:
                    first = (se - phase) % se
                    for f, b in enumerate(out["bins"]):
                        k = k0 + first + f * se + 1
                        blk = min(4 * (k - 1) // max(n_steps, 1), 3)
                        st["ke_bin"][blk] += b[:, 3]
                        st["n_bin"][blk] += b[:, 4]
                        st["n_samples"] += 1
                        if self.local_flux:
                            st["j_bin"][blk] += self.n_bins * (b[:, :3] @ self.normal)
                            st["n_jsamp"][blk] += 1
                if th is not None:
                    th.flush()
                st["pos"], st["vel"] = out["positions"], out["velocities"]
                st["device_state"] = out["state"]
                st["step"] = end
                if ce > 0 and end % ce == 0 and end < total:
                    self._save(checkpoint, st)
        finally:
            if th is not None:
                th.close()
        if checkpoint:
            self._save(checkpoint, st)
        st["E_minus_Q_drift_eV"] = (float(last["e_pot"][-1] + last["e_kin"][-1]) - st["E0"]
                                    - st["Q_hot"] - st["Q_cold"])
        self.state = st
        return st

    # ------------------------------------------------------------------
    def results(self, fit_fraction=0.5, interfaces=(), kapitza_window=0.1, kapitza_skip=1):
        """T(z) per bin and block, bath heat rates, the flux J = Qdot_hot /
        (2 A) (and from the cold bath), the per-bin local flux along +axis,
        kappa from linear fits over the central ``fit_fraction`` of the two
        segments between the baths with block errors, steady-state
        indicators, and per interface the temperature jump and Kapitza
        resistance from fits within ``kapitza_window`` on each side, leaving
        ``kapitza_skip`` bins next to it out (NEMDSimulation.results,
        solvers/nemd.py:259-337; the same keys)."""
        st = self.state
        if st is None:
            raise RuntimeError("NEMDSimulation.results: call run() first.")
        nb = self.n_bins
        z = (np.arange(nb) + 0.5) / nb
        with np.errstate(invalid="ignore", divide="ignore"):
            T_blk = 2.0 * st["ke_bin"] / (3.0 * K_B * st["n_bin"])
            T = 2.0 * st["ke_bin"].sum(0) / (3.0 * K_B * st["n_bin"].sum(0))
            n_j = st["n_jsamp"].sum()
            j_loc = st["j_bin"].sum(0) / n_j if n_j else np.full(nb, np.nan)
        t_prod = st["n_prod"] * self.dt
        rate_hot = st["Q_hot_prod"] / t_prod
        rate_cold = st["Q_cold_prod"] / t_prod
        J_hot = rate_hot / (2.0 * self.area)
        J_cold = -rate_cold / (2.0 * self.area)
        J = 0.5 * (J_hot + J_cold)

        def fit_slope(x, Tv, idx):
            ok = np.isfinite(Tv[idx])
            return np.polyfit(x[ok], Tv[idx][ok], 1)[0] if ok.sum() >= 2 else np.nan

        segments = []
        for name, a_end, b_start, sign in (
                ("hot->cold (+axis)", self.hot_iv[1], self.cold_iv[0], -1),
                ("cold->hot (+axis)", self.cold_iv[1], self.hot_iv[0], +1)):
            span = (b_start - a_end) % 1.0
            lo = a_end + 0.5 * (1.0 - fit_fraction) * span
            idx = np.flatnonzero((z - lo) % 1.0 < fit_fraction * span)
            x = ((z[idx] - lo) % 1.0) * self.length
            g = fit_slope(x, T, idx)
            g_blk = np.array([fit_slope(x, T_blk[k], idx) for k in range(4)])
            kap = J / abs(g) * EV_ANG_K_FS_TO_W_MK if np.isfinite(g) and g else np.nan
            kap_blk = J / np.abs(g_blk) * EV_ANG_K_FS_TO_W_MK
            segments.append({"segment": name, "bins": idx.tolist(),
                             "gradient_K_per_A": float(g),
                             "gradient_sign_expected": sign,
                             "gradient_sign_ok": bool(np.isfinite(g) and np.sign(g) == sign),
                             "kappa_W_mK": float(kap),
                             "kappa_block_W_mK": kap_blk.tolist(),
                             "kappa_err_W_mK": float(np.std(kap_blk, ddof=1) / 2.0)})
        kapitza = []
        w = int(round(kapitza_window * nb))
        for x0 in interfaces:
            i0 = int(np.floor(x0 * nb + 1e-9))
            left = [(i0 - kapitza_skip - 1 - k) % nb for k in range(w)]
            right = [(i0 + kapitza_skip + k) % nb for k in range(w)]

            def at_interface(ix):
                xs = ((z[ix] - x0 + 0.5) % 1.0 - 0.5) * self.length
                return np.polyval(np.polyfit(xs, T[ix], 1), 0.0)

            dT = float(at_interface(left) - at_interface(right))
            jj = abs(J)
            kapitza.append({"interface_frac": float(x0), "delta_T_K": dT,
                            "R_m2K_per_W": abs(dT) / (jj * EV_ANG2_FS_TO_W_M2)
                            if jj else float("nan")})
        prof = [np.nanmax(np.abs(T_blk[k] - T_blk[k - 1])) for k in (2, 3)]
        return {
            "z_frac": z, "z_A": z * self.length, "T_K": T, "T_block_K": T_blk,
            "atoms_per_bin": st["n_bin"].sum(0) / max(st["n_samples"], 1),
            "local_flux_eV_A2_fs": j_loc,
            "local_flux_W_m2": j_loc * EV_ANG2_FS_TO_W_M2,
            "Q_hot_rate_eV_fs": rate_hot, "Q_cold_rate_eV_fs": rate_cold,
            "flux_hot_W_m2": J_hot * EV_ANG2_FS_TO_W_M2,
            "flux_cold_W_m2": J_cold * EV_ANG2_FS_TO_W_M2,
            "flux_W_m2": J * EV_ANG2_FS_TO_W_M2,
            "bath_balance": float(abs(rate_hot + rate_cold) / max(abs(rate_hot), 1e-300)),
            "profile_change_last_blocks_K": [float(p) for p in prof],
            "segments": segments,
            "kappa_W_mK": float(np.nanmean([s["kappa_W_mK"] for s in segments])),
            "kapitza": kapitza,
            "area_A2": self.area, "length_A": self.length,
            "production_fs": t_prod,
            "E_minus_Q_drift_eV": st.get("E_minus_Q_drift_eV"),
        }
