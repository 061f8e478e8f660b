"""Non-equilibrium MD with two Langevin heat baths on the device
(NEMDSimulation, solvers/nemd.py:1-337): temperature profile, local heat
flux from the per-atom heat currents, thermal conductivity from the
gradient and the Kapitza resistance of given interfaces.

The velocity Verlet step, the Ornstein-Uhlenbeck update of the bath atoms
and the per-bin sums of the samples (kinetic energy, atom count, per-atom
heat currents) run in one captured hipGraph step (``ak_nemd_run``).  The
host keeps the production sums in ``_Production`` (four consecutive blocks
of the production steps) and derives ``results`` from them; ``run`` returns
them as the reference's accumulator dictionary.  The initial velocities are
the reference's draw from numpy's generator (the same numbers for the same
seed); the bath noise is the Philox4x32-10 stream keyed by the seed (the
reference continues numpy's generator: same process, other realisations),
so a checkpoint needs no host generator: it holds the state, the sums and
the device's random-stream position, and a run continued from it ends where
one written with the same checkpoint spacing does, bit for bit.
"""
from __future__ import annotations

import os

import numpy as np

from .md import EV_ANG_TO_AMU_ANG_FS2, K_B, STATE_LEN, masses_of, nemd_run
from .simulations import _potential_of

__all__ = ["NEMDSimulation"]

# eV / (A K fs) -> W / (m K) and eV / (A^2 fs) -> W / m^2 (solvers/nemd.py:26-27)
EV_ANG_K_FS_TO_W_MK = 1.602176634e-19 / (1e-10 * 1e-15)
EV_ANG2_FS_TO_W_M2 = 1.602176634e-19 / (1e-20 * 1e-15)
_AXIS = {"a": 0, "b": 1, "c": 2, 0: 0, 1: 1, 2: 2}
N_BLOCKS = 4


def _in_interval(f, lo, hi):
    """Membership of fractional coordinates f in [lo, hi) taken on the unit
    circle: an interval with lo > hi wraps through 0, lo == hi is the whole
    circle."""
    if lo < hi:
        return (f >= lo) & (f < hi)
    return ~((f >= hi) & (f < lo))


def _bath_members(frac_axis, interval):
    """Atoms of a bath given as a fractional slab [lo, hi) of the initial
    positions.  Coordinates are shifted by 1e-9 before wrapping, so a layer
    sitting at 0 (or a rounding error below it) counts as 0, not as ~1."""
    lo, hi = (float(v) % 1.0 for v in interval)
    return np.flatnonzero(_in_interval((frac_axis + 1e-9) % 1.0, lo, hi))


def _slope(x, y):
    """Least-squares slope of y(x) over its finite entries (NaN below two)."""
    ok = np.isfinite(y)
    if np.count_nonzero(ok) < 2:
        return np.nan
    return np.polyfit(x[ok], y[ok], 1)[0]


def _line_at(x, y, x0):
    """Value at x0 of the least-squares line through (x, y)."""
    return np.polyval(np.polyfit(x, y, 1), x0)


class _Production:
    """Sums over the production phase.  Step k >= 1 of n production steps
    belongs to block min(4 (k - 1) // n, 3); per block: the bath heats and
    time, and per bin the sampled kinetic energy, atom count and local flux
    (with its sample count).  ``as_dict`` gives the reference's accumulator
    names; ``arrays`` / ``from_arrays`` the checkpoint form."""

    SCALARS = ("step", "n_prod", "n_samples", "Q_hot", "Q_cold", "Q_hot_prod", "Q_cold_prod",
               "E0", "E_last")

    def __init__(self, n_bins):
        self.step = self.n_prod = self.n_samples = 0
        self.Q_hot = self.Q_cold = self.Q_hot_prod = self.Q_cold_prod = 0.0
        self.E0 = self.E_last = None
        self.bins = np.zeros((3, N_BLOCKS, n_bins))       # ke, atom count, local flux
        self.n_jsamp = np.zeros(N_BLOCKS)
        self.blocks = np.zeros((N_BLOCKS, 3))             # Q_hot, Q_cold, time

    @staticmethod
    def block_of(k, n_prod_total):
        return min(N_BLOCKS * (k - 1) // max(n_prod_total, 1), N_BLOCKS - 1)

    def add_step(self, dh, dc, k, n_prod_total, dt):
        self.Q_hot += dh
        self.Q_cold += dc
        if k > 0:
            b = self.block_of(k, n_prod_total)
            self.Q_hot_prod += dh
            self.Q_cold_prod += dc
            self.n_prod += 1
            self.blocks[b] += (dh, dc, dt)

    def add_sample(self, b, ke, count, flux=None):
        self.bins[0, b] += ke
        self.bins[1, b] += count
        self.n_samples += 1
        if flux is not None:
            self.bins[2, b] += flux
            self.n_jsamp[b] += 1

    def as_dict(self):
        d = {k: getattr(self, k) for k in self.SCALARS}
        d.update(ke_bin=self.bins[0], n_bin=self.bins[1], j_bin=self.bins[2],
                 n_jsamp=self.n_jsamp, q_block=self.blocks[:, :2], t_block=self.blocks[:, 2])
        return d

    def arrays(self):
        nan = lambda v: np.nan if v is None else float(v)
        return {"scalars": np.array([nan(getattr(self, k)) for k in self.SCALARS]),
                "bins": self.bins, "n_jsamp": self.n_jsamp, "blocks": self.blocks}

    @classmethod
    def from_arrays(cls, d):
        p = cls(d["bins"].shape[2])
        for k, v in zip(cls.SCALARS, d["scalars"]):
            if k in ("step", "n_prod", "n_samples"):
                setattr(p, k, int(v))
            else:
                setattr(p, k, None if np.isnan(v) else float(v))
        p.bins, p.n_jsamp, p.blocks = d["bins"].copy(), d["n_jsamp"].copy(), d["blocks"].copy()
        return p


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
        self.lattice = np.array(lattice, float)
        self.pos = np.array(positions, float)
        self.types = list(atom_types)
        self.N = len(self.types)
        self.masses = (np.asarray(masses, float) if masses is not None
                       else masses_of(self.types))
        self.T_hot, self.T_cold = float(T_hot), float(T_cold)
        self.axis = _AXIS[axis]
        self.damping, self.dt = float(damping), float(timestep)
        self.n_bins = int(n_bins)
        self.local_flux = bool(local_flux)
        self.seed = int(seed)
        self.skin_width = float(skin_width)
        self.use_graph = bool(use_graph)
        if self.local_flux and not hasattr(self.potential, "site_terms"):
            raise ValueError("NEMDSimulation: local_flux needs a potential with a site-energy "
                             "decomposition")
        along = self._frac(self.pos)[:, self.axis]
        self.hot = (np.asarray(hot_atoms, int) if hot_atoms is not None
                    else _bath_members(along, hot))
        self.cold = (np.asarray(cold_atoms, int) if cold_atoms is not None
                     else _bath_members(along, cold))
        if self.hot.size == 0 or self.cold.size == 0:
            raise ValueError("NEMDSimulation: a heat bath has no atoms.")
        if np.intersect1d(self.hot, self.cold).size:
            raise ValueError("NEMDSimulation: the hot and cold baths overlap.")
        self.bath = np.zeros(self.N, np.uint8)
        self.bath[self.hot], self.bath[self.cold] = 1, 2
        self.hot_iv = (along[self.hot].min(), along[self.hot].max())
        self.cold_iv = (along[self.cold].min(), along[self.cold].max())
        # cross-section spanned by the other two lattice vectors
        u, w = (self.lattice[k] for k in range(3) if k != self.axis)
        normal = np.cross(u, w)
        self.area = float(np.linalg.norm(normal))
        self.volume = abs(float(np.linalg.det(self.lattice)))
        self.length = self.volume / self.area      # width of the cell along the normal
        direction = 1.0 if normal @ self.lattice[self.axis] > 0 else -1.0
        self.normal = direction * normal / self.area
        self.state = None

    def _frac(self, pos):
        f = pos @ np.linalg.inv(self.lattice)
        return f - np.floor(f)

    def _bath_temperature(self, k_bath, n_atoms):
        return 2.0 * k_bath / (3.0 * n_atoms * K_B)

    # ------------------------------------------------------------------
    # state and checkpoints
    def _start(self):
        """Positions, velocities (the reference's draw: N(0, k_B T0 / m) per
        component from default_rng(seed), T0 the mean bath temperature,
        centre-of-mass momentum removed) and empty sums."""
        T0 = 0.5 * (self.T_hot + self.T_cold)
        scale = np.sqrt(K_B * T0 * EV_ANG_TO_AMU_ANG_FS2 / self.masses)
        vel = np.random.default_rng(self.seed).standard_normal((self.N, 3)) * scale[:, None]
        p_tot = self.masses @ vel
        vel -= p_tot / self.masses.sum()
        return self.pos.copy(), vel, np.zeros(STATE_LEN), _Production(self.n_bins)

    def _write_checkpoint(self, path, pos, vel, dev, prod):
        tmp = path + ".tmp.npz"
        np.savez(tmp, positions=pos, velocities=vel, device_state=dev, **prod.arrays())
        os.replace(tmp, path)

    def _read_checkpoint(self, path):
        with np.load(path, allow_pickle=False) as d:
            needed = {"positions", "velocities", "device_state", "scalars", "bins"}
            if not needed <= set(d.files):
                raise ValueError(f"NEMDSimulation: {path} is not a checkpoint of this engine; "
                                 "start a fresh run")
            return (d["positions"].copy(), d["velocities"].copy(), d["device_state"].copy(),
                    _Production.from_arrays(d))

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
        resume = bool(checkpoint) and os.path.isfile(checkpoint)
        pos, vel, dev, prod = self._read_checkpoint(checkpoint) if resume else self._start()
        if verbose and prod.step:
            print(f"  NEMD: restarting from {checkpoint} at step {prod.step}")
        n_eq, n_prod = int(equil_steps), int(n_steps)
        every = int(sample_every)
        total = n_eq + n_prod
        spacing = int(checkpoint_every) if checkpoint else 0
        log_every = max(10 * every, 1)
        log = open(thermo_path, "a") if thermo_path else None
        n_hot, n_cold = self.hot.size, self.cold.size

        def log_row(step, k_hot, k_cold, e_tot):
            drift = e_tot - prod.E0 - prod.Q_hot - prod.Q_cold
            log.write(f"{step} {self._bath_temperature(k_hot, n_hot):.3f} "
                      f"{self._bath_temperature(k_cold, n_cold):.3f} {prod.Q_hot:.8e} "
                      f"{prod.Q_cold:.8e} {drift:.6e}\n")

        try:
            # one device call per chunk: up to the next checkpoint, and the
            # equilibration and production phases in separate calls
            while prod.step < total or prod.E_last is None:
                start = prod.step
                stop = total
                if spacing > 0:
                    stop = min(stop, (start // spacing + 1) * spacing)
                if start < n_eq:
                    stop = min(stop, n_eq)
                n = stop - start
                sampling = start >= n_eq and every > 0
                k0 = start - n_eq                        # production steps done
                phase = (k0 + 1) % every if sampling else 0
                out = nemd_run(self.potential, self.lattice, pos, vel, self.types, self.bath,
                               self.T_hot, self.T_cold, self.damping, dt=self.dt, n_steps=n,
                               seed=self.seed, skin_width=self.skin_width, masses=self.masses,
                               use_graph=self.use_graph, state=dev,
                               sample_every=every if sampling and n > 0 else 0,
                               sample_phase=phase, axis=self.axis, n_bins=self.n_bins)
                e_tot = out["e_pot"] + out["e_kin"]
                rec = out["bath"]                        # dQ_hot, dQ_cold, K_hot, K_cold
                if prod.E0 is None:
                    prod.E0 = float(e_tot[0])
                    if log is not None:
                        log.write("# step T_hot_K T_cold_K Q_hot_eV Q_cold_eV "
                                  "E_minus_Q_drift_eV\n")
                        log_row(0, rec[0, 2], rec[0, 3], float(e_tot[0]))
                for s in range(1, n + 1):
                    step = start + s
                    prod.add_step(float(rec[s, 0]), float(rec[s, 1]), step - n_eq, n_prod,
                                  self.dt)
                    if log is not None and step % log_every == 0:
                        log_row(step, rec[s, 2], rec[s, 3], float(e_tot[s]))
                if sampling and n > 0:
                    # the device sampled steps k0 + first + f * every + 1
                    first = (every - phase) % every
                    for f, b in enumerate(out["bins"]):
                        blk = _Production.block_of(k0 + first + f * every + 1, n_prod)
                        flux = (self.n_bins * (b[:, :3] @ self.normal)
                                if self.local_flux else None)
                        prod.add_sample(blk, b[:, 3], b[:, 4], flux)
                if log is not None:
                    log.flush()
                pos, vel, dev = out["positions"], out["velocities"], out["state"]
                prod.step = stop
                prod.E_last = float(e_tot[-1])
                if spacing > 0 and stop % spacing == 0 and stop < total:
                    self._write_checkpoint(checkpoint, pos, vel, dev, prod)
        finally:
            if log is not None:
                log.close()
        if checkpoint:
            self._write_checkpoint(checkpoint, pos, vel, dev, prod)
        st = prod.as_dict()
        st.update(pos=pos, vel=vel, device_state=dev,
                  E_minus_Q_drift_eV=prod.E_last - prod.E0 - prod.Q_hot - prod.Q_cold)
        self.state = st
        return st

    # ------------------------------------------------------------------
    def _segment(self, name, leave, reach, sign, z, T, T_blk, J, fit_fraction):
        """Bins of the central fit_fraction of the stretch from the end of
        one bath to the start of the next along +axis; the temperature
        gradient over them (all blocks, and per block) and kappa = J / |g|."""
        span = (reach - leave) % 1.0
        lo = leave + 0.5 * (1.0 - fit_fraction) * span
        offset = (z - lo) % 1.0
        idx = np.flatnonzero(offset < fit_fraction * span)
        x = offset[idx] * self.length
        g = _slope(x, T[idx])
        g_blk = np.array([_slope(x, T_blk[b][idx]) for b in range(N_BLOCKS)])
        kappa = J / abs(g) * EV_ANG_K_FS_TO_W_MK if np.isfinite(g) and g != 0 else np.nan
        kappa_blk = J / np.abs(g_blk) * EV_ANG_K_FS_TO_W_MK
        return {"segment": name, "bins": idx.tolist(), "gradient_K_per_A": float(g),
                "gradient_sign_expected": sign,
                "gradient_sign_ok": bool(np.isfinite(g) and np.sign(g) == sign),
                "kappa_W_mK": float(kappa), "kappa_block_W_mK": kappa_blk.tolist(),
                "kappa_err_W_mK": float(np.std(kappa_blk, ddof=1) / np.sqrt(N_BLOCKS))}

    def _interface(self, x0, z, T, J, window, skip):
        """Temperature jump at fractional position x0 from lines fitted to
        `window` bins on each side (`skip` bins next to x0 left out), each
        extrapolated to x0; R = |dT| / |J|."""
        nb = self.n_bins
        first_right = int(np.floor(x0 * nb + 1e-9))
        left = [(first_right - skip - 1 - k) % nb for k in range(window)]
        right = [(first_right + skip + k) % nb for k in range(window)]
        rel = lambda ix: ((z[ix] - x0 + 0.5) % 1.0 - 0.5) * self.length
        dT = float(_line_at(rel(left), T[left], 0.0) - _line_at(rel(right), T[right], 0.0))
        flux = abs(J) * EV_ANG2_FS_TO_W_M2
        return {"interface_frac": float(x0), "delta_T_K": dT,
                "R_m2K_per_W": abs(dT) / flux if flux else float("nan")}

    def results(self, fit_fraction=0.5, interfaces=(), kapitza_window=0.1, kapitza_skip=1):
        """T(z) per bin and block, bath heat rates, the flux J = Qdot_hot /
        (2 A) (and from the cold bath), the per-bin local flux along +axis,
        kappa from linear fits over the central ``fit_fraction`` of the two
        segments between the baths with block errors, steady-state
        indicators, and per interface the temperature jump and Kapitza
        resistance from fits within ``kapitza_window`` on each side, leaving
        ``kapitza_skip`` bins next to it out (the outputs of
        NEMDSimulation.results, solvers/nemd.py:259-337, under its keys)."""
        st = self.state
        if st is None:
            raise RuntimeError("NEMDSimulation.results: call run() first.")
        nb = self.n_bins
        z = (np.arange(nb) + 0.5) / nb
        ke, cnt = np.asarray(st["ke_bin"]), np.asarray(st["n_bin"])
        with np.errstate(invalid="ignore", divide="ignore"):
            T_blk = 2.0 * ke / (3.0 * K_B * cnt)
            T = 2.0 * ke.sum(0) / (3.0 * K_B * cnt.sum(0))
        n_j = float(np.sum(st["n_jsamp"]))
        j_loc = np.asarray(st["j_bin"]).sum(0) / n_j if n_j else np.full(nb, np.nan)
        t_prod = st["n_prod"] * self.dt
        rate_hot, rate_cold = st["Q_hot_prod"] / t_prod, st["Q_cold_prod"] / t_prod
        # both ways round the periodic cell: Qdot / (2 A)
        J_hot, J_cold = rate_hot / (2.0 * self.area), -rate_cold / (2.0 * self.area)
        J = 0.5 * (J_hot + J_cold)
        segments = [self._segment("hot->cold (+axis)", self.hot_iv[1], self.cold_iv[0], -1,
                                  z, T, T_blk, J, fit_fraction),
                    self._segment("cold->hot (+axis)", self.cold_iv[1], self.hot_iv[0], +1,
                                  z, T, T_blk, J, fit_fraction)]
        window = int(round(kapitza_window * nb))
        kapitza = [self._interface(x0, z, T, J, window, kapitza_skip) for x0 in interfaces]
        change = [float(np.nanmax(np.abs(T_blk[b] - T_blk[b - 1]))) for b in (2, 3)]
        to_si = EV_ANG2_FS_TO_W_M2
        return {
            "z_frac": z, "z_A": z * self.length, "T_K": T, "T_block_K": T_blk,
            "atoms_per_bin": cnt.sum(0) / max(st["n_samples"], 1),
            "local_flux_eV_A2_fs": j_loc, "local_flux_W_m2": j_loc * to_si,
            "Q_hot_rate_eV_fs": rate_hot, "Q_cold_rate_eV_fs": rate_cold,
            "flux_hot_W_m2": J_hot * to_si, "flux_cold_W_m2": J_cold * to_si,
            "flux_W_m2": J * to_si,
            "bath_balance": float(abs(rate_hot + rate_cold) / max(abs(rate_hot), 1e-300)),
            "profile_change_last_blocks_K": change,
            "segments": segments,
            "kappa_W_mK": float(np.nanmean([s["kappa_W_mK"] for s in segments])),
            "kapitza": kapitza,
            "area_A2": self.area, "length_A": self.length,
            "production_fs": t_prod,
            "E_minus_Q_drift_eV": st.get("E_minus_Q_drift_eV"),
        }
