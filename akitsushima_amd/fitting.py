"""Tersoff fitting on the device with the reference's fitting interface.

``TersoffParameterSet`` maps the optimisation vector x to Tersoff
parameters as fitting/parameter_sets.py:899-1016 does: the keys of
``LOG_KEYS`` are optimised as log(value), h directly; R and D are held at
2.85 and 0.15 and lambda3 follows lambda2 unless fixed; ``bounds`` are the
reference's search box.  ``generic_loss_function`` and ``generic_residuals``
keep the signatures and semantics of fitting/loss.py:51-319 (per-structure
"weight" and "scale", optional reference forces and stress, the Huber
option, 1e30 / 1e15 for a failed evaluation), evaluated for all structures
in one batched device pass (``TersoffFit``).  Added to the reference:
``generic_loss_and_grad`` returns dL/dx from the engine's hand-written
backward pass, chained through the decode (d theta / dx = theta for a log
key, the lambda3 tie folded into lambda2), for gradient-based optimisers
(scipy ``minimize(..., jac=True)``); ``residual_jacobian`` is the forward
difference the reference's least-squares driver builds
(fitting/driver.py:802-827), with every shifted x evaluated on the device.
"""
from __future__ import annotations

import hashlib

import numpy as np

from ._lib import EngineError
from .fit import THETA_KEYS, TersoffFit
from .potentials import PotentialEngine, TersoffPotential

__all__ = ["TersoffParameterSet", "generic_loss_function", "generic_loss_and_grad",
           "generic_residuals", "residual_length", "residual_jacobian"]

FAILED_LOSS = 1e30
FAILED_RESIDUAL = 1e15


class TersoffParameterSet:
    """Optimisation vector <-> flat Tersoff parameters (the single-species
    record of engine/tersoff.py:1071-1229: m = 3, gamma = 1).

    initial_t  : starting values of the Tersoff keys;
    species    : element symbols;
    fix_params : values to hold (default lambda3 = None, meaning "equal to
                 lambda2", R = 2.85, D = 0.15);
    fit_keys   : keys to optimise (default A, B, lambda1, lambda2, beta, n,
                 c, d, h; fixed keys are dropped)."""

    DEFAULT_FIX = {"lambda3": None, "R": 2.85, "D": 0.15}
    DEFAULT_FIT = ("A", "B", "lambda1", "lambda2", "beta", "n", "c", "d", "h")
    LOG_KEYS = frozenset({"A", "B", "lambda1", "lambda2", "beta", "n", "c", "d"})
    # search box of x (log of the value for a log key)
    BOUNDS = {
        "A": (np.log(1.0), np.log(1e5)),
        "B": (np.log(0.1), np.log(1e4)),
        "lambda1": (np.log(0.5), np.log(10.0)),
        "lambda2": (np.log(0.1), np.log(8.0)),
        "beta": (np.log(1e-10), np.log(1.0)),
        "n": (np.log(0.1), np.log(10.0)),
        "c": (np.log(1.0), np.log(1e7)),
        "d": (np.log(0.5), np.log(1e4)),
        "h": (-1.0, 0.5),
        "R": (1.5, 5.0),
        "D": (0.05, 0.5),
    }

    def __init__(self, initial_t, species, fix_params=None, fit_keys=None):
        self.species = list(species)
        self.fixed = dict(self.DEFAULT_FIX)
        self.fixed.update(fix_params or {})
        keys = list(fit_keys) if fit_keys else list(self.DEFAULT_FIT)
        self.fit_keys = [k for k in keys if k not in self.fixed]
        self._initial = dict(initial_t)
        self._x0 = np.array([np.log(max(float(initial_t[k]), 1e-15)) if k in self.LOG_KEYS
                             else float(initial_t[k]) for k in self.fit_keys])
        self._bounds = [self.BOUNDS[k] for k in self.fit_keys]

    @property
    def x0(self):
        return self._x0.copy()

    @property
    def bounds(self):
        return self._bounds

    @property
    def lambda3_tied(self):
        return self.fixed.get("lambda3") is None

    def decode(self, x):
        """{"tersoff": parameter dict} for the optimisation vector x."""
        p = {k: float(v) for k, v in self.fixed.items() if k != "lambda3"}
        for k, v in zip(self.fit_keys, np.asarray(x, float)):
            p[k] = float(np.exp(v)) if k in self.LOG_KEYS else float(v)
        p["lambda3"] = (p.get("lambda2", self._initial.get("lambda2", 1.0)) if self.lambda3_tied
                        else float(self.fixed["lambda3"]))
        return {"tersoff": p}

    def build_engine(self, params):
        return PotentialEngine().add(TersoffPotential(params["tersoff"], species=self.species))

    def theta(self, x):
        """The engine's 14 parameters (THETA_KEYS) for x."""
        p = dict(self.decode(x)["tersoff"], m=3.0, gamma=1.0)
        return np.array([float(p[k]) for k in THETA_KEYS])

    def grad_x(self, x, grad_theta):
        """dL/dx from dL/dtheta: d theta_k / d x_k = theta_k for a log key,
        1 otherwise; with the lambda3 tie, lambda2's component also collects
        dL/dlambda3."""
        p = self.decode(x)["tersoff"]
        g = np.zeros(len(self.fit_keys))
        for i, k in enumerate(self.fit_keys):
            d = grad_theta[THETA_KEYS.index(k)]
            if k == "lambda2" and self.lambda3_tied:
                d += grad_theta[THETA_KEYS.index("lambda3")]
            g[i] = d * (p[k] if k in self.LOG_KEYS else 1.0)
        return g


# ---------------------------------------------------------------------------
# Device loss sessions: one uploaded batch per (structures, weights, loss type)
# ---------------------------------------------------------------------------

_SESSIONS = {}
_MAX_SESSIONS = 4


def _fingerprint(structures, weights, loss_type):
    h = hashlib.blake2b(digest_size=16)
    h.update(repr((loss_type, sorted((k, float(v)) for k, v in weights.items()))).encode())
    for s in structures:
        for key in ("lattice", "positions", "ref_forces", "ref_stress"):
            v = s.get(key)
            h.update(b"-" if v is None else np.ascontiguousarray(v, float).tobytes())
        h.update(repr([float(s.get(k, d)) for k, d in (("ref_energy", 0.0), ("weight", 1.0),
                                                       ("scale", 1.0))]).encode())
        h.update(repr(s.get("volume")).encode())
        h.update(repr(list(s.get("atom_types", []))).encode())
    return h.digest()


def _session(param_set, structures, weights, loss_type):
    if not isinstance(param_set, TersoffParameterSet):
        raise TypeError("the device loss is defined for a TersoffParameterSet")
    if len(set(param_set.species)) != 1:
        raise ValueError("the device loss is defined for a single-species Tersoff set")
    sp = param_set.species[0]
    for s in structures:
        if any(t != sp for t in s.get("atom_types", [sp])):
            raise ValueError(f"every atom must be {sp!r} for this parameter set")
        if "natoms" in s and int(s["natoms"]) != len(s["positions"]):
            raise ValueError("natoms does not match the positions")
    key = _fingerprint(structures, weights, loss_type)
    fit = _SESSIONS.get(key)
    if fit is None:
        fit = TersoffFit(structures, weights=weights, loss_type=loss_type)
        if len(_SESSIONS) >= _MAX_SESSIONS:
            _SESSIONS.pop(next(iter(_SESSIONS)))
        _SESSIONS[key] = fit
    return fit


def _theta_or_none(param_set, x):
    """The engine's parameters for x, or None when x does not decode (the
    reference scores such a point as a failed evaluation)."""
    try:
        theta = param_set.theta(x)
    except Exception:
        return None
    return theta if np.all(np.isfinite(theta)) else None


def _check_loss_type(loss_type, weights):
    """The reference's argument checks (fitting/loss.py:150-158)."""
    if loss_type not in ("squared", "huber"):
        raise ValueError(f"generic_loss_function: unknown loss_type '{loss_type}'. "
                         "Choose 'squared' or 'huber'.")
    if loss_type == "huber":
        for key in ("huber_delta_energy", "huber_delta_forces", "huber_delta_stress"):
            if key not in weights:
                raise ValueError(f"generic_loss_function: loss_type='huber' requires "
                                 f"weights['{key}'] (eV) -- no silent default is guessed.")


def generic_loss_and_grad(x, param_set, structures, weights, r_cut=None, ewald_alpha=None,
                          coulomb_method="ewald", loss_type="squared"):
    """(L, dL/dx): the loss of ``generic_loss_function`` and its gradient in
    the optimisation vector.  A failed evaluation gives (1e30, zeros)."""
    _check_loss_type(loss_type, weights)
    theta = _theta_or_none(param_set, x)
    fit = _session(param_set, structures, weights, loss_type)
    try:
        L, g = fit.loss_and_grad(theta) if theta is not None else (None, None)
    except (ValueError, EngineError):
        L = None
    if L is None:
        return FAILED_LOSS, np.zeros(len(param_set.fit_keys))
    if not np.isfinite(L):
        return L, np.zeros(len(param_set.fit_keys))
    return float(L), param_set.grad_x(x, g)


def generic_loss_function(x, param_set, structures, weights, r_cut=None, ewald_alpha=None,
                          coulomb_method="ewald", loss_type="squared"):
    """Weighted sum of squared (or Huber) residuals of energies, forces and
    stresses (fitting/loss.py:51-248), evaluated on the device for every
    structure at once.  ``r_cut``, ``ewald_alpha`` and ``coulomb_method``
    belong to the reference's Coulomb term, which a Tersoff set has none
    of; they are accepted and unused."""
    _check_loss_type(loss_type, weights)
    theta = _theta_or_none(param_set, x)
    fit = _session(param_set, structures, weights, loss_type)
    if theta is None:
        return FAILED_LOSS
    try:
        return float(fit.loss(theta))
    except (ValueError, EngineError):
        return FAILED_LOSS


def residual_length(structures):
    """Length of the vector generic_residuals returns (fitting/loss.py:251-260)."""
    n = 0
    for s in structures:
        n += 1
        if s.get("ref_forces") is not None:
            n += 3 * len(s["positions"])
        if s.get("ref_stress") is not None:
            n += 9
    return n


def generic_residuals(x, param_set, structures, weights, r_cut=None, ewald_alpha=None,
                      coulomb_method="ewald"):
    """Residual vector r with sum(r**2) == generic_loss_function(...) for the
    squared loss (fitting/loss.py:263-319): per structure, in order,
    sqrt(w_E) N dE_per_atom; sqrt(w_F / 3N) r0 dF (3N); sqrt(w_S / 9) V eps0
    dsigma (9), each times sqrt(weight) scale.  Every entry is 1e15 when the
    evaluation fails."""
    n_total = residual_length(structures)
    theta = _theta_or_none(param_set, x)
    fit = _session(param_set, structures, weights, "squared")
    if theta is None:
        return np.full(n_total, FAILED_RESIDUAL)
    try:
        _, _, (E, F, S) = fit.evaluate(theta, want_grad=False, want_outputs=True)
    except (ValueError, EngineError):
        return np.full(n_total, FAILED_RESIDUAL)
    w_E, w_F = weights.get("energy", 1.0), weights.get("forces", 1.0)
    w_S = weights.get("stress", 0.1)
    r0, eps0 = float(weights.get("r0", 1.0)), float(weights.get("eps0", 1.0))
    out = np.empty(n_total)
    at = 0
    for k, s in enumerate(structures):
        N = len(s["positions"])
        a0, a1 = fit.off[k], fit.off[k + 1]
        c = np.sqrt(float(s.get("weight", 1.0))) * float(s.get("scale", 1.0))
        out[at] = c * np.sqrt(w_E) * N * (E[k] / N - float(s["ref_energy"]))
        at += 1
        if s.get("ref_forces") is not None:
            d = (F[a0:a1] - np.asarray(s["ref_forces"], float)).ravel()
            out[at:at + 3 * N] = c * np.sqrt(w_F / (3.0 * N)) * r0 * d
            at += 3 * N
        if s.get("ref_stress") is not None:
            vol = float(s.get("volume", abs(np.linalg.det(np.asarray(s["lattice"], float)))))
            d = (S[k] - np.asarray(s["ref_stress"], float)).ravel()
            out[at:at + 9] = c * np.sqrt(w_S / 9.0) * vol * eps0 * d
            at += 9
    return out


def residual_jacobian(x, param_set, structures, weights, r_cut=None, ewald_alpha=None,
                      coulomb_method="ewald", r0=None):
    """d r / d x by forward differences, step 1e-7 max(1, |x_i|), backward at
    an upper bound (the Jacobian of the reference's least-squares driver,
    fitting/driver.py:802-827); r0 = the residuals at x when already known."""
    x = np.asarray(x, float)
    base = (generic_residuals(x, param_set, structures, weights) if r0 is None
            else np.asarray(r0, float))
    hi = np.array([b[1] for b in param_set.bounds], float)
    J = np.empty((len(base), len(x)))
    for i in range(len(x)):
        h = 1e-7 * max(1.0, abs(x[i]))
        if x[i] + h > hi[i]:
            h = -h
        xs = x.copy()
        xs[i] += h
        J[:, i] = (generic_residuals(xs, param_set, structures, weights) - base) / h
    return J
