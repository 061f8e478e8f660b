"""Differentiable Tersoff fitting loss.

``TersoffFit`` holds a batch of structures with reference energies (eV/atom),
forces and stresses and evaluates the squared loss of fitting/loss.py:212-242
(default weights w_E 1, w_F 1, w_S 0.1, r0 1, eps0 1, fitting/loss.py:184-188)
for the 14 Tersoff parameters

    theta = A, B, lambda1, lambda2, lambda3, beta, n, c, d, h, R, D, m, gamma

together with dL/dtheta from the engine's hand-written backward kernel (the
reference has no parameter gradient: its fits use derivative-free or
finite-difference optimisers, fitting/driver.py:520-541).  All structures go
through the device in one batched pass.  ``tersoff_loss(theta, fit)`` wraps
it as a ``torch.autograd.Function``.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _lib
from ._lib import check, lib
from .potentials import TersoffPotential

THETA_KEYS = ("A", "B", "lambda1", "lambda2", "lambda3", "beta", "n", "c", "d", "h",
              "R", "D", "m", "gamma")
DEFAULT_WEIGHTS = {"energy": 1.0, "forces": 1.0, "stress": 0.1, "r0": 1.0, "eps0": 1.0}


def theta_from_params(params, m=3.0, gamma=1.0):
    p = dict(params)
    p.setdefault("m", m)
    p.setdefault("gamma", gamma)
    return np.array([float(p[k]) for k in THETA_KEYS])


def params_from_theta(theta):
    return {k: float(v) for k, v in zip(THETA_KEYS, theta)}


class TersoffFit:
    """structures: list of dicts with lattice, positions, ref_energy (eV per
    atom), ref_forces (N, 3) and ref_stress (3, 3) (either may be None or
    absent: its term is left out), and optionally the per-structure keys of
    fitting/loss.py:212-242: "weight" (trust, linear), "scale" (unit bridge,
    squared) and "volume" (of the stress scale; default |det lattice|).
    weights: w_E ("energy"), w_F ("forces"), w_S ("stress"), "r0", "eps0";
    with loss_type "huber" also "huber_delta_energy", "huber_delta_forces"
    and "huber_delta_stress" (fitting/loss.py:27-48, 166-175)."""

    def __init__(self, structures, weights=None, autotune=True, loss_type="squared"):
        w = dict(DEFAULT_WEIGHTS)
        w.update(weights or {})
        self.weights = np.array([w["energy"], w["forces"], w["stress"], w["r0"], w["eps0"]],
                                float)
        if loss_type not in ("squared", "huber"):
            raise ValueError(f"TersoffFit: unknown loss_type '{loss_type}'. Choose 'squared' "
                             "or 'huber'.")
        self.loss_type = loss_type
        self.huber = None
        if loss_type == "huber":
            keys = ("huber_delta_energy", "huber_delta_forces", "huber_delta_stress")
            for k in keys:
                if k not in w:
                    raise ValueError(f"TersoffFit: loss_type='huber' requires weights['{k}'] "
                                     "(eV) -- no silent default is guessed.")
            self.huber = np.array([float(w[k]) for k in keys])
        S = len(structures)
        self.S = S
        self.off = np.zeros(S + 1, np.int64)
        for s, st in enumerate(structures):
            self.off[s + 1] = self.off[s] + len(st["positions"])
        self.lat = np.ascontiguousarray(np.stack([np.asarray(st["lattice"], float)
                                                  for st in structures]))
        self.pos = np.ascontiguousarray(np.concatenate([np.asarray(st["positions"], float)
                                                        for st in structures]))
        self.ref_e = np.ascontiguousarray([float(st["ref_energy"]) for st in structures])
        has_f = [st.get("ref_forces") is not None for st in structures]
        has_s = [st.get("ref_stress") is not None for st in structures]
        self.ref_f = np.ascontiguousarray(np.concatenate(
            [np.asarray(st["ref_forces"], float) if h else np.zeros((len(st["positions"]), 3))
             for st, h in zip(structures, has_f)]))
        self.ref_s = np.ascontiguousarray(np.stack(
            [np.asarray(st["ref_stress"], float) if h else np.zeros((3, 3))
             for st, h in zip(structures, has_s)]))
        self.terms = np.ascontiguousarray(np.array(
            [[float(st.get("weight", 1.0)), float(st.get("scale", 1.0)),
              float(st.get("volume", abs(np.linalg.det(np.asarray(st["lattice"], float))))),
              float((1 if hf else 0) | (2 if hs else 0))]
             for st, hf, hs in zip(structures, has_f, has_s)], float).reshape(S, 4))
        self._pot = TersoffPotential({"A": 1.0, "B": 1.0, "lambda1": 1.0, "lambda2": 1.0,
                                      "lambda3": 0.0, "beta": 1.0, "n": 1.0, "c": 1.0,
                                      "d": 1.0, "h": 0.0, "R": 2.85, "D": 0.15})
        self._tls = threading.local()   # per thread: the handle holding the batch
        self._autotune = bool(autotune) and S > 1
        self.launch_shape = None

    def _handle(self):
        """This thread's potential handle, with the batch uploaded once."""
        h = self._pot.handle()
        if getattr(self._tls, "loaded", None) != h:
            P = _lib._pf64
            check(lib().ak_tersoff_fit_load(
                h, self.S, self.off.ctypes.data_as(_lib._pi64), self.lat.ctypes.data_as(P),
                self.pos.ctypes.data_as(P), self.ref_e.ctypes.data_as(P),
                self.ref_f.ctypes.data_as(P), self.ref_s.ctypes.data_as(P)))
            deltas = self.huber if self.huber is not None else np.zeros(3)
            check(lib().ak_tersoff_fit_terms(h, self.S, self.terms.ctypes.data_as(P),
                                             1 if self.huber is not None else 0,
                                             deltas.ctypes.data_as(P)))
            self._tls.loaded = h
            if self._autotune and self.launch_shape is None:
                # launch shape of the forward force pass, timed on this batch
                cells = [(self.lat[k], self.pos[self.off[k]:self.off[k + 1]],
                          ["Si"] * int(self.off[k + 1] - self.off[k])) for k in range(self.S)]
                self.launch_shape = self._pot.autotune(cells, reps=2)["best"]
        return h

    def evaluate(self, theta, want_grad=True, want_outputs=False):
        """(loss, grad (14,) or None[, (E (S,), F (sum N, 3), sigma (S, 3, 3))])."""
        th = np.ascontiguousarray(theta, float)
        if th.shape != (14,):
            raise ValueError("theta must hold the 14 Tersoff parameters")
        P = _lib._pf64
        loss = ctypes.c_double(0.0)
        grad = np.zeros(14) if want_grad else None
        E = np.zeros(self.S) if want_outputs else None
        F = np.zeros_like(self.pos) if want_outputs else None
        S = np.zeros((self.S, 3, 3)) if want_outputs else None
        ptr = lambda a: a.ctypes.data_as(P) if a is not None else None
        check(lib().ak_tersoff_fit_eval(self._handle(), ptr(th), ptr(self.weights),
                                        ctypes.byref(loss), ptr(grad), ptr(E), ptr(F), ptr(S)))
        if want_outputs:
            return loss.value, grad, (E, F, S)
        return loss.value, grad

    def loss(self, theta):
        return self.evaluate(theta, want_grad=False)[0]

    def loss_and_grad(self, theta):
        return self.evaluate(theta, want_grad=True)


def _function():
    import torch

    class TersoffLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, fit):
            L, g = fit.loss_and_grad(theta.detach().cpu().numpy().astype(np.float64))
            ctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype, device=theta.device))
            return theta.new_tensor(L)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g, None

    return TersoffLoss


_TL = None


def tersoff_loss(theta, fit):
    """Differentiable loss: a torch scalar whose backward is the engine's
    dL/dtheta.  ``theta`` is a float64 tensor of the 14 parameters."""
    global _TL
    if _TL is None:
        _TL = _function()
    return _TL.apply(theta, fit)
