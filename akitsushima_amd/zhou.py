"""Host-side construction of the Zhou-Johnson-Wadley alloy EAM.

Follows engine/eam_zhou.py:47-207: ``continuous`` re-derives Fn0, Fn1, F0
and F_e so that F and dF/drho are continuous at rho_n = 0.85 rho_e and
rho_0 = 1.15 rho_e; ``zhou_cutoff`` scans r_e..5 r_e on a 40001-point grid
for the distance beyond which |phi_aa| < tol max(A, B) and f < tol f_e; every
function is multiplied by a C2 quintic switch over the last ``width`` before
r_c; the cross pair phi_ab = chi/2 [f_b/f_a phi_aa + f_a/f_b phi_bb] is
formed from unswitched functions, then switched.  The result is a table of
function rows for the GPU kernels.
"""
from __future__ import annotations

import math

import numpy as np

from .potentials import EAMPotential
from .published import ZHOU_2004_KEYS

KEYS = ZHOU_2004_KEYS
ROW = 44  # [code, r_c, width, chi, pa[20], pb[20]]
CODE_DENSITY, CODE_PAIR, CODE_CROSS, CODE_EMBED = 3, 4, 5, 6


def continuous(p):
    """Copy of an element's parameters with Fn0, Fn1, F0, F_e re-derived
    for continuity of F and F' at rho_n and rho_0 (F1 as given)."""
    q = {k: float(v) for k, v in p.items()}
    re, rs, eta = q["rho_e"], q["rho_s"], q["eta"]
    F1, F2, F3 = q["F1"], q["F2"], q["F3"]
    y, x = 0.15, -0.15
    t0 = 1.15 * re / rs
    slope0 = (F1 + 2 * F2 * y + 3 * F3 * y * y) / re
    dg = -eta * eta * math.log(t0) * t0 ** (eta - 1.0) / rs
    if dg == 0.0:
        raise ValueError("Zhou EAM: rho_0 = rho_s (or eta = 0) leaves F_e undetermined "
                         "by the slope condition")
    q["F_e"] = slope0 / dg
    q["F0"] = q["F_e"] * (1 - eta * math.log(t0)) * t0 ** eta - (F1 * y + F2 * y * y + F3 * y ** 3)
    q["Fn0"] = q["F0"] + F1 * x + F2 * x * x + F3 * x ** 3
    q["Fn1"] = 0.85 * (F1 + 2 * F2 * x + 3 * F3 * x * x)
    return q


def _decay(r, re, pre, expo, shift):
    x = r / re
    e = pre * np.exp(-expo * (x - 1.0))
    u = x - shift
    return e / (1.0 + u ** 20)


def zhou_cutoff(p, tol=1e-6):
    r = np.linspace(p["r_e"], 5.0 * p["r_e"], 40001)
    pair = _decay(r, p["r_e"], p["A"], p["alpha"], p["kappa"]) - \
        _decay(r, p["r_e"], p["B"], p["beta"], p["lambda"])
    dens = _decay(r, p["r_e"], p["f_e"], p["beta"], p["lambda"])
    ok = (np.abs(pair) < tol * max(p["A"], p["B"])) & (dens < tol * p["f_e"])
    bad = np.nonzero(~ok)[0]
    if len(bad) == len(r):
        raise ValueError("Zhou EAM: functions do not decay within 5 r_e")
    return float(r[bad[-1] + 1]) if len(bad) else float(r[0])


def function_table(P, species, r_cut, chi=None, width=0.5):
    """Rows and the index map [rho[ns], emb[ns], phi[ns*ns]]."""
    chi = chi or {}
    rows = []

    def row(code, pa, pb=None, k=1.0):
        v = np.zeros(ROW)
        v[0], v[1], v[2], v[3] = code, r_cut, width, k
        v[4:24] = [pa[kk] for kk in KEYS]
        if pb is not None:
            v[24:44] = [pb[kk] for kk in KEYS]
        rows.append(v)
        return len(rows) - 1

    ns = len(species)
    rho = [row(CODE_DENSITY, P[s]) for s in species]
    emb = [row(CODE_EMBED, P[s]) for s in species]
    phi = np.zeros((ns, ns), np.int32)
    for a in range(ns):
        for b in range(a, ns):
            sa, sb = species[a], species[b]
            if a == b:
                k = row(CODE_PAIR, P[sa])
            else:
                cf = float(chi.get(f"{sa}-{sb}", chi.get(f"{sb}-{sa}", 1.0)))
                k = row(CODE_CROSS, P[sa], P[sb], cf)
            phi[a, b] = phi[b, a] = k
    index_map = np.concatenate([np.array(rho, np.int32), np.array(emb, np.int32),
                                phi.reshape(-1)]).astype(np.int32)
    return np.array(rows), index_map


def zhou_eam(elements, species, r_cut=None, chi=None, width=0.5, tol=1e-6,
             enforce_continuity=True):
    """EAMPotential over the GPU engine from Zhou (2004) element parameters
    (engine/eam_zhou.py:128-207)."""
    species = list(species)
    P = {}
    for s in species:
        if s not in elements:
            raise ValueError(f"Zhou EAM: no parameters for '{s}'")
        missing = [k for k in KEYS if k not in elements[s]]
        if missing:
            raise ValueError(f"Zhou EAM: '{s}' lacks {missing}")
        P[s] = continuous(elements[s]) if enforce_continuity else {
            k: float(v) for k, v in elements[s].items()}
    if r_cut is None:
        r_cut = max(zhou_cutoff(P[s], tol) for s in species)
    rows, index_map = function_table(P, species, float(r_cut), chi, width)
    pot = EAMPotential(rows, index_map, float(r_cut), species)
    pot.zhou_params = P
    return pot
