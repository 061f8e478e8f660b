"""setfl (LAMMPS ``eam/alloy``) files for the GPU EAM.

``read_setfl`` follows engine/eam_zhou.py:257-315: every table becomes a
scipy ``CubicSpline`` (not-a-knot ends, as the reference builds it), the
pair tables hold r*phi and phi = s(r) / r; the spline coefficients go to the
device, where the kernels evaluate them as engine/_eam_kernels.py:128-145
does (interval from searchsorted 'right', end polynomials beyond the knots,
Horner).  ``write_setfl`` follows engine/eam_zhou.py:216-254; the functions
are tabulated on the device (``EAMPotential.evaluate``).
"""
from __future__ import annotations

import numpy as np

from .potentials import EAMPotential
from .published import ATOMIC_MASS, ATOMIC_NUMBER

ROW = 44
CODE_SPLINE, CODE_SPLINE_OVER_R = 1, 2


def _block(vals):
    vals = list(vals)
    return "\n".join(" ".join(f"{v: .16e}" for v in vals[i:i + 5])
                     for i in range(0, len(vals), 5)) + "\n"


def write_setfl(path, pot, species, n_rho=5000, rho_max=None, n_r=5000, comments=("", "", ""),
                lattice=None):
    """Tabulate ``pot`` (an :class:`EAMPotential`) in the setfl format: F on
    [0, rho_max], rho(r) and r*phi(r) on [0, r_cut].  rho_max defaults to
    twice the largest rho_e (Zhou) or 100; lattice = {symbol: (a0, "fcc")}
    for the element lines."""
    r_cut = float(pot.r_cut)
    if rho_max is None:
        zp = getattr(pot, "zhou_params", None)
        rho_max = 2.0 * max(p["rho_e"] for p in zp.values()) if zp else 100.0
    drho = rho_max / (n_rho - 1)
    dr = r_cut / (n_r - 1)
    rho = np.arange(n_rho) * drho
    r = np.arange(n_r) * dr
    with open(path, "w") as fh:
        for c in (list(comments) + ["", "", ""])[:3]:
            fh.write(c.replace("\n", " ") + "\n")
        fh.write(f"{len(species)} " + " ".join(species) + "\n")
        fh.write(f"{n_rho} {drho:.16e} {n_r} {dr:.16e} {r_cut:.16e}\n")
        for s in species:
            a0, kind = (lattice or {}).get(s, (0.0, "fcc"))
            fh.write(f"{ATOMIC_NUMBER.get(s, 0)} {ATOMIC_MASS[s]:.6f} {a0:.6f} {kind}\n")
            fh.write(_block(pot.evaluate(pot.function_index("F", s), rho)[0]))
            fh.write(_block(pot.evaluate(pot.function_index("rho", s), r)[0]))
        for i, a in enumerate(species):
            for b in species[:i + 1]:
                fh.write(_block(r * pot.evaluate(pot.function_index("phi", a, b), r)[0]))


def read_setfl(path, species=None):
    """An :class:`EAMPotential` from a setfl file; ``species`` (default: the
    file's elements in order) names the species the potential indexes."""
    from scipy.interpolate import CubicSpline

    with open(path) as fh:
        lines = fh.read().splitlines()
    head = lines[3].split()
    n_el = int(head[0])
    names = head[1:1 + n_el]
    n_rho, drho, n_r, dr, r_cut = lines[4].split()[:5]
    n_rho, n_r = int(n_rho), int(n_r)
    drho, dr, r_cut = float(drho), float(dr), float(r_cut)
    pos = 5
    tables = {}

    def take(n):
        nonlocal pos
        out = []
        while len(out) < n:
            out.extend(float(v) for v in lines[pos].split())
            pos += 1
        return np.array(out[:n])

    for s in names:
        pos += 1                                   # element header line
        tables[("F", s)] = take(n_rho)
        tables[("rho", s)] = take(n_r)
    for i, a in enumerate(names):
        for b in names[:i + 1]:
            tables[("rphi", a, b)] = take(n_r)
    rho_g = np.arange(n_rho) * drho
    r_g = np.arange(n_r) * dr

    rows, knots, coeffs = [], [], []
    n_knots = 0

    def add(code, x, y):
        nonlocal n_knots
        cs = CubicSpline(x, y)
        row = np.zeros(ROW)
        row[0], row[1], row[2] = code, n_knots, len(cs.x)
        c = np.zeros((len(cs.x), 4))
        c[:-1] = cs.c.T
        rows.append(row)
        knots.append(np.asarray(cs.x, float))
        coeffs.append(c)
        n_knots += len(cs.x)
        return len(rows) - 1

    F_f = {s: add(CODE_SPLINE, rho_g, tables[("F", s)]) for s in names}
    rho_f = {s: add(CODE_SPLINE, r_g, tables[("rho", s)]) for s in names}
    phi_f = {}
    for i, a in enumerate(names):
        for b in names[:i + 1]:
            phi_f[frozenset((a, b))] = add(CODE_SPLINE_OVER_R, r_g, tables[("rphi", a, b)])
    sp = list(species) if species else list(names)
    for s in sp:
        if s not in names:
            raise KeyError(f"read_setfl: species {s!r} is not in the file ({names})")
    index_map = ([rho_f[s] for s in sp] + [F_f[s] for s in sp]
                 + [phi_f[frozenset((a, b))] for a in sp for b in sp])
    return EAMPotential(np.array(rows), np.array(index_map, np.int32), r_cut, sp,
                        knots=np.concatenate(knots), coeffs=np.concatenate(coeffs))
