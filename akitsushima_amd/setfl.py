"""setfl (LAMMPS ``eam/alloy``) files for the GPU EAM.

The file layout is the interface shared with LAMMPS and the reference
(engine/eam_zhou.py:216-315): three comment lines; ``N el_1 .. el_N``;
``n_rho drho n_r dr r_cut``; per element a line ``Z mass a0 lattice``
followed by F(rho) at rho = k drho and rho(r) at r = k dr; then r*phi(r) for
every pair (i, j <= i).  Numbers are written as ``% .16e``, five per line.

Reading treats everything after the grid line as one stream of tokens
(four header tokens per element, then the tables), so the line breaks inside
a table do not matter.  Every table becomes a scipy ``CubicSpline``
(not-a-knot ends), as the reference builds it, and phi = (r phi) / r; the
spline coefficients go to the device, where the kernels evaluate them as
engine/_eam_kernels.py:128-145 does (interval from searchsorted 'right', end
polynomials beyond the knots, Horner).  Writing tabulates the functions on
the device (``EAMPotential.evaluate``).
"""
from __future__ import annotations

import numpy as np

from .potentials import EAMPotential
from .published import ATOMIC_MASS, ATOMIC_NUMBER

ROW = 44
CODE_SPLINE, CODE_SPLINE_OVER_R = 1, 2
PER_LINE = 5


def _pairs(names):
    """The (i, j <= i) element pairs in file order."""
    return [(a, b) for i, a in enumerate(names) for b in names[:i + 1]]


def _table_lines(values):
    cells = [format(float(x), " .16e") for x in np.ravel(values)]
    return [" ".join(cells[k:k + PER_LINE]) for k in range(0, len(cells), PER_LINE)]


def write_setfl(path, pot, species, n_rho=5000, rho_max=None, n_r=5000, comments=("", "", ""),
                lattice=None):
    """Tabulate ``pot`` (an :class:`EAMPotential`) in the setfl format: F on
    [0, rho_max], rho(r) and r*phi(r) on [0, r_cut].  rho_max defaults to
    twice the largest rho_e (Zhou) or 100; lattice = {symbol: (a0, "fcc")}
    for the element lines."""
    r_cut = float(pot.r_cut)
    if rho_max is None:
        zhou = getattr(pot, "zhou_params", None)
        rho_max = 2.0 * max(p["rho_e"] for p in zhou.values()) if zhou else 100.0
    drho, dr = rho_max / (n_rho - 1), r_cut / (n_r - 1)
    rho_grid = drho * np.arange(n_rho)
    r_grid = dr * np.arange(n_r)
    lattice = lattice or {}

    def values(kind, a, b=None, x=None):
        return pot.evaluate(pot.function_index(kind, a, b), x)[0]

    text = [str(c).replace("\n", " ") for c in tuple(comments)[:3]]
    text += [""] * (3 - len(text))
    text.append(" ".join([str(len(species))] + list(species)))
    text.append(f"{n_rho} {drho:.16e} {n_r} {dr:.16e} {r_cut:.16e}")
    for el in species:
        a0, kind = lattice.get(el, (0.0, "fcc"))
        text.append(f"{ATOMIC_NUMBER.get(el, 0)} {ATOMIC_MASS[el]:.6f} {a0:.6f} {kind}")
        text += _table_lines(values("F", el, x=rho_grid))
        text += _table_lines(values("rho", el, x=r_grid))
    for a, b in _pairs(list(species)):
        text += _table_lines(r_grid * values("phi", a, b, x=r_grid))
    with open(path, "w") as fh:
        fh.write("\n".join(text) + "\n")


class _Tokens:
    """Whitespace-separated tokens of the file body, consumed in order."""

    def __init__(self, lines):
        self.tok = " ".join(lines).split()
        self.at = 0

    def skip(self, n):
        self.at += n

    def floats(self, n):
        if self.at + n > len(self.tok):
            raise ValueError("setfl file ends inside a table")
        out = np.array(self.tok[self.at:self.at + n], dtype=float)
        self.at += n
        return out


def _spline_tables(grids_values):
    """Device function rows for [(code, x, y), ...]: one CubicSpline per
    table; row = [code, first knot, knot count, 0...], knots concatenated,
    coefficients (knots, 4) with the end knot's row zero."""
    from scipy.interpolate import CubicSpline

    rows = np.zeros((len(grids_values), ROW))
    knots, coeffs, start = [], [], 0
    for k, (code, x, y) in enumerate(grids_values):
        cs = CubicSpline(x, y)
        m = len(cs.x)
        rows[k, :3] = code, start, m
        c = np.zeros((m, 4))
        c[:m - 1] = cs.c.T
        knots.append(np.asarray(cs.x, float))
        coeffs.append(c)
        start += m
    return rows, np.concatenate(knots), np.concatenate(coeffs)


def read_setfl(path, species=None):
    """An :class:`EAMPotential` from a setfl file; ``species`` (default: the
    file's elements in order) names the species the potential indexes."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    head = lines[3].split()
    names = head[1:1 + int(head[0])]
    grid = lines[4].split()
    n_rho, drho, n_r, dr, r_cut = int(grid[0]), float(grid[1]), int(grid[2]), float(grid[3]), \
        float(grid[4])
    body = _Tokens(lines[5:])
    emb, dens = {}, {}
    for el in names:
        body.skip(4)  # Z, mass, lattice constant, lattice type
        emb[el] = body.floats(n_rho)
        dens[el] = body.floats(n_r)
    rphi = {frozenset(p): body.floats(n_r) for p in _pairs(names)}

    rho_grid = drho * np.arange(n_rho)
    r_grid = dr * np.arange(n_r)
    order = ([(CODE_SPLINE, rho_grid, emb[el]) for el in names]
             + [(CODE_SPLINE, r_grid, dens[el]) for el in names]
             + [(CODE_SPLINE_OVER_R, r_grid, rphi[frozenset(p)]) for p in _pairs(names)])
    rows, knots, coeffs = _spline_tables(order)
    ne = len(names)
    f_emb = {el: k for k, el in enumerate(names)}
    f_dens = {el: ne + k for k, el in enumerate(names)}
    f_pair = {frozenset(p): 2 * ne + k for k, p in enumerate(_pairs(names))}

    sp = list(species) if species else list(names)
    missing = [s for s in sp if s not in names]
    if missing:
        raise KeyError(f"read_setfl: species {missing[0]!r} is not in the file ({names})")
    # index map of the engine: densities, embeddings, then every ordered pair
    index_map = ([f_dens[s] for s in sp] + [f_emb[s] for s in sp]
                 + [f_pair[frozenset((a, b))] for a in sp for b in sp])
    return EAMPotential(rows, np.array(index_map, np.int32), r_cut, sp, knots=knots,
                        coeffs=coeffs)
