"""Harmonic force constants and Gamma-point phonons from batched force calls.

``force_constants`` follows PhononCalculator.compute_ifc
(solvers/phonon.py:412-482) and _finish_ifc (:484-...): column (j, b) of the
interatomic force constants is the central difference -(F(+delta) -
F(-delta)) / (2 delta) of the forces when atom j moves along axis b, then
the matrix is symmetrised and the acoustic sum rule is imposed on the
diagonal blocks.  The reference evaluates the displaced cells one force call
at a time (a thread pool); here all 6 N displaced cells go to the device in
one batched call.
"""
from __future__ import annotations

import numpy as np

from .md import EV_ANG_TO_AMU_ANG_FS2

# sqrt(eV / (A^2 amu)) in rad/fs -> THz
_OMEGA_TO_THZ = np.sqrt(EV_ANG_TO_AMU_ANG_FS2) * 1e15 / (2.0 * np.pi) / 1e12


def displaced_cells(lattice, positions, atom_types, delta, atoms=None):
    """The 6 N (or 6 len(atoms)) cells of the central differences, in the
    order (atom, axis, +delta then -delta); the minus cell is reached the way
    the reference reaches it, by moving the plus cell back by 2 delta."""
    pos0 = np.asarray(positions, float)
    idx = range(len(pos0)) if atoms is None else atoms
    cells, tasks = [], []
    for j in idx:
        for b in range(3):
            p = pos0.copy()
            p[j, b] += delta
            cells.append((lattice, p.copy(), atom_types))
            p[j, b] -= 2.0 * delta
            cells.append((lattice, p, atom_types))
            tasks.append((j, b))
    return cells, tasks


def finish_ifc(ifc):
    """Symmetrise and impose the acoustic sum rule at the level of the force
    constants (solvers/phonon.py:484-520): Phi <- (Phi + Phi^T)/2, then each
    diagonal block Phi_ii <- Phi_ii - sum_j Phi_ij."""
    n3 = ifc.shape[0]
    n = n3 // 3
    ifc = 0.5 * (ifc + ifc.T)
    f4 = ifc.reshape(n, 3, n, 3)
    row = f4.sum(axis=2)
    for i in range(n):
        f4[i, :, i, :] -= row[i]
    return f4.reshape(n3, n3)


def force_constants(potential, lattice, positions, atom_types, delta=0.01, finish=True):
    """(3N, 3N) force constants [eV/A^2] of the cell by central differences
    of the forces, all displaced cells in one batched device call."""
    cells, tasks = displaced_cells(lattice, positions, atom_types, delta)
    _, F, _ = potential.compute_batched(cells)
    n = len(positions)
    ifc = np.zeros((3 * n, 3 * n))
    for t, (j, b) in enumerate(tasks):
        ifc[:, 3 * j + b] = -(np.asarray(F[2 * t]) - np.asarray(F[2 * t + 1])).ravel() / (
            2.0 * delta)
    return finish_ifc(ifc) if finish else ifc


def gamma_frequencies(ifc, masses):
    """Gamma-point frequencies [THz] of the cell's dynamical matrix
    D = Phi / sqrt(m_i m_j); imaginary modes come out negative."""
    m = np.repeat(np.asarray(masses, float), 3)
    D = ifc / np.sqrt(np.outer(m, m))
    w2 = np.linalg.eigvalsh(0.5 * (D + D.T))
    return np.sign(w2) * np.sqrt(np.abs(w2)) * _OMEGA_TO_THZ
