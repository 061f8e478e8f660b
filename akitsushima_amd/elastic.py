"""Elastic constants from batched strained-cell force calls.

Strain points, fits and sign conventions follow ElasticConstantsCalculator
(solvers/elastic.py:45-77, 222-236, 300-350): each Voigt mode j is applied
at +-s for every amplitude s with fractional coordinates fixed, and column
j of the clamped-ion tensor is the least-squares slope of the standard
(tensile-positive) stress -sigma against eta over those points and the
unstrained cell, in GPa; the tensor is then symmetrised.  All 1 + 12 n
strained cells are evaluated in one batched device call.

Relaxed ions.  The reference relaxes the ions of every strained cell with
FIRE (GeometryOptimiser); here the relaxation is taken in the harmonic
limit instead: with the force response Lambda = dF/deta at fixed fractional
coordinates (the slope of the forces over the same strain points) and the
force constants Phi of the unstrained cell (one more batched call),
relaxation u = Phi^+ Lambda eta lowers the tensor by
    C = C_clamped - Lambda^T Phi^+ Lambda / V,
Phi^+ the pseudo-inverse on the complement of the three translations.  This
is the eta -> 0 limit of the reference's relaxed fit.
"""
from __future__ import annotations

import numpy as np

from .phonon import force_constants

EV_ANG3_TO_GPA = 160.2176621          # constants.py:51
VOIGT_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
DEFAULT_STRAINS = (0.0025, 0.005, 0.01)


def strain_matrix(j, s):
    """Symmetric strain of Voigt mode j (engineering shear: the tensor entry
    is s / 2), solvers/elastic.py:60-72."""
    e = np.zeros((3, 3))
    a, b = VOIGT_PAIRS[j]
    if a == b:
        e[a, b] = s
    else:
        e[a, b] = e[b, a] = s / 2.0
    return e


def voigt(sigma):
    return np.array([sigma[a, b] for a, b in VOIGT_PAIRS])


def strain_points(strains=DEFAULT_STRAINS):
    """(mode, signed strain) in the reference's order; mode None is the
    unstrained cell (solvers/elastic.py:222-230)."""
    pts = [(None, 0.0)]
    for j in range(6):
        for s in sorted(abs(float(x)) for x in strains):
            pts += [(j, +s), (j, -s)]
    return pts


def elastic_constants(potential, lattice, positions, atom_types, strains=DEFAULT_STRAINS,
                      relax_ions=False, delta=0.01):
    """{"C": (6, 6) GPa (relaxed ions when asked, else clamped), "C_clamped",
    "internal_strain" (N, 3, 6) or None, "bulk_modulus", "strains"}."""
    lat0 = np.asarray(lattice, float)
    pos0 = np.asarray(positions, float)
    frac0 = pos0 @ np.linalg.inv(lat0)
    pts = strain_points(strains)
    cells = []
    for j, s in pts:
        lat = lat0.copy() if j is None else lat0 @ (np.eye(3) + strain_matrix(j, s))
        cells.append((lat, frac0 @ lat, atom_types))
    _, F, S = potential.compute_batched(cells)
    n = len(pos0)
    C0 = np.zeros((6, 6))
    Lam = np.zeros((3 * n, 6))
    for j in range(6):
        idx = [0] + [k for k, (jj, _) in enumerate(pts) if jj == j]
        eta = np.array([pts[k][1] for k in idx])
        A = np.column_stack([eta, np.ones_like(eta)])
        y = np.array([-voigt(S[k]) for k in idx]) * EV_ANG3_TO_GPA
        C0[:, j] = np.linalg.lstsq(A, y, rcond=None)[0][0]
        if relax_ions:
            f = np.array([np.asarray(F[k]).ravel() for k in idx])
            Lam[:, j] = np.linalg.lstsq(A, f, rcond=None)[0][0]
    C0 = 0.5 * (C0 + C0.T)
    C = C0.copy()
    u = None
    if relax_ions:
        phi = force_constants(potential, lat0, pos0, atom_types, delta=delta)
        w, V = np.linalg.eigh(0.5 * (phi + phi.T))
        keep = np.argsort(np.abs(w))[3:]            # drop the three translations
        phi_p = (V[:, keep] / w[keep]) @ V[:, keep].T
        vol = abs(np.linalg.det(lat0))
        C = C0 - (Lam.T @ phi_p @ Lam) / vol * EV_ANG3_TO_GPA
        C = 0.5 * (C + C.T)
        u = (phi_p @ Lam).reshape(n, 3, 6)
    bulk = (C[0, 0] + C[1, 1] + C[2, 2] + 2.0 * (C[0, 1] + C[1, 2] + C[0, 2])) / 9.0
    return {"C": C, "C_clamped": C0, "internal_strain": u, "bulk_modulus": float(bulk),
            "strains": tuple(sorted(abs(float(x)) for x in strains))}
