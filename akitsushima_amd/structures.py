"""Synthetic workloads: crystal supercells, rattles, strains and batches.

``cubic`` builds a supercell the way the reference's force benchmark does
(benchmarks/bench_force.py:152-158): n^3 cubic cells of a fractional basis,
cells (i, j, k) outer and basis atoms inner, Cartesian positions displaced by
``default_rng(seed).normal(0, rattle)``.  The other builders produce the
batched and fitting workloads (uneven diamond supercells under a seeded
symmetric strain) and the alloy species draw.
"""
from __future__ import annotations

import numpy as np

DIAMOND = np.array([
    [0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0],
    [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25],
], dtype=float)
FCC = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]], dtype=float)


def cubic(basis, a, n, types, rattle=0.05, seed=0):
    """(lattice, positions, types) of n^3 cubic cells of edge ``a``."""
    rng = np.random.default_rng(seed)
    cells = np.array([[i, j, k] for i in range(n) for j in range(n) for k in range(n)])
    frac = (np.asarray(basis, float)[None] + cells[:, None]).reshape(-1, 3)
    pos = frac * a + rng.normal(0.0, rattle, frac.shape)
    return np.eye(3) * a * n, pos, list(types) * len(cells)


def supercell(basis, a, reps, types):
    """Unrattled n1 x n2 x n3 supercell of a cubic basis."""
    n1, n2, n3 = reps
    cells = np.array([[i, j, k] for i in range(n1) for j in range(n2) for k in range(n3)])
    frac = (np.asarray(basis, float)[None] + cells[:, None]).reshape(-1, 3)
    lat = np.diag([n1 * a, n2 * a, n3 * a]).astype(float)
    return lat, frac * a, list(types) * len(cells)


def strained(lattice, positions, strain):
    """Affine strain in the row-vector convention: lattice @ (I + eps),
    positions follow their fractional coordinates."""
    T = np.eye(3) + np.asarray(strain, float)
    frac = positions @ np.linalg.inv(lattice)
    lat2 = lattice @ T
    return lat2, frac @ lat2


def random_strain(rng, amp):
    """Symmetric strain with independent components uniform in +-amp."""
    e = rng.uniform(-amp, amp, (3, 3))
    return np.triu(e) + np.triu(e, 1).T


def diamond_batch(n_struct, min_atoms, max_atoms, strain_amp, rattle, seed,
                  a=5.431, species="Si", rep_range=(2, 8)):
    """Uneven diamond supercells (n1, n2, n3) drawn from rep_range^3 with
    min_atoms <= 8 n1 n2 n3 <= max_atoms, each strained by a symmetric strain
    with components uniform in +-strain_amp and rattled by ``rattle``;
    structure k uses the generator seeded with (seed, k)."""
    out = []
    lo, hi = rep_range
    for k in range(n_struct):
        rng = np.random.default_rng([seed, k])
        while True:
            reps = rng.integers(lo, hi + 1, size=3)
            nat = 8 * int(np.prod(reps))
            if min_atoms <= nat <= max_atoms:
                break
        lat, pos, types = supercell(DIAMOND, a, reps, [species] * 8)
        lat, pos = strained(lat, pos, random_strain(rng, strain_amp))
        pos = pos + rng.normal(0.0, rattle, pos.shape)
        out.append((lat, pos, types))
    return out


def random_species(rng, choices, n, p=None):
    return [str(x) for x in rng.choice(list(choices), n, p=p)]
