"""Small fitting sets shared by the fit tests: rattled (0.15 A), strained
(+-3 %) diamond Si cells with targets from the oracle at parameters
perturbed by +-5 % (m held at 3)."""
import numpy as np

import oracle as O
import torch_tersoff as TT
from akitsushima_amd.fit import theta_from_params
from akitsushima_amd.structures import diamond_batch

from _cases import TERSOFF_SI


def make_set(n_struct=3, seed=5, lambda3=1.3, gamma=1.05, max_atoms=128):
    structs = diamond_batch(n_struct, 64, max_atoms, 0.03, 0.15, seed=seed, rep_range=(2, 3))
    theta = theta_from_params(dict(TERSOFF_SI, lambda3=lambda3), gamma=gamma)
    pert = theta * (1 + np.random.default_rng(seed + 1).uniform(-0.05, 0.05, 14))
    pert[12] = 3.0
    recp = {("Si", "Si", "Si"): dict(zip(TT.KEYS, pert))}
    data = []
    for lat, pos, t in structs:
        e, f, s = O.tersoff_multi(lat, pos, t, recp, ["Si"])
        data.append({"lattice": lat, "positions": pos, "atom_types": t,
                     "ref_energy": e / len(pos), "ref_forces": f, "ref_stress": s,
                     "_topo": TT.topology(lat, pos, theta[10] + theta[11] + 0.3)})
    return theta, data


def oracle_loss(theta, data):
    return O.tersoff_loss(data, {("Si", "Si", "Si"): dict(zip(TT.KEYS, theta))}, ["Si"])
