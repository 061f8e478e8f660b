"""Torch FP64 reference of the Tersoff fitting loss, differentiated by
autograd (test helper, CPU).

Energy of the 14-parameter single-species form (site sum, engine/tersoff.py
:528-855 semantics), forces by autograd of E with respect to positions,
stress by autograd with respect to a homogeneous deformation T of the cell
(sigma_ab = -(1/V) dE/dT_ab at T = I), loss of fitting/loss.py:212-242, and
dL/dtheta by a second autograd pass.  (r_ij - r_ik)^m is sign(x)|x|^m.
"""
import numpy as np
import torch

import oracle as O

KEYS = ("A", "B", "lambda1", "lambda2", "lambda3", "beta", "n", "c", "d", "h",
        "R", "D", "m", "gamma")


def topology(lat, pos, r_list):
    """Directed bonds (i, j, image) within r_list and ordered triplets
    (bond ij, bond ik) sharing the centre."""
    i, j, _, _ = O.nl_build(lat, pos, r_list)
    inv = np.linalg.inv(lat)
    fr = (pos[j] - pos[i]) @ inv
    img = -np.round(fr)
    src = np.concatenate([i, j])
    dst = np.concatenate([j, i])
    im = np.concatenate([img, -img])
    order = np.argsort(src, kind="stable")
    src, dst, im = src[order], dst[order], im[order]
    trip = []
    start = np.searchsorted(src, np.arange(len(pos) + 1))
    for a in range(len(pos)):
        for x in range(start[a], start[a + 1]):
            for y in range(start[a], start[a + 1]):
                if x != y:
                    trip.append((x, y))
    trip = np.array(trip, dtype=np.int64).reshape(-1, 2)
    return src, dst, im, trip


def energy(theta, lat, pos, T, topo):
    src, dst, im, trip = topo
    A, B, l1, l2, l3, beta, n, c, d, h, R, D, m, gam = (theta[k] for k in range(14))
    src_t, dst_t = torch.as_tensor(src), torch.as_tensor(dst)
    im_t = torch.as_tensor(im, dtype=torch.float64)
    rij = (pos[dst_t] - pos[src_t] + im_t @ lat) @ T
    r = torch.linalg.norm(rij, dim=1)
    rc = R + D
    act = (r < rc).to(r.dtype)

    def fc(x):
        mid = 0.5 - 0.5 * torch.sin(0.5 * np.pi * (x - R) / D)
        return torch.where(x < R - D, torch.ones_like(x), torch.where(x > R + D,
                                                                      torch.zeros_like(x), mid))

    fcr = fc(r) * act
    x, y = trip[:, 0], trip[:, 1]
    cs = (rij[x] * rij[y]).sum(1) / (r[x] * r[y])
    g = 1 + c * c / (d * d) - c * c / (d * d + (h - cs) ** 2)
    u = l3 * (r[x] - r[y])
    um = torch.sign(u) * torch.abs(u) ** m
    z = gam * fcr[y] * g * torch.exp(um) * act[x]
    zeta = torch.zeros_like(r).index_add(0, torch.as_tensor(x), z)
    pos_z = zeta > 0
    zs = torch.where(pos_z, zeta, torch.ones_like(zeta))
    b = torch.where(pos_z, (1 + (beta * zs) ** n) ** (-0.5 / n), torch.ones_like(zeta))
    e = 0.5 * fcr * (A * torch.exp(-l1 * r) - b * B * torch.exp(-l2 * r))
    return e.sum()


def loss_and_grad(theta_np, structures, weights=(1.0, 1.0, 0.1, 1.0, 1.0)):
    wE, wF, wS, r0, eps0 = weights
    theta = torch.tensor(theta_np, dtype=torch.float64, requires_grad=True)
    total = torch.zeros((), dtype=torch.float64)
    for s in structures:
        lat = torch.tensor(s["lattice"], dtype=torch.float64)
        pos = torch.tensor(s["positions"], dtype=torch.float64, requires_grad=True)
        T = torch.eye(3, dtype=torch.float64, requires_grad=True)
        topo = s["_topo"]
        E = energy(theta, lat, pos, T, topo)
        gpos, gT = torch.autograd.grad(E, (pos, T), create_graph=True)
        V = abs(float(np.linalg.det(s["lattice"])))
        F = -gpos
        sig = -gT / V
        N = len(s["positions"])
        total = total + wE * (E - N * s["ref_energy"]) ** 2
        total = total + wF * r0 ** 2 * ((F - torch.tensor(s["ref_forces"])) ** 2).mean()
        total = total + wS * (V * eps0) ** 2 * ((sig - torch.tensor(s["ref_stress"])) ** 2).mean()
    (g,) = torch.autograd.grad(total, theta)
    return float(total.detach()), g.numpy()
