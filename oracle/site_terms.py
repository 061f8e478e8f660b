"""Oracle: site-energy decomposition and heat flux (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the reference's site terms (engine/site_terms.py):
for E = sum_i U_i with U_i a function of the bond vectors v_m = r_nbr - r_c
of its entries m, the terms are e_site (N), centre, nbr (M), vec (M, 3) and
dUdv (M, 3) = dU_centre / dv_m, from which

    F_k = sum_m dUdv_m (delta(k, centre) - delta(k, nbr)),
    sigma = -(1/V) sum_m dUdv_m (x) v_m,
    J V = sum_i E_i v_i - sum_m v_m (dUdv_m . v_nbr),   E_i = K_i + U_i

(Fan et al., Phys. Rev. B 92, 094301 (2015), eq. 21).  The energy split of
each potential is the engine's: SW U_i = V2/2 per bond + every ordered
angle at i with weight 1/2 (engine/stillinger_weber.py), Tersoff U_i =
fc (fR + b_ij fA) / 2 per bond (engine/tersoff.py), EAM U_i = F_i + phi/2
per bond (engine/eam.py).  Pinned by tests/test_oracle_site_terms.py: the
forces, stress and energy rebuilt from the terms equal the oracle's force
path, and J equals d/dt sum_i r_i E_i by central differences of a
velocity-Verlet step on isolated clusters.
"""
from __future__ import annotations

import numpy as np

from . import EV_ANG_TO_AMU_ANG_FS2, candidate_pairs, sw_tables, tersoff_tables


class Terms:
    def __init__(self, e_site, centre, nbr, vec, dUdv):
        self.e_site, self.centre, self.nbr, self.vec, self.dUdv = e_site, centre, nbr, vec, dUdv

    @property
    def energy(self):
        return float(self.e_site.sum())

    def forces(self):
        F = np.zeros((len(self.e_site), 3))
        np.add.at(F, self.centre, self.dUdv)
        np.add.at(F, self.nbr, -self.dUdv)
        return F

    def stress(self, volume):
        S = -(self.dUdv.T @ self.vec) / volume
        return 0.5 * (S + S.T)

    def heat_flux(self, velocities, masses, volume):
        v = np.asarray(velocities, float)
        E = 0.5 * masses * np.sum(v * v, axis=1) / EV_ANG_TO_AMU_ANG_FS2 + self.e_site
        conv = (E[:, None] * v).sum(axis=0)
        c = np.sum(self.dUdv * v[self.nbr], axis=1)
        vir = -(c[:, None] * self.vec).sum(axis=0)
        return (conv + vir) / volume

    def atom_heat_current(self, velocities, masses):
        """Per-atom heat current J_i [eV A/fs] with sum_i J_i = V J
        (ThermalConductivityCalculator._per_atom_heat_current,
        solvers/thermal_conductivity.py:305-337): E_i v_i, and each entry's
        term -vec_m (dU/dv_m . v_nbr) to the atom whose velocity it carries."""
        v = np.asarray(velocities, float)
        E = 0.5 * masses * np.sum(v * v, axis=1) / EV_ANG_TO_AMU_ANG_FS2 + self.e_site
        Ji = E[:, None] * v
        c = np.sum(self.dUdv * v[self.nbr], axis=1)
        np.add.at(Ji, self.nbr, -c[:, None] * self.vec)
        return Ji


def group_labels(lattice, positions, atom_types, decompose):
    """{name: (labels (N,), names)} of a decomposition (_group_labels,
    solvers/thermal_conductivity.py:339-363): "species" in order of first
    appearance; "regions" by the fractional coordinate along axis a/b/c
    between the bounds (searchsorted right - 1, clipped)."""
    out = {}
    if decompose.get("species"):
        names = sorted(set(atom_types), key=list(atom_types).index)
        out["species"] = (np.array([names.index(t) for t in atom_types]), names)
    reg = decompose.get("regions")
    if reg:
        ax = {"a": 0, "b": 1, "c": 2}.get(str(reg.get("axis", "c")), reg.get("axis"))
        bounds = np.asarray(reg["bounds"], dtype=float)
        f = (np.asarray(positions, float) @ np.linalg.inv(lattice))[:, int(ax)]
        f -= np.floor(f)
        lab = np.clip(np.searchsorted(bounds, f, side="right") - 1, 0, len(bounds) - 2)
        out["regions"] = (lab, [f"[{bounds[k]:.3f},{bounds[k + 1]:.3f})"
                                for k in range(len(bounds) - 1)])
    return out


def entries(lattice, positions, rc):
    """Bonds within rc under the minimum image, both directions, grouped by
    centre (engine/site_terms.py:84-110): centre, nbr, vec, r."""
    lattice = np.asarray(lattice, float)
    positions = np.asarray(positions, float)
    ii, jj = candidate_pairs(lattice, positions, rc)
    ii = np.asarray(ii, np.int64)
    jj = np.asarray(jj, np.int64)
    s = (positions[jj] - positions[ii]) @ np.linalg.inv(lattice)
    s -= np.round(s)
    v = s @ lattice
    r = np.linalg.norm(v, axis=1)
    keep = (r < rc) & (r > 1e-12)
    ii, jj, v, r = ii[keep], jj[keep], v[keep], r[keep]
    c = np.concatenate([ii, jj])
    o = np.argsort(c, kind="stable")
    return (c[o], np.concatenate([jj, ii])[o], np.concatenate([v, -v])[o],
            np.concatenate([r, r])[o])


def _angle_pairs(centre, n):
    """Ordered pairs (m1, m2), m1 != m2, of entries of one centre."""
    cnt = np.bincount(centre, minlength=n)
    start = np.concatenate([[0], np.cumsum(cnt)])
    m1, m2 = [], []
    for a in range(n):
        idx = np.arange(start[a], start[a + 1])
        if len(idx) < 2:
            continue
        A, B = np.meshgrid(idx, idx, indexing="ij")
        off = A != B
        m1.append(A[off])
        m2.append(B[off])
    if not m1:
        z = np.zeros(0, np.int64)
        return z, z
    return np.concatenate(m1), np.concatenate(m2)


def _cos(u, w, ru, rw):
    cos = np.sum(u * w, axis=1) / (ru * rw)
    dcu = w / (ru * rw)[:, None] - (cos / ru ** 2)[:, None] * u
    dcw = u / (ru * rw)[:, None] - (cos / rw ** 2)[:, None] * w
    return cos, dcu, dcw


def sw(lattice, positions, types, records, species):
    """SW site terms (engine/site_terms.py:179-281 semantics) from records;
    the flat form is the one-species record (oracle.sw_flat_as_records)."""
    tb, leg, tri, rc = sw_tables(records, species)
    idx = {s: k for k, s in enumerate(species)}
    t = np.array([idx[x] for x in types], np.int64)
    n = len(positions)
    centre, nbr, vec, r = entries(lattice, positions, rc)
    e = np.zeros(n)
    g = np.zeros_like(vec)
    # two-body: the record of the lower-indexed atom, half to each end
    lo, hi = np.minimum(centre, nbr), np.maximum(centre, nbr)
    P = tb[t[lo], t[hi]]
    valid, eps, sig, a, A, B, p, q = (P[:, k] for k in range(8))
    x = r / sig
    ok = (valid != 0) & (x < a)
    xs, d = x[ok], x[ok] - a[ok]
    inv_d = 1.0 / d
    ex = np.exp(inv_d)
    poly = B[ok] * xs ** (-p[ok]) - xs ** (-q[ok])
    V2 = eps[ok] * A[ok] * poly * ex
    dV2 = eps[ok] * A[ok] * ex * ((-p[ok] * B[ok] * xs ** (-p[ok] - 1)
                                   + q[ok] * xs ** (-q[ok] - 1)) - poly * inv_d ** 2)
    np.add.at(e, centre[ok], 0.5 * V2)
    g[ok] += (0.5 * dV2 / sig[ok] / r[ok])[:, None] * vec[ok]
    # three-body: every ordered pair of legs, weight 1/2
    m1, m2 = _angle_pairs(centre, n)
    ts, tj, tk = t[centre[m1]], t[nbr[m1]], t[nbr[m2]]
    T3 = tri[ts, tj, tk]
    Lj, Lk = leg[ts, tj], leg[ts, tk]
    xj, xk = r[m1] / Lj[:, 0], r[m2] / Lk[:, 0]
    ok = (T3[:, 0] != 0) & (xj < Lj[:, 1]) & (xk < Lk[:, 1])
    m1, m2, T3, Lj, Lk, xj, xk = m1[ok], m2[ok], T3[ok], Lj[ok], Lk[ok], xj[ok], xk[ok]
    el = 0.5 * T3[:, 1] * T3[:, 2]
    u, w, ru, rw = vec[m1], vec[m2], r[m1], r[m2]
    cos, dcu, dcw = _cos(u, w, ru, rw)
    dj, dk = xj - Lj[:, 1], xk - Lk[:, 1]
    E = np.exp(Lj[:, 2] / dj + Lk[:, 2] / dk)
    D = cos - T3[:, 3]
    h = el * E * D ** 2
    lnj = -Lj[:, 2] / dj ** 2 / Lj[:, 0]
    lnk = -Lk[:, 2] / dk ** 2 / Lk[:, 0]
    np.add.at(e, centre[m1], h)
    np.add.at(g, m1, (h * lnj / ru)[:, None] * u + (2 * el * E * D)[:, None] * dcu)
    np.add.at(g, m2, (h * lnk / rw)[:, None] * w + (2 * el * E * D)[:, None] * dcw)
    return Terms(e, centre, nbr, vec, g)


def _fc(r, R, D):
    arg = 0.5 * np.pi * (r - R) / D
    inner = (r >= R - D) & (r <= R + D)
    fc = np.where(r < R - D, 1.0, np.where(inner, 0.5 - 0.5 * np.sin(arg), 0.0))
    return fc, np.where(inner, -0.25 * np.pi / D * np.cos(arg), 0.0)


def tersoff(lattice, positions, types, records, species):
    """Tersoff site terms (engine/site_terms.py:284-411 semantics): bond (i,
    j) exists when the lower-indexed atom's record (s_lo, s_hi, s_hi) does;
    b_ij with beta, n of (s_i, s_j, s_j); angular terms of (s_i, s_j, s_k)
    with the cutoff of (s_i, s_k, s_k)."""
    tb, tri, rc = tersoff_tables(records, species)
    idx = {s: k for k, s in enumerate(species)}
    t = np.array([idx[x] for x in types], np.int64)
    n = len(positions)
    centre, nbr, vec, r = entries(lattice, positions, rc)
    M = len(centre)
    tc, tn = t[centre], t[nbr]
    lo, hi = np.minimum(centre, nbr), np.maximum(centre, nbr)
    Plo = tb[t[lo], t[hi]]
    bond_ok = Plo[:, 0] != 0
    m1, m2 = _angle_pairs(centre, n)
    ts, tj, tk = tc[m1], tn[m1], tn[m2]
    T3 = tri[ts, tj, tk]
    Lk = tb[ts, tk]
    ok = bond_ok[m1] & (T3[:, 0] != 0) & (Lk[:, 0] != 0)
    m1, m2, T3, Lk = m1[ok], m2[ok], T3[ok], Lk[ok]
    fck, dfck = _fc(r[m2], Lk[:, 5], Lk[:, 6])
    keep = (fck != 0) | (dfck != 0)
    m1, m2, T3, fck, dfck = m1[keep], m2[keep], T3[keep], fck[keep], dfck[keep]
    c, d, hh, l3, mm, gam = (T3[:, k] for k in range(1, 7))
    u, w, ru, rw = vec[m1], vec[m2], r[m1], r[m2]
    cos, dcu, dcw = _cos(u, w, ru, rw)
    X = d ** 2 + (hh - cos) ** 2
    gth = 1.0 + c ** 2 / d ** 2 - c ** 2 / X
    dg = -2.0 * c ** 2 * (hh - cos) / X ** 2
    dr = ru - rw
    ex = np.exp(l3 ** mm * dr ** mm)
    dex = ex * mm * l3 ** mm * dr ** (mm - 1.0)
    z = gam * fck * gth * ex
    dz_du = (gam * fck * ex * dg)[:, None] * dcu + (gam * fck * gth * dex / ru)[:, None] * u
    dz_dw = ((gam * dfck * gth * ex / rw)[:, None] * w + (gam * fck * ex * dg)[:, None] * dcw
             - (gam * fck * gth * dex / rw)[:, None] * w)
    zeta = np.zeros(M)
    np.add.at(zeta, m1, z)
    Pb = tb[tc, tn]
    beta, nn = Pb[:, 7], Pb[:, 8]
    b = np.ones(M)
    dbdz = np.zeros(M)
    act = (zeta > 0) & (beta != 0) & (Pb[:, 0] != 0)
    bz = (beta[act] * zeta[act]) ** nn[act]
    b[act] = (1.0 + bz) ** (-0.5 / nn[act])
    dbdz[act] = -0.5 * b[act] * bz / (zeta[act] * (1.0 + bz))
    A, B, l1, l2, R, D = (Plo[:, k] for k in range(1, 7))
    fc, dfc = _fc(r, R, D)
    fR = A * np.exp(-l1 * r)
    fA = -B * np.exp(-l2 * r)
    on = bond_ok & ((fc != 0) | (dfc != 0))
    fc, dfc, fR, fA = (np.where(on, q, 0.0) for q in (fc, dfc, fR, fA))
    e = np.zeros(n)
    np.add.at(e, centre, 0.5 * fc * (fR + b * fA))
    dE = 0.5 * (dfc * (fR + b * fA) + fc * (-l1 * fR - l2 * b * fA))
    g = (dE / r)[:, None] * vec
    pref = 0.5 * fc * fA * dbdz
    np.add.at(g, m1, pref[m1][:, None] * dz_du)
    np.add.at(g, m2, pref[m1][:, None] * dz_dw)
    return Terms(e, centre, nbr, vec, g)


def eam(lattice, positions, types, tab, species):
    """EAM site terms (engine/site_terms.py:418-475 semantics) over an
    oracle.EamTable."""
    idx = {s: k for k, s in enumerate(species)}
    t = np.array([idx[x] for x in types], np.int64)
    n = len(positions)
    ns = len(species)
    centre, nbr, vec, r = entries(lattice, positions, tab.r_cut)
    rho = np.zeros(len(r))
    drho = np.zeros(len(r))
    for s in range(ns):
        m = t[nbr] == s
        if m.any():
            rho[m], drho[m] = tab.feval(tab.rho_f[s], r[m])
    rb = np.zeros(n)
    np.add.at(rb, centre, rho)
    rb = np.maximum(rb, 0.0)
    F = np.zeros(n)
    dF = np.zeros(n)
    for s in range(ns):
        m = t == s
        if m.any():
            F[m], dF[m] = tab.feval(tab.emb_f[s], rb[m])
    phi = np.zeros(len(r))
    dphi = np.zeros(len(r))
    for a in range(ns):
        for b in range(ns):
            m = (t[centre] == a) & (t[nbr] == b)
            if m.any():
                phi[m], dphi[m] = tab.feval(tab.phi_f[a * ns + b], r[m])
    e = F.copy()
    np.add.at(e, centre, 0.5 * phi)
    coef = dF[centre] * drho + 0.5 * dphi
    return Terms(e, centre, nbr, vec, (coef / r)[:, None] * vec)
