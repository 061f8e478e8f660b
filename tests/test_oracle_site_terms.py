"""Pins of the oracle's site terms and heat flux (oracle/site_terms.py), as
the reference pins its own (tests/test_site_terms_heat_flux.py):

* energy, forces and stress rebuilt from the site terms equal the oracle's
  force path (itself pinned by ledgers, brute force and finite differences);
* the heat flux equals J V = d/dt sum_i r_i E_i, E_i = K_i + U_i, by central
  differences over one velocity-Verlet step each way, on isolated clusters
  in a 40 A box (no bond crosses the boundary, so sum_i r_i E_i is defined).
"""
import numpy as np
import pytest

import oracle as O
import oracle.site_terms as ST
from akitsushima_amd.published import ZHOU_2004
from akitsushima_amd.structures import DIAMOND, FCC

from _cases import GESN_SPECIES, SW_GESN, SW_SI, TERSOFF_SI, TERSOFF_XY, rattled, sheared_si

C = O.EV_ANG_TO_AMU_ANG_FS2
T3L = dict(TERSOFF_SI, lambda3=1.7322)
CUNI = {s: dict(ZHOU_2004[s]) for s in ("Cu", "Ni")}


def _gesn(seed=2):
    lat, pos, _ = sheared_si(2, shear=0.04, rattle=0.1, seed=seed)
    f = 5.9 / 5.431
    t = [str(x) for x in np.random.default_rng(3).choice(GESN_SPECIES, len(pos))]
    return lat * f, pos * f, t


def _cuni(seed=4):
    from akitsushima_amd.structures import cubic
    lat, pos, _ = cubic(FCC, 3.6, 4, ["Cu"] * 4, rattle=0.08, seed=seed)
    t = [str(x) for x in np.random.default_rng(seed).choice(["Cu", "Ni"], len(pos))]
    return lat, pos, t


CASES = {
    "sw": (lambda l, p, t: O.sw_flat(l, p, SW_SI),
           lambda l, p, t: ST.sw(l, p, t, O.sw_flat_as_records(SW_SI), ["Si"]),
           lambda: sheared_si(2, shear=0.05, rattle=0.1, seed=1)),
    "sw_gesn": (lambda l, p, t: O.sw_multi(l, p, t, SW_GESN, GESN_SPECIES),
                lambda l, p, t: ST.sw(l, p, t, SW_GESN, GESN_SPECIES), _gesn),
    "tersoff": (lambda l, p, t: O.tersoff_flat(l, p, T3L),
                lambda l, p, t: ST.tersoff(l, p, t, O.tersoff_flat_as_records(T3L), ["Si"]),
                lambda: sheared_si(2, shear=0.05, rattle=0.1, seed=1)),
    "tersoff_xy": (lambda l, p, t: O.tersoff_multi(l, p, t, TERSOFF_XY, ["X", "Y"]),
                   lambda l, p, t: ST.tersoff(l, p, t, TERSOFF_XY, ["X", "Y"]),
                   lambda: (lambda s: (s[0], s[1], (["X"] * 4 + ["Y"] * 4) * 8))(
                       sheared_si(2, shear=0.05, rattle=0.1, seed=4))),
    "eam_cuni": (lambda l, p, t: O.eam_zhou(l, p, t, CUNI, ["Cu", "Ni"]),
                 lambda l, p, t: ST.eam(l, p, t, O.zhou_table(CUNI, ["Cu", "Ni"]), ["Cu", "Ni"]),
                 _cuni),
}


@pytest.mark.parametrize("name", list(CASES))
def test_site_terms_rebuild_the_force_path(name):
    force, terms, cell = CASES[name]
    lat, pos, t = cell()
    E, F, S = force(lat, pos, t)
    st = terms(lat, pos, t)
    assert abs(st.energy - E) <= 1e-12 * abs(E)
    assert np.abs(st.forces() - F).max() < 1e-9 * max(1.0, np.abs(F).max())
    assert np.abs(st.stress(abs(np.linalg.det(lat))) - S).max() < 1e-11


def _cluster(n_cells, a, basis, seed, L=40.0, jitter=0.08):
    lat = np.eye(3) * L
    pts = np.concatenate([(np.asarray(basis) + [i, j, k]) * a for i in range(n_cells)
                          for j in range(n_cells) for k in range(n_cells)])
    rng = np.random.default_rng(seed)
    pos = pts - pts.mean(0) + L / 2 + rng.normal(scale=jitter, size=pts.shape)
    return lat, pos, rng


def flux_vs_moment_derivative(force, terms, lat, pos, t, m, rng):
    vel = rng.normal(scale=np.sqrt(8.617e-5 * 300 * C / m[:, None]), size=pos.shape)
    vel -= (m[:, None] * vel).sum(0) / m.sum()
    V = abs(np.linalg.det(lat))

    def moment(p, v):
        E = 0.5 * m * np.sum(v ** 2, axis=1) / C + terms(lat, p, t).e_site
        return (E[:, None] * p).sum(axis=0)

    def verlet(p, v, dt):
        v = v + 0.5 * dt * force(lat, p, t)[1] * C / m[:, None]
        p = p + dt * v
        v = v + 0.5 * dt * force(lat, p, t)[1] * C / m[:, None]
        return p, v

    J = terms(lat, pos, t).heat_flux(vel, m, V)
    dt = 1e-3
    J_fd = (moment(*verlet(pos, vel, dt)) - moment(*verlet(pos, vel, -dt))) / (2 * dt) / V
    return J, J_fd


@pytest.mark.parametrize("name", ["sw", "sw_gesn", "tersoff_xy", "eam_cuni"])
def test_heat_flux_is_the_time_derivative_of_the_energy_moment(name):
    force, terms, _ = CASES[name]
    if name in ("sw", "sw_gesn", "tersoff_xy"):
        a = 5.9 if name == "sw_gesn" else 5.431
        lat, pos, rng = _cluster(2, a, DIAMOND, {"sw": 0, "sw_gesn": 1, "tersoff_xy": 2}[name])
        if name == "sw":
            t = ["Si"] * len(pos)
            m = np.full(len(pos), 28.0855)
        elif name == "sw_gesn":
            t = [str(x) for x in rng.choice(GESN_SPECIES, len(pos))]
            m = np.array([72.63 if x == "Ge" else 118.71 for x in t])
        else:
            t = (["X"] * 4 + ["Y"] * 4) * 8
            m = np.full(len(pos), 28.0855)
    else:
        lat, pos, rng = _cluster(2, 3.6, FCC, 3)
        t = [str(x) for x in rng.choice(["Cu", "Ni"], len(pos))]
        m = np.array([63.546 if x == "Cu" else 58.693 for x in t])
    J, J_fd = flux_vs_moment_derivative(force, terms, lat, pos, t, m, rng)
    scale = np.abs(J_fd).max()
    assert np.abs(J - J_fd).max() < 1e-5 * scale, (J, J_fd)
