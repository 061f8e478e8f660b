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


@pytest.mark.parametrize("name", list(CASES))
def test_per_atom_heat_currents_add_up_to_the_heat_flux(name):
    """sum_i J_i = V J (the reference's
    test_per_atom_heat_currents_add_up_to_the_heat_flux, tests/test_nemd.py:
    41-67), J itself pinned by the moment derivative above; and with one
    atom moving the whole current sits on that atom (each entry's term goes
    to the atom whose velocity it carries), which an assignment to the
    centre atom fails."""
    _, terms, cell = CASES[name]
    lat, pos, t = cell()
    rng = np.random.default_rng(7)
    m = rng.uniform(20.0, 80.0, len(pos))
    vel = rng.normal(0.0, 0.01, pos.shape)
    st = terms(lat, pos, t)
    V = abs(np.linalg.det(lat))
    Ji = st.atom_heat_current(vel, m)
    assert np.allclose(Ji.sum(axis=0) / V, st.heat_flux(vel, m, V), rtol=1e-12, atol=1e-16)
    one = np.zeros_like(vel)
    one[5] = vel[5]
    Ji = st.atom_heat_current(one, m)
    assert np.abs(np.delete(Ji, 5, axis=0)).max() == 0.0
    assert np.allclose(Ji[5], st.heat_flux(one, m, V) * V, rtol=1e-12, atol=0.0)
    assert np.abs(Ji[5] - (0.5 * m[5] * one[5] @ one[5] / C + st.e_site[5]) * one[5]).max() > \
        1e-6 * np.abs(Ji[5]).max()


def test_group_labels_follow_the_reference_rules():
    """species in order of first appearance; regions by the wrapped
    fractional coordinate, searchsorted right - 1 (a point on an edge joins
    the region above), clipped into range (_group_labels,
    solvers/thermal_conductivity.py:339-363)."""
    # a sheared cell whose inverse is exact in binary, so the point on the
    # edge 0.25 lands on it exactly
    lat = np.array([[4.0, 2.0, 0.0], [0.0, 8.0, 0.0], [0.0, 0.0, 2.0]])
    fc = np.array([0.1, 0.35, 0.6, 0.95, 1.2, -0.1, 0.25])
    frac = np.stack([np.full(7, 0.25), fc, np.full(7, 0.5)], axis=1)
    pos = frac @ lat
    t = ["B", "A", "B", "C", "A", "A", "C"]
    g = ST.group_labels(lat, pos, t, {"species": True,
                                     "regions": {"axis": "b", "bounds": [0, 0.25, 0.7, 1.0]}})
    assert g["species"][1] == ["B", "A", "C"]
    assert g["species"][0].tolist() == [0, 1, 0, 2, 1, 1, 2]
    assert g["regions"][0].tolist() == [0, 1, 1, 2, 0, 2, 1]
    assert g["regions"][1] == ["[0.000,0.250)", "[0.250,0.700)", "[0.700,1.000)"]
    g = ST.group_labels(lat, pos, t, {"regions": {"axis": 1, "bounds": [0.2, 0.5, 0.8]}})
    assert set(g) == {"regions"}
    assert g["regions"][0].tolist() == [0, 0, 1, 1, 0, 1, 0]


def test_kappa_decomposition_is_the_brute_force_correlation_integral():
    """compute_kappa_decomposition on synthetic group series: each group's
    tensor is V / (k_B T^2) times the trapezoid integral of the
    mean-subtracted <G_a(t + k) J_b(t)> summed directly over t, and the
    groups add up to the total (the integral is linear in the group
    flux)."""
    from akitsushima_amd import StillingerWeberPotential
    from akitsushima_amd.thermal import EV_ANG_K_FS_TO_W_MK, K_B, ThermalConductivityCalculator
    lat, pos, _ = sheared_si(1, shear=0.0, rattle=0.0, seed=0)
    tc = ThermalConductivityCalculator(lat, pos, ["Si"] * len(pos),
                                       StillingerWeberPotential(SW_SI, species=["Si"]),
                                       temperature=400.0)
    with pytest.raises(RuntimeError):
        tc.compute_kappa_decomposition(verbose=False)
    rng = np.random.default_rng(11)
    n, dt = 96, 2.0
    G = np.cumsum(rng.normal(size=(n, 3, 3)), axis=0) * 1e-4 + rng.normal(size=(n, 3, 3)) * 1e-4
    J = G.sum(axis=1)
    tc._J_series, tc._dt_hf = J, dt
    tc._J_groups = {"species": (np.array(["A", "B", "C"]), G)}
    out = tc.compute_kappa_decomposition(tau_max=0.05, verbose=False)
    n_lag = min(int(0.05 * 1e3 / dt), n // 2)
    V = abs(np.linalg.det(lat))
    pref = V / (K_B * 400.0 ** 2) * EV_ANG_K_FS_TO_W_MK

    def brute(A, B):
        A = A - A.mean(0)
        B = B - B.mean(0)
        k = np.zeros((3, 3))
        for a in range(3):
            for b in range(3):
                c = [np.dot(A[j:, a], B[:n - j, b]) / (n - j) for j in range(n_lag)]
                k[a, b] = sum(0.5 * dt * (c[j - 1] + c[j]) for j in range(1, n_lag))
        return k * pref

    total = brute(J, J)
    assert np.allclose(out["total"], total, rtol=1e-10, atol=1e-12 * np.abs(total).max())
    for q, g in enumerate("ABC"):
        want = brute(G[:, q], J)
        assert np.allclose(out["species"][g], want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
    assert out["sum_rule_residual"]["species"] < 1e-12 * np.abs(total).max()


def test_save_results_writes_the_reference_csv_layouts(tmp_path):
    """compute_kappa on a synthetic flux series, then save_results: the
    three CSV files with the reference's headers, one row per lag, and the
    tensor rows equal to the returned kappa (save_results,
    solvers/thermal_conductivity.py:1017-1069)."""
    from akitsushima_amd import StillingerWeberPotential
    from akitsushima_amd.thermal import ThermalConductivityCalculator
    lat, pos, _ = sheared_si(1, shear=0.0, rattle=0.0, seed=0)
    tc = ThermalConductivityCalculator(lat, pos, ["Si"] * len(pos),
                                       StillingerWeberPotential(SW_SI, species=["Si"]),
                                       temperature=300.0)
    J = np.random.default_rng(1).normal(size=(400, 3)) * 1e-6
    kappa = tc.compute_kappa(J_series=J, dt_hf=2.0, tau_max=0.05, n_blocks=2)
    tc.save_results(str(tmp_path), verbose=False)
    h = open(tmp_path / "hcacf.csv").read().splitlines()
    assert h[0] == "tau_ps,Cxx,Cyy,Czz,Cxy,Cxz,Cyz" and len(h) == 1 + 25
    r = open(tmp_path / "kappa_running.csv").read().splitlines()
    assert r[0].startswith("tau_ps,kappa_xx,kappa_yy,kappa_zz,kappa_avg") and len(r) == 26
    t = [ln.split(",") for ln in open(tmp_path / "kappa_tensor.csv").read().splitlines()[2:]]
    assert len(t) == 9
    got = np.array([float(x[2]) for x in t]).reshape(3, 3)
    assert np.allclose(got, kappa, atol=6e-5)
