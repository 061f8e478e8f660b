"""Site terms, heat flux, Nose-Hoover NVT and Green-Kubo on the HIP engine.

Against the oracle (pinned in tests/test_oracle_site_terms.py and
tests/test_oracle_md.py): per-atom site energies and every bond's dU/dv to
1e-10 relative, for each kind and row length; the heat flux of one state to
1e-10; the flux recorded inside a device MD run and the NVT trajectory
(energies, H_NHC, positions) against the oracle integrators.  Independently
of the oracle: the device heat flux equals d/dt sum_i r_i E_i on isolated
clusters (the reference's own check, tests/test_site_terms_heat_flux.py).
"""
import numpy as np
import pytest

import oracle as O
import oracle.site_terms as ST
from akitsushima_amd import (NeighbourList, StillingerWeberPotential, TersoffPotential,
                             ThermalConductivityCalculator, read_setfl, zhou_eam)
from akitsushima_amd.md import csvr_run, langevin_run, maxwell_boltzmann, md_run, nvt_run
from akitsushima_amd.published import ATOMIC_MASS, ZHOU_2004
from akitsushima_amd.structures import DIAMOND, FCC, cubic

from _cases import GESN_SPECIES, SW_GESN, SW_SI, TERSOFF_SI, TERSOFF_XY, rel_err, sheared_si

pytestmark = pytest.mark.gpu

C = O.EV_ANG_TO_AMU_ANG_FS2
T3L = dict(TERSOFF_SI, lambda3=1.7322)
CUNI = {s: dict(ZHOU_2004[s]) for s in ("Cu", "Ni")}


def _gesn(n=2, seed=2):
    lat, pos, _ = sheared_si(n, shear=0.04, rattle=0.1, seed=seed)
    f = 5.9 / 5.431
    t = [str(x) for x in np.random.default_rng(3).choice(GESN_SPECIES, len(pos))]
    return lat * f, pos * f, t


def _cuni(seed=4):
    lat, pos, _ = cubic(FCC, 3.6, 4, ["Cu"] * 4, rattle=0.08, seed=seed)
    t = [str(x) for x in np.random.default_rng(seed).choice(["Cu", "Ni"], len(pos))]
    return lat, pos, t


def _xy(seed=4):
    lat, pos, _ = sheared_si(3, shear=0.05, rattle=0.1, seed=seed)
    return lat, pos, (["X"] * 4 + ["Y"] * 4) * (len(pos) // 8)


CASES = {
    "sw": (lambda: StillingerWeberPotential(SW_SI, species=["Si"]),
           lambda l, p, t: ST.sw(l, p, t, O.sw_flat_as_records(SW_SI), ["Si"]),
           lambda: sheared_si(3, shear=0.05, rattle=0.1, seed=1)),
    "sw_gesn": (lambda: StillingerWeberPotential(SW_GESN, species=GESN_SPECIES),
                lambda l, p, t: ST.sw(l, p, t, SW_GESN, GESN_SPECIES), lambda: _gesn(3)),
    "tersoff": (lambda: TersoffPotential(T3L, species=["Si"]),
                lambda l, p, t: ST.tersoff(l, p, t, O.tersoff_flat_as_records(T3L), ["Si"]),
                lambda: sheared_si(3, shear=0.05, rattle=0.1, seed=1)),
    "tersoff_xy": (lambda: TersoffPotential(TERSOFF_XY, species=["X", "Y"]),
                   lambda l, p, t: ST.tersoff(l, p, t, TERSOFF_XY, ["X", "Y"]), _xy),
    "eam_cuni": (lambda: zhou_eam(CUNI, ["Cu", "Ni"]),
                 lambda l, p, t: ST.eam(l, p, t, O.zhou_table(CUNI, ["Cu", "Ni"]), ["Cu", "Ni"]),
                 _cuni),
}


def _sorted(st):
    o = np.lexsort((st.nbr, st.centre))
    return st.centre[o], st.nbr[o], st.vec[o], st.dUdv[o]


def _compare_terms(a, b, label):
    assert rel_err(a.e_site, b.e_site) < 1e-10, label
    ca, na, va, ga = _sorted(a)
    cb, nb, vb, gb = _sorted(b)
    assert np.array_equal(ca, cb) and np.array_equal(na, nb), label
    assert rel_err(va, vb) < 1e-12, label
    assert rel_err(ga, gb) < 1e-10, label


@pytest.mark.parametrize("skin", [None, 0.8])
@pytest.mark.parametrize("name", list(CASES))
def test_site_terms_match_the_oracle(name, skin):
    make, terms, cell = CASES[name]
    pot = make()
    lat, pos, t = cell()
    nl = None
    if skin is not None:                       # long rows: masked bond kernels
        nl = NeighbourList(pot.r_cut, skin_width=skin)
        nl.build(lat, pos)
    st = pot.site_terms(lat, pos, t, neigh_list=nl)
    ref = terms(lat, pos, t)
    _compare_terms(st, ref, f"{name} skin {skin}")
    E, F, S = pot.compute(lat, pos, t)
    assert abs(st.energy - E) < 1e-11 * abs(E)
    assert rel_err(st.forces(), F) < 1e-10
    assert rel_err(st.stress(abs(np.linalg.det(lat))), S) < 1e-10


def test_setfl_site_terms_match_the_oracle(tmp_path):
    el = {s: dict(ZHOU_2004[s]) for s in ("Al", "Ni")}
    path = str(tmp_path / "AlNi.eam.alloy")
    O.write_setfl_zhou(path, el, ["Al", "Ni"], ATOMIC_MASS, n_rho=1500, n_r=1500)
    pot = read_setfl(path)
    tab, names = O.read_setfl(path)
    lat, pos, _ = cubic(FCC, 3.646, 4, ["Al"] * 4, rattle=0.07, seed=2)
    t = [str(x) for x in np.random.default_rng(1).choice(names, len(pos))]
    _compare_terms(pot.site_terms(lat, pos, t), ST.eam(lat, pos, t, tab, names), "setfl")


@pytest.mark.parametrize("name", list(CASES))
def test_heat_flux_matches_the_oracle(name):
    make, terms, cell = CASES[name]
    pot = make()
    lat, pos, t = cell()
    m = np.random.default_rng(7).uniform(20.0, 120.0, len(pos))
    vel = maxwell_boltzmann(600.0, m, 11)
    J, e = pot.heat_flux(lat, pos, vel, t, m)
    ref = terms(lat, pos, t)
    J0 = ref.heat_flux(vel, m, abs(np.linalg.det(lat)))
    assert rel_err(J, J0) < 1e-10, (J, J0)
    assert rel_err(e, ref.e_site) < 1e-10


def _cluster(n_cells, a, basis, seed, L=40.0, jitter=0.08):
    lat = np.eye(3) * L
    pts = np.concatenate([(np.asarray(basis) + [i, j, k]) * a for i in range(n_cells)
                          for j in range(n_cells) for k in range(n_cells)])
    rng = np.random.default_rng(seed)
    pos = pts - pts.mean(0) + L / 2 + rng.normal(scale=jitter, size=pts.shape)
    return lat, pos, rng


@pytest.mark.parametrize("name", ["sw_gesn", "tersoff_xy", "eam_cuni"])
def test_device_heat_flux_is_the_time_derivative_of_the_energy_moment(name):
    make = CASES[name][0]
    pot = make()
    if name == "eam_cuni":
        lat, pos, rng = _cluster(2, 3.6, FCC, 3)
        t = [str(x) for x in rng.choice(["Cu", "Ni"], len(pos))]
        m = np.array([63.546 if x == "Cu" else 58.693 for x in t])
    else:
        lat, pos, rng = _cluster(2, 5.9 if name == "sw_gesn" else 5.431, DIAMOND, 1)
        if name == "sw_gesn":
            t = [str(x) for x in rng.choice(GESN_SPECIES, len(pos))]
            m = np.array([72.63 if x == "Ge" else 118.71 for x in t])
        else:
            t = (["X"] * 4 + ["Y"] * 4) * 8
            m = np.full(len(pos), 28.0855)
    vel = rng.normal(scale=np.sqrt(8.617e-5 * 300 * C / m[:, None]), size=pos.shape)
    vel -= (m[:, None] * vel).sum(0) / m.sum()
    V = abs(np.linalg.det(lat))

    def moment(p, v):
        E = 0.5 * m * np.sum(v ** 2, axis=1) / C + pot.site_terms(lat, p, t).e_site
        return (E[:, None] * p).sum(axis=0)

    def verlet(p, v, dt):
        v = v + 0.5 * dt * pot.compute(lat, p, t)[1] * C / m[:, None]
        p = p + dt * v
        v = v + 0.5 * dt * pot.compute(lat, p, t)[1] * C / m[:, None]
        return p, v

    J, _ = pot.heat_flux(lat, pos, vel, t, m)
    dt = 1e-3
    J_fd = (moment(*verlet(pos, vel, dt)) - moment(*verlet(pos, vel, -dt))) / (2 * dt) / V
    assert np.abs(J - J_fd).max() < 1e-5 * np.abs(J_fd).max(), (J, J_fd)


THERMOSTATS = {
    "nve": None,
    "nhc": ("nhc", 700.0, 40.0, 3),
    "langevin": ("langevin", 700.0, 40.0, 12345),
    "csvr": ("csvr", 700.0, 40.0, (1 << 40) + 99),
}


@pytest.mark.parametrize("thermo", list(THERMOSTATS))
@pytest.mark.parametrize("name", ["sw", "tersoff_xy", "eam_cuni"])
def test_md_flux_series_and_thermostats_match_the_oracle(name, thermo):
    make, terms, cell = CASES[name]
    pot = make()
    lat, pos, t = cell()
    m = np.full(len(pos), 28.086) if name != "eam_cuni" else np.array(
        [63.546 if x == "Cu" else 58.693 for x in t])
    vel = maxwell_boltzmann(500.0, m, 4)
    n_steps, every, skin = 40, 3, 0.3
    th = THERMOSTATS[thermo]
    res = md_run(pot, lat, pos, vel, t, dt=1.0, n_steps=n_steps, skin_width=skin, masses=m,
                 thermostat=th, flux_every=every)
    if name == "sw":
        force = lambda l, p, pr: O.sw_flat(l, p, SW_SI, pairs=pr)
    elif name == "tersoff_xy":
        force = lambda l, p, pr: O.tersoff_multi(l, p, t, TERSOFF_XY, ["X", "Y"], pairs=pr)
    else:
        force = lambda l, p, pr: O.eam_zhou(l, p, t, CUNI, ["Cu", "Ni"], pairs=pr)
    V = abs(np.linalg.det(lat))
    flux = lambda p, v: terms(lat, p, t).heat_flux(v, m, V)
    kw = {}
    if th is not None:
        kw = {"nhc": th[1:]} if th[0] == "nhc" else {th[0]: th[1:]}
    ep, ek, ec, p1, v1, nreb, J = O.md_run(force, lat, pos, vel, m, 1.0, n_steps, pot.r_cut,
                                           skin, flux=flux, every=every, **kw)
    assert np.abs(res["e_pot"] - ep).max() < 1e-9 * np.abs(ep).max()
    assert np.abs(res["e_kin"] - ek).max() < 1e-9 * np.abs(ek).max()
    assert np.abs(res["e_cons"] - ec).max() < 1e-9 * np.abs(ec).max()
    assert res["n_rebuilds"] == nreb
    d = res["positions"] - p1
    d -= np.round(d @ np.linalg.inv(lat)) @ lat
    assert np.abs(d).max() < 1e-8
    assert np.abs(res["velocities"] - v1).max() < 1e-9 * np.abs(v1).max()
    assert res["flux"].shape == J.shape == ((n_steps + every - 1) // every, 3)
    assert rel_err(res["flux"], J) < 1e-8


def test_nvt_conserves_h_and_reaches_the_temperature():
    lat, pos, t = cubic(DIAMOND, 5.431, 4, ["Si"] * 8, rattle=0.02, seed=1)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    m = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(300.0, m, 2)
    res = nvt_run(pot, lat, pos, vel, t, 900.0, dt=0.5, n_steps=6000, tau_nhc=50.0,
                  nhc_chain=3)
    et = res["e_tot"]
    ec = res["e_cons"]
    assert ec.max() - ec.min() < 2e-3 * (et.max() - et.min())
    assert abs(res["T"][3000:].mean() - 900.0) < 0.05 * 900.0


@pytest.mark.parametrize("run", ["langevin", "csvr"])
def test_stochastic_thermostats_sample_the_temperature(run):
    lat, pos, t = cubic(DIAMOND, 5.431, 4, ["Si"] * 8, rattle=0.02, seed=1)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    m = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(300.0, m, 2)
    fn = langevin_run if run == "langevin" else csvr_run
    res = fn(pot, lat, pos, vel, t, 900.0, dt=1.0, n_steps=4000, damping=50.0, seed=3)
    et, ec = res["e_tot"], res["e_cons"]
    assert ec.max() - ec.min() < 5e-3 * (et.max() - et.min())
    assert abs(res["T"][1000:].mean() - 900.0) < 0.03 * 900.0
    P = (m[:, None] * res["velocities"]).sum(axis=0)
    assert np.abs(P).max() < 1e-9 * np.abs(m[:, None] * res["velocities"]).sum()


def test_green_kubo_run_uses_the_device_flux():
    lat, pos, t = cubic(DIAMOND, 5.431, 3, ["Si"] * 8, rattle=0.0)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    tc = ThermalConductivityCalculator(lat, pos, t, pot, temperature=500.0)
    out = tc.run_md(n_steps=400, timestep=1.0, equil_steps=200, heat_flux_every=1, seed=3)
    J = out["J_series"]
    assert J.shape == (400, 3) and np.all(np.isfinite(J))
    # the last frame is the final state: the same flux from a single call
    J1, _ = pot.heat_flux(lat, out["positions"], out["velocities"], t, tc.masses)
    assert rel_err(J[-1], J1) < 1e-10
    kappa = tc.compute_kappa(tau_max=0.1, n_blocks=2)
    assert kappa.shape == (3, 3) and np.all(np.isfinite(kappa))


# --------------------------------------------------------------------------
#  NPT (isotropic barostat) on the device
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["sw", "tersoff_xy", "eam_cuni"])
def test_npt_matches_the_oracle(name):
    from akitsushima_amd.md import npt_run

    make, _, cell = CASES[name]
    pot = make()
    lat, pos, t = cell()
    lat, pos = lat * 0.985, pos * 0.985          # compressed: the cell moves
    m = np.full(len(pos), 28.086) if name != "eam_cuni" else np.array(
        [63.546 if x == "Cu" else 58.693 for x in t])
    vel = maxwell_boltzmann(500.0, m, 4)
    n_steps, skin = 40, 0.5
    res = npt_run(pot, lat, pos, vel, t, 500.0, pressure=0.5, dt=1.0, n_steps=n_steps,
                  tau_nhc=40.0, nhc_chain=3, tau_baro=150.0, skin_width=skin, masses=m)
    if name == "sw":
        force = lambda l, p, pr: O.sw_flat(l, p, SW_SI, pairs=pr)
    elif name == "tersoff_xy":
        force = lambda l, p, pr: O.tersoff_multi(l, p, t, TERSOFF_XY, ["X", "Y"], pairs=pr)
    else:
        force = lambda l, p, pr: O.eam_zhou(l, p, t, CUNI, ["Cu", "Ni"], pairs=pr)
    r = O.npt_run(force, lat, pos, vel, m, 1.0, n_steps, pot.r_cut, skin, 500.0, 40.0, 3, 0.5,
                  150.0)
    for a, b in (("e_pot", "ep"), ("e_kin", "ek"), ("e_cons", "h"), ("volume", "v"),
                 ("pressure_gpa", "p")):
        assert np.abs(res[a] - r[b]).max() < 1e-9 * np.abs(r[b]).max(), a
    assert res["n_rebuilds"] == r["rebuilds"]
    assert np.abs(res["lattice"] - r["lattice"]).max() < 1e-10 * np.abs(r["lattice"]).max()
    d = res["positions"] - r["positions"]
    d -= np.round(d @ np.linalg.inv(res["lattice"])) @ res["lattice"]
    assert np.abs(d).max() < 1e-8
    assert np.abs(res["velocities"] - r["velocities"]).max() < 1e-9 * np.abs(r["velocities"]).max()
    assert abs(res["volume"][-1] - res["volume"][0]) > 1e-4 * res["volume"][0]


def test_npt_with_a_frozen_barostat_is_the_device_nvt():
    from akitsushima_amd.md import npt_run

    lat, pos, t = cubic(DIAMOND, 5.431, 3, ["Si"] * 8, rattle=0.03, seed=1)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    m = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(600.0, m, 2)
    a = npt_run(pot, lat, pos, vel, t, 600.0, dt=1.0, n_steps=300, tau_nhc=50.0, tau_baro=1e12)
    b = nvt_run(pot, lat, pos, vel, t, 600.0, dt=1.0, n_steps=300, tau_nhc=50.0)
    assert np.abs(a["e_pot"] - b["e_pot"]).max() < 1e-9 * np.abs(b["e_pot"]).max()
    assert np.abs(a["positions"] - b["positions"]).max() < 1e-8
    assert np.array_equal(a["lattice"], lat)


def test_npt_relaxes_to_the_target_pressure():
    from akitsushima_amd.md import npt_run

    lat, pos, t = cubic(DIAMOND, 5.40, 4, ["Si"] * 8, rattle=0.02, seed=1)
    pot = StillingerWeberPotential(SW_SI, species=["Si"])
    m = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(600.0, m, 2)
    r = npt_run(pot, lat, pos, vel, t, 600.0, pressure=0.0, dt=1.0, n_steps=4000, tau_nhc=50.0,
                tau_baro=300.0)
    et, h = r["e_tot"], r["e_cons"]
    assert h.max() - h.min() < 3e-3 * (et.max() - et.min())
    assert r["pressure_gpa"][0] > 2.0 and r["volume"][-1] > 1.02 * r["volume"][0]
    assert abs(r["pressure_gpa"][2000:].mean()) < 0.2


# --------------------------------------------------------------------------
#  Heat flux per group of atoms (run_md decompose)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["sw_gesn", "tersoff_xy", "eam_cuni"])
def test_group_heat_fluxes_match_the_oracle(name):
    make, terms, cell = CASES[name]
    pot = make()
    lat, pos, t = cell()
    m = np.array([ATOMIC_MASS.get(x, 28.086) for x in t])
    vel = maxwell_boltzmann(500.0, m, 5)
    species = sorted(set(t), key=list(t).index)
    labels = np.array([species.index(x) for x in t])
    bounds = [0.0, 0.3, 0.55, 1.0]
    n_steps, every, skin = 12, 4, 0.3
    res = md_run(pot, lat, pos, vel, t, dt=1.0, n_steps=n_steps, skin_width=skin, masses=m,
                 flux_every=every, flux_groups=(labels, len(species), (1, bounds)),
                 atom_flux=True)
    Jg = res["flux_groups"]
    assert Jg.shape == (3, len(species), 3, 3)
    # the groups add up to the total flux
    assert rel_err(Jg.sum(axis=(1, 2)), res["flux"]) < 1e-12
    # each group against the oracle's per-atom currents on the same states
    if name == "sw_gesn":
        force = lambda l, p, pr: O.sw_multi(l, p, t, SW_GESN, GESN_SPECIES, pairs=pr)
    elif name == "tersoff_xy":
        force = lambda l, p, pr: O.tersoff_multi(l, p, t, TERSOFF_XY, ["X", "Y"], pairs=pr)
    else:
        force = lambda l, p, pr: O.eam_zhou(l, p, t, CUNI, ["Cu", "Ni"], pairs=pr)
    V = abs(np.linalg.det(lat))
    frames, atoms = [], []

    def flux(p, v):
        Ji = terms(lat, p, t).atom_heat_current(v, m)
        atoms.append(Ji / V)
        reg = ST.group_labels(lat, p, t, {"regions": {"axis": "b", "bounds": bounds}})
        rl = reg["regions"][0]
        g = np.zeros((len(species), 3, 3))
        np.add.at(g, (labels, rl), Ji)
        frames.append(g / V)
        return Ji.sum(axis=0) / V

    O.md_run(force, lat, pos, vel, m, 1.0, n_steps, pot.r_cut, skin, flux=flux, every=every)
    want = np.array(frames)
    assert rel_err(Jg, want) < 1e-8
    assert np.count_nonzero(np.abs(want).max(axis=(0, 3)) > 0) >= 4
    # the per-atom currents of every frame, rounded to float32
    Ja, wa = res["atom_flux"], np.array(atoms)
    assert Ja.dtype == np.float32 and Ja.shape == (3, len(t), 3)
    assert np.abs(Ja - wa).max() < 1e-6 * np.abs(wa).max()


def test_green_kubo_decomposition_sums_to_the_total(tmp_path):
    lat, pos, t = _gesn(2)
    pot = StillingerWeberPotential(SW_GESN, species=GESN_SPECIES)
    tc = ThermalConductivityCalculator(lat, pos, t, pot, temperature=500.0)
    out = tc.run_md(n_steps=400, timestep=1.0, equil_steps=50, heat_flux_every=2, seed=3,
              output_dir=str(tmp_path), thermo_every=50, traj_every=100,
              decompose={"species": True, "save_atom_flux": True,
                         "regions": {"axis": "c", "bounds": [0, 0.5, 1]}})
    # the NVE logs in the reference's formats, from the device records
    rows = [ln.split() for ln in open(tmp_path / "thermo_nve.log")][1:]
    assert [int(r[0]) for r in rows] == list(range(0, 401, 50))
    assert np.allclose([float(r[3]) for r in rows], out["e_tot"][::50], atol=2e-6)
    p0 = out["equil"]["positions"]
    _, _, S0 = pot.compute(lat, p0, t)
    assert abs(float(rows[0][4]) - np.trace(S0) / 3 * 160.2176621) < 2e-4
    xyz = open(tmp_path / "trajectory_tc.xyz").read().split("\n")
    n = len(t)
    steps = [int(ln.rsplit("step=", 1)[1]) for ln in xyz if "step=" in ln]
    assert steps == [0, 100, 200, 300, 400]
    last = np.array([[float(x) for x in ln.split()[1:]] for ln in xyz[-n - 1:-1]])
    assert np.abs(last - out["positions"]).max() < 1e-7
    flux = np.load(tmp_path / "atom_heat_flux.npy")
    assert flux.shape == (200, len(t), 3) and flux.dtype == np.float32
    assert np.allclose(flux.sum(axis=1), tc._J_series, rtol=1e-5,
                       atol=1e-6 * np.abs(tc._J_series).max())
    tc.compute_kappa(tau_max=0.1, n_blocks=2)
    out = tc.compute_kappa_decomposition(tau_max=0.1, verbose=False)
    assert set(out["species"]) == set(GESN_SPECIES) and len(out["regions"]) == 2
    scale = np.abs(out["total"]).max()
    assert max(out["sum_rule_residual"].values()) < 1e-10 * scale
    assert np.allclose(np.diag(out["total"]), np.diag(tc._kappa_run[-1]), rtol=1e-9)


@pytest.mark.parametrize("every", [3, 0])
def test_md_with_the_bond_pair_site_form_matches_the_oracle(every):
    """Tersoff with the lane-per-(bond, partner) site form: a captured MD run
    that records the flux every 3 steps alternates it (plain steps) with the
    lane-per-bond form (recording steps, which need site energies); and the
    same run with the flux off."""
    make, terms, cell = CASES["tersoff_xy"]
    pot = make()
    pot.set_bond_tile(2)
    lat, pos, t = cell()
    m = np.full(len(pos), 28.086)
    vel = maxwell_boltzmann(500.0, m, 6)
    n_steps, skin = 40, 0.3
    res = md_run(pot, lat, pos, vel, t, dt=1.0, n_steps=n_steps, skin_width=skin, masses=m,
                 flux_every=every)
    force = lambda l, p, pr: O.tersoff_multi(l, p, t, TERSOFF_XY, ["X", "Y"], pairs=pr)
    V = abs(np.linalg.det(lat))
    flux = (lambda p, v: terms(lat, p, t).heat_flux(v, m, V)) if every else None
    ep, ek, ec, p1, v1, nreb, J = O.md_run(force, lat, pos, vel, m, 1.0, n_steps, pot.r_cut,
                                           skin, flux=flux, every=every)
    assert np.abs(res["e_pot"] - ep).max() < 1e-9 * np.abs(ep).max()
    assert np.abs(res["e_kin"] - ek).max() < 1e-9 * np.abs(ek).max()
    assert res["n_rebuilds"] == nreb
    d = res["positions"] - p1
    d -= np.round(d @ np.linalg.inv(lat)) @ lat
    assert np.abs(d).max() < 1e-8
    if every:
        assert rel_err(res["flux"], J) < 1e-8
