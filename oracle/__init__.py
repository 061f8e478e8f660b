"""Oracle: CPU restatement of the reference's FP64 force path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker (or the timed CPU baseline).  The product package ``akitsushima_amd``
never imports it and has no CPU fallback.

The compiled kernels live in ``refcpu.c`` (built by ``oracle/Makefile`` into
``oracle/liboracle.so``); this module is the NumPy glue around them, written
after the reference's Python glue:

* ``cell_metrics``/``check_min_cell``: engine/neighborlist.py:170-267, 489-572
* ``nl_build``: NeighbourList._build_cell_linked, engine/neighborlist.py:935-1024,
  plus _canonical_pair_order :1166-1181
* ``candidate_pairs``: owned_neighbour_list, engine/neighborlist.py:1103-1163
  (skin 0.3, all pairs below 100 atoms or in thin cells)
* ``tersoff_flat``: TersoffPotential._compute_single_species, engine/tersoff.py:1233-1513
* ``tersoff_multi``: TersoffPotential._compute_multi_species, engine/tersoff.py:1518-1951
* ``sw_flat``: StillingerWeberPotential._compute_single_species,
  engine/stillinger_weber.py:927-1159
* ``sw_multi``: StillingerWeberPotential._compute_multi_species (tables
  :1255-1315, kernels _sw_multi_2body_numpy :414-447 and
  _sw_multi_3body_numpy :476-560)
* ``zhou_*`` and ``eam_zhou``: engine/eam_zhou.py:47-207 and the numba kernel
  engine/_eam_kernels.py:231-307
* ``read_setfl`` / ``write_setfl_zhou`` / ``eam_table``: setfl tables as
  engine/eam_zhou.py:210-315 reads and writes them, evaluated as
  engine/_eam_kernels.py:128-145, 199-205 (scipy CubicSpline coefficients)
* ``tersoff_loss``: fitting/loss.py:51-248 (squared loss)
* ``nve_run``, ``md_run``: NVESimulation.run, solvers/molecular_dynamics.py:1969-1996,
  and NVTSimulation.run with the Nose-Hoover chain (:635-727, 1113-1160)
* ``npt_run``: NPTSimulation.run, every cell mode and both barostat
  thermostats (:2141-2692, 2871-2895)
* ``site_terms``: engine/site_terms.py (site energies, per-bond gradients,
  heat flux)
* ``nemd_run``: NEMDSimulation.run's step with the two heat baths,
  solvers/nemd.py:113-117, 184-193

What pins each function (none is left unpinned):

* cell list: brute-force O(N^2) pair sets (tests/test_oracle_derivatives.py)
* Tersoff flat / multi, SW flat / multi, Zhou EAM: the reference's ledger energies
  (benchmarks/adr0047.jsonl, tests/golden/), published cohesive energies,
  brute-force double sums, 4th-order finite differences of forces and stress
  (tests/test_oracle_golden.py, tests/test_oracle_derivatives.py); SW multi
  also one-species records == flat, and invariance under atom relabelling
* tersoff_loss: closed-form normalisation on hand-set residuals, and an
  independent torch-autograd restatement (tests/test_fit_reference.py)
* mb_velocities, nve_run, md_run: zero COM momentum and exact temperature;
  energy conservation and time reversibility; with the chain, conservation
  of H_NHC, time reversibility and the canonical mean temperature
* md_run with the chain and npt_run: the reference's recorded 64-atom Si
  NVT and NPT runs (tests/golden/nhc_legacy_si64.npz: energies, P, V of
  every step, the chain and barostat state at step 150, the step-300
  state), from the start and continued from its step-150 checkpoint
  (tests/test_oracle_recorded_md.py); the flexible-cell modes: H_NPT
  conserved to O(dt^2), the frozen-barostat limit = the NVT run, axes
  without a generator left bit-identical, relaxation of an anisotropic
  strain to the target pressure tensor (tests/test_oracle_npt_cell.py)
* site_terms: rebuilt forces/stress/energy = the force path; heat flux =
  d/dt sum_i r_i E_i on isolated clusters (tests/test_oracle_site_terms.py)
* nemd_run: E_tot - Q conserved to O(dt^2) (the reference's own check,
  tests/test_nemd.py:113-121); with the baths off (damping -> infinity)
  the NVE run bit for bit; free bath atoms relax to the exact
  Ornstein-Uhlenbeck variance k_B T / m per component
  (tests/test_oracle_nemd.py)
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile refcpu.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_nl_build.restype = ctypes.c_int64
        L.orc_nl_build.argtypes = [ctypes.c_int32, _i32p, _f64p, _f64p, _f64p,
                                   ctypes.c_double, _i32p, _i32p, _f64p, _f64p]
        common = [ctypes.c_int32, _f64p, _f64p, _f64p, ctypes.c_double,
                  ctypes.c_int64, _i32p, _i32p]
        L.orc_tersoff_flat.restype = ctypes.c_int
        L.orc_tersoff_flat.argtypes = common + [_f64p, ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_tersoff_multi.restype = ctypes.c_int
        L.orc_tersoff_multi.argtypes = common + [_i32p, ctypes.c_int32, _f64p, _f64p,
                                                 ctypes.c_double, ctypes.c_int,
                                                 _f64p, _f64p, _f64p]
        L.orc_sw_flat.restype = ctypes.c_int
        L.orc_sw_flat.argtypes = common + [_f64p, ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_sw_multi.restype = ctypes.c_int
        L.orc_sw_multi.argtypes = common + [ctypes.c_int32, _i32p, _f64p, _f64p, _f64p,
                                            ctypes.c_double, ctypes.c_int,
                                            _f64p, _f64p, _f64p]
        L.orc_eam.restype = ctypes.c_int
        _i64p = ctypes.POINTER(ctypes.c_int64)
        L.orc_eam.argtypes = common + [_i32p, ctypes.c_int32, _i32p, _i32p, _i32p,
                                       _f64p, _i64p, _i64p, _f64p, _f64p, ctypes.c_int64,
                                       ctypes.c_double, ctypes.c_int,
                                       _f64p, _f64p, _f64p]
        L.orc_eam_feval.restype = None
        L.orc_eam_feval.argtypes = [_f64p, _i64p, _i64p, _f64p, _f64p, ctypes.c_int64,
                                    ctypes.c_int32, ctypes.c_int64, _f64p, _f64p, _f64p]
        L.orc_zhou_feval.restype = None
        L.orc_zhou_feval.argtypes = [_f64p, ctypes.c_int64, _f64p, _f64p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------
#  Cell geometry (engine/neighborlist.py:170-267, 98-132, 489-572)
# ---------------------------------------------------------------------------

def cell_metrics(lattice):
    lat = np.asarray(lattice, dtype=float)
    volume = abs(float(np.linalg.det(lat)))
    cross = np.array([np.linalg.norm(np.cross(lat[(i + 1) % 3], lat[(i + 2) % 3]))
                      for i in range(3)])
    if volume <= 0.0 or not np.all(cross > 0.0):
        raise ValueError("degenerate (singular) lattice")
    perp = volume / cross
    sigma_min = float(np.linalg.svd(lat, compute_uv=False)[-1])
    return perp, sigma_min


def check_min_cell(lattice, r_list):
    perp, _ = cell_metrics(lattice)
    if r_list >= float(np.min(perp)) / 2.0:
        raise ValueError(
            f"cutoff radius ({r_list:.4f} A) must be < half the smallest "
            f"perpendicular cell width ({float(np.min(perp)) / 2.0:.4f} A)")


# ---------------------------------------------------------------------------
#  Neighbour list
# ---------------------------------------------------------------------------

def nl_build(lattice, positions, r_list):
    """Candidate pairs i < j with r^2 < r_list^2 in canonical (i, j) order:
    (idx_i, idx_j, rij, r), as NeighbourList.build with the CLL back-end
    followed by _canonical_pair_order."""
    lattice = np.ascontiguousarray(lattice, dtype=float)
    positions = np.ascontiguousarray(positions, dtype=float)
    check_min_cell(lattice, r_list)
    inv = np.ascontiguousarray(np.linalg.inv(lattice))
    perp, _ = cell_metrics(lattice)
    nc = np.maximum(1, np.floor(perp / r_list).astype(int)).astype(np.int32)
    n = len(positions)
    rl2 = r_list ** 2
    L = lib()
    npair = L.orc_nl_build(n, _p(nc, _i32p), _p(positions, _f64p), _p(lattice, _f64p),
                           _p(inv, _f64p), rl2, None, None, None, None)
    oi = np.empty(npair, np.int32)
    oj = np.empty(npair, np.int32)
    orij = np.empty((npair, 3))
    orr = np.empty(npair)
    L.orc_nl_build(n, _p(nc, _i32p), _p(positions, _f64p), _p(lattice, _f64p),
                   _p(inv, _f64p), rl2, _p(oi, _i32p), _p(oj, _i32p),
                   _p(orij, _f64p), _p(orr, _f64p))
    return oi, oj, orij, orr


def brute_force_pairs(lattice, positions, r_list):
    """O(N^2) minimum-image pair set, independent of the cell list
    (tests/test_neighborlist.py:27-41 style), vectorised per row."""
    lattice = np.asarray(lattice, float)
    positions = np.asarray(positions, float)
    inv = np.linalg.inv(lattice)
    out_i, out_j = [], []
    n = len(positions)
    for i in range(n - 1):
        d = positions[i + 1:] - positions[i]
        s = d[:, 0:1] * inv[0] + d[:, 1:2] * inv[1] + d[:, 2:3] * inv[2]
        s = s - np.round(s)
        v = s[:, 0:1] * lattice[0] + s[:, 1:2] * lattice[1] + s[:, 2:3] * lattice[2]
        r2 = v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]
        m = np.nonzero((r2 < r_list * r_list) & (r2 > 1e-20))[0]
        out_i.append(np.full(len(m), i, np.int32))
        out_j.append((m + i + 1).astype(np.int32))
    if not out_i:
        return np.empty(0, np.int32), np.empty(0, np.int32)
    return np.concatenate(out_i), np.concatenate(out_j)


OWNED_SKIN = 0.3
OWNED_MIN_ATOMS = 100


def candidate_pairs(lattice, positions, r_cut):
    """The candidates a force function called without a list uses
    (owned_neighbour_list, engine/neighborlist.py:1103-1163)."""
    n = len(positions)
    perp, _ = cell_metrics(lattice)
    if n < OWNED_MIN_ATOMS or r_cut + OWNED_SKIN >= float(np.min(perp)) / 2.0:
        check_min_cell(lattice, r_cut)
        ii, jj = np.triu_indices(n, k=1)
        return ii.astype(np.int32), jj.astype(np.int32)
    i, j, _, _ = nl_build(lattice, positions, r_cut + OWNED_SKIN)
    return i, j


def _prep(lattice, positions, pairs, r_cut):
    lattice = np.ascontiguousarray(lattice, dtype=float)
    positions = np.ascontiguousarray(positions, dtype=float)
    if pairs is None:
        pairs = candidate_pairs(lattice, positions, r_cut)
    ci = np.ascontiguousarray(pairs[0], dtype=np.int32)
    cj = np.ascontiguousarray(pairs[1], dtype=np.int32)
    inv = np.ascontiguousarray(np.linalg.inv(lattice))
    vol = float(np.abs(np.linalg.det(lattice)))
    return lattice, positions, ci, cj, inv, vol


# ---------------------------------------------------------------------------
#  Tersoff
# ---------------------------------------------------------------------------

TERSOFF_KEYS = ("A", "B", "lambda1", "lambda2", "lambda3", "beta", "n",
                "c", "d", "h", "R", "D")

# benchmarks/bench_force.py:81-98 (T3 Si)
TERSOFF_SI = {"A": 1830.8, "B": 471.18, "lambda1": 2.4799, "lambda2": 1.7322,
              "lambda3": 0.0, "beta": 1.1e-6, "n": 0.78734, "c": 100390.0,
              "d": 16.217, "h": -0.59825, "R": 2.85, "D": 0.15}


def tersoff_flat(lattice, positions, params=TERSOFF_SI, pairs=None, nthreads=1):
    prm = np.array([float(params[k]) for k in TERSOFF_KEYS])
    rc = prm[10] + prm[11]
    lattice, positions, ci, cj, inv, vol = _prep(lattice, positions, pairs, rc)
    n = len(positions)
    e = ctypes.c_double()
    f = np.zeros((n, 3))
    s = np.zeros((3, 3))
    lib().orc_tersoff_flat(n, _p(positions, _f64p), _p(lattice, _f64p), _p(inv, _f64p),
                           vol, len(ci), _p(ci, _i32p), _p(cj, _i32p), _p(prm, _f64p),
                           int(nthreads), ctypes.byref(e), _p(f, _f64p), _p(s, _f64p))
    return e.value, f, s


def tersoff_tables(records, species):
    """Dense (n_sp, n_sp) pair table and (n_sp)^3 triple table as built in
    engine/tersoff.py:1606-1653.  records: {(i, j, k): dict}."""
    ns = len(species)
    tb = np.zeros((ns, ns, 9))
    tb[:, :, 5] = 1.0   # R default
    tb[:, :, 6] = 1.0   # D default
    tb[:, :, 8] = 1.0   # n default
    tri = np.zeros((ns, ns, ns, 7))
    tri[:, :, :, 2] = 1.0   # d default
    tri[:, :, :, 5] = 3.0   # m default
    tri[:, :, :, 6] = 1.0   # gamma default
    for a, sa in enumerate(species):
        for b, sb in enumerate(species):
            t = records.get((sa, sb, sb))
            if t is not None:
                tb[a, b] = [1.0, float(t.get("A", 0.0)), float(t.get("B", 0.0)),
                            float(t.get("lambda1", 0.0)), float(t.get("lambda2", 0.0)),
                            float(t.get("R", 1.0)), float(t.get("D", 1.0)),
                            float(t.get("beta", 0.0)), float(t.get("n", 1.0))]
            for c, sc in enumerate(species):
                t3 = records.get((sa, sb, sc))
                if t3 is not None:
                    tri[a, b, c] = [1.0, float(t3["c"]), float(t3["d"]), float(t3["h"]),
                                    float(t3["lambda3"]), float(t3.get("m", 3.0)),
                                    float(t3.get("gamma", 1.0))]
    rc = 0.0
    for (i, j, k), v in records.items():
        if j == k and "R" in v and "D" in v:
            rc = max(rc, float(v["R"]) + float(v["D"]))
    return tb, tri, (rc if rc > 0 else 3.0)


def tersoff_multi(lattice, positions, types, records, species, pairs=None, nthreads=1):
    tb, tri, rc = tersoff_tables(records, species)
    lattice, positions, ci, cj, inv, vol = _prep(lattice, positions, pairs, rc)
    idx = {s: k for k, s in enumerate(species)}
    tid = np.array([idx[t] for t in types], np.int32)
    n = len(positions)
    e = ctypes.c_double()
    f = np.zeros((n, 3))
    s = np.zeros((3, 3))
    tb = np.ascontiguousarray(tb)
    tri = np.ascontiguousarray(tri)
    lib().orc_tersoff_multi(n, _p(positions, _f64p), _p(lattice, _f64p), _p(inv, _f64p),
                            vol, len(ci), _p(ci, _i32p), _p(cj, _i32p), _p(tid, _i32p),
                            len(species), _p(tb, _f64p), _p(tri, _f64p), rc, int(nthreads),
                            ctypes.byref(e), _p(f, _f64p), _p(s, _f64p))
    return e.value, f, s


def tersoff_flat_as_records(params, sp="Si", m=3.0, gamma=1.0):
    """The single-species flat dict as one (sp, sp, sp) record with explicit
    m and gamma (the 14-parameter form of the multi-species path)."""
    rec = {k: float(params[k]) for k in TERSOFF_KEYS}
    rec["m"] = float(m)
    rec["gamma"] = float(gamma)
    return {(sp, sp, sp): rec}


# ---------------------------------------------------------------------------
#  Stillinger-Weber (defaults: engine/stillinger_weber.py:933-946)
# ---------------------------------------------------------------------------

SW_KEYS = ("epsilon", "sigma", "A", "B", "p", "q", "a", "lambda", "gamma", "costheta0")
SW_DEFAULTS = {"epsilon": 2.1682, "sigma": 2.0951, "A": 7.049556, "B": 0.6022246,
               "p": 4.0, "q": 0.0, "a": 1.80, "lambda": 21.0, "gamma": 1.20,
               "costheta0": -1.0 / 3.0}
# benchmarks/bench_force.py:64-79
SW_SI = {"epsilon": 2.1683, "sigma": 2.0951, "a": 1.80, "gamma": 1.20, "lambda": 21.0,
         "A": 7.049556277, "B": 0.6022245584, "p": 4.0, "q": 0.0, "costheta0": -1 / 3}


def sw_prm(params):
    return np.array([float(params.get(k, SW_DEFAULTS[k])) for k in SW_KEYS])


def sw_flat(lattice, positions, params=SW_SI, pairs=None, nthreads=1):
    prm = sw_prm(params)
    rc = prm[6] * prm[1]
    lattice, positions, ci, cj, inv, vol = _prep(lattice, positions, pairs, rc)
    n = len(positions)
    e = ctypes.c_double()
    f = np.zeros((n, 3))
    s = np.zeros((3, 3))
    lib().orc_sw_flat(n, _p(positions, _f64p), _p(lattice, _f64p), _p(inv, _f64p), vol,
                      len(ci), _p(ci, _i32p), _p(cj, _i32p), _p(prm, _f64p), int(nthreads),
                      ctypes.byref(e), _p(f, _f64p), _p(s, _f64p))
    return e.value, f, s


def sw_tables(records, species):
    """Dense tables of the multi-species SW records
    (engine/stillinger_weber.py:1255-1315, :874-890): tb (ns, ns, 8) =
    valid, epsilon, sigma, a, A, B, p, q of (i, j, j) when it exists with
    A != 0; leg (ns, ns, 3) = sigma, a, gamma of (i, j, j); tri (ns, ns, ns,
    4) = valid, epsilon, lambda, costheta0 of (i, j, k) when (i, j, k),
    (i, j, j) and (i, k, k) all exist.  r_cut = max a sigma over records
    with j == k, else 4 (:860-871)."""
    ns = len(species)
    tb = np.zeros((ns, ns, 8))
    tb[:, :, 2] = 1.0
    leg = np.zeros((ns, ns, 3))
    leg[:, :, 0] = 1.0
    tri = np.zeros((ns, ns, ns, 4))
    tri[..., 3] = -1.0 / 3.0
    for a, sa in enumerate(species):
        for b, sb in enumerate(species):
            t = records.get((sa, sb, sb))
            if t is None:
                continue
            leg[a, b] = [float(t["sigma"]), float(t["a"]), float(t["gamma"])]
            if float(t.get("A", 0.0)) != 0.0:
                tb[a, b] = [1.0, float(t.get("epsilon", 1.0)), float(t["sigma"]),
                            float(t["a"]), float(t["A"]), float(t.get("B", 0.0)),
                            float(t.get("p", 4)), float(t.get("q", 0))]
    for a, sa in enumerate(species):
        for b, sb in enumerate(species):
            for c, sc in enumerate(species):
                t3 = records.get((sa, sb, sc))
                if (t3 is None or records.get((sa, sb, sb)) is None
                        or records.get((sa, sc, sc)) is None):
                    continue
                tri[a, b, c] = [1.0, float(t3.get("epsilon", 1.0)), float(t3["lambda"]),
                                float(t3.get("costheta0", -1.0 / 3.0))]
    rc = 0.0
    for (i, j, k), v in records.items():
        if j == k:
            rc = max(rc, float(v["a"]) * float(v["sigma"]))
    return tb, leg, tri, (rc if rc > 0 else 4.0)


def sw_multi(lattice, positions, types, records, species, pairs=None, nthreads=1):
    tb, leg, tri, rc = sw_tables(records, species)
    lattice, positions, ci, cj, inv, vol = _prep(lattice, positions, pairs, rc)
    idx = {s: k for k, s in enumerate(species)}
    tid = np.array([idx[t] for t in types], np.int32)
    n = len(positions)
    e = ctypes.c_double()
    f = np.zeros((n, 3))
    s = np.zeros((3, 3))
    lib().orc_sw_multi(n, _p(positions, _f64p), _p(lattice, _f64p), _p(inv, _f64p), vol,
                       len(ci), _p(ci, _i32p), _p(cj, _i32p), len(species), _p(tid, _i32p),
                       _p(tb, _f64p), _p(leg, _f64p), _p(tri, _f64p), rc, int(nthreads),
                       ctypes.byref(e), _p(f, _f64p), _p(s, _f64p))
    return e.value, f, s


def sw_flat_as_records(params, sp="Si"):
    """The single-species flat dict as one (sp, sp, sp) record (the
    multi-species form of the same potential)."""
    rec = {k: float(params.get(k, SW_DEFAULTS[k])) for k in SW_KEYS}
    return {(sp, sp, sp): rec}


# ---------------------------------------------------------------------------
#  Zhou 2004 EAM (engine/eam_zhou.py:47-207)
# ---------------------------------------------------------------------------

ZHOU_KEYS = ("r_e", "f_e", "rho_e", "rho_s", "alpha", "beta", "A", "B", "kappa",
             "lambda", "Fn0", "Fn1", "Fn2", "Fn3", "F0", "F1", "F2", "F3", "eta", "F_e")

# Zhou, Johnson and Wadley, PRB 69, 144113 (2004), Table III rows used by the
# checks (engine/published_parameters.py:116-180).
ZHOU_CU = dict(zip(ZHOU_KEYS, (2.556162, 1.554485, 21.175871, 21.175395, 8.127620, 4.334731,
                               0.396620, 0.548085, 0.308782, 0.756515, -2.170269, -0.263788,
                               1.088878, -0.817603, -2.19, 0.0, 0.561830, -2.100595, 0.310490,
                               -2.186568)))
ZHOU_NI = dict(zip(ZHOU_KEYS, (2.488746, 2.007018, 27.562015, 27.930410, 8.383453, 4.471175,
                               0.429046, 0.633531, 0.443599, 0.820658, -2.693513, -0.076445,
                               0.241442, -2.375626, -2.70, 0.0, 0.265390, -0.152856, 0.469000,
                               -2.699486)))


def zhou_continuous(p):
    """engine/eam_zhou.py:47-64."""
    q = {k: float(v) for k, v in p.items()}
    re, rs, eta = q["rho_e"], q["rho_s"], q["eta"]
    F1, F2, F3 = q["F1"], q["F2"], q["F3"]
    y, x = 0.15, -0.15
    t0 = 1.15 * re / rs
    slope0 = (F1 + 2 * F2 * y + 3 * F3 * y * y) / re
    dg = -eta * eta * math.log(t0) * t0 ** (eta - 1.0) / rs
    q["F_e"] = slope0 / dg
    q["F0"] = q["F_e"] * (1 - eta * math.log(t0)) * t0 ** eta - (F1 * y + F2 * y * y + F3 * y ** 3)
    q["Fn0"] = q["F0"] + F1 * x + F2 * x * x + F3 * x ** 3
    q["Fn1"] = 0.85 * (F1 + 2 * F2 * x + 3 * F3 * x * x)
    return q


def _zdecay(r, re, pre, expo, shift):
    x = r / re
    e = pre * np.exp(-expo * (x - 1.0))
    u = x - shift
    den = 1.0 + u ** 20
    return e / den


def zhou_cutoff(p, tol=1e-6):
    """engine/eam_zhou.py:117-125."""
    r = np.linspace(p["r_e"], 5.0 * p["r_e"], 40001)
    pair = _zdecay(r, p["r_e"], p["A"], p["alpha"], p["kappa"]) - \
        _zdecay(r, p["r_e"], p["B"], p["beta"], p["lambda"])
    dens = _zdecay(r, p["r_e"], p["f_e"], p["beta"], p["lambda"])
    ok = (np.abs(pair) < tol * max(p["A"], p["B"])) & (dens < tol * p["f_e"])
    bad = np.nonzero(~ok)[0]
    return float(r[bad[-1] + 1]) if len(bad) else float(r[0])


ZW = 44


def zhou_function_table(elements, species, r_cut=None, chi=None, width=0.5):
    """Rows [code, rc, width, chi, pa[20], pb[20]] and the index maps
    RHO[s], EMB[s], PHI[s, t] of the numba kernel tables
    (engine/_eam_kernels.py:77-121, engine/eam_zhou.py:128-207)."""
    P = {s: zhou_continuous(elements[s]) for s in species}
    if r_cut is None:
        r_cut = max(zhou_cutoff(P[s]) for s in species)
    chi = chi or {}
    rows = []

    def row(code, pa, pb=None, k=1.0):
        v = np.zeros(ZW)
        v[0], v[1], v[2], v[3] = code, r_cut, width, k
        v[4:24] = [pa[kk] for kk in ZHOU_KEYS]
        if pb is not None:
            v[24:44] = [pb[kk] for kk in ZHOU_KEYS]
        rows.append(v)
        return len(rows) - 1

    ns = len(species)
    rho_f = np.array([row(3, P[s]) for s in species], np.int32)
    emb_f = np.array([row(6, P[s]) for s in species], np.int32)
    phi_f = np.zeros((ns, ns), np.int32)
    for a in range(ns):
        for b in range(a, ns):
            sa, sb = species[a], species[b]
            if a == b:
                k = row(4, P[sa])
            else:
                cf = float(chi.get(f"{sa}-{sb}", chi.get(f"{sb}-{sa}", 1.0)))
                k = row(5, P[sa], P[sb], cf)
            phi_f[a, b] = phi_f[b, a] = k
    return np.array(rows), rho_f, emb_f, phi_f, float(r_cut), P


class EamTable:
    """Function rows plus the spline tables of build_tables
    (engine/_eam_kernels.py:77-121): xo, nx per function, X (sum nx), C (4,
    sum nx) with a zero column at each function's last knot."""

    def __init__(self, rows, rho_f, emb_f, phi_f, r_cut, splines=None):
        self.rows = np.ascontiguousarray(rows, float)
        nf = len(self.rows)
        self.rho_f = np.ascontiguousarray(rho_f, np.int32)
        self.emb_f = np.ascontiguousarray(emb_f, np.int32)
        self.phi_f = np.ascontiguousarray(np.asarray(phi_f, np.int32).reshape(-1))
        self.r_cut = float(r_cut)
        self.xo = np.zeros(nf, np.int64)
        self.nx = np.zeros(nf, np.int64)
        Xs, Cs, off = [], [], 0
        for f, (x, c) in (splines or {}).items():
            x = np.asarray(x, float)
            cc = np.zeros((4, len(x)))
            cc[:, :len(x) - 1] = c
            self.xo[f], self.nx[f] = off, len(x)
            Xs.append(x)
            Cs.append(cc)
            off += len(x)
        self.X = np.ascontiguousarray(np.concatenate(Xs) if Xs else np.zeros(1))
        self.C = np.ascontiguousarray(np.concatenate(Cs, axis=1) if Cs else np.zeros((4, 1)))

    def args(self):
        i64 = ctypes.POINTER(ctypes.c_int64)
        return (_p(self.rows, _f64p), _p(self.xo, i64), _p(self.nx, i64), _p(self.X, _f64p),
                _p(self.C, _f64p), self.C.shape[1])

    def feval(self, f, r):
        r = np.ascontiguousarray(r, dtype=float)
        v = np.empty_like(r)
        dv = np.empty_like(r)
        lib().orc_eam_feval(*self.args(), int(f), len(r), _p(r, _f64p), _p(v, _f64p),
                            _p(dv, _f64p))
        return v, dv


def eam_table(lattice, positions, types, tab, species, pairs=None, nthreads=1):
    """engine/_eam_kernels.py:231-307 over an EamTable."""
    lattice, positions, ci, cj, inv, vol = _prep(lattice, positions, pairs, tab.r_cut)
    idx = {s: k for k, s in enumerate(species)}
    spc = np.array([idx[t] for t in types], np.int32)
    n = len(positions)
    e = ctypes.c_double()
    f = np.zeros((n, 3))
    s = np.zeros((3, 3))
    rows, xo, nx, X, C, NX = tab.args()
    lib().orc_eam(n, _p(positions, _f64p), _p(lattice, _f64p), _p(inv, _f64p), vol,
                  len(ci), _p(ci, _i32p), _p(cj, _i32p), _p(spc, _i32p), len(species),
                  _p(tab.rho_f, _i32p), _p(tab.emb_f, _i32p), _p(tab.phi_f, _i32p), rows,
                  xo, nx, X, C, NX, tab.r_cut, int(nthreads), ctypes.byref(e), _p(f, _f64p),
                  _p(s, _f64p))
    return e.value, f, s


def zhou_table(elements, species, r_cut=None, chi=None):
    fn, rho_f, emb_f, phi_f, rc, _ = zhou_function_table(elements, species, r_cut, chi)
    return EamTable(fn, rho_f, emb_f, phi_f, rc)


def eam_zhou(lattice, positions, types, elements, species, pairs=None, nthreads=1,
             r_cut=None, chi=None):
    return eam_table(lattice, positions, types, zhou_table(elements, species, r_cut, chi),
                     species, pairs=pairs, nthreads=nthreads)


# setfl files (engine/eam_zhou.py:210-315)

def _setfl_block(vals):
    vals = list(vals)
    return "\n".join(" ".join(f"{v: .16e}" for v in vals[i:i + 5])
                     for i in range(0, len(vals), 5)) + "\n"


def write_setfl_zhou(path, elements, species, masses, n_rho=5000, rho_max=None, n_r=5000,
                     r_cut=None, chi=None):
    """write_setfl (engine/eam_zhou.py:216-254) of the Zhou potential, the
    functions tabulated by this oracle."""
    fn, rho_f, emb_f, phi_f, rc, P = zhou_function_table(elements, species, r_cut, chi)
    if rho_max is None:
        rho_max = 2.0 * max(p["rho_e"] for p in P.values())
    drho = rho_max / (n_rho - 1)
    dr = rc / (n_r - 1)
    rho = np.arange(n_rho) * drho
    r = np.arange(n_r) * dr
    with open(path, "w") as fh:
        fh.write("\n\n\n")
        fh.write(f"{len(species)} " + " ".join(species) + "\n")
        fh.write(f"{n_rho} {drho:.16e} {n_r} {dr:.16e} {rc:.16e}\n")
        for k, s in enumerate(species):
            fh.write(f"0 {masses[s]:.6f} {0.0:.6f} fcc\n")
            fh.write(_setfl_block(zhou_feval(fn[emb_f[k]], rho)[0]))
            fh.write(_setfl_block(zhou_feval(fn[rho_f[k]], r)[0]))
        for i in range(len(species)):
            for j in range(i + 1):
                fh.write(_setfl_block(r * zhou_feval(fn[phi_f[i, j]], r)[0]))


def read_setfl(path):
    """read_setfl (engine/eam_zhou.py:257-315): each table a scipy
    CubicSpline (default not-a-knot ends), r*phi stored, phi = (r phi) / r.
    Returns (EamTable, species)."""
    from scipy.interpolate import CubicSpline

    with open(path) as fh:
        lines = fh.read().splitlines()
    head = lines[3].split()
    n_el = int(head[0])
    names = head[1:1 + n_el]
    n_rho, drho, n_r, dr, r_cut = lines[4].split()[:5]
    n_rho, n_r = int(n_rho), int(n_r)
    drho, dr, r_cut = float(drho), float(dr), float(r_cut)
    pos = 5
    tables = {}

    def take(n):
        nonlocal pos
        out = []
        while len(out) < n:
            out.extend(float(v) for v in lines[pos].split())
            pos += 1
        return np.array(out[:n])

    for s in names:
        pos += 1
        tables[("F", s)] = take(n_rho)
        tables[("rho", s)] = take(n_r)
    for i, a in enumerate(names):
        for b in names[:i + 1]:
            tables[("rphi", a, b)] = take(n_r)
    rho_g = np.arange(n_rho) * drho
    r_g = np.arange(n_r) * dr
    ns = len(names)
    rows, splines = [], {}

    def add(code, x, y):
        cs = CubicSpline(x, y)
        row = np.zeros(ZW)
        row[0] = code
        rows.append(row)
        splines[len(rows) - 1] = (cs.x, cs.c)
        return len(rows) - 1

    rho_f = [add(1, r_g, tables[("rho", s)]) for s in names]
    emb_f = [add(1, rho_g, tables[("F", s)]) for s in names]
    phi_f = np.zeros((ns, ns), np.int32)
    for i, a in enumerate(names):
        for j, b in enumerate(names[:i + 1]):
            phi_f[i, j] = phi_f[j, i] = add(2, r_g, tables[("rphi", a, b)])
    return EamTable(np.array(rows), rho_f, emb_f, phi_f, r_cut, splines), names


def zhou_feval(row, r):
    r = np.ascontiguousarray(r, dtype=float)
    v = np.empty_like(r)
    dv = np.empty_like(r)
    row = np.ascontiguousarray(row, dtype=float)
    lib().orc_zhou_feval(_p(row, _f64p), len(r), _p(r, _f64p), _p(v, _f64p), _p(dv, _f64p))
    return v, dv


# ---------------------------------------------------------------------------
#  Fit loss (fitting/loss.py:184-242, squared loss, default weights)
# ---------------------------------------------------------------------------

DEFAULT_WEIGHTS = {"energy": 1.0, "forces": 1.0, "stress": 0.1, "r0": 1.0, "eps0": 1.0}


def tersoff_loss(structures, records, species, weights=None, nthreads=1):
    """Sum over structures of w_E (N dE/N)^2 + w_F r0^2 mean dF^2
    + w_S (V eps0)^2 mean dsigma^2."""
    w = dict(DEFAULT_WEIGHTS)
    w.update(weights or {})
    total = 0.0
    for s in structures:
        lat, pos, types = s["lattice"], s["positions"], s["atom_types"]
        N = len(pos)
        vol = abs(np.linalg.det(lat))
        e, f, st = tersoff_multi(lat, pos, types, records, species, nthreads=nthreads)
        tot = w["energy"] * (N * (e / N - s["ref_energy"])) ** 2
        tot += w["forces"] * w["r0"] ** 2 * float(np.mean((f - s["ref_forces"]) ** 2))
        tot += w["stress"] * (vol * w["eps0"]) ** 2 * float(np.mean((st - s["ref_stress"]) ** 2))
        total += tot
    return total


def _huber_of_squared(q, delta):
    """fitting/loss.py:27-48: q if sqrt(q) <= delta else 2 delta sqrt(q) -
    delta^2 (q a squared residual, or a mean of squared residuals)."""
    r = np.sqrt(max(q, 0.0))
    return q if r <= delta else 2.0 * delta * r - delta ** 2


def generic_loss(params, structures, weights, loss_type="squared", nthreads=1):
    """generic_loss_function (fitting/loss.py:166-248) for a flat Tersoff
    parameter dict: per structure c_s^2 w_s [w_E rho((N dE)^2) + w_F rho(r0^2
    mean dF^2) + w_S rho((V eps0)^2 mean dsigma^2)], rho the identity or the
    Huber transform of each term, terms without a reference left out, V the
    structure's "volume" when given; an evaluation that raises gives 1e30."""
    w_E, w_F = weights.get("energy", 1.0), weights.get("forces", 1.0)
    w_S = weights.get("stress", 0.1)
    r0, eps0 = float(weights.get("r0", 1.0)), float(weights.get("eps0", 1.0))
    rho = ((lambda q, key: _huber_of_squared(q, weights[key])) if loss_type == "huber"
           else (lambda q, key: q))
    total = 0.0
    for s in structures:
        N = s["natoms"]
        vol = float(s.get("volume", abs(np.linalg.det(s["lattice"]))))
        try:
            e, f, st = tersoff_flat(s["lattice"], s["positions"], params, nthreads=nthreads)
        except Exception:
            return 1e30
        part = w_E * rho((N * (e / N - s["ref_energy"])) ** 2, "huber_delta_energy")
        if s.get("ref_forces") is not None:
            part += w_F * rho(r0 ** 2 * float(np.mean((f - s["ref_forces"]) ** 2)),
                              "huber_delta_forces")
        if s.get("ref_stress") is not None:
            part += w_S * rho((vol * eps0) ** 2 * float(np.mean((st - s["ref_stress"]) ** 2)),
                              "huber_delta_stress")
        total += float(s.get("weight", 1.0)) * float(s.get("scale", 1.0)) ** 2 * part
    return total


def generic_residuals(params, structures, weights, nthreads=1):
    """generic_residuals (fitting/loss.py:263-319): per structure
    sqrt(w_s) c_s [sqrt(w_E) N dE/N, sqrt(w_F / 3N) r0 dF (3N), sqrt(w_S / 9)
    V eps0 dsigma (9)], the force and stress parts only with a reference."""
    w_E, w_F = weights.get("energy", 1.0), weights.get("forces", 1.0)
    w_S = weights.get("stress", 0.1)
    r0, eps0 = float(weights.get("r0", 1.0)), float(weights.get("eps0", 1.0))
    out = []
    for s in structures:
        N = s["natoms"]
        vol = float(s.get("volume", abs(np.linalg.det(s["lattice"]))))
        e, f, st = tersoff_flat(s["lattice"], s["positions"], params, nthreads=nthreads)
        c = np.sqrt(float(s.get("weight", 1.0))) * float(s.get("scale", 1.0))
        out.append([c * np.sqrt(w_E) * N * (e / N - s["ref_energy"])])
        if s.get("ref_forces") is not None:
            out.append(c * np.sqrt(w_F / (3.0 * N)) * r0 * (f - s["ref_forces"]).ravel())
        if s.get("ref_stress") is not None:
            out.append(c * np.sqrt(w_S / 9.0) * vol * eps0 * (st - s["ref_stress"]).ravel())
    return np.concatenate(out) if out else np.zeros(0)


# ---------------------------------------------------------------------------
#  NVE velocity Verlet (solvers/molecular_dynamics.py:195-221, 1969-1996)
# ---------------------------------------------------------------------------

K_B = 1.380649e-23 * 6.241509074e18
EV_ANG_TO_AMU_ANG_FS2 = (1.60217663e-19 / 1.0e-10) / (1.66053906660e-27) * 1.0e-20


def mb_velocities(T, masses, seed):
    rng = np.random.default_rng(seed)
    N = len(masses)
    sigma = np.sqrt(K_B * T * EV_ANG_TO_AMU_ANG_FS2 / masses)
    v = rng.standard_normal((N, 3)) * sigma[:, None]
    v -= (np.sum(masses[:, None] * v, axis=0) / masses.sum())
    e_kin = 0.5 * float(np.sum(masses[:, None] * v ** 2)) / EV_ANG_TO_AMU_ANG_FS2
    n_dof = max(3 * N - 3, 1)
    T_act = 2.0 * e_kin / (n_dof * K_B)
    if T_act > 1e-12:
        v *= np.sqrt(T / T_act)
    return v


def nhc_half(state, dt, kT, n_dof, vel, masses):
    """One M-link chain half-step (_nhc_step, solvers/molecular_dynamics.py:
    635-727): top-link quarter kick; down the chain damp, kick, damp; scale
    the velocities by exp(-vxi_1 dt/2); advance xi; E_kin recomputed; up
    the chain with the same damping factors; top-link quarter kick.  state =
    (Q, xi, vxi) lists, changed in place; returns the scaled velocities."""
    Q, xi, v = state
    M = len(Q)
    ek = 0.5 * float(np.sum(masses[:, None] * vel ** 2)) / EV_ANG_TO_AMU_ANG_FS2

    def G(j):
        return (2.0 * ek - n_dof * kT) / Q[0] if j == 0 else (Q[j - 1] * v[j - 1] ** 2 - kT) / Q[j]

    v[M - 1] += G(M - 1) * (dt * 0.25)
    s = [1.0] * (M - 1)
    for j in range(M - 2, -1, -1):
        s[j] = float(np.exp(np.clip(-v[j + 1] * dt * 0.125, -350.0, 350.0)))
        v[j] = s[j] * (s[j] * v[j] + G(j) * dt * 0.25)
    vel = vel * float(np.exp(np.clip(-v[0] * dt * 0.5, -350.0, 350.0)))
    for j in range(M):
        xi[j] += v[j] * dt * 0.5
    ek = 0.5 * float(np.sum(masses[:, None] * vel ** 2)) / EV_ANG_TO_AMU_ANG_FS2
    for j in range(M - 1):
        v[j] = s[j] * (s[j] * v[j] + G(j) * dt * 0.25)
    v[M - 1] += G(M - 1) * (dt * 0.25)
    return vel


# Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11), the counter-based
# generator of the device's stochastic thermostats; numpy over uint64 lanes.
_PM0, _PM1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_PW0, _PW1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_M32 = np.uint64(0xFFFFFFFF)


def philox4x32(c, k0, k1):
    """c: (4, ...) counters; returns the (4, ...) output block."""
    c = [np.asarray(x, np.uint64) & _M32 for x in c]
    k0 = np.uint64(k0) & _M32
    k1 = np.uint64(k1) & _M32
    for _ in range(10):
        p0 = _PM0 * c[0]
        p1 = _PM1 * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & _M32, (p0 >> np.uint64(32)) ^ c[3] ^ k1,
             p0 & _M32]
        k0 = (k0 + _PW0) & _M32
        k1 = (k1 + _PW1) & _M32
    return c


def _u53(a, b):
    return ((a >> np.uint64(5)).astype(float) * 67108864.0
            + (b >> np.uint64(6)).astype(float)) * (1.0 / 9007199254740992.0)


def normal2(q):
    """Box-Muller of one Philox block: r = sqrt(-2 ln(1 - u1)), angle 2 pi u2."""
    u1, u2 = _u53(q[0], q[1]), _u53(q[2], q[3])
    r = np.sqrt(-2.0 * np.log(1.0 - u1))
    t = 6.283185307179586 * u2
    return r * np.cos(t), r * np.sin(t)


def gamma_mt(a, h, k0, k1):
    """Marsaglia-Tsang Gamma(a >= 1) with the device's draw order."""
    d = a - 1.0 / 3.0
    c = 1.0 / math.sqrt(9.0 * d)
    for it in range(64):
        x = float(normal2(philox4x32([2 * it + 1, h, 3, 0], k0, k1))[0])
        v = 1.0 + c * x
        if v <= 0.0:
            continue
        v = v * v * v
        q = philox4x32([2 * it + 2, h, 3, 0], k0, k1)
        u = float(_u53(q[0], q[1]))
        if u < 1.0 - 0.0331 * (x * x) * (x * x):
            return d * v
        if math.log(u) < 0.5 * x * x + d * (1.0 - v + math.log(v)):
            return d * v
    return d


def csvr_alpha(K, c, nf, kT, h, k0, k1):
    """CSVRSimulation._rescale (solvers/molecular_dynamics.py:1648-1667) with
    R_1 and S drawn from the Philox stream h: (alpha, alpha^2 K - K)."""
    if not K > 0.0 or nf < 1:
        return 1.0, 0.0
    Kt = 0.5 * nf * kT
    r1 = float(normal2(philox4x32([0, h, 3, 0], k0, k1))[0])
    sq = 2.0 * gamma_mt(0.5 * (nf - 1.0), h, k0, k1) if nf > 1 else 0.0
    Kn = c * K + (1.0 - c) * Kt * (r1 * r1 + sq) / nf + 2.0 * r1 * math.sqrt(
        c * (1.0 - c) * K * Kt / nf)
    alpha = math.sqrt(max(Kn, 0.0) / K)
    if r1 + math.sqrt(c * nf * K / ((1.0 - c) * Kt)) < 0.0:
        alpha = -alpha
    return alpha, alpha * alpha * K - K


def langevin_noise(n, step, k0, k1):
    """(n, 3) standard normals of the Langevin O step at `step`: Philox
    blocks (a, step, 1, 0) and (a, step, 1, 1), Box-Muller."""
    a = np.arange(n, dtype=np.uint64)
    z0, z1 = normal2(philox4x32([a, np.full(n, step), np.ones(n), np.zeros(n)], k0, k1))
    z2, _ = normal2(philox4x32([a, np.full(n, step), np.ones(n), np.ones(n)], k0, k1))
    return np.stack([z0, z1, z2], axis=1)


def md_run(force, lattice, positions, velocities, masses, dt, n_steps, r_cut, skin,
           nhc=None, flux=None, every=0, langevin=None, csvr=None, state_out=None,
           nhc_state=None, fix=None, rescale_steps=0):
    """Velocity Verlet with PBC wrap and the Verlet-skin rebuild test of
    engine/neighborlist.py:613-673 (fixed cell: rebuild when r_list - 2 d_max
    <= r_cut); ``force(lattice, pos, (i, j))`` returns (E, F, S).  nhc =
    (T, tau, M): NVTSimulation.run (solvers/molecular_dynamics.py:1113-1160),
    the step between two chain half-steps, Q_1 = n_dof kT tau^2, Q_j = kT
    tau^2, n_dof = 3N - 3.  langevin = (T, damping, seed): BAOAB
    (LangevinSimulation._step, :1552-1587); csvr = (T, damping, seed): R B A
    F B R (CSVRSimulation._step, :1645-1687); both with the Philox streams of
    the device.  flux(pos, vel) -> J of the state after steps s with s %
    every == 0.  Returns E_pot, E_kin, conserved quantity (E_tot, H_NHC
    (:947-969) or E_tot - Q), final positions and velocities, rebuilds, J
    frames; a dict passed as state_out receives the final chain positions
    "xi" and velocities "vxi"; nhc_state = (xi, vxi) starts the chain there
    (a checkpoint restart) instead of at rest.  fix (N, 3) bool: frozen
    components, zeroed at the start and after every kick, n_dof = 3N minus
    their number, no centre-of-mass removal in the Langevin O step
    (_build_fix_mask, _compute_n_dof :248-267, 341-361, _o_step :1552-1565);
    rescale_steps: with the chain, v *= sqrt(T / T_inst) after the
    end-of-step half-step of the first that many steps (:1154-1167)."""
    lattice = np.asarray(lattice, float)
    pos = np.array(positions, float)
    vel = np.array(velocities, float)
    inv = np.linalg.inv(lattice)
    r_list = r_cut + skin
    ref = pos.copy()
    pairs = nl_build(lattice, pos, r_list)[:2]
    e, f, _ = force(lattice, pos, pairs)
    acc = f * EV_ANG_TO_AMU_ANG_FS2 / masses[:, None]
    ek = lambda v: 0.5 * float(np.sum(masses[:, None] * v ** 2)) / EV_ANG_TO_AMU_ANG_FS2
    keep = None
    n_fixed = int(np.sum(fix)) if fix is not None else 0
    if n_fixed:
        keep = ~np.asarray(fix, dtype=bool)
        vel = vel * keep
    n_dof = max(3 * len(pos) - 3, 1) if n_fixed == 0 else max(3 * len(pos) - n_fixed, 1)
    mask = (lambda v: v * keep) if keep is not None else (lambda v: v)
    state = None
    heat = [0.0]
    if nhc is not None:
        T, tau, M = nhc
        kT = K_B * T
        state = ([n_dof * K_B * T * tau ** 2] + [K_B * T * tau ** 2] * (M - 1), [0.0] * M,
                 [0.0] * M)
        if nhc_state is not None:
            state = (state[0], [float(x) for x in nhc_state[0]],
                     [float(x) for x in nhc_state[1]])
    stoch = langevin if langevin is not None else csvr
    if stoch is not None:
        T, damping, seed = stoch
        kT = K_B * T
        k0, k1 = int(seed) & 0xFFFFFFFF, int(seed) >> 32
        M_tot = float(masses.sum())
    half = [0]

    def rescale(v):
        a, dq = csvr_alpha(ek(v), math.exp(-0.5 * dt / damping), n_dof, kT, half[0], k0, k1)
        half[0] += 1
        heat[0] += dq
        return v * a

    def wrap(p):
        f = p @ inv
        f -= np.floor(f)
        f[f >= 1.0] = 0.0
        return f @ lattice

    def vel_rescale(v):
        e_now = ek(v)
        T_now = 2.0 * e_now / (n_dof * K_B) if e_now > 0 else 0.0
        if T_now > 1e-12:
            v = mask(v * np.sqrt(T / T_now))
        return v

    def cons(e_pot, e_kin):
        if stoch is not None:
            return e_kin + e_pot - heat[0]
        if state is None:
            return e_pot + e_kin
        Q, xi, v = state
        t = 0.5 * Q[0] * v[0] ** 2
        for j in range(1, len(Q)):
            t = t + 0.5 * Q[j] * v[j] ** 2
        u = n_dof * kT * xi[0]
        for j in range(1, len(Q)):
            u = u + kT * xi[j]
        return e_kin + e_pot + t + u

    epot, ekin, econs, J = [e], [ek(vel)], [cons(e, ek(vel))], []
    rebuilds = 0
    for step in range(n_steps):
        if state is not None:
            vel = nhc_half(state, dt, kT, n_dof, vel, masses)
        if csvr is not None:
            vel = rescale(vel)
        if langevin is not None:
            c = math.exp(-dt / damping)
            vel = mask(vel + 0.5 * dt * acc)
            pos = pos + 0.5 * dt * vel
            before = ek(vel)
            sig = np.sqrt((1.0 - c * c) * kT * EV_ANG_TO_AMU_ANG_FS2 / masses)
            vel = c * vel + sig[:, None] * langevin_noise(len(pos), step, k0, k1)
            if keep is not None:
                vel = vel * keep
            else:
                vel = vel - np.sum(masses[:, None] * vel, axis=0) / M_tot
            heat[0] += ek(vel) - before
            pos = wrap(pos + 0.5 * dt * vel)
        else:
            vel = mask(vel + 0.5 * dt * acc)
            pos = wrap(pos + dt * vel)
        dr = pos - ref
        fr = dr @ inv
        fr -= np.round(fr)
        dmax = float(np.max(np.linalg.norm(fr @ lattice, axis=1)))
        if r_list - 2.0 * dmax <= r_cut:
            pairs = nl_build(lattice, pos, r_list)[:2]
            ref = pos.copy()
            rebuilds += 1
        e, f, _ = force(lattice, pos, pairs)
        acc = f * EV_ANG_TO_AMU_ANG_FS2 / masses[:, None]
        vel = mask(vel + 0.5 * dt * acc)
        if state is not None:
            vel = nhc_half(state, dt, kT, n_dof, vel, masses)
            if step < rescale_steps:
                vel = vel_rescale(vel)
        if csvr is not None:
            vel = rescale(vel)
        epot.append(e)
        ekin.append(ek(vel))
        econs.append(cons(e, ek(vel)))
        if flux is not None and every > 0 and step % every == 0:
            J.append(flux(pos, vel))
    if state_out is not None and state is not None:
        state_out.update(xi=list(state[1]), vxi=list(state[2]))
    return (np.array(epot), np.array(ekin), np.array(econs), pos, vel, rebuilds,
            np.array(J).reshape(-1, 3))


def cell_generators(mode):
    """Strain generators E_k of an NPT cell mode (_cell_generators,
    solvers/molecular_dynamics.py:2116-2138): a diagonal unit lets one axis
    breathe, e_ab + e_ba is a shear, the sum of two diagonal units couples
    those axes."""
    def unit(a, b):
        e = np.zeros((3, 3))
        e[a, b] = e[b, a] = 1.0
        return e
    xx, yy, zz = unit(0, 0), unit(1, 1), unit(2, 2)
    table = {"aniso": [xx, yy, zz], "tri": [xx, yy, zz, unit(1, 2), unit(0, 2), unit(0, 1)],
             "xy": [xx, yy], "xy_iso": [xx + yy], "z": [zz]}
    if mode not in table:
        raise ValueError(f"cell_dof must be one of iso | {' | '.join(table)}, got {mode!r}")
    return np.array(table[mode])


def baro_noise(n, draw, k0, k1):
    """n standard normals of barostat draw ``draw``: Philox blocks
    (k, draw, 5, 0), Box-Muller, first normal."""
    k = np.arange(n, dtype=np.uint64)
    z0, _ = normal2(philox4x32([k, np.full(n, draw), np.full(n, 5), np.zeros(n)], k0, k1))
    return np.asarray(z0, float)


def npt_run(force, lattice, positions, velocities, masses, dt, n_steps, r_cut, skin, T, tau, M,
            p_gpa, tau_baro, start=None, cell_dof="iso", baro="nhc", seed=0, generators=None,
            fix=None, rescale_steps=0):
    """NPTSimulation.run (solvers/molecular_dynamics.py:2871-2895): particle
    chain and barostat thermostat half-steps, barostat half-step, half kick,
    drift with the cell scaling, forces, half kick, barostat half-step at the
    new pressure, barostat thermostat, particle chain.

    cell_dof "iso": p_eps += dt/2 [3 V (P_inst - P) + (3/n_dof) 2 E_kin] with
    P_inst of the velocities before the particle chain's half-step, the cell
    and positions scaled by exp(eps_dot dt) and the velocities by
    exp(-(1 + 3/n_dof) eps_dot dt) (_barostat_half_step, _scale_step
    :2496-2555).  Other modes (ADR-0036, :2566-2692): one momentum p_k per
    strain generator E_k (``cell_generators``, or ``generators`` given),
    p_k += dt/2 [V Tr(E_k (P_int - P I)) + Tr(E_k) 2 E_kin / n_dof] with the
    pressure tensor P_int = sum m v v^T / V + sym(sigma) of the current
    velocities, G = sum_k (p_k / W) E_k, L -> L expm(G dt), r -> r expm(G dt),
    v -> v expm(-(G + Tr(G)/n_dof) dt) (elementwise for a diagonal G,
    through eigh otherwise).  baro "nhc": a 2-link chain on the barostat
    momenta (Q1 = n_k kT tau_b^2, target n_k kT); "langevin": p <- c p +
    sqrt((1 - c^2) W kT) xi, c = exp(-dt / (2 tau_b)), its heat subtracted
    from H_NPT (_baro_langevin :2596-2617), xi from the Philox stream of the
    device (``baro_noise``, keyed by seed).  The list follows
    check_and_rebuild with cell changes (engine/neighborlist.py:613-677):
    sigma_min(inv(L_b) L), 1 when L is allclose to L_b, and the non-affine
    displacement from the build positions mapped by inv(L_b) L.
    ``force(lattice, pos, pairs)`` -> (E, F, sigma).  Returns E_pot, E_kin,
    H_NPT (:2425-2454, 2648-2657), V, P_inst [GPa], the cell ("lat") and the
    pressure tensor [GPa] ("ptensor", _pressure_tensor :2570-2577) per step, final positions,
    velocities and lattice, rebuilds, the final particle chain ("xi", "vxi")
    and barostat state ("p_eps" (iso) or "p_g", "xi_b", "vxi_b", "heat"); a
    dict with those keys passed as ``start`` begins the run from that state
    (a checkpoint restart).  fix and rescale_steps as in ``md_run``
    (:2360-2373, 2886-2920)."""
    lattice = np.array(lattice, float)
    pos = np.array(positions, float)
    vel = np.array(velocities, float)
    r_list = r_cut + skin
    n_fixed = int(np.sum(fix)) if fix is not None else 0
    keep = ~np.asarray(fix, dtype=bool) if n_fixed else None
    mask = (lambda v: v * keep) if keep is not None else (lambda v: v)
    vel = mask(vel)
    n_dof = max(3 * len(pos) - 3, 1) if n_fixed == 0 else max(3 * len(pos) - n_fixed, 1)
    kT = K_B * T
    ek = lambda v: 0.5 * float(np.sum(masses[:, None] * v ** 2)) / EV_ANG_TO_AMU_ANG_FS2
    state = ([n_dof * kT * tau ** 2] + [kT * tau ** 2] * (M - 1), [0.0] * M, [0.0] * M)
    W = (n_dof + 3) * kT * tau_baro ** 2
    iso = cell_dof == "iso" and generators is None
    if iso:
        gens, n_b = None, 1
    else:
        gens = np.asarray(generators, float) if generators is not None else \
            cell_generators(cell_dof)
        n_b = len(gens)
        gtr = np.trace(gens, axis1=1, axis2=2)
    Qb1, Qb2 = n_b * kT * tau_baro ** 2, kT * tau_baro ** 2
    P_ext = p_gpa / 160.2176621
    if baro not in ("nhc", "langevin"):
        raise ValueError("baro must be nhc | langevin")
    b = {"p": 0.0, "xi1": 0.0, "xi2": 0.0, "v1": 0.0, "v2": 0.0}
    pg = np.zeros(n_b)
    heat = [0.0]
    draw = [0]
    k0, k1 = int(seed) & 0xFFFFFFFF, int(seed) >> 32
    if start is not None:
        state = (state[0], [float(x) for x in start["xi"]], [float(x) for x in start["vxi"]])
        b = {"p": float(start.get("p_eps", 0.0)), "xi1": float(start["xi_b"][0]),
             "xi2": float(start["xi_b"][1]), "v1": float(start["vxi_b"][0]),
             "v2": float(start["vxi_b"][1])}
        if not iso:
            pg = np.array(start["p_g"], float)
        heat[0] = float(start.get("heat", 0.0))
        draw[0] = int(start.get("draws", 0))
    vol = lambda L: float(abs(np.linalg.det(L)))

    def p_inst(v, S, L):
        return 2.0 * ek(v) / (3.0 * vol(L)) + np.trace(S) / 3.0

    def baro_nhc():
        # _nhc_baro_step / _nhc_baro_step_g (:2456-2494, 2619-2646)
        pb = np.array([b["p"]]) if iso else pg
        e1 = 0.5 * float(np.sum(pb ** 2)) / W
        G1 = (2.0 * e1 - n_b * kT) / Qb1
        G2 = (Qb1 * b["v1"] ** 2 - kT) / Qb2
        b["v2"] += G2 * (dt * 0.25)
        sc = float(np.exp(np.clip(-b["v2"] * dt * 0.125, -350.0, 350.0)))
        b["v1"] = sc * (sc * b["v1"] + G1 * dt * 0.25)
        f = float(np.exp(np.clip(-b["v1"] * dt * 0.5, -350.0, 350.0)))
        if iso:
            b["p"] *= f
            pb = np.array([b["p"]])
        else:
            pg[:] = pg * f
            pb = pg
        b["xi1"] += b["v1"] * dt * 0.5
        b["xi2"] += b["v2"] * dt * 0.5
        e1 = 0.5 * float(np.sum(pb ** 2)) / W
        G1 = (2.0 * e1 - n_b * kT) / Qb1
        b["v1"] = sc * (sc * b["v1"] + G1 * dt * 0.25)
        G2 = (Qb1 * b["v1"] ** 2 - kT) / Qb2
        b["v2"] += G2 * (dt * 0.25)

    def baro_lv():
        c = math.exp(-0.5 * dt / tau_baro)
        p = np.array([b["p"]]) if iso else pg.copy()
        new = c * p + math.sqrt((1.0 - c * c) * W * kT) * baro_noise(len(p), draw[0], k0, k1)
        draw[0] += 1
        heat[0] += 0.5 * float(new @ new - p @ p) / W
        if iso:
            b["p"] = float(new[0])
        else:
            pg[:] = new

    baro_thermo = baro_lv if baro == "langevin" else baro_nhc

    def baro_half_iso(P, v, L):
        b["p"] += 0.5 * dt * (3.0 * vol(L) * (P - P_ext) + (3.0 / n_dof) * 2.0 * ek(v))

    def baro_half_g(v, S, L):
        V = vol(L)
        kin = (v.T * masses) @ v / EV_ANG_TO_AMU_ANG_FS2
        D = kin / V + 0.5 * (S + S.T) - P_ext * np.eye(3)
        f = V * np.einsum("kab,ab->k", gens, D) + gtr * (2.0 * ek(v)) / n_dof
        pg[:] = pg + 0.5 * dt * f

    def wrap(p, L):
        fr = p @ np.linalg.inv(L)
        fr -= np.floor(fr)
        fr[fr >= 1.0] = 0.0
        return fr @ L

    def H(e, v, L):
        Q, xi, w = state
        t = 0.5 * Q[0] * w[0] ** 2
        for j in range(1, M):
            t = t + 0.5 * Q[j] * w[j] ** 2
        u = n_dof * kT * xi[0]
        for j in range(1, M):
            u = u + kT * xi[j]
        if iso:
            ed = b["p"] / W
            hb = 0.5 * Qb1 * b["v1"] ** 2 + 0.5 * Qb2 * b["v2"] ** 2 + kT * b["xi1"] + \
                kT * b["xi2"]
            return ek(v) + e + P_ext * vol(L) + 0.5 * W * ed ** 2 + t + u + hb - heat[0]
        hb = 0.5 * Qb1 * b["v1"] ** 2 + 0.5 * Qb2 * b["v2"] ** 2 + n_b * kT * b["xi1"] + \
            kT * b["xi2"]
        return (ek(v) + e + P_ext * vol(L) + 0.5 * float(np.sum(pg ** 2)) / W + (t + u) + hb) \
            - heat[0]

    ref = pos.copy()
    lat_b = lattice.copy()
    pairs = nl_build(lattice, pos, r_list)[:2]
    e, f, S = force(lattice, pos, pairs)
    acc = f * EV_ANG_TO_AMU_ANG_FS2 / masses[:, None]
    def ptensor(v, S, L):
        kin = (v.T * masses) @ v / EV_ANG_TO_AMU_ANG_FS2
        return (kin / vol(L) + 0.5 * (S + S.T)) * 160.2176621

    rec = {"ep": [e], "ek": [ek(vel)], "h": [H(e, vel, lattice)], "v": [vol(lattice)],
           "p": [p_inst(vel, S, lattice) * 160.2176621], "lat": [lattice.copy()],
           "ptensor": [ptensor(vel, S, lattice)]}
    rebuilds = 0
    for step in range(n_steps):
        P = p_inst(vel, S, lattice)
        vel = nhc_half(state, dt, kT, n_dof, vel, masses)
        baro_thermo()
        if iso:
            baro_half_iso(P, vel, lattice)
        else:
            baro_half_g(vel, S, lattice)
        vel = mask(vel + 0.5 * dt * acc)
        if iso:
            ed = b["p"] / W
            s = float(np.exp(ed * dt))
            pos = wrap(pos + dt * vel, lattice)
            fr = pos @ np.linalg.inv(lattice)
            lattice = lattice * s
            pos = fr @ lattice
            vel = vel * float(np.exp(-(1.0 + 3.0 / n_dof) * ed * dt))
        else:
            G = np.einsum("k,kab->ab", pg / W, gens)
            pos = wrap(pos + dt * vel, lattice)
            a = np.trace(G) / n_dof
            if not np.any(G - np.diag(np.diag(G))):
                g = np.diag(G)
                Mx = np.diag(np.exp(g * dt))
                Mv = np.diag(np.exp(-(g + a) * dt))
            else:
                w, U = np.linalg.eigh(G)
                Mx = (U * np.exp(w * dt)) @ U.T
                Mv = (U * np.exp(-(w + a) * dt)) @ U.T
            lattice = lattice @ Mx
            pos = pos @ Mx
            vel = vel @ Mv
        Mm = np.linalg.inv(lat_b) @ lattice
        if np.allclose(lattice, lat_b, atol=1e-12):
            sig = 1.0
        else:
            sig = float(np.linalg.svd(Mm, compute_uv=False)[-1])
        dr = pos - ref @ Mm
        fr = dr @ np.linalg.inv(lattice)
        fr -= np.round(fr)
        dmax = float(np.max(np.linalg.norm(fr @ lattice, axis=1)))
        if sig * r_list - 2.0 * dmax <= r_cut:
            pairs = nl_build(lattice, pos, r_list)[:2]
            ref = pos.copy()
            lat_b = lattice.copy()
            rebuilds += 1
        e, f, S = force(lattice, pos, pairs)
        acc = f * EV_ANG_TO_AMU_ANG_FS2 / masses[:, None]
        vel = mask(vel + 0.5 * dt * acc)
        Pn = p_inst(vel, S, lattice)
        if iso:
            baro_half_iso(Pn, vel, lattice)
        else:
            baro_half_g(vel, S, lattice)
        baro_thermo()
        vel = nhc_half(state, dt, kT, n_dof, vel, masses)
        if step < rescale_steps:
            e_now = ek(vel)
            T_now = 2.0 * e_now / (n_dof * K_B) if e_now > 0 else 0.0
            if T_now > 1e-12:
                vel = mask(vel * np.sqrt(T / T_now))
        rec["ep"].append(e)
        rec["ek"].append(ek(vel))
        rec["h"].append(H(e, vel, lattice))
        rec["v"].append(vol(lattice))
        rec["p"].append(Pn * 160.2176621)
        rec["lat"].append(lattice.copy())
        rec["ptensor"].append(ptensor(vel, S, lattice))
    out = {k: np.array(v) for k, v in rec.items()}
    out.update(positions=pos, velocities=vel, lattice=lattice, rebuilds=rebuilds,
               xi=list(state[1]), vxi=list(state[2]), p_eps=b["p"], p_g=pg.copy(),
               xi_b=[b["xi1"], b["xi2"]], vxi_b=[b["v1"], b["v2"]], heat=heat[0],
               draws=draw[0], stress=S)
    return out


def nve_run(force, lattice, positions, velocities, masses, dt, n_steps, r_cut, skin):
    """md_run without a thermostat: E_pot, E_kin, final state, rebuilds."""
    ep, ek, _, pos, vel, rb, _ = md_run(force, lattice, positions, velocities, masses, dt,
                                        n_steps, r_cut, skin)
    return ep, ek, pos, vel, rb

def bath_noise(n, draw, k0, k1):
    """(n, 3) standard normals of the NEMD bath update at random draw
    `draw`: Philox blocks (a, draw, 7, 0) and (a, draw, 7, 1), Box-Muller."""
    a = np.arange(n, dtype=np.uint64)
    z0, z1 = normal2(philox4x32([a, np.full(n, draw), np.full(n, 7), np.zeros(n)], k0, k1))
    z2, _ = normal2(philox4x32([a, np.full(n, draw), np.full(n, 7), np.ones(n)], k0, k1))
    return np.stack([z0, z1, z2], axis=1)


def nemd_run(force, lattice, positions, velocities, masses, dt, n_steps, r_cut, skin, bath,
             T_hot, T_cold, damping, seed, draw0=0, heat0=0.0, sample=None, every=0, phase=0):
    """NEMDSimulation.run's step (solvers/nemd.py:184-193): velocity Verlet
    (with the wrap and skin-test rebuild of md_run), then the exact
    Ornstein-Uhlenbeck update of the hot bath's atoms and then the cold
    bath's, v <- c v + sqrt((1 - c^2) k_B T EV_ANG_TO_AMU_ANG_FS2 / m) xi,
    c = exp(-dt / damping) (_o_step :113-117), the heat of each bath its
    kinetic-energy change; xi from the Philox stream of the device (draw0 +
    step).  bath (N,): 0 none, 1 hot, 2 cold.  sample(pos, vel) after the
    steps s with (s + phase) % every == 0.  Returns a dict: ep, ek, econs
    (E_tot - Q) and rows (dQ_hot, dQ_cold, K_hot, K_cold) per record
    (n_steps + 1, record 0 the initial state), positions, velocities,
    rebuilds, samples."""
    lattice = np.asarray(lattice, float)
    pos = np.array(positions, float)
    vel = np.array(velocities, float)
    bath = np.asarray(bath)
    inv = np.linalg.inv(lattice)
    r_list = r_cut + skin
    ref = pos.copy()
    pairs = nl_build(lattice, pos, r_list)[:2]
    e, f, _ = force(lattice, pos, pairs)
    conv = EV_ANG_TO_AMU_ANG_FS2
    acc = f * conv / masses[:, None]
    ek = lambda v, idx=slice(None): 0.5 * float(np.sum(masses[idx, None] * v[idx] ** 2)) / conv
    hot, cold = np.flatnonzero(bath == 1), np.flatnonzero(bath == 2)
    k0, k1 = int(seed) & 0xFFFFFFFF, int(seed) >> 32
    c = math.exp(-dt / damping)
    Q = float(heat0)
    ep, ekin, ec = [e], [ek(vel)], [e + ek(vel) - Q]
    rows = [(0.0, 0.0, ek(vel, hot), ek(vel, cold))]
    samples = []
    rebuilds = 0

    def wrap(p):
        fr = p @ inv
        fr -= np.floor(fr)
        fr[fr >= 1.0] = 0.0
        return fr @ lattice

    for step in range(n_steps):
        vel = vel + 0.5 * dt * acc
        pos = wrap(pos + dt * vel)
        dr = (pos - ref) @ inv
        dr -= np.round(dr)
        if r_list - 2.0 * float(np.max(np.linalg.norm(dr @ lattice, axis=1))) <= r_cut:
            pairs = nl_build(lattice, pos, r_list)[:2]
            ref = pos.copy()
            rebuilds += 1
        e, f, _ = force(lattice, pos, pairs)
        acc = f * conv / masses[:, None]
        vel = vel + 0.5 * dt * acc
        z = bath_noise(len(pos), draw0 + step, k0, k1)
        dq = []
        for idx, T in ((hot, T_hot), (cold, T_cold)):
            before = ek(vel, idx)
            sig = np.sqrt((1.0 - c * c) * K_B * T * conv / masses[idx])
            vel[idx] = c * vel[idx] + sig[:, None] * z[idx]
            dq.append(ek(vel, idx) - before)
        Q += dq[0] + dq[1]
        ep.append(e)
        ekin.append(ek(vel))
        ec.append(e + ek(vel) - Q)
        rows.append((dq[0], dq[1], ek(vel, hot), ek(vel, cold)))
        if sample is not None and every > 0 and (step + phase) % every == 0:
            samples.append(sample(pos, vel))
    return {"ep": np.array(ep), "ek": np.array(ekin), "econs": np.array(ec),
            "rows": np.array(rows), "positions": pos, "velocities": vel, "rebuilds": rebuilds,
            "samples": samples, "heat": Q}
