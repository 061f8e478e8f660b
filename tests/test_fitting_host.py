"""Host side of the fitting interface (no GPU): TersoffParameterSet's decode
(log keys, the lambda3 := lambda2 tie, fixed R and D, fit keys, bounds) and
the chain rule of its x-gradient against central differences; the oracle's
restatement of generic_loss_function / generic_residuals against the
reference's own loss tests (fitting/loss.py; tests/test_fitting.py:133-230):
a scale of 2 quadruples a structure's term, a weight of 2 doubles it, a
Huber delta above every residual gives the squared loss, a zero delta gives
zero, and the residuals' sum of squares is the loss."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import TersoffParameterSet, residual_length
from akitsushima_amd.fit import THETA_KEYS
from akitsushima_amd.fitting import _check_loss_type
from akitsushima_amd.structures import diamond_batch

T3 = {"A": 1830.8, "B": 471.18, "lambda1": 2.4799, "lambda2": 1.7322, "beta": 1.0999e-6,
      "n": 0.78734, "c": 100390.0, "d": 16.217, "h": -0.59825}


def test_decode_follows_the_reference_semantics():
    ps = TersoffParameterSet(T3, species=["Si"])
    assert ps.fit_keys == ["A", "B", "lambda1", "lambda2", "beta", "n", "c", "d", "h"]
    x0 = ps.x0
    assert np.isclose(x0[0], np.log(1830.8)) and x0[-1] == -0.59825
    p = ps.decode(x0)["tersoff"]
    assert set(p) == set(T3) | {"lambda3", "R", "D"}
    assert p["R"] == 2.85 and p["D"] == 0.15
    assert p["lambda3"] == p["lambda2"] and np.isclose(p["A"], 1830.8, rtol=1e-15)
    assert ps.bounds[0] == (np.log(1.0), np.log(1e5)) and ps.bounds[-1] == (-1.0, 0.5)
    # fixed lambda3, and a fit key that is fixed is dropped
    ps2 = TersoffParameterSet(T3, ["Si"], fix_params={"lambda3": 0.0, "h": -0.5},
                              fit_keys=["A", "h", "n"])
    assert ps2.fit_keys == ["A", "n"]
    p2 = ps2.decode(ps2.x0)["tersoff"]
    assert p2["lambda3"] == 0.0 and p2["h"] == -0.5
    th = ps.theta(x0)
    assert th[THETA_KEYS.index("m")] == 3.0 and th[THETA_KEYS.index("gamma")] == 1.0
    assert th[THETA_KEYS.index("lambda3")] == th[THETA_KEYS.index("lambda2")]


@pytest.mark.parametrize("fix", [None, {"lambda3": 1.1}])
def test_x_gradient_chain_rule_matches_central_differences(fix):
    ps = TersoffParameterSet(T3, ["Si"], fix_params=fix)
    a = np.linspace(0.3, 1.7, 14)

    def L(theta):  # a smooth function of every engine parameter
        return float(np.sum(a * np.sin(theta / (1.0 + np.abs(theta)))))

    def dL(theta):
        u = theta / (1.0 + np.abs(theta))
        du = 1.0 / (1.0 + np.abs(theta)) ** 2
        return a * np.cos(u) * du

    x = ps.x0 + 0.01
    g = ps.grad_x(x, dL(ps.theta(x)))
    fd = np.zeros_like(x)
    for i in range(len(x)):
        h = 1e-6 * max(1.0, abs(x[i]))
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (L(ps.theta(xp)) - L(ps.theta(xm))) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-7, atol=1e-10)


def test_loss_argument_checks_and_residual_length():
    with pytest.raises(ValueError, match="unknown loss_type"):
        _check_loss_type("l1", {})
    with pytest.raises(ValueError, match="huber_delta"):
        _check_loss_type("huber", {"huber_delta_energy": 1.0})
    structs = [{"positions": np.zeros((5, 3)), "ref_forces": np.zeros((5, 3)),
                "ref_stress": np.zeros((3, 3))},
               {"positions": np.zeros((7, 3)), "ref_forces": None},
               {"positions": np.zeros((2, 3)), "ref_stress": np.eye(3)}]
    assert residual_length(structs) == (1 + 15 + 9) + 1 + (1 + 9)


def _oracle_set(n=3, seed=4):
    params = dict(T3, lambda3=T3["lambda2"], R=2.85, D=0.15)
    target = dict(params, A=1.03 * params["A"], c=0.97 * params["c"])
    out = []
    for lat, pos, t in diamond_batch(n, 64, 64, 0.02, 0.1, seed=seed, rep_range=(2, 2)):
        e, f, s = O.tersoff_flat(lat, pos, target)
        out.append({"lattice": lat, "positions": pos, "atom_types": t, "natoms": len(pos),
                    "ref_energy": e / len(pos), "ref_forces": f, "ref_stress": s})
    return params, out


def test_oracle_loss_matches_the_references_loss_tests():
    params, structs = _oracle_set()
    w = {"energy": 1.0, "forces": 1.0, "stress": 0.1}
    base = O.generic_loss(params, structs[:1], w)
    assert base > 0.0
    assert np.isclose(O.generic_loss(params, [dict(structs[0], scale=2.0)], w), 4.0 * base,
                      rtol=1e-12)
    assert np.isclose(O.generic_loss(params, [dict(structs[0], weight=2.0)], w), 2.0 * base,
                      rtol=1e-12)
    huge = dict(w, huber_delta_energy=1e6, huber_delta_forces=1e6, huber_delta_stress=1e6)
    full = O.generic_loss(params, structs, w)
    assert O.generic_loss(params, structs, huge, "huber") == pytest.approx(full, rel=1e-12)
    zero = dict(w, huber_delta_energy=0.0, huber_delta_forces=0.0, huber_delta_stress=0.0)
    assert O.generic_loss(params, structs, zero, "huber") == 0.0
    # an outlier's influence is tempered
    out = structs + [dict(structs[0], ref_energy=structs[0]["ref_energy"] + 50.0)]
    half = dict(w, huber_delta_energy=0.5, huber_delta_forces=0.5, huber_delta_stress=0.5)
    assert O.generic_loss(params, out, half, "huber") < O.generic_loss(params, out, w)
    # the residuals' sum of squares is the loss, missing references included
    mixed = [structs[0], dict(structs[1], ref_forces=None), dict(structs[2], ref_stress=None,
                                                               weight=0.5, scale=1.5)]
    r = O.generic_residuals(params, mixed, w)
    assert len(r) == residual_length(mixed)
    assert np.isclose(r @ r, O.generic_loss(params, mixed, w), rtol=1e-12)


def test_oracle_huber_terms_against_closed_forms():
    """Energy term (N dE/N)^2 and force term r0^2 mean(dF^2) with known
    offsets, through both branches of rho(q) = q (sqrt q <= delta) and
    2 delta sqrt(q) - delta^2 (fitting/loss.py:27-48, 212-227)."""
    params, structs = _oracle_set(n=1)
    s = structs[0]
    N = s["natoms"]
    e, f, _ = O.tersoff_flat(s["lattice"], s["positions"], params)
    d_e = 0.25 / N  # so that N dE = 0.25 eV
    s_e = dict(s, ref_energy=e / N + d_e, ref_forces=None, ref_stress=None)
    w = {"energy": 1.0, "forces": 0.0, "stress": 0.0}
    assert O.generic_loss(params, [s_e], w) == pytest.approx(0.0625, rel=1e-9)
    hub = dict(w, huber_delta_energy=0.1, huber_delta_forces=1.0, huber_delta_stress=1.0)
    # r = 0.25 > delta = 0.1: 2 (0.1)(0.25) - 0.01 = 0.04
    assert O.generic_loss(params, [s_e], hub, "huber") == pytest.approx(0.04, rel=1e-9)
    hub_in = dict(hub, huber_delta_energy=0.3)  # r <= delta: the square
    assert O.generic_loss(params, [s_e], hub_in, "huber") == pytest.approx(0.0625, rel=1e-9)
    # forces offset by 0.2 eV/A in every component: q = r0^2 0.04, r0 = 1.5
    s_f = dict(s, ref_energy=e / N, ref_forces=f + 0.2, ref_stress=None)
    wf = {"energy": 0.0, "forces": 2.0, "stress": 0.0, "r0": 1.5}
    q = 1.5 ** 2 * 0.04
    assert O.generic_loss(params, [s_f], wf) == pytest.approx(2.0 * q, rel=1e-9)
    hf = dict(wf, huber_delta_energy=1.0, huber_delta_forces=0.1, huber_delta_stress=1.0)
    r = np.sqrt(q)
    assert O.generic_loss(params, [s_f], hf, "huber") == pytest.approx(
        2.0 * (2 * 0.1 * r - 0.01), rel=1e-9)
