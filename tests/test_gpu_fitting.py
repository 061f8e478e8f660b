"""The reference's fitting interface on the device (akitsushima_amd.fitting;
fitting/parameter_sets.py:899-1016, fitting/loss.py:51-319) against the
oracle's restatement of the loss: loss and residuals at 1e-10 (squared and
Huber, per-structure weight and scale, structures without reference forces
or stress), the x-gradient against central differences of the oracle loss
through the same decode, sum(r^2) = loss, the 1e30 / 1e15 failure values,
the reference's own loss tests (tests/test_fitting.py:133-230), and scipy
L-BFGS-B with the device gradient converging on a small fit set."""
import numpy as np
import pytest

import oracle as O
from akitsushima_amd import (TersoffParameterSet, generic_loss_and_grad, generic_loss_function,
                             generic_residuals, residual_jacobian, residual_length)
from akitsushima_amd.structures import diamond_batch

pytestmark = pytest.mark.gpu

T3 = {"A": 1830.8, "B": 471.18, "lambda1": 2.4799, "lambda2": 1.7322, "beta": 1.0999e-6,
      "n": 0.78734, "c": 100390.0, "d": 16.217, "h": -0.59825}
W = {"energy": 1.0, "forces": 1.0, "stress": 0.1}


def _data(n=4, seed=9, max_atoms=128, rattle=0.15):
    ps = TersoffParameterSet(T3, species=["Si"])
    x_true = ps.x0 + np.random.default_rng(seed).uniform(-0.04, 0.04, len(ps.x0))
    target = ps.decode(x_true)["tersoff"]
    out = []
    for lat, pos, t in diamond_batch(n, 64, max_atoms, 0.03, rattle, seed=seed,
                                     rep_range=(2, 3)):
        e, f, s = O.tersoff_flat(lat, pos, target)
        out.append({"lattice": lat, "positions": pos, "atom_types": t, "natoms": len(pos),
                    "ref_energy": e / len(pos), "ref_forces": f, "ref_stress": s})
    return ps, x_true, out


def _mixed(structs):
    s = [dict(x) for x in structs]
    s[0]["weight"] = 2.5
    s[1]["scale"] = 0.7
    s[1]["ref_forces"] = None
    s[2]["ref_stress"] = None
    s[2]["volume"] = 1.1 * abs(np.linalg.det(s[2]["lattice"]))
    return s


@pytest.mark.parametrize("loss_type", ["squared", "huber"])
def test_device_loss_matches_the_oracle(loss_type):
    ps, _, structs = _data()
    structs = _mixed(structs)
    w = dict(W, huber_delta_energy=3.0, huber_delta_forces=0.05, huber_delta_stress=0.02)
    x = ps.x0
    L = generic_loss_function(x, ps, structs, w, r_cut=3.0, ewald_alpha=None,
                              loss_type=loss_type)
    L0 = O.generic_loss(ps.decode(x)["tersoff"], structs, w, loss_type)
    assert abs(L - L0) <= 1e-10 * abs(L0), (L, L0)


@pytest.mark.parametrize("loss_type", ["squared", "huber"])
def test_x_gradient_matches_central_differences_of_the_oracle(loss_type):
    ps, _, structs = _data(n=3, seed=3)
    structs = _mixed(structs)
    w = dict(W, huber_delta_energy=3.0, huber_delta_forces=0.05, huber_delta_stress=0.02)
    x = ps.x0 + 0.01
    L, g = generic_loss_and_grad(x, ps, structs, w, loss_type=loss_type)
    ref = lambda xx: O.generic_loss(ps.decode(xx)["tersoff"], structs, w, loss_type)
    assert abs(L - ref(x)) <= 1e-10 * abs(L)
    fd = np.zeros_like(x)
    for i in range(len(x)):
        h = 1e-6 * max(1.0, abs(x[i]))
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (ref(xp) - ref(xm)) / (2 * h)
    scale = np.abs(fd).max()
    assert np.abs(g - fd).max() <= 1e-5 * scale, (g, fd)


def test_residuals_match_the_oracle_and_sum_to_the_loss():
    ps, _, structs = _data()
    structs = _mixed(structs)
    x = ps.x0
    r = generic_residuals(x, ps, structs, W, r_cut=3.0, ewald_alpha=None)
    assert r.shape == (residual_length(structs),)
    r0 = O.generic_residuals(ps.decode(x)["tersoff"], structs, W)
    assert np.abs(r - r0).max() <= 1e-10 * np.abs(r0).max()
    L = generic_loss_function(x, ps, structs, W)
    assert abs(r @ r - L) <= 1e-12 * L


def test_failed_evaluation_values():
    ps, _, structs = _data(n=2)
    bad = TersoffParameterSet(T3, ["Si"], fix_params={"R": 6.0})   # r_c beyond the cells
    assert generic_loss_function(bad.x0, bad, structs, W) == 1e30
    L, g = generic_loss_and_grad(bad.x0, bad, structs, W)
    assert L == 1e30 and not np.any(g)
    r = generic_residuals(bad.x0, bad, structs, W)
    assert np.all(r == 1e15) and len(r) == residual_length(structs)
    assert generic_loss_function(ps.x0 + np.inf, ps, structs, W) == 1e30


def test_the_references_loss_tests_on_the_device():
    ps, _, structs = _data(n=4, seed=12)
    x = ps.x0
    one = structs[:1]
    base = generic_loss_function(x, ps, one, W)
    assert base > 0.0
    assert generic_loss_function(x, ps, [dict(one[0], scale=2.0)], W) == \
        pytest.approx(4.0 * base, rel=1e-10)
    assert generic_loss_function(x, ps, [dict(one[0], weight=2.0)], W) == \
        pytest.approx(2.0 * base, rel=1e-10)
    with pytest.raises(ValueError, match="huber_delta"):
        generic_loss_function(x, ps, structs, W, loss_type="huber")
    out = structs + [dict(structs[0], ref_energy=structs[0]["ref_energy"] + 50.0)]
    sq = generic_loss_function(x, ps, out, W)
    half = dict(W, huber_delta_energy=0.5, huber_delta_forces=0.5, huber_delta_stress=0.5)
    assert generic_loss_function(x, ps, out, half, loss_type="huber") < sq
    huge = dict(W, huber_delta_energy=1e6, huber_delta_forces=1e6, huber_delta_stress=1e6)
    assert generic_loss_function(x, ps, out, huge, loss_type="huber") == \
        pytest.approx(sq, rel=1e-8)


def test_lbfgsb_with_the_device_gradient_converges():
    from scipy.optimize import minimize

    ps, x_true, structs = _data(n=8, seed=21, rattle=0.1)
    f = lambda x: generic_loss_and_grad(x, ps, structs, W)
    L0 = f(ps.x0)[0]
    res = minimize(f, ps.x0, jac=True, method="L-BFGS-B", bounds=ps.bounds,
                   options={"maxiter": 500, "ftol": 1e-15, "gtol": 1e-10})
    print(f"L-BFGS-B: {res.nit} iterations, loss {L0:.3e} -> {res.fun:.3e}", flush=True)
    assert res.fun < 1e-8 * L0
    # the pair terms are well determined; c and d trade off against each
    # other in g(theta) over the sampled angles
    p, pt = ps.decode(res.x)["tersoff"], ps.decode(x_true)["tersoff"]
    for k in ("A", "B", "lambda1", "lambda2"):
        assert abs(p[k] - pt[k]) <= 1e-3 * abs(pt[k]), k


def test_least_squares_jacobian_matches_the_gradient():
    # J^T r = grad / 2 for the squared loss
    ps, _, structs = _data(n=3, seed=5)
    x = ps.x0 + 0.02
    r = generic_residuals(x, ps, structs, W)
    J = residual_jacobian(x, ps, structs, W, r0=r)
    L, g = generic_loss_and_grad(x, ps, structs, W)
    assert np.abs(2.0 * J.T @ r - g).max() <= 1e-4 * np.abs(g).max()
