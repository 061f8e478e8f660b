"""Hand-written backward kernel of the Tersoff fitting loss against the
torch-autograd FP64 reference (all 14 parameters, including the mixed
d2E/dtheta dr terms through forces and stress) and central finite
differences of the oracle loss."""
import numpy as np
import pytest

import torch_tersoff as TT
from _fitdata import make_set, oracle_loss
from akitsushima_amd.fit import TersoffFit, tersoff_loss

pytestmark = pytest.mark.gpu


def _gpu(theta, data):
    fit = TersoffFit(data)
    return fit, fit.loss_and_grad(theta)


@pytest.mark.parametrize("seed,lambda3,gamma", [(5, 1.3, 1.05), (9, 0.0, 1.0), (13, 1.7322, 0.95)])
def test_loss_and_gradient_match_torch_reference(seed, lambda3, gamma):
    theta, data = make_set(n_struct=3, seed=seed, lambda3=lambda3, gamma=gamma)
    fit, (L, g) = _gpu(theta, data)
    L0 = oracle_loss(theta, data)
    assert abs(L - L0) <= 1e-10 * abs(L0)
    Lt, gt = TT.loss_and_grad(theta, data)
    scale = np.abs(gt).max()
    for p in range(14):
        if lambda3 == 0.0 and p == 12:
            continue  # d/dm of 0^m is undefined in the autograd reference
        assert abs(g[p] - gt[p]) <= 1e-9 * max(abs(gt[p]), 1e-6 * scale), (TT.KEYS[p], g[p], gt[p])


def test_gradient_with_a_linear_exponent_m():
    """m = 1 takes the multiplied-out power in the rows of the other
    parameters and the general power in the row of m itself."""
    theta, data = make_set(n_struct=2, seed=17, lambda3=1.1, gamma=1.0)
    theta = theta.copy()
    theta[12] = 1.0
    fit, (L, g) = _gpu(theta, data)
    L0 = oracle_loss(theta, data)
    assert abs(L - L0) <= 1e-10 * abs(L0)
    Lt, gt = TT.loss_and_grad(theta, data)
    scale = np.abs(gt).max()
    for p in range(14):
        assert abs(g[p] - gt[p]) <= 1e-9 * max(abs(gt[p]), 1e-6 * scale), (TT.KEYS[p], g[p], gt[p])


def test_gradient_matches_central_differences():
    theta, data = make_set(n_struct=2, seed=21)
    _, (L, g) = _gpu(theta, data)
    for p in (0, 3, 4, 6, 8, 9, 10, 11, 13):
        h = 1e-6 * abs(theta[p])
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (oracle_loss(tp, data) - oracle_loss(tm, data)) / (2 * h)
        assert abs(g[p] - fd) <= 2e-7 * abs(fd) + 1e-9 * np.abs(g).max(), (TT.KEYS[p], g[p], fd)


def test_batched_gradient_is_sum_of_single_structures():
    theta, data = make_set(n_struct=4, seed=3)
    _, (L, g) = _gpu(theta, data)
    Ls, gs = 0.0, np.zeros(14)
    for d in data:
        _, (l1, g1) = _gpu(theta, [d])
        Ls += l1
        gs += g1
    assert abs(L - Ls) <= 1e-12 * abs(L)
    assert np.abs(g - gs).max() <= 1e-12 * np.abs(g).max()


def test_autograd_function_backward_is_engine_gradient():
    import torch

    theta, data = make_set(n_struct=2, seed=6)
    fit = TersoffFit(data)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    L = tersoff_loss(th, fit)
    L.backward()
    _, g = fit.loss_and_grad(theta)
    assert np.array_equal(th.grad.numpy(), g)
    assert abs(float(L.detach()) - fit.loss(theta)) == 0.0


def test_replayed_evaluations_equal_fresh_ones():
    """From the second evaluation of one shape on (same cutoff, weights and
    outputs), the device work replays a captured graph with the new
    parameters copied in: every replay equals a fresh, uncaptured evaluation
    bit for bit (R and D held, so every fit builds the same list; the launch
    shape fixed, since each fit's autotuner may pick another kernel form,
    which rounds differently).  A change of R moves the cutoff: other
    launches, evaluated and captured anew."""
    theta, data = make_set(n_struct=3, seed=29, lambda3=1.2, gamma=1.0)
    rng = np.random.default_rng(30)
    thetas = []
    for _ in range(4):
        t = theta * (1 + rng.uniform(-0.02, 0.02, 14))
        t[10:13] = theta[10:13]  # R, D, m held
        thetas.append(t)
    ref = [TersoffFit(data, autotune=False).loss_and_grad(t) for t in thetas]
    ref_loss = [TersoffFit(data, autotune=False).loss(t) for t in thetas]
    fit = TersoffFit(data, autotune=False)
    for k in list(range(4)) * 2:  # eager, capture, then replays
        L, g = fit.loss_and_grad(thetas[k])
        assert L == ref[k][0] and np.array_equal(g, ref[k][1]), k
    for k in (0, 1, 2):  # the loss alone: a graph of its own
        assert fit.loss(thetas[k]) == ref_loss[k]
    moved = thetas[0].copy()
    moved[10] *= 1.01
    La, ga = fit.loss_and_grad(moved)
    Lb, gb = fit.loss_and_grad(moved)  # captured at the new cutoff
    assert La == Lb and np.array_equal(ga, gb)
    for k in (1, 2, 3):  # back at the first cutoff
        L, g = fit.loss_and_grad(thetas[k])
        assert L == ref[k][0] and np.array_equal(g, ref[k][1]), k
