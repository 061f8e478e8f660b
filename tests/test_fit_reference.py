"""The fitting-gradient references agree with each other on the CPU: the
torch-autograd FP64 reference loss equals the oracle loss
(fitting/loss.py:212-242), and its dL/dtheta equals central finite
differences of the oracle loss (h = 1e-6 relative) for the 13 parameters the
reference's pow admits (m is held at its integer value)."""
import numpy as np
import pytest

import torch_tersoff as TT
from _fitdata import make_set, oracle_loss


@pytest.mark.slow
def test_torch_reference_matches_oracle_loss_and_finite_differences():
    theta, data = make_set(n_struct=2, seed=7)
    L0 = oracle_loss(theta, data)
    Lt, gt = TT.loss_and_grad(theta, data)
    assert abs(Lt - L0) <= 1e-12 * abs(L0)
    for p in range(14):
        if p == 12:
            continue
        h = 1e-6 * abs(theta[p])
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (oracle_loss(tp, data) - oracle_loss(tm, data)) / (2 * h)
        assert abs(gt[p] - fd) <= 2e-7 * abs(fd) + 1e-9 * np.abs(gt).max(), (TT.KEYS[p], gt[p], fd)
