"""Pins the numpy network restatement (tests/mlp_ref.py) by central finite
differences, in float64, before it is used to check the HIP kernels."""
import numpy as np

import mlp_ref


def test_mlp_ref_gradient_matches_finite_differences():
    F, H, A, B = 7, 5, 3, 9
    P = H * F + H + H * H + H + A * H + A
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.5, P)
    X = rng.normal(0, 1, (B, F))
    T = rng.normal(0, 1, (B, A))
    M = (rng.random((B, A)) < 0.8).astype(np.float64)
    loss, g = mlp_ref.loss_and_grad(p, X, T, M, F, H, A)
    eps = 1e-6
    num = np.empty(P)
    for i in range(P):
        q = p.copy(); q[i] += eps
        lp, _ = mlp_ref.loss_and_grad(q, X, T, M, F, H, A)
        q[i] -= 2 * eps
        lm, _ = mlp_ref.loss_and_grad(q, X, T, M, F, H, A)
        num[i] = (lp - lm) / (2 * eps)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-7), float(np.abs(g - num).max())
    assert loss > 0
