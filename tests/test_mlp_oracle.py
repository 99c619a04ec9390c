"""Pins the estimator's CPU oracle (oracle/mlp_ref.py): its analytic gradients against central
finite differences, and its bfloat16 rounding against the definition."""
import numpy as np

from oracle import mlp_ref


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 1]
    n_params = sum(o * i + o for o, i in mlp_ref.layer_shapes(dims))
    p = rng.normal(0, 0.5, n_params)
    x = rng.normal(0, 1, (11, 5))
    y = rng.normal(0, 1, 11)
    loss, g = mlp_ref.loss_and_grads(p, dims, x, y)
    eps = 1e-6
    num = np.zeros_like(p)
    for i in range(n_params):
        d = np.zeros_like(p)
        d[i] = eps
        num[i] = (mlp_ref.loss_and_grads(p + d, dims, x, y)[0] - mlp_ref.loss_and_grads(p - d, dims, x, y)[0]) / (2 * eps)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-7)
    assert loss > 0


def test_zero_network_is_constant():
    dims = [3, 4, 1]
    n_params = sum(o * i + o for o, i in mlp_ref.layer_shapes(dims))
    p = np.zeros(n_params)
    p[-1] = 0.25  # output bias
    out, _ = mlp_ref.forward(p, dims, np.ones((6, 3)))
    assert np.all(out == 0.25)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = mlp_ref.bf16_round(x)
    # ties to even at the 8-bit mantissa boundary
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 2 ** -8 * 3.0e-3  # half a unit in the last of 8 significant bits
