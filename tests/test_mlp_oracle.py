"""The CPU restatement of the precision classifier (oracle/mlp.py), pinned by mathematics: its
analytic gradient equals central differences of its loss, probabilities are a softmax (rows sum
to one, invariant to a shift of the output bias), and a network whose output layer is zero
predicts the uniform distribution with loss log 3."""
import numpy as np
import pytest

import mlp


def _problem(seed, n=40):
    rng = np.random.default_rng(seed)
    params = rng.standard_normal(mlp.NPARAMS) * 0.3
    x = rng.standard_normal((n, mlp.NI))
    y = rng.integers(0, mlp.NO, n)
    cw = rng.uniform(0.5, 2.0, mlp.NO)
    return params, x, y, cw


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_central_differences(seed):
    params, x, y, cw = _problem(seed)
    _, g = mlp.loss_grad(params, x, y, cw)
    rng = np.random.default_rng(100 + seed)
    # every bias and a sample of weights of each layer
    idx = np.unique(np.concatenate([rng.choice(mlp.NPARAMS, 120, replace=False),
                                    np.arange(mlp.NPARAMS - 3, mlp.NPARAMS)]))
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (mlp.loss_grad(p1, x, y, cw)[0] - mlp.loss_grad(p2, x, y, cw)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_softmax_properties():
    params, x, y, cw = _problem(3)
    p, _ = mlp.forward(params, x)
    np.testing.assert_allclose(p.sum(axis=1), 1.0, rtol=1e-14)
    shifted = params.copy()
    shifted[-3:] += 5.0
    p2, _ = mlp.forward(shifted, x)
    np.testing.assert_allclose(p, p2, rtol=1e-12)


def test_zero_output_layer_is_uniform():
    params, x, y, _ = _problem(4)
    P = mlp.unpack(params)
    off = mlp.NPARAMS - (mlp.NH * mlp.NO + mlp.NO)
    params[off:] = 0.0
    loss, g = mlp.loss_grad(params, x, y)
    assert loss == pytest.approx(np.log(3.0), rel=1e-14)
    p, _ = mlp.forward(params, x)
    np.testing.assert_allclose(p, 1.0 / 3.0, rtol=1e-14)
