"""Pins the FP64 step-model oracle: layout and finite-difference gradients."""
import numpy as np
import pytest

from aieplace_amd import api


def test_blob_size_matches_product(oracle):
    assert oracle.mlp_nweights() == api.MLP_NWEIGHTS == 5347


def test_mlp_grad_finite_differences(oracle):
    rng = np.random.default_rng(0)
    w = oracle.mlp_random_weights(rng)
    x = rng.standard_normal((12, 16))
    y = rng.standard_normal(12)
    g, loss = oracle.mlp_grad(w, x, y)
    assert np.all(g[:34] == 0)  # normalisation is not trained
    # loss definition: mean squared error on the normalised target
    pred = (oracle.mlp_forward(w, x) - w[32]) / w[33]
    assert loss == pytest.approx(np.mean((pred - (y - w[32]) / w[33]) ** 2), rel=1e-12)
    h = 1e-6
    for p in list(rng.integers(34, len(w), 60)) + [len(w) - 1, 34, 34 + 1024]:
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (oracle.mlp_grad(wp, x, y)[1] - oracle.mlp_grad(wm, x, y)[1]) / (2 * h)
        assert g[p] == pytest.approx(fd, rel=1e-4, abs=1e-8)


def test_mlp_forward_linear_limit(oracle):
    """With zero hidden weights the output is the bias: t = b3 * std + mean."""
    rng = np.random.default_rng(1)
    w = oracle.mlp_random_weights(rng)
    w[34:] = 0.0
    w[-1] = 0.7
    out = oracle.mlp_forward(w, rng.standard_normal((5, 16)))
    np.testing.assert_allclose(out, 0.7 * w[33] + w[32], rtol=1e-14)
