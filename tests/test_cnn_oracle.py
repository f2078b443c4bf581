"""Pins of the CPU oracle of the warm-start network (oracle/cnn.py), all on the CPU:
its convolutions against torch's conv2d and autograd in float64, its analytic loss
gradient against central finite differences, its BF16 rounding against torch's."""
import numpy as np
import pytest
import torch

from oracle import cnn


def _params(seed, scale=0.1):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, cnn.NPARAM)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-5, 1000), rng.normal(0, 1e5, 1000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(cnn.bf16(v), ref)
    np.testing.assert_array_equal(cnn.from_bits(cnn.bf16_bits(v)), ref)


def test_conv_and_its_gradients_match_torch():
    rng = np.random.default_rng(1)
    a = rng.normal(0, 1, (2, 5, 7, cnn.C))
    W = rng.normal(0, 0.1, (cnn.C, cnn.C, 3, 3))
    dz = rng.normal(0, 1, (2, 5, 7, cnn.C))
    at = torch.tensor(a.transpose(0, 3, 1, 2), requires_grad=True)
    Wt = torch.tensor(W, requires_grad=True)
    zt = torch.nn.functional.conv2d(at, Wt, padding=1)
    np.testing.assert_allclose(cnn.conv(a, W), zt.detach().numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-12)
    zt.backward(torch.tensor(dz.transpose(0, 3, 1, 2)))
    np.testing.assert_allclose(cnn.conv_data_grad(dz, W), at.grad.numpy().transpose(0, 2, 3, 1), rtol=1e-12,
                               atol=1e-12)
    np.testing.assert_allclose(cnn.conv_weight_grad(a, dz), Wt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    p = _params(3)
    x = rng.normal(0, 1, (1, 4, 16, cnn.C))
    t = rng.normal(0, 1, (1, 4, 16, cnn.OUT))
    w = (rng.uniform(0, 1, (1, 4, 16)) > 0.2).astype(np.float64)
    _, g = cnn.loss_grad(p, x, t, w, emulate=False)
    idx = rng.choice(cnn.NPARAM, 40, replace=False)
    # the last layer's unused output rows have no effect on the loss
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (cnn.loss_grad(pp, x, t, w, emulate=False)[0] - cnn.loss_grad(pm, x, t, w, emulate=False)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_adam_step_is_the_published_update():
    rng = np.random.default_rng(4)
    p, g = rng.normal(0, 1, 100), rng.normal(0, 1, 100)
    p1, m1, v1 = cnn.adam(p, np.zeros(100), np.zeros(100), g, 1, 1e-3)
    # first step of Adam (Kingma and Ba 2015): m_hat = g, v_hat = g^2 -> p - lr g / (|g| + eps)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
