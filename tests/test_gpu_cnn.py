"""The warm-start network's gfx950 kernels (MFMA forward, backward, Adam) against the CPU
oracle (oracle/cnn.py) on the same seeded parameters and inputs.

Tolerances: the oracle reproduces the device's BF16 roundings with exact sums, so the
remaining differences come from FP32 accumulation order (a rounding of a stored BF16
activation or gradient can flip): forward 1e-3, loss 1e-5, gradient 5e-3 relative L2
per layer, Adam 1e-6.
"""
import numpy as np
import pytest

from oracle import cnn

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _dev(a):
    from nextsimdg_amd.learned import DeviceArray

    return DeviceArray.from_numpy(a)


def _setup(seed, n=2, ny=8, nx=32):
    from nextsimdg_amd.learned import WarmStartNet

    rng = np.random.default_rng(seed)
    net = WarmStartNet(seed)
    p = net.get_params().astype(np.float64)
    p[[l * cnn.PSZ + cnn.WSZ + c for l in range(4) for c in range(cnn.C)]] = rng.normal(0, 0.05, 4 * cnn.C)
    net.set_params(p.astype(np.float32))
    p = net.get_params().astype(np.float64)
    xb = cnn.bf16_bits(rng.normal(0, 1, (n, ny, nx, cnn.C)).astype(np.float32))
    t = rng.normal(0, 0.5, (n, ny, nx, cnn.OUT)).astype(np.float32)
    w = (rng.uniform(0, 1, (n, ny, nx)) > 0.1).astype(np.float32)
    return net, p, _dev(xb), cnn.from_bits(xb).astype(np.float64), t, w


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_forward_matches_oracle():
    net, p, x, xh, _, _ = _setup(1)
    y = _dev(np.zeros((2, 8, 32, cnn.OUT), np.float32))
    net.forward(x, y)
    ref, _ = cnn.forward(p, xh, emulate=True)
    e = _rel(y.numpy(), ref)
    _say(f"  forward rel L2 {e:.2e}")
    assert e < 1e-3


def test_loss_and_gradients_match_oracle():
    net, p, x, xh, t, w = _setup(2)
    loss, g = net.loss_grad(x, _dev(t), _dev(w), want_grad=True)
    rl, rg = cnn.loss_grad(p, xh, t, w, emulate=True)
    _say(f"  loss gpu {loss:.8e} oracle {rl:.8e}")
    assert loss == pytest.approx(rl, rel=1e-5)
    for l in range(cnn.LAYERS):
        sl = slice(l * cnn.PSZ, (l + 1) * cnn.PSZ)
        e = _rel(g[sl], rg[sl])
        _say(f"  layer {l} gradient rel L2 {e:.2e}")
        assert e < 5e-3, (l, e)
    # unused output rows of the last layer get no gradient
    gl = g[3 * cnn.PSZ:3 * cnn.PSZ + cnn.WSZ].reshape(cnn.C, -1)
    assert np.all(gl[cnn.OUT:] == 0.0)


def test_adam_step_matches_oracle():
    net, p, x, _, t, w = _setup(3)
    _, g = net.loss_grad(x, _dev(t), _dev(w), want_grad=True)
    net.adam(1e-3)
    ref, _, _ = cnn.adam(p, np.zeros_like(p), np.zeros_like(p), g.astype(np.float64), 1, 1e-3)
    np.testing.assert_allclose(net.get_params(), ref, rtol=1e-6, atol=1e-7)


def test_training_fits_a_teacher_network():
    """End to end: the kernels learn the map of a random teacher network of the same shape."""
    from nextsimdg_amd.learned import WarmStartNet

    rng = np.random.default_rng(5)
    n, ny, nx = 8, 16, 32
    teacher = rng.normal(0, 0.05, cnn.NPARAM)
    xb = cnn.bf16_bits(rng.normal(0, 1, (n, ny, nx, cnn.C)).astype(np.float32))
    t, _ = cnn.forward(teacher, cnn.from_bits(xb).astype(np.float64), emulate=False)
    t = (t / np.std(t)).astype(np.float32)
    net = WarmStartNet(11)
    x, tt = _dev(xb), _dev(t)
    first = loss = None
    for _ in range(300):
        loss, _ = net.loss_grad(x, tt)
        net.adam(3e-3)
        first = loss if first is None else first
    _say(f"  loss {first:.4e} -> {loss:.4e}")
    assert loss < 0.1 * first


def test_device_arrays_views_and_gather():
    from nextsimdg_amd.learned import DeviceArray

    a = np.arange(5 * 3 * 4, dtype=np.float32).reshape(5, 3, 4)
    d = DeviceArray.from_numpy(a)
    np.testing.assert_array_equal(d[2].numpy(), a[2])
    np.testing.assert_array_equal(d[1:4].numpy(), a[1:4])
    np.testing.assert_array_equal(d.gather([4, 0, 0, 3]).numpy(), a[[4, 0, 0, 3]])
    d[3].upload(np.full((3, 4), -1.0, np.float32))
    a[3] = -1.0
    np.testing.assert_array_equal(d.numpy(), a)
