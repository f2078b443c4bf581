"""CPU oracle of the warm-start network (test infrastructure only: tests/ and the checks
of __graft_entry__ use it to check the gfx950 kernels, nothing in the product calls it).

The network has no counterpart in the reference (it accelerates the solver the
reference runs with a fixed subcycle count). Its definition is the one of
include/nsdg_cnn.h: 4 layers of 3x3 'same' convolutions with 32 channels, ReLU between
layers, linear output of which channels 0..7 are used; NHWC activations. Parameters in
the flat order [layer][W[co][ci][ky][kx] ..., b[co] ...].

With emulate=True the BF16 roundings of the device path are reproduced: the weights,
the post-ReLU activations and the pre-activation gradients are rounded to BF16 (round
to nearest even) where the device stores them, every sum is exact (float64) here.
Pinned by tests/test_cnn_oracle.py: the convolution against torch's CPU conv2d in
float64, the analytic gradient against central finite differences, the BF16 rounding
against torch's float32 -> bfloat16 conversion.
"""
import numpy as np

C, OUT, LAYERS, TAPS = 32, 8, 4, 9
WSZ = C * C * TAPS
PSZ = WSZ + C
NPARAM = LAYERS * PSZ


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(a):
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def unpack(params):
    p = np.asarray(params, dtype=np.float64).reshape(LAYERS, PSZ)
    return [(p[l, :WSZ].reshape(C, C, 3, 3), p[l, WSZ:]) for l in range(LAYERS)]


def conv(a, W):
    """z[n,y,x,co] = sum_{ci,ky,kx} a[n, y+ky-1, x+kx-1, ci] W[co,ci,ky,kx] (zero padding)."""
    n, ny, nx, _ = a.shape
    ap = np.zeros((n, ny + 2, nx + 2, a.shape[3]))
    ap[:, 1:-1, 1:-1] = a
    z = np.zeros((n, ny, nx, W.shape[0]))
    for ky in range(3):
        for kx in range(3):
            z += np.einsum("nyxc,oc->nyxo", ap[:, ky:ky + ny, kx:kx + nx], W[:, :, ky, kx])
    return z


def conv_data_grad(dz, W):
    """da[n,y,x,ci] = sum_{co,ky,kx} dz[n, y-ky+1, x-kx+1, co] W[co,ci,ky,kx]."""
    n, ny, nx, _ = dz.shape
    dp = np.zeros((n, ny + 2, nx + 2, dz.shape[3]))
    dp[:, 1:-1, 1:-1] = dz
    da = np.zeros((n, ny, nx, W.shape[1]))
    for ky in range(3):
        for kx in range(3):
            da += np.einsum("nyxo,oc->nyxc", dp[:, 2 - ky:2 - ky + ny, 2 - kx:2 - kx + nx], W[:, :, ky, kx])
    return da


def conv_weight_grad(a, dz):
    """dW[co,ci,ky,kx] = sum_{n,y,x} a[n, y+ky-1, x+kx-1, ci] dz[n,y,x,co]."""
    n, ny, nx, _ = a.shape
    ap = np.zeros((n, ny + 2, nx + 2, a.shape[3]))
    ap[:, 1:-1, 1:-1] = a
    dW = np.zeros((dz.shape[3], a.shape[3], 3, 3))
    for ky in range(3):
        for kx in range(3):
            dW[:, :, ky, kx] = np.einsum("nyxc,nyxo->oc", ap[:, ky:ky + ny, kx:kx + nx], dz)
    return dW


def forward(params, x, emulate=True):
    """Output y[n,y,x,8] and the activations [x, a1, a2, a3]."""
    layers = unpack(params)
    rnd = bf16 if emulate else (lambda v: v)
    acts = [np.asarray(x, dtype=np.float64)]
    a = acts[0]
    for l, (W, b) in enumerate(layers):
        Wq = rnd(W.astype(np.float32)).astype(np.float64) if emulate else W
        z = conv(a, Wq)
        if emulate:
            z = z.astype(np.float32).astype(np.float64)
        z = z + b
        if l < LAYERS - 1:
            a = np.maximum(z, 0.0)
            if emulate:
                a = bf16(a.astype(np.float32)).astype(np.float64)
            acts.append(a)
        else:
            return z[..., :OUT], acts
    raise AssertionError


def loss_grad(params, x, t, w=None, emulate=True):
    """Loss sum_p w_p sum_{c<8} (y - t)^2 / (8 npix) and its gradient (flat order)."""
    layers = unpack(params)
    y, acts = forward(params, x, emulate)
    n, ny, nx, _ = y.shape
    npix = n * ny * nx
    wp = np.ones((n, ny, nx)) if w is None else np.asarray(w, dtype=np.float64).reshape(n, ny, nx)
    d = y - np.asarray(t, dtype=np.float64).reshape(y.shape)
    if emulate:
        d = d.astype(np.float32).astype(np.float64)
    loss = float(np.sum(wp[..., None] * d * d) / (8.0 * npix))
    dz = np.zeros((n, ny, nx, C))
    dz[..., :OUT] = (2.0 / (8.0 * npix)) * wp[..., None] * d
    if emulate:
        dz = bf16(dz.astype(np.float32)).astype(np.float64)
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        W, _ = layers[l]
        Wq = bf16(W.astype(np.float32)).astype(np.float64) if emulate else W
        grads[l] = np.concatenate([conv_weight_grad(acts[l], dz).ravel(), dz.sum(axis=(0, 1, 2))])
        if l > 0:
            da = conv_data_grad(dz, Wq)
            dz = da * (acts[l] > 0)
            if emulate:
                dz = bf16(dz.astype(np.float32)).astype(np.float64)
    return loss, np.concatenate(grads)


def adam(params, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
