"""Learned warm start of the mEVP solve (include/nsdg_cnn.h), Python side.

No counterpart in the reference, which runs mEVP from the previous timestep's velocity
and stress for a fixed subcycle count (VPCGDynamicsKernel.hpp:63-94). Here a small
convolutional network on the element grid predicts the velocity that mEVP converges to
in a timestep, from the state and forcing at its start; the solve starts from that
velocity and its VP stress when the momentum residual says it is the better start.

Everything runs in libnsdg_hip.so: the network's forward, backward and Adam kernels
(BF16 on MFMA, FP32 accumulation), the feature, target and residual kernels, and the
device buffers of the training data; the training pairs come from the GPU solver itself.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import NsdgError, check

CHANNELS, OUTPUTS = 32, 8
BF16 = np.uint16  # BF16 values are held as their bit patterns


class DeviceArray:
    """A C-contiguous array in device memory allocated by libnsdg_hip.so (no arithmetic:
    the kernels read and write it through its address). Indexing the first axis with an
    int or a slice gives a view that shares the memory."""

    def __init__(self, shape, dtype, _base=None, _offset=0):
        self.lib = _lib.load()
        self.shape = tuple(int(n) for n in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._base = _base
        if _base is None:
            p = ctypes.c_void_p()
            check(self.lib.nsdg_device_alloc(self.nbytes, ctypes.byref(p)))
            self._addr = p.value or 0
        else:
            self._addr = _base._addr + _offset

    def __del__(self):
        if getattr(self, "_base", 1) is None and getattr(self, "_addr", 0):
            self.lib.nsdg_device_free(ctypes.c_void_p(self._addr))
            self._addr = 0

    @property
    def ptr(self):
        return ctypes.c_void_p(self._addr)

    @property
    def item_bytes(self):
        return self.nbytes // self.shape[0]

    def __getitem__(self, i):
        root = self._base if self._base is not None else self
        if isinstance(i, slice):
            a, b, st = i.indices(self.shape[0])
            assert st == 1
            return DeviceArray((b - a,) + self.shape[1:], self.dtype, root, self._addr - root._addr + a * self.item_bytes)
        i = int(i) + (self.shape[0] if int(i) < 0 else 0)
        assert 0 <= i < self.shape[0]
        return DeviceArray(self.shape[1:], self.dtype, root, self._addr - root._addr + i * self.item_bytes)

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(self.lib.nsdg_device_upload(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(self.lib.nsdg_device_download(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes))
        return out

    def gather(self, index, out=None):
        """out[i] = self[index[i]] (device to device)."""
        idx = np.ascontiguousarray(index, dtype=np.int64)
        assert idx.size == 0 or (idx.min() >= 0 and idx.max() < self.shape[0])
        out = DeviceArray((idx.size,) + self.shape[1:], self.dtype) if out is None else out
        assert out.shape == (idx.size,) + self.shape[1:] and out.dtype == self.dtype
        check(self.lib.nsdg_device_gather(out.ptr, self.ptr, idx.ctypes.data_as(ctypes.c_void_p), idx.size,
                                          self.item_bytes))
        return out


class WarmStartNet:
    """The 4-layer 3x3 convolutional network of include/nsdg_cnn.h."""

    def __init__(self, seed=0):
        self.lib = _lib.load()
        h = self.lib.nsdg_cnn_create(int(seed))
        if not h:
            raise NsdgError(-1, (self.lib.nsdg_last_error() or b"").decode())
        self.h = ctypes.c_void_p(h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.nsdg_cnn_destroy(self.h)
            self.h = None

    @property
    def nparam(self):
        n = ctypes.c_size_t()
        check(self.lib.nsdg_cnn_param_count(self.h, ctypes.byref(n)))
        return n.value

    def get_params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        check(self.lib.nsdg_cnn_get_params(self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.nparam
        check(self.lib.nsdg_cnn_set_params(self.h, a.ctypes.data_as(ctypes.c_void_p)))

    def forward(self, x, y):
        """x: BF16 [n, ny, nx, 32], y: float32 [n, ny, nx, 8] (DeviceArray)."""
        n, ny, nx, c = x.shape
        assert c == CHANNELS and x.dtype == BF16
        assert y.shape == (n, ny, nx, OUTPUTS) and y.dtype == np.float32
        check(self.lib.nsdg_cnn_forward(self.h, x.ptr, n, ny, nx, y.ptr))

    def loss_grad(self, x, t, w=None, want_grad=False):
        n, ny, nx, _ = x.shape
        assert x.dtype == BF16 and t.shape == (n, ny, nx, OUTPUTS) and t.dtype == np.float32
        if w is not None:
            assert w.shape == (n, ny, nx) and w.dtype == np.float32
        loss = ctypes.c_double()
        g = np.zeros(self.nparam, dtype=np.float32) if want_grad else None
        check(self.lib.nsdg_cnn_loss_grad(self.h, x.ptr, t.ptr, w.ptr if w is not None else None, n, ny, nx,
                                          ctypes.byref(loss), g.ctypes.data_as(ctypes.c_void_p) if want_grad else None))
        return loss.value, g

    def adam(self, lr):
        check(self.lib.nsdg_cnn_adam(self.h, ctypes.c_float(lr)))

    def set_profiling(self, on):
        check(self.lib.nsdg_cnn_set_profiling(self.h, int(bool(on))))

    def kernel_stats(self, name):
        c, t = ctypes.c_int(), ctypes.c_double()
        check(self.lib.nsdg_cnn_kernel_stats(self.h, name.encode(), ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value


# ---------------------------------------------------------------- coupling to the solver
def features(k, x):
    """Network input of the current timestep of kernel k into x (BF16 [ny, nx, 32])."""
    check(k.lib.nsdg_dyn_features(k.h, x.ptr))


def velocity_target(k, t, w):
    check(k.lib.nsdg_dyn_velocity_target(k.h, t.ptr, w.ptr))


def apply_velocity(k, y):
    check(k.lib.nsdg_dyn_apply_velocity(k.h, y.ptr))


def warm_start(k, net):
    """Predicted start with the residual gate; returns (r_predicted, r_cold, used)."""
    rp, rc, used = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    check(k.lib.nsdg_dyn_warm_start(k.h, net.h, ctypes.byref(rp), ctypes.byref(rc), ctypes.byref(used)))
    return rp.value, rc.value, bool(used.value)


def subcycles_to_residual(k, r_target, max_subcycles, check_every, r_start=-1.0):
    """Subcycles until r <= r_target; r_start is the residual of the current state when
    known (a negative value has it computed). Returns (subcycles run, final r)."""
    done, r = ctypes.c_int(), ctypes.c_double()
    check(k.lib.nsdg_dyn_subcycles_to_residual(k.h, float(r_target), int(max_subcycles), int(check_every),
                                               float(r_start), ctypes.byref(done), ctypes.byref(r)))
    return done.value, r.value


# ---------------------------------------------------------------- training pairs
KINDS = ("plain", "heterogeneous", "leads")


def new_track(nx, seed, kind, precision=0):
    """A benchmark box with the seeded cyclone track and ice of the given kind."""
    from . import dynamics, inputs

    s = inputs.benchmark_setup(nx, nx, seed, kind=kind)
    k = dynamics.MEVPDynamicsKernel(2, 6)
    k.initialise(s["coords"], False, s["mask"], nx, nx)
    for name, key in (("hice", "h"), ("cice", "A"), ("uocean", "uocean"), ("vocean", "vocean")):
        k.setData(name, s[key], 1)
    k.setData("ssh", 0 * s["h"], 1)
    k.set_precision(precision)
    return k, inputs.cyclone_offset(seed)


def begin_step(k, off, step, dt=120.0):
    """The start of model step `step`: DG0 reset, wind, advection, prepareIteration, u0 = u."""
    from . import inputs

    k.reset_dg0("hice")
    k.reset_dg0("cice")
    k.benchmark_wind(inputs.L_BOX, step * dt, off[0], off[1], 30.0)
    k.advectionAndLimits(dt)
    k.prepareIteration()
    k.begin_iteration(dt)


def plain_step(k, off, step, dt=120.0):
    from . import inputs

    k.reset_dg0("hice")
    k.reset_dg0("cice")
    k.benchmark_wind(inputs.L_BOX, step * dt, off[0], off[1], 30.0)
    k.update(dt)


def generate_pairs(nx, seeds, samples_per_track, steps_between=3, spinup=2, converge=10000, r_stop=0.0,
                   check_every=1000, dt=120.0, log=None):
    """Pairs (network input, converged velocity change, loss weight) from the GPU solver.

    For each seed (a cyclone track; the ice kind cycles through plain, heterogeneous and
    leads) the plain model runs; at every sampled step the start state gives the input,
    mEVP then runs up to `converge` subcycles (stopping early at r <= r_stop) and the
    velocity it reaches gives the target; the state is restored and the plain model
    continues."""
    n = len(seeds) * samples_per_track
    X = DeviceArray((n, nx, nx, CHANNELS), BF16)
    T = DeviceArray((n, nx, nx, OUTPUTS), np.float32)
    W = DeviceArray((n, nx, nx), np.float32)
    stats = []
    i = 0
    for si, seed in enumerate(seeds):
        kind = KINDS[si % len(KINDS)]
        k, off = new_track(nx, seed, kind)
        step = 0
        for _ in range(spinup):
            plain_step(k, off, step, dt)
            step += 1
        for _ in range(samples_per_track):
            blob = k.save_state()
            begin_step(k, off, step, dt)
            features(k, X[i])
            r0 = k.momentum_residual()[0]
            done, r = subcycles_to_residual(k, r_stop, converge, check_every)
            velocity_target(k, T[i], W[i])
            stats.append((seed, kind, step, r0, done, r))
            k.load_state(blob)
            for _ in range(steps_between):
                plain_step(k, off, step, dt)
                step += 1
            i += 1
        if log:
            log(f"track {seed} ({kind}): r0 {stats[-1][3]:.3e} -> {stats[-1][5]:.3e} after {stats[-1][4]}")
        del k
    return X, T, W, stats


def train(net, X, T, W, epochs, batch, lr, seed=0, idx=None, log=None):
    """Adam on mini-batches drawn without replacement each epoch; returns per-epoch mean losses."""
    idx = np.arange(X.shape[0]) if idx is None else np.asarray(idx)
    rng = np.random.default_rng(seed)
    xb = DeviceArray((batch,) + X.shape[1:], X.dtype)
    tb = DeviceArray((batch,) + T.shape[1:], T.dtype)
    wb = DeviceArray((batch,) + W.shape[1:], W.dtype)
    hist = []
    for ep in range(epochs):
        perm = idx[rng.permutation(idx.size)]
        losses = []
        for b in range(0, perm.size - batch + 1, batch):
            sel = perm[b:b + batch]
            X.gather(sel, xb)
            T.gather(sel, tb)
            W.gather(sel, wb)
            loss, _ = net.loss_grad(xb, tb, wb)
            net.adam(lr)
            losses.append(loss)
        hist.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.5e}")
    return hist


def evaluate(net, X, T, W, idx):
    """Mean loss of the network and of the zero prediction (the cold start) on samples idx."""
    ln, lz = [], []
    for i in idx:
        loss, _ = net.loss_grad(X[i:i + 1], T[i:i + 1], W[i:i + 1])
        ln.append(loss)
        t, w = T[i].numpy(), W[i].numpy()
        lz.append(float(np.sum(w[..., None] * t.astype(np.float64) ** 2) / t.size))
    return float(np.mean(ln)), float(np.mean(lz))
