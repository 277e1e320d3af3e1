"""The learned constitutive surrogate: an MLP 12 -> 64 -> 64 -> 7 (ReLU) that maps the
state at the return map -- the trial deviator, the plastic strain, the temperature, the
step's rate scale, the density -- to the returned deviator and the plastic strain
increment, evaluated and trained on the GPU's bf16 matrix cores (`sph_mlp_*` in
libsigmasph.so; the kernels are sph_mlp.h).

`Network` owns the parameters, the Adam state and the normalisation on the device and
trains on tuples harvested from plain runs (`HypoElasticSolver.harvest`); the solid step
evaluates it for every yielding particle and keeps its answer only where the
yield-surface residual is within tolerance, falling back to the plain return map
elsewhere.
"""
import ctypes
import math

import numpy as np
import torch

from . import _native as N

NIN, NH, NOUT, PIN, POUT = 12, 64, 7, 16, 16
W1, B1 = 0, NH * PIN
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + POUT * NH
NPARAM = B3 + POUT
assert NPARAM == N.SPH_MLP_NPARAM


def init_params(seed=0):
    """He-uniform matrices, zero biases, zero padding (so padded inputs and outputs stay
    inert through training: their gradients are zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    w1 = np.zeros((NH, PIN), np.float32)
    w1[:, :NIN] = rng.uniform(-1.0, 1.0, (NH, NIN)) * math.sqrt(6.0 / NIN)
    w2 = (rng.uniform(-1.0, 1.0, (NH, NH)) * math.sqrt(6.0 / NH)).astype(np.float32)
    w3 = np.zeros((POUT, NH), np.float32)
    w3[:NOUT] = rng.uniform(-1.0, 1.0, (NOUT, NH)) * math.sqrt(6.0 / NH) * 0.5
    p[W1:B1] = w1.reshape(-1)
    p[W2:B2] = w2.reshape(-1)
    p[W3:B3] = w3.reshape(-1)
    return p


class Network:
    """sph_mlp on the device: parameters, bf16 copies, gradient, Adam moments, norms."""

    def __init__(self, device="cuda", seed=0, params=None):
        self.device = torch.device(device)
        f = lambda n: torch.zeros(n, dtype=torch.float32, device=self.device)
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        self.w = torch.from_numpy(p.copy()).to(self.device)
        self.wb = torch.zeros(N.SPH_MLP_NBF, dtype=torch.int16, device=self.device)
        self.grad, self.m1, self.m2 = f(NPARAM), f(NPARAM), f(NPARAM)
        self.norm = f(64)
        self.norm[16:32] = 1.0
        self.norm[48:64] = 1.0
        self.loss = f(1)
        self.steps = 0
        self.s = N.SphMlp(self.w.data_ptr(), self.wb.data_ptr(), self.grad.data_ptr(),
                          self.m1.data_ptr(), self.m2.data_ptr(), self.norm.data_ptr())
        self._lib = N.lib()
        self.sync()

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def sync(self):
        N.check(self._lib.sph_mlp_sync(ctypes.byref(self.s), self._stream()), None, "sph_mlp_sync")

    def set_normalisation(self, x, y):
        """Means and standard deviations of the training set (floored at 1e-6)."""
        x = torch.as_tensor(x, dtype=torch.float32, device=self.device)[:, :NIN]
        y = torch.as_tensor(y, dtype=torch.float32, device=self.device)[:, :NOUT]
        self.norm.zero_()
        self.norm[0:NIN] = x.mean(0)
        self.norm[16:32] = 1.0
        self.norm[16:16 + NIN] = 1.0 / x.std(0).clamp_min(1e-6)
        self.norm[32:32 + NOUT] = y.mean(0)
        self.norm[48:64] = 1.0
        # an answer that hardly varies (the direction correction of a radial return) keeps
        # a unit-order scale, so its near-zero targets do not dominate the loss
        self.norm[48:48 + NOUT] = y.std(0).clamp_min(1e-3)

    def forward(self, x):
        """x [n][16] float32 on the device -> y [n][8]."""
        n = x.shape[0]
        y = torch.empty((n, 8), dtype=torch.float32, device=self.device)
        N.check(self._lib.sph_mlp_forward(ctypes.byref(self.s), x.data_ptr(), y.data_ptr(), n, self._stream()),
                None, "sph_mlp_forward")
        return y

    def gradient(self, x, t):
        """grad += d loss / d parameters over (x, t); returns the loss (synchronises)."""
        self.loss.zero_()
        N.check(self._lib.sph_mlp_gradient(ctypes.byref(self.s), x.data_ptr(), t.data_ptr(), x.shape[0],
                                           self.loss.data_ptr(), self._stream()), None, "sph_mlp_gradient")
        return float(self.loss.item())

    def train(self, x, t, steps=2000, batch=4096, lr=3e-3, lr_end=3e-4, seed=0, log_every=0):
        """Adam on random mini-batches of the device-resident training set, learning rate
        decaying geometrically from `lr` to `lr_end`; returns the mean loss of every
        `log_every` steps (or of the last 100)."""
        n = x.shape[0]
        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        hist = []
        acc = torch.zeros(1, dtype=torch.float32, device=self.device)
        window = log_every or 100
        for k in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), device=self.device, generator=g)
            xb, tb = x[idx].contiguous(), t[idx].contiguous()
            self.steps += 1
            rate = lr * (lr_end / lr) ** (k / max(1, steps - 1))
            self.loss.zero_()
            N.check(self._lib.sph_mlp_train_step(ctypes.byref(self.s), xb.data_ptr(), tb.data_ptr(), xb.shape[0],
                                                 ctypes.c_float(rate), self.steps, self.loss.data_ptr(),
                                                 self._stream()), None, "sph_mlp_train_step")
            acc += self.loss
            if (k + 1) % window == 0:
                hist.append(float(acc.item()) / window)
                acc.zero_()
        return hist

    def state(self):
        return {"w": self.w.detach().cpu().numpy().copy(), "norm": self.norm.detach().cpu().numpy().copy()}

    def load(self, state):
        self.w.copy_(torch.from_numpy(np.asarray(state["w"], dtype=np.float32)))
        self.norm.copy_(torch.from_numpy(np.asarray(state["norm"], dtype=np.float32)))
        self.sync()
