"""The surrogate network's bf16 matrix-core kernels (sph_mlp_*) against the CPU
restatement in the device's arithmetic (oracle/mlp.py: bf16 operands, f32 sums).

Forward and gradient on ragged sample counts (not multiples of the 16-sample wave tile
or the 64-sample workgroup); the f32 sum order differs from numpy's, which can move an
activation across a bf16 rounding boundary, so the bar is 1e-2 of each block's scale at
the worst element and 1e-4 at the median.  Training on a smooth target must cut the loss
by ten.
"""
import numpy as np
import pytest
import torch

from oracle import mlp as M
from sigmasph_amd.surrogate import Network, init_params

pytestmark = pytest.mark.gpu


def _net(gpu, seed=4):
    rng = np.random.default_rng(seed)
    p = init_params(seed)
    p[M.B1:M.W2] = rng.normal(size=M.NH) * 0.1
    p[M.B2:M.W3] = rng.normal(size=M.NH) * 0.1
    p[M.B3:M.B3 + M.NOUT] = rng.normal(size=M.NOUT) * 0.1
    net = Network(device=gpu, params=p)
    mean = rng.normal(size=M.NIN).astype(np.float32)
    std = rng.uniform(0.5, 2.0, M.NIN).astype(np.float32)
    omean = rng.normal(size=M.NOUT).astype(np.float32)
    ostd = rng.uniform(0.5, 2.0, M.NOUT).astype(np.float32)
    net.norm.zero_()
    net.norm[0:M.NIN] = torch.from_numpy(mean)
    net.norm[16:32] = 1.0
    net.norm[16:16 + M.NIN] = torch.from_numpy(1.0 / std)
    net.norm[32:32 + M.NOUT] = torch.from_numpy(omean)
    net.norm[48:64] = 1.0
    net.norm[48:48 + M.NOUT] = torch.from_numpy(ostd)
    return net, p, (mean, std, omean, ostd)


def _inputs(n, norm, seed=5):
    mean, std, _, _ = norm
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 16), np.float32)
    x[:, :M.NIN] = mean + std * rng.normal(size=(n, M.NIN))
    return x


def _close(got, want, what):
    err = np.abs(got - want)
    scale = max(np.abs(want).max(), 1e-12)
    assert err.max() <= 1e-2 * scale, "%s: worst %g of %g" % (what, err.max(), scale)
    assert np.median(err) <= 1e-4 * scale, "%s: median %g of %g" % (what, np.median(err), scale)


@pytest.mark.parametrize("n", [1, 37, 200, 1000])
def test_forward_matches_the_oracle(gpu, n):
    net, p, norm = _net(gpu)
    x = _inputs(n, norm)
    y = net.forward(torch.from_numpy(x).to(gpu)).cpu().numpy()[:, :M.NOUT]
    mean, std, omean, ostd = norm
    xn = ((x[:, :M.NIN] - mean) * np.float32(1.0) / std).astype(np.float32)
    yn, _ = M.forward(p, xn, device=True)
    _close(y, yn * ostd + omean, "forward")


def test_gradient_matches_the_oracle(gpu):
    net, p, norm = _net(gpu)
    n = 300
    x = _inputs(n, norm)
    mean, std, omean, ostd = norm
    rng = np.random.default_rng(6)
    t = np.zeros((n, 8), np.float32)
    t[:, :M.NOUT] = omean + ostd * rng.normal(size=(n, M.NOUT))
    loss = net.gradient(torch.from_numpy(x).to(gpu), torch.from_numpy(t).to(gpu))
    g = net.grad.cpu().numpy()
    xn = ((x[:, :M.NIN] - mean) / std).astype(np.float32)
    tn = ((t[:, :M.NOUT] - omean) / ostd).astype(np.float32)
    lo, go = M.loss_and_grad(p, xn, tn, device=True)
    assert loss == pytest.approx(lo, rel=1e-3)
    for a, b, name in ((M.W1, M.B1, "W1"), (M.B1, M.W2, "b1"), (M.W2, M.B2, "W2"), (M.B2, M.W3, "b2"),
                       (M.W3, M.B3, "W3"), (M.B3, M.NPARAM, "b3")):
        _close(g[a:b], go[a:b], name)


def test_training_fits_a_smooth_map(gpu):
    rng = np.random.default_rng(7)
    n = 20000
    x = np.zeros((n, 16), np.float32)
    x[:, :M.NIN] = rng.uniform(-1, 1, (n, M.NIN))
    t = np.zeros((n, 8), np.float32)
    for k in range(M.NOUT):
        t[:, k] = np.sin(2 * x[:, k]) + 0.5 * x[:, (k + 3) % M.NIN] * x[:, (k + 5) % M.NIN]
    net = Network(device=gpu, seed=1)
    X, T = torch.from_numpy(x).to(gpu), torch.from_numpy(t).to(gpu)
    net.set_normalisation(X, T)
    hist = net.train(X, T, steps=1500, batch=2048, lr=3e-3, lr_end=3e-4, log_every=100)
    assert hist[-1] < 0.1 * hist[0], hist
    assert np.all(np.isfinite(net.w.cpu().numpy()))
