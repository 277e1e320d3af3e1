"""The CPU restatement of the surrogate network (oracle/mlp.py), pinned before it judges
the GPU: its float64 loss gradient against central finite differences on every parameter
block; its bf16 rounding against torch's float32 -> bfloat16 conversion (round to nearest
even) on random and tie values; the padded inputs and outputs carry zero gradient."""
import numpy as np
import pytest
import torch

from oracle import mlp as M
from sigmasph_amd.surrogate import init_params


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = init_params(3).astype(np.float64)
    p[M.B1:M.W2] = rng.normal(size=M.NH) * 0.1
    p[M.B2:M.W3] = rng.normal(size=M.NH) * 0.1
    x = rng.normal(size=(37, M.NIN))
    t = rng.normal(size=(37, M.NOUT))
    loss, g = M.loss_and_grad(p, x, t)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False)
                          for a, b in ((M.W1, M.B1), (M.B1, M.W2), (M.W2, M.B2), (M.B2, M.W3),
                                       (M.W3, M.W3 + M.NOUT * M.NH), (M.B3, M.B3 + M.NOUT))])
    for k in idx:
        if k < M.B1 and (k - M.W1) % M.PIN >= M.NIN:
            continue
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        fd = (M.loss_and_grad(pp, x, t)[0] - M.loss_and_grad(pm, x, t)[0]) / (2 * e)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9)
    # the padding: input columns 12..15 of W1, output rows 7..15 of W3, b3[7:]
    w1 = g[M.W1:M.B1].reshape(M.NH, M.PIN)
    assert np.all(w1[:, M.NIN:] == 0.0)
    assert np.all(g[M.W3 + M.NOUT * M.NH:M.B3] == 0.0) and np.all(g[M.B3 + M.NOUT:] == 0.0)


def test_bf16_rounding_is_round_to_nearest_even():
    rng = np.random.default_rng(2)
    a = (rng.normal(size=4096) * 10.0 ** rng.uniform(-6, 6, 4096)).astype(np.float32)
    ties = (np.arange(1, 65, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    for v in (a, ties, -ties):
        want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
        assert np.array_equal(M.bf16(v), want)
