"""CPU restatement of the surrogate network (test infrastructure: only tests/ load it).

The same 12 -> 64 -> 64 -> 7 ReLU network as sph_mlp.h, on the same packed parameter
vector, in two modes: exact float64 (pinned by finite differences of the loss), and the
device's arithmetic -- every matrix operand (the normalised inputs, the weights, each
layer's activations, the output error) rounded to bf16 (round to nearest even) and the
products summed in float32 -- which is what the GPU's results are compared with.
"""
import numpy as np

NIN, NH, NOUT, PIN, POUT = 12, 64, 7, 16, 16
W1 = 0
B1 = W1 + NH * PIN
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + POUT * NH
NPARAM = B3 + POUT


def unpack(p):
    p = np.asarray(p)
    return (p[W1:B1].reshape(NH, PIN), p[B1:W2], p[W2:B2].reshape(NH, NH), p[B2:W3],
            p[W3:B3].reshape(POUT, NH), p[B3:NPARAM])


def bf16(a):
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _mm(a, b, device):
    if device:
        return (bf16(a).astype(np.float32) @ bf16(b).astype(np.float32)).astype(np.float32)
    return a @ b


def forward(p, xn, device=False):
    """xn: normalised inputs [n][12]; returns (y, cache) with y normalised [n][7]."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    dt = np.float32 if device else np.float64
    x = np.zeros((len(xn), PIN), dtype=dt)
    x[:, :NIN] = xn
    z1 = _mm(x, w1.T.astype(dt), device) + b1.astype(dt)
    h1 = np.maximum(z1, 0)
    z2 = _mm(h1, w2.T.astype(dt), device) + b2.astype(dt)
    h2 = np.maximum(z2, 0)
    y = _mm(h2, w3.T.astype(dt), device) + b3.astype(dt)
    return y[:, :NOUT], (x, h1, h2)


def loss_and_grad(p, xn, tn, device=False):
    """Mean square error over n x 7 outputs and its gradient on the packed vector."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    dt = np.float32 if device else np.float64
    y, (x, h1, h2) = forward(p, xn, device)
    n = len(xn)
    d = y - tn
    loss = float(np.sum(d.astype(np.float64) ** 2) / (n * NOUT))
    dy = np.zeros((n, POUT), dtype=dt)
    dy[:, :NOUT] = d * (2.0 / (n * NOUT))
    g = np.zeros(NPARAM, dtype=np.float64)
    g[W3:B3] = _mm(dy.T, h2, device).reshape(-1)
    g[B3:NPARAM] = dy.sum(0)
    dz2 = _mm(dy, w3.astype(dt), device) * (h2 > 0)
    g[W2:B2] = _mm(dz2.T, h1, device).reshape(-1)
    g[B2:W3] = dz2.sum(0)
    dz1 = _mm(dz2, w2.astype(dt), device) * (h1 > 0)
    g[W1:B1] = _mm(dz1.T, x, device).reshape(-1)
    g[B1:W2] = dz1.sum(0)
    return loss, g
