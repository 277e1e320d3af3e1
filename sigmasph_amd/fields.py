"""Taichi-field-shaped views of device tensors.

The reference's callers read particle state as `ps.x.to_numpy()[:n]`, write it with
`from_numpy`, and read 0-d scalars as `solver.dt[None]`.  `Field` and `Scalar` keep
that surface over torch tensors that live on the GPU, so scripts written against the
reference run unchanged.
"""
import numpy as np
import torch


class Field:
    """Per-particle field backed by the live buffer's tensor.

    `getter()` returns the tensor that currently holds the field (the counting sort
    swaps buffers, so the tensor behind a name changes every step).  Vector fields are
    stored [n][4] with a padding lane; `ncomp` is the number of meaningful components.
    """

    def __init__(self, getter, ncomp=1, name=""):
        self._get = getter
        self.ncomp = ncomp
        self.name = name

    @property
    def tensor(self):
        return self._get()

    @property
    def shape(self):
        t = self._get()
        return (t.shape[0],) if self.ncomp == 1 else (t.shape[0], self.ncomp)

    def _view(self):
        t = self._get()
        return t if self.ncomp == 1 or t.dim() == 1 else t[:, : self.ncomp]

    def to_numpy(self):
        return self._view().detach().cpu().numpy().copy()

    def to_torch(self):
        return self._view().clone()

    def from_numpy(self, arr):
        t = self._get()
        src = torch.as_tensor(np.asarray(arr), dtype=t.dtype)
        if self.ncomp == 1 or t.dim() == 1:
            t.copy_(src.to(t.device).reshape(t.shape))
        else:
            t[:, : self.ncomp].copy_(src.to(t.device).reshape(t.shape[0], self.ncomp))

    def fill(self, value):
        self._view().fill_(value)

    def __getitem__(self, idx):
        v = self._view()[idx]
        return v.detach().cpu().numpy() if v.dim() else v.item()


class Scalar:
    """A 0-d value read and written as `s[None]`: a slot of a device f64 tensor, or a
    host value.  Every read of a device slot is a device-to-host sync, as a Taichi 0-d
    field read is."""

    def __init__(self, tensor=None, index=0, value=0.0, dtype=float):
        self._t = tensor
        self._i = index
        self._v = value
        self._dtype = dtype

    def __getitem__(self, key):
        if key is not None:
            raise IndexError("0-d scalar: index with [None]")
        if self._t is None:
            return self._dtype(self._v)
        return self._dtype(self._t[self._i].item())

    def __setitem__(self, key, value):
        if key is not None:
            raise IndexError("0-d scalar: index with [None]")
        if self._t is None:
            self._v = self._dtype(value)
        else:
            self._t[self._i] = float(value)

    def __float__(self):
        return float(self[None])

    def __int__(self):
        return int(self[None])


# the stored order of a symmetric deviator row: xx, yy, xy, zz, xz, yz
_SYM = ((0, 2, 4), (2, 1, 5), (4, 5, 3))
# the stored order of a velocity-gradient row: G00 G01 G10 G11 G22 G02 G12 G20 G21
_GV = ((0, 1, 5), (2, 3, 6), (7, 8, 4))


class MatrixField:
    """A per-particle 3x3 field (the reference's `ti.Matrix.field(3, 3)`), stored
    compactly: `layout` maps (row, col) to the column of the stored row."""

    def __init__(self, getter, layout, name=""):
        self._get = getter
        self._layout = layout
        self.name = name

    @property
    def shape(self):
        return (self._get().shape[0], 3, 3)

    def to_numpy(self):
        t = self._get().detach().cpu().numpy()
        out = np.empty((t.shape[0], 3, 3), dtype=t.dtype)
        for a in range(3):
            for b in range(3):
                out[:, a, b] = t[:, self._layout[a][b]]
        return out

    def from_numpy(self, arr):
        arr = np.asarray(arr)
        t = self._get()
        host = t.detach().cpu().numpy().copy()
        for a in range(3):
            for b in range(3):
                host[:, self._layout[a][b]] = arr[:, a, b]
        t.copy_(torch.as_tensor(host, dtype=t.dtype).to(t.device))

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


def sym_field(getter, name=""):
    return MatrixField(getter, _SYM, name)


def grad_field(getter, name=""):
    return MatrixField(getter, _GV, name)
