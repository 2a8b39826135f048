"""Field objects with the part of the ti.field surface the reference drivers use.

The reference hands ``ti.field`` objects between its classes and reads them
with ``.to_numpy()`` (main_moving.py:125-128).  Here the data live in
torch-ROCm device tensors in the library's own layout ([z][y][x], x
fastest); the field object converts to and from the reference's host layout
``(Nx, Ny, Nz)`` (C order, z fastest) only when the host asks.
"""

from __future__ import annotations

import numpy as np


def _torch():
    import torch
    return torch


class Field:
    """A float32 scalar field of shape (Nx, Ny, Nz) stored as a device tensor [Nz, Ny, Nx].

    ``refresh`` (optional) is called before any read so that lazily computed
    fields (the macroscopic readout) are brought up to date first."""

    def __init__(self, shape, tensor, refresh=None, owner=None, box=None):
        self.shape = tuple(int(s) for s in shape)
        self._t = tensor
        self._refresh = refresh
        self.owner = owner      # the object whose state this field shows
        self.box = box          # device address of the int32[6] force box of a spread force

    @property
    def tensor(self):
        if self._refresh is not None:
            self._refresh()
        return self._t

    def data_ptr(self) -> int:
        return self.tensor.data_ptr()

    def to_numpy(self) -> np.ndarray:
        t = self.tensor.detach().to("cpu")
        return np.ascontiguousarray(t.numpy().transpose(2, 1, 0))

    def to_torch(self):
        """The device tensor in the library layout [Nz, Ny, Nx] (no copy)."""
        return self.tensor

    def from_numpy(self, arr) -> None:
        torch = _torch()
        a = np.asarray(arr, dtype=np.float32)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != field shape {self.shape}")
        # through ``tensor``: the owner orders its own pending work first
        self.tensor.copy_(torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))))

    def fill(self, value: float) -> None:
        self.tensor.fill_(float(value))

    def __repr__(self) -> str:
        return f"Field(shape={self.shape})"


class MarkerField:
    """A float32 per-marker array of shape (N,) (the reference's (N,) ti.fields).

    ``sync`` (optional) is called before any access so that work queued on
    another stream (the owner's marker kernels) is ordered first."""

    def __init__(self, tensor, sync=None):
        self._t = tensor
        self._sync = sync
        self.shape = (int(tensor.shape[0]),)

    @property
    def tensor(self):
        if self._sync is not None:
            self._sync()
        return self._t

    def data_ptr(self) -> int:
        return self.tensor.data_ptr()

    def to_numpy(self) -> np.ndarray:
        return self.tensor.detach().to("cpu").numpy().copy()

    def to_torch(self):
        return self.tensor

    def from_numpy(self, arr) -> None:
        torch = _torch()
        a = np.asarray(arr, dtype=np.float32).reshape(self.shape)
        self.tensor.copy_(torch.from_numpy(a))

    def __repr__(self) -> str:
        return f"MarkerField(shape={self.shape})"


class HostConstant:
    """A constant host array exposed with ``to_numpy`` (e.g. CLBMSolver.M_inv)."""

    def __init__(self, arr):
        self._a = np.asarray(arr)
        self.shape = self._a.shape

    def to_numpy(self) -> np.ndarray:
        return self._a.copy()


def as_device_ptr(x, shape=None) -> int:
    """Device address of a Field / MarkerField / torch tensor, 0 for None."""
    if x is None:
        return 0
    if isinstance(x, (Field, MarkerField)):
        return x.data_ptr()
    torch = _torch()
    if isinstance(x, torch.Tensor):
        if not x.is_cuda:
            raise TypeError("force/velocity tensors must live on the GPU")
        if x.dtype != torch.float32 or not x.is_contiguous():
            raise TypeError("tensors must be contiguous float32 in [Nz, Ny, Nx] layout")
        if shape is not None and x.numel() != int(np.prod(shape)):
            raise ValueError("tensor size does not match the lattice")
        return x.data_ptr()
    raise TypeError(f"cannot pass {type(x).__name__} to the device; use a Field or a CUDA tensor")
