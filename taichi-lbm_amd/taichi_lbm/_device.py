"""Device plumbing: torch-ROCm tensors as raw HIP buffers, field views in the
reference's shapes, and the lattice/fields structs handed to the C-ABI."""
from __future__ import annotations

from typing import Callable, Optional

import numpy as np

from . import _hip as H

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("taichi_lbm needs a HIP GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    return t.device("cuda", t.cuda.current_device())


def stream() -> int:
    return int(torch().cuda.current_stream().cuda_stream)


def plane_stride(n: int) -> int:
    """Elements from one plane of a field to the next: planes start on
    256-byte boundaries (64 elements), an odd number of 256-byte blocks
    apart.  At a power-of-two stride (the 8192^2 lattice: 2^26 elements) the
    nine population planes of a node fall on the same memory channel and
    cache sets; one block more made the two-step MRT step 4 % faster
    (0.593 against 0.619 ms/step, means of two runs each)."""
    ps = (n + 63) // 64 * 64
    if (ps // 64) % 2 == 0:
        ps += 64
    return ps


class _Scalar:
    """Stand-in for a 0-d Taichi field: supports x[None] reads and writes."""

    def __init__(self, value):
        self.value = value

    def __getitem__(self, key):
        return self.value

    def __setitem__(self, key, v):
        self.value = v

    def __repr__(self):
        return f"{self.value}"


class HostField:
    """A setup-time field kept on the host (mask, neighbour table).

    Indexing is bounds-checked: the reference reads out of range silently
    (cases/Circular cylinder.py:96-98); here such a read raises IndexError.
    """

    def __init__(self, arr: np.ndarray, on_change: Optional[Callable] = None):
        self.arr = arr
        self._on_change = on_change

    @property
    def shape(self):
        return self.arr.shape

    def _check(self, key):
        key = key if isinstance(key, tuple) else (key,)
        for ax, k in enumerate(key):
            if k is None or isinstance(k, slice) or k is Ellipsis:
                continue
            a = np.asarray(k)
            if a.dtype == bool:
                continue
            if a.size and (a.min() < 0 or a.max() >= self.arr.shape[ax]):
                raise IndexError(f"index out of range on axis {ax} of a field of shape {self.arr.shape}")

    def __getitem__(self, key):
        self._check(key)
        return self.arr[key]

    def __setitem__(self, key, v):
        self._check(key)
        self.arr[key] = v
        if self._on_change:
            self._on_change()

    def to_numpy(self):
        return self.arr.copy()

    def from_numpy(self, a):
        self.arr[...] = a
        if self._on_change:
            self._on_change()

    def fill(self, v):
        self.arr.fill(v)
        if self._on_change:
            self._on_change()


class DeviceField:
    """A device field exposed with the reference's Taichi shape.

    `view()` returns a strided torch view of the plane-layout buffer in the
    reference's shape; `before_read` / `after_write` keep the owner's lazy
    state (materialisation, derived values) consistent.
    """

    def __init__(self, view: Callable, before_read: Optional[Callable] = None,
                 after_write: Optional[Callable] = None, readonly: bool = False, name: str = ""):
        self._view = view
        self._before_read = before_read
        self._after_write = after_write
        self._readonly = readonly
        self.name = name

    def view(self, for_write: bool = False):
        if self._before_read:
            self._before_read()
        return self._view()

    @property
    def shape(self):
        return tuple(self.view().shape)

    def to_numpy(self):
        return self.view().detach().cpu().numpy().copy()

    def to_torch(self):
        return self.view().clone()

    def _write(self, fn):
        if self._readonly:
            raise TypeError(f"field '{self.name}' is derived and read-only")
        v = self.view(for_write=True)
        fn(v)
        if self._after_write:
            self._after_write()

    def from_numpy(self, a):
        t = torch()
        self._write(lambda v: v.copy_(t.as_tensor(np.asarray(a), dtype=v.dtype)))

    def fill(self, val):
        self._write(lambda v: v.fill_(val))

    def __getitem__(self, key):
        v = self.view()[key]
        return v.item() if v.dim() == 0 else v.detach().cpu().numpy()

    def __setitem__(self, key, val):
        t = torch()

        def w(v):
            v[key] = t.as_tensor(np.asarray(val), dtype=v.dtype, device=v.device)
        self._write(w)


def planes_view(buf, lead: tuple, NX: int, NY: int):
    """(P, ps) plane buffer -> reference-shaped view (NX, NY, *lead)."""
    N = NX * NY
    P = int(np.prod(lead)) if lead else 1
    v = buf[:P, :N].reshape(*lead, NX, NY) if lead else buf[0, :N].reshape(NX, NY)
    nl = len(lead)
    return v.permute(nl, nl + 1, *range(nl)) if nl else v
