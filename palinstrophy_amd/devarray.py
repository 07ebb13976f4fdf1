"""CuPy-style device arrays over torch CUDA tensors.

The reference's callers treat the DNS state's arrays as CuPy arrays and its
``xp`` attribute as the CuPy module (turbo_main.py:900-921, 1280-1281,
1413-1445, 1947-1959, 2015-2073).  ``DeviceArray`` is a torch tensor subclass
that keeps the CuPy/NumPy idioms those callers use:

  * ``a.get()``            -> host ``numpy.ndarray`` (CuPy)
  * ``numpy.asarray(a)``   -> host copy (NumPy protocol)
  * ``a.size``             -> number of elements as an int (NumPy/CuPy); it is
                              still callable, so torch's ``a.size()`` and
                              ``a.size(dim)`` keep working
  * ``a.astype(dtype)``    -> converted copy

``xp`` (this module, as ``DnsState.xp``) is the subset of the CuPy namespace the
callers use, on the current CUDA device: ``asarray`` places host data on the
device (torch's own ``asarray`` would leave it on the host), ``sum`` takes a
``dtype``, ``transpose`` takes an axes tuple, ``stack`` promotes mixed dtypes,
``get_default_memory_pool().free_all_blocks()`` releases cached blocks.
Results of torch operations on a ``DeviceArray`` are ``DeviceArray`` again.
"""
from __future__ import annotations

import numpy as np
import torch

float32 = torch.float32
float64 = torch.float64
int32 = torch.int32
int64 = torch.int64
uint8 = torch.uint8
complex64 = torch.complex64
complex128 = torch.complex128
bool_ = torch.bool
pi = float(np.pi)

_NP_TO_TORCH = {
    np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
    np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
    np.dtype(np.uint8): torch.uint8, np.dtype(np.complex64): torch.complex64,
    np.dtype(np.complex128): torch.complex128, np.dtype(np.bool_): torch.bool,
}


def _dtype(dt):
    """torch dtype of a torch / NumPy dtype, NumPy scalar type or dtype name."""
    if dt is None or isinstance(dt, torch.dtype):
        return dt
    return _NP_TO_TORCH[np.dtype(dt)]


class _Size(int):
    """Element count as an int (``a.size`` in NumPy / CuPy) that is also
    callable like ``torch.Tensor.size``."""

    def __new__(cls, shape):
        n = 1
        for s in shape:
            n *= int(s)
        obj = super().__new__(cls, n)
        obj._shape = shape
        return obj

    def __call__(self, dim=None):
        return self._shape if dim is None else self._shape[dim]


class DeviceArray(torch.Tensor):
    """A torch CUDA tensor with the CuPy idioms the reference's callers use."""

    def get(self) -> np.ndarray:
        return torch.Tensor.numpy(torch.Tensor.cpu(self.detach().as_subclass(torch.Tensor)).resolve_conj())

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    @property
    def size(self):
        return _Size(self.shape)

    def astype(self, dtype, copy=True):
        return self.to(_dtype(dtype), copy=copy)


def wrap(t: torch.Tensor) -> DeviceArray:
    """View a tensor as a DeviceArray (no copy)."""
    return t if isinstance(t, DeviceArray) else t.as_subclass(DeviceArray)


def asarray(a, dtype=None) -> DeviceArray:
    """cupy.asarray: device array of `a` (a view when `a` already is one of that dtype)."""
    dt = _dtype(dtype)
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.to("cuda")
        return wrap(t if dt is None or t.dtype == dt else t.to(dt))
    arr = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(arr)).to("cuda")
    return wrap(t if dt is None else t.to(dt))


def asnumpy(a) -> np.ndarray:
    return a.get() if isinstance(a, DeviceArray) else np.asarray(a)


def zeros(shape, dtype=float64) -> DeviceArray:
    return wrap(torch.zeros(shape, dtype=_dtype(dtype), device="cuda"))


def empty(shape, dtype=float64) -> DeviceArray:
    return wrap(torch.empty(shape, dtype=_dtype(dtype), device="cuda"))


def sum(a, axis=None, dtype=None):  # noqa: A001 - the CuPy name
    a = asarray(a)
    dt = _dtype(dtype)
    if axis is None:
        return wrap(torch.sum(a, dtype=dt))
    return wrap(torch.sum(a, dim=axis, dtype=dt))


def transpose(a, axes=None) -> DeviceArray:
    a = asarray(a)
    if axes is None:
        axes = tuple(range(a.dim() - 1, -1, -1))
    return wrap(a.permute(*axes))


def stack(arrays, axis=0) -> DeviceArray:
    ts = [asarray(x) for x in arrays]
    dt = ts[0].dtype
    for t in ts[1:]:
        dt = torch.promote_types(dt, t.dtype)
    return wrap(torch.stack([t.to(dt) for t in ts], dim=axis))


def sqrt(a):
    return wrap(torch.sqrt(asarray(a)))


def floor(a):
    return wrap(torch.floor(asarray(a)))


def abs(a):  # noqa: A001 - the CuPy name
    return wrap(torch.abs(asarray(a)))


def clip(a, lo, hi):
    return wrap(torch.clamp(asarray(a), lo, hi))


def where(cond, x, y):
    c = asarray(cond)
    x = x if isinstance(x, torch.Tensor) else torch.as_tensor(x, device="cuda")
    y = y if isinstance(y, torch.Tensor) else torch.as_tensor(y, device="cuda")
    return wrap(torch.where(c, x, y))


def bincount(x, weights=None, minlength=0):
    w = None if weights is None else asarray(weights)
    return wrap(torch.bincount(asarray(x), weights=w, minlength=minlength))


def meshgrid(*xs, indexing="xy"):
    return tuple(wrap(t) for t in torch.meshgrid(*[asarray(x) for x in xs], indexing=indexing))


class _Fft:
    """cupy.fft subset (device transforms through torch.fft)."""

    @staticmethod
    def fft2(a):
        return wrap(torch.fft.fft2(asarray(a)))

    @staticmethod
    def ifft2(a):
        return wrap(torch.fft.ifft2(asarray(a)))

    @staticmethod
    def rfft2(a):
        return wrap(torch.fft.rfft2(asarray(a)))

    @staticmethod
    def irfft2(a, s=None):
        return wrap(torch.fft.irfft2(asarray(a), s=s))

    @staticmethod
    def fftfreq(n, d=1.0):
        return wrap(torch.fft.fftfreq(n, d=d, device="cuda", dtype=torch.float64))


fft = _Fft()


class _Pool:
    def free_all_blocks(self):
        torch.cuda.empty_cache()

    def used_bytes(self):
        return int(torch.cuda.memory_allocated())


def get_default_memory_pool():
    return _Pool()
