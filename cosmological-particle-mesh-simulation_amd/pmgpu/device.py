"""Minimal device array: an HBM allocation from libpm with a NumPy shape and dtype.

It exists so that the reference-layout functions can keep their data on the device between
calls.  ``numpy.asarray(devarr)`` (or ``.numpy()``) copies it back to the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    __slots__ = ("ptr", "shape", "dtype", "nbytes", "_owner", "__weakref__")

    def __init__(self, shape, dtype, ptr: int | None = None, owner=None):
        _lib.init()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        if ptr is None:
            p = C.c_void_p()
            _lib.call("pm_malloc", C.byref(p), self.nbytes)
            self.ptr = p.value
            self._owner = None
        else:
            self.ptr = int(ptr)
            self._owner = owner if owner is not None else False

    # -- construction ---------------------------------------------------------------------
    @classmethod
    def empty(cls, shape, dtype):
        return cls(shape, dtype)

    @classmethod
    def zeros(cls, shape, dtype):
        a = cls(shape, dtype)
        _lib.call("pm_memset", a.ptr, 0, a.nbytes)
        return a

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        a = cls(arr.shape, arr.dtype)
        if a.nbytes:
            _lib.call("pm_memcpy_htod", a.ptr, arr.ctypes.data, a.nbytes)
        return a

    # -- transfer -------------------------------------------------------------------------
    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _lib.call("pm_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def copy_from(self, arr) -> None:
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.shape != self.shape:
            raise ValueError(f"shape {arr.shape} != {self.shape}")
        if self.nbytes:
            _lib.call("pm_memcpy_htod", self.ptr, arr.ctypes.data, self.nbytes)

    def view(self, index: int) -> "DeviceArray":
        """Row ``index`` of the leading axis, sharing memory."""
        if not 0 <= index < self.shape[0]:
            raise IndexError(index)
        row = self.nbytes // self.shape[0]
        return DeviceArray(self.shape[1:], self.dtype, self.ptr + index * row, owner=self)

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)

    def __len__(self):
        return self.shape[0]

    @property
    def ndim(self):
        return len(self.shape)

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype}, ptr=0x{self.ptr:x})"

    def __del__(self):
        try:
            if self._owner is None and self.ptr:
                _lib.load().pm_free(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass


def as_device(x, dtype) -> tuple[DeviceArray, bool]:
    """Return (device array, uploaded?) for a NumPy array or an existing DeviceArray."""
    if isinstance(x, DeviceArray):
        if x.dtype != np.dtype(dtype):
            raise TypeError(f"device array has dtype {x.dtype}, expected {np.dtype(dtype)}")
        return x, False
    return DeviceArray.from_numpy(np.asarray(x, dtype=dtype)), True
