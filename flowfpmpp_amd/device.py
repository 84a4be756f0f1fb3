"""Device arrays in HBM, allocated and copied through the library's own C ABI
(one HIP runtime per process: the library's).

A :class:`DeviceArray` holds Fortran-ordered (column-major) data of a given
Fortran shape, exactly the layout the solver entry points take.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceArray:
    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ptr = C.c_void_p()
        check(lib().ffpm_malloc(C.byref(self._ptr), C.c_size_t(self.nbytes)))

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.asfortranarray(a, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        check(lib().ffpm_memcpy_htod(self._ptr, a.ctypes.data_as(C.c_void_p),
                                     C.c_size_t(self.nbytes)))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        check(lib().ffpm_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self._ptr,
                                     C.c_size_t(self.nbytes)))
        return out

    def zero_(self) -> "DeviceArray":
        check(lib().ffpm_memset_zero(self._ptr, C.c_size_t(self.nbytes)))
        return self

    def copy_(self, other: "DeviceArray") -> "DeviceArray":
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(lib().ffpm_memcpy_dtod(self._ptr, other._ptr, C.c_size_t(self.nbytes)))
        return self

    def free(self) -> None:
        if self._ptr and self._ptr.value:
            check(lib().ffpm_free(self._ptr))
            self._ptr = C.c_void_p()

    def __del__(self):
        try:
            if self._ptr and self._ptr.value:
                lib().ffpm_free(self._ptr)
        except Exception:
            pass


def synchronize() -> None:
    check(lib().ffpm_synchronize())


def device_count() -> int:
    n = C.c_int()
    check(lib().ffpm_device_count(C.byref(n)))
    return n.value
