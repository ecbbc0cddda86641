"""HBM buffers owned through the library's own C ABI (pg_dev_*).

The library links the system HIP runtime; a caller that brings its own device
memory (PETSc, hipSPARSE, another framework built on the same runtime) passes
raw device pointers to `AssemblerGPU.assemble_device` instead.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import check, lib


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().pg_dev_alloc(C.byref(p), self.nbytes), "pg_dev_alloc")
        self._p = p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @classmethod
    def zeros(cls, n: int, dtype=np.float64) -> "DeviceBuffer":
        return cls.from_numpy(np.zeros(n, dtype=dtype))

    @property
    def ptr(self) -> C.c_void_p:
        return self._p

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        check(lib().pg_dev_upload(self._p, C.c_void_p(a.ctypes.data), a.nbytes), "pg_dev_upload")

    def download(self, n: int, dtype=np.float64) -> np.ndarray:
        out = np.empty(int(n), dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("read past the buffer")
        check(lib().pg_dev_download(C.c_void_p(out.ctypes.data), self._p, out.nbytes), "pg_dev_download")
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None) -> None:
        n = other.nbytes if nbytes is None else int(nbytes)
        if n > self.nbytes or n > other.nbytes:
            raise ValueError("copy past a buffer")
        check(lib().pg_dev_copy(self._p, other._p, n), "pg_dev_copy")

    def free(self) -> None:
        if self._p:
            lib().pg_dev_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    check(lib().pg_dev_synchronize(), "pg_dev_synchronize")
