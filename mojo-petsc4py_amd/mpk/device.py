"""Device memory owned from Python: a hipMalloc'd block with its size."""

from __future__ import annotations

import numpy as np

from ._native import check, lib


class DeviceBuffer:
    """`nbytes` of device memory, freed when the object goes away."""

    __slots__ = ("ptr", "nbytes", "__weakref__")

    def __init__(self, nbytes: int):
        self.nbytes = int(max(nbytes, 16))
        self.ptr = lib().mpk_malloc(self.nbytes)
        if not self.ptr:
            check()
            raise MemoryError(f"device allocation of {self.nbytes} bytes failed")

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes + offset > self.nbytes:
            raise ValueError("upload larger than the device buffer")
        if arr.nbytes:
            lib().mpk_memcpy_htod(self.ptr + offset, arr.ctypes.data, arr.nbytes)
            check()

    def download(self, out: np.ndarray, offset: int = 0) -> np.ndarray:
        if out.nbytes + offset > self.nbytes:
            raise ValueError("download larger than the device buffer")
        if out.nbytes:
            lib().mpk_memcpy_dtoh(out.ctypes.data, self.ptr + offset, out.nbytes)
            check()
        return out

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        return self.download(out)

    def free(self) -> None:
        if self.ptr:
            try:
                lib().mpk_free(self.ptr)
            except Exception:
                pass
            self.ptr = 0

    def __del__(self):
        self.free()
