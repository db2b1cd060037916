"""Real-to-real transforms on the GPU: the reference's DctPlan (engine/numerics/hesap-fft/include/crd/hesap/fft/
dct.hpp:34-170) with its four methods dct2 / dct3 / dst2 / dst3 (scipy norm=None conventions), for batches of rows
of a power-of-two length, through libcrdfft's crdfft_dct_* C-ABI.

DctPlan works on device buffers (float32 rows); dct2 / dct3 / dst2 / dst3 below take and return host arrays.
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import check, lib
from .fft import DeviceBuffer, _ptr, _stream, synchronize

DCT2, DCT3, DST2, DST3 = N.DCT2, N.DCT3, N.DST2, N.DST3


class DctPlan:
    """`batch` rows of n reals (n a power of two >= 2), rows idist / odist floats apart (0 = n).
    passes=True runs the separate passes around the real FFT plans even where the one-kernel form applies."""

    def __init__(self, n, batch=1, idist=0, odist=0, passes=False):
        self.n = int(n)
        self.batch = int(batch)
        self.handle = ctypes.c_void_p()
        flags = N.FLAG_SEPARATE_PASSES if passes else 0
        check(lib().crdfft_dct_plan_create(ctypes.byref(self.handle), self.n, self.batch, int(idist), int(odist),
                                           flags), f"crdfft_dct_plan_create(n={n}, batch={batch})")

    def size(self):
        return self.n

    def info(self):
        i = N.PlanInfo()
        check(lib().crdfft_dct_plan_get_info(self.handle, ctypes.byref(i)), "crdfft_dct_plan_get_info")
        return i.as_dict()

    def execute(self, kind, x, y, stream=None):
        check(lib().crdfft_dct_execute(self.handle, int(kind), _ptr(x), _ptr(y), _stream(stream)),
              "crdfft_dct_execute")

    def dct2(self, x, y, stream=None):
        self.execute(DCT2, x, y, stream)

    def dct3(self, x, y, stream=None):
        self.execute(DCT3, x, y, stream)

    def dst2(self, x, y, stream=None):
        self.execute(DST2, x, y, stream)

    def dst3(self, x, y, stream=None):
        self.execute(DST3, x, y, stream)

    def __del__(self):
        try:
            if self.handle:
                lib().crdfft_dct_plan_destroy(self.handle)
        except Exception:
            pass


def _host(kind, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[-1]
    rows = x.size // n if x.size else 0
    plan = DctPlan(n, rows)
    src = DeviceBuffer.from_array(x)
    dst = DeviceBuffer(x.nbytes)
    plan.execute(kind, src, dst)
    synchronize()
    return dst.download(x.shape, np.float32)


def dct2(x):
    """DCT-II of every row of x (last axis, a power of two), float32."""
    return _host(DCT2, x)


def dct3(x):
    """DCT-III of every row (the inverse of dct2 up to 2n)."""
    return _host(DCT3, x)


def dst2(x):
    """DST-II of every row."""
    return _host(DST2, x)


def dst3(x):
    """DST-III of every row (the inverse of dst2 up to 2n)."""
    return _host(DST3, x)
