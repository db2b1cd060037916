"""Mixed-size batches in one persistent launch (include/crdfft.h, crdfft_many_*) and the radix-plan autotuner.

The reference runs a batch of many lengths as one plan per length, every length that is not a power of two through
BluesteinPlan (engine/numerics/hesap-fft/include/crd/hesap/fft/bluestein.hpp), one job per contiguous range
(detail/batch_arms.hpp:59-77). Here one launch covers every transform; each length runs a direct radix plan.
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import check
from .fft import _ptr, _stream

from ._native import ManyInfo, RadixPlan, TuneRecord  # noqa: F401


def lib():
    return N.lib()


def radix_candidates(n):
    cnt = ctypes.c_uint32(0)
    buf = (RadixPlan * 32)()
    check(lib().crdfft_radix_candidates(n, buf, 32, ctypes.byref(cnt)), f"crdfft_radix_candidates({n})")
    return [buf[i] for i in range(min(cnt.value, 32))]


def _u64_array(values):
    a = np.ascontiguousarray(values, dtype=np.uint64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class ManyPlan:
    """Transform t of lengths[t] at element offset offsets[t] (packed when None); choices: RadixPlan per length."""

    def __init__(self, lengths, offsets=None, choices=None):
        self._lengths, lp = _u64_array(lengths)
        self._offsets, op = (None, None) if offsets is None else _u64_array(offsets)
        ch = None
        nch = 0
        if choices:
            ch = (RadixPlan * len(choices))(*choices)
            nch = len(choices)
        self.handle = ctypes.c_void_p()
        check(lib().crdfft_many_plan_create(ctypes.byref(self.handle), lp, op, len(self._lengths), ch, nch, 0),
              "crdfft_many_plan_create")

    def info(self):
        i = ManyInfo()
        check(lib().crdfft_many_plan_get_info(self.handle, ctypes.byref(i)), "crdfft_many_plan_get_info")
        return {name: getattr(i, name) for name, _ in ManyInfo._fields_}

    def execute(self, data_in, data_out=None, direction=N.FORWARD, stream=None):
        out = data_in if data_out is None else data_out
        check(lib().crdfft_many_execute(self.handle, _ptr(data_in), _ptr(out), int(direction), _stream(stream)),
              "crdfft_many_execute")

    def capture_graph(self, data_in, data_out, direction, stream):
        check(lib().crdfft_many_graph_capture(self.handle, _ptr(data_in), _ptr(data_out), int(direction),
                                              _stream(stream)), "crdfft_many_graph_capture")

    def launch_graph(self, stream=None):
        check(lib().crdfft_many_graph_launch(self.handle, _stream(stream)), "crdfft_many_graph_launch")

    def __del__(self):
        try:
            if self.handle:
                lib().crdfft_many_plan_destroy(self.handle)
        except Exception:
            pass


def autotune(lengths, reps=5):
    """Times every radix-plan candidate of every distinct length; returns (choices, records)."""
    arr, lp = _u64_array(lengths)
    distinct = len(set(int(x) for x in arr))
    choices = (RadixPlan * distinct)()
    nch = ctypes.c_uint64(0)
    cap = distinct * 32
    recs = (TuneRecord * cap)()
    nrec = ctypes.c_uint64(0)
    check(lib().crdfft_many_autotune(lp, len(arr), reps, choices, distinct, ctypes.byref(nch), recs, cap,
                                     ctypes.byref(nrec)), "crdfft_many_autotune")
    return [choices[i] for i in range(nch.value)], [recs[i] for i in range(nrec.value)]
