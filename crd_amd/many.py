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

MAX_STAGES = 16


class RadixPlan(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("nstages", ctypes.c_uint32), ("radix", ctypes.c_uint32 * MAX_STAGES),
                ("group", ctypes.c_uint32)]

    def radices(self):
        return [int(self.radix[i]) for i in range(self.nstages)]

    def as_dict(self):
        return {"n": int(self.n), "radices": self.radices(), "group": int(self.group)}


class ManyInfo(ctypes.Structure):
    _fields_ = [("transforms", ctypes.c_uint64), ("lengths", ctypes.c_uint64), ("tiles", ctypes.c_uint64),
                ("points", ctypes.c_uint64), ("workgroups", ctypes.c_uint32), ("imbalance", ctypes.c_double)]


class TuneRecord(ctypes.Structure):
    _fields_ = [("plan", RadixPlan), ("candidate", ctypes.c_uint32), ("transforms", ctypes.c_uint32),
                ("us_per_transform", ctypes.c_double)]


def _bind():
    L = N.lib()
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
    sigs = {
        "crdfft_radix_candidates": (i32, [u64, ctypes.POINTER(RadixPlan), u32, ctypes.POINTER(u32)]),
        "crdfft_many_plan_create": (i32, [ctypes.POINTER(vp), ctypes.POINTER(u64), ctypes.POINTER(u64), u64,
                                          ctypes.POINTER(RadixPlan), u64, u32]),
        "crdfft_many_plan_destroy": (i32, [vp]),
        "crdfft_many_plan_get_info": (i32, [vp, ctypes.POINTER(ManyInfo)]),
        "crdfft_many_execute": (i32, [vp, vp, vp, i32, vp]),
        "crdfft_many_graph_capture": (i32, [vp, vp, vp, i32, vp]),
        "crdfft_many_graph_launch": (i32, [vp, vp]),
        "crdfft_many_autotune": (i32, [ctypes.POINTER(u64), u64, u32, ctypes.POINTER(RadixPlan), u64,
                                       ctypes.POINTER(u64), ctypes.POINTER(TuneRecord), u64, ctypes.POINTER(u64)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


_L = None


def lib():
    global _L
    if _L is None:
        _L = _bind()
    return _L


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
    cap = distinct * 16
    recs = (TuneRecord * cap)()
    nrec = ctypes.c_uint64(0)
    check(lib().crdfft_many_autotune(lp, len(arr), reps, choices, distinct, ctypes.byref(nch), recs, cap,
                                     ctypes.byref(nrec)), "crdfft_many_autotune")
    return [choices[i] for i in range(nch.value)], [recs[i] for i in range(nrec.value)]
