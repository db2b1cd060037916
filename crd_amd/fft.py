"""Python mirror of the reference's FFT plan interface over the crdfft C-ABI.

Names, argument meaning and error behaviour follow engine/numerics/hesap-fft/include/crd/hesap/fft/:
  FftPlan(n)                  fft.hpp:220-304  (a device plan; `batch` transforms per execute)
  FftPlan.execute             fft.hpp:409-413  (device buffers; in place when out is in)
  FftPlan.execute_batched     fft.hpp:462-470  (element-major, in place)
  FftPlan.execute_reference   fft.hpp:415-460  (radix-2 reference, in place or out of place)
  FftPlan.size                fft.hpp:405
  RealFftPlan.rfft / .irfft   real_fft.hpp:41-93 (irfft scales by 1/n)
  fft(), ifft_normalized()    fft.hpp:1005-1029 (host arrays in, host arrays out)
Contract violations raise CrdfftError with the library's status (the reference asserts instead).
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import CrdfftError, check

FftDirection = type("FftDirection", (), {"Forward": N.FORWARD, "Inverse": N.INVERSE})


def device_count():
    c = ctypes.c_int32(0)
    st = N.lib().crdfft_device_count(ctypes.byref(c))
    return c.value if st == N.OK else 0


def device_props(device=0):
    p = N.DeviceProps()
    check(N.lib().crdfft_device_props_get(device, ctypes.byref(p)), "crdfft_device_props_get")
    return {
        "name": p.name.decode(),
        "arch": p.arch.decode(),
        "compute_units": p.compute_units,
        "clock_khz": p.clock_khz,
        "memory_clock_khz": p.memory_clock_khz,
        "memory_bus_width": p.memory_bus_width,
        "global_mem_bytes": p.global_mem_bytes,
        "l2_bytes": p.l2_bytes,
        "max_lds_per_block": p.max_lds_per_block,
    }


class DeviceBuffer:
    """A device allocation of `nbytes` bytes (hipMalloc through the library)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(N.lib().crdfft_malloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "crdfft_malloc")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    @property
    def address(self):
        return self.ptr.value

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        check(N.lib().crdfft_memcpy_h2d(ctypes.c_void_p(self.address + offset), a.ctypes.data_as(ctypes.c_void_p),
                                        a.nbytes), "crdfft_memcpy_h2d")

    def download(self, shape, dtype, offset=0):
        out = np.empty(shape, dtype=dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download past the end of the buffer")
        check(N.lib().crdfft_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.address + offset),
                                        out.nbytes), "crdfft_memcpy_d2h")
        return out

    def fill(self, value=0, stream=None):
        check(N.lib().crdfft_fill_bytes(self.ptr, int(value), self.nbytes, _stream(stream)), "crdfft_fill_bytes")

    def copy_from(self, other, nbytes=None, stream=None):
        nb = self.nbytes if nbytes is None else int(nbytes)
        check(N.lib().crdfft_memcpy_d2d(self.ptr, other.ptr, nb, _stream(stream)), "crdfft_memcpy_d2d")

    def free(self):
        if self.ptr is not None and self.ptr.value:
            N.lib().crdfft_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(N.lib().crdfft_stream_create(ctypes.byref(self.handle)), "crdfft_stream_create")

    def synchronize(self):
        check(N.lib().crdfft_stream_synchronize(self.handle), "crdfft_stream_synchronize")

    def __del__(self):
        try:
            if self.handle:
                N.lib().crdfft_stream_destroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(N.lib().crdfft_event_create(ctypes.byref(self.handle)), "crdfft_event_create")

    def record(self, stream=None):
        check(N.lib().crdfft_event_record(self.handle, _stream(stream)), "crdfft_event_record")

    def synchronize(self):
        check(N.lib().crdfft_event_synchronize(self.handle), "crdfft_event_synchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(N.lib().crdfft_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle), "crdfft_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                N.lib().crdfft_event_destroy(self.handle)
        except Exception:
            pass


def synchronize():
    check(N.lib().crdfft_device_synchronize(), "crdfft_device_synchronize")


def _stream(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


def _ptr(buf):
    if isinstance(buf, DeviceBuffer):
        return buf.ptr
    return ctypes.c_void_p(int(buf))


_TYPES = {"c2c": N.C2C, "r2c": N.R2C, "c2r": N.C2R}
_PRECISIONS = {"fp32": N.FP32, "split_bf16": N.SPLIT_BF16, "split_fp16": N.SPLIT_FP16}


def describe(n, batch=1, type="c2c", precision="fp32"):
    """The planner's choice for a shape, without a device."""
    info = N.PlanInfo()
    check(N.lib().crdfft_plan_describe(n, batch, _TYPES[type], _PRECISIONS[precision], ctypes.byref(info)),
          "crdfft_plan_describe")
    return info.as_dict()


class FftPlan:
    """A device plan of `batch` length-n transforms (contiguous at distance idist/odist, 0 = packed)."""

    def __init__(self, n, batch=1, type="c2c", precision="fp32", idist=0, odist=0, self_check=True, tuning=None,
                 passes=False, flags=0):
        """passes=True (R2C/C2R): the separate recombine pass even where the one-kernel form applies; flags: more
        crdfft plan flags (CRDFFT_FLAG_NORMALIZE_INVERSE)."""
        self.n = int(n)
        self.batch = int(batch)
        self.type = type
        self.handle = ctypes.c_void_p()
        flags = int(flags) | (N.FLAG_NONE if self_check else N.FLAG_NO_SELF_CHECK)
        if passes:
            flags |= N.FLAG_SEPARATE_PASSES
        if tuning is None:
            st = N.lib().crdfft_plan_create(ctypes.byref(self.handle), self.n, self.batch, idist, odist,
                                            _TYPES[type], _PRECISIONS[precision], flags)
        else:
            t = N.Tuning(**{k: v for k, v in tuning.items()})
            st = N.lib().crdfft_plan_create_tuned(ctypes.byref(self.handle), self.n, self.batch, idist, odist,
                                                  _TYPES[type], _PRECISIONS[precision], flags, ctypes.byref(t))
        check(st, f"crdfft_plan_create(n={n}, batch={batch}, {type}, {precision})")

    def size(self):
        return self.n

    def info(self):
        info = N.PlanInfo()
        check(N.lib().crdfft_plan_get_info(self.handle, ctypes.byref(info)), "crdfft_plan_get_info")
        return info.as_dict()

    def execute(self, data_in, data_out=None, direction=N.FORWARD, stream=None):
        out = data_in if data_out is None else data_out
        check(N.lib().crdfft_execute(self.handle, _ptr(data_in), _ptr(out), int(direction), _stream(stream)),
              "crdfft_execute")

    def execute_profiled(self, data_in, data_out=None, direction=N.FORWARD, stream=None, acc=None):
        """execute() with hipEvents around every launch; returns the accumulated per-phase times."""
        acc = N.PhaseTimes() if acc is None else acc
        out = data_in if data_out is None else data_out
        check(N.lib().crdfft_execute_profiled(self.handle, _ptr(data_in), _ptr(out), int(direction), _stream(stream),
                                              ctypes.byref(acc)), "crdfft_execute_profiled")
        return acc

    def execute_batched(self, data, batch, direction=N.FORWARD, stream=None):
        check(N.lib().crdfft_execute_batched(self.handle, _ptr(data), int(batch), int(direction), _stream(stream)),
              "crdfft_execute_batched")

    def execute_reference(self, data_in, data_out=None, direction=N.FORWARD, stream=None):
        """The radix-2 reference transform (fft.hpp:415-460); data_out None runs it in place."""
        out = data_in if data_out is None else data_out
        check(N.lib().crdfft_execute_reference(self.handle, _ptr(data_in), _ptr(out), int(direction),
                                               _stream(stream)), "crdfft_execute_reference")

    def self_check(self):
        return N.lib().crdfft_self_check(self.handle)

    def destroy(self):
        if self.handle is not None and self.handle.value:
            N.lib().crdfft_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BluesteinPlan(FftPlan):
    """Any-length C2C (bluestein.hpp:29-136): chirp-z over a power-of-two plan of m >= 2n-1; inverse scaled 1/n."""

    def __init__(self, n, batch=1, idist=0, odist=0, passes=False, chirp_z=False):
        """passes=True: the five separate chirp-z passes even where the one-kernel form (m <= 16384) applies;
        chirp_z=True: the chirp-z algorithm even for the 2/3/5/7-smooth n that otherwise run as one mixed-radix
        launch."""
        self.n = int(n)
        self.batch = int(batch)
        self.type = "c2c"
        self.handle = ctypes.c_void_p()
        flags = (N.FLAG_SEPARATE_PASSES if passes else 0) | (N.FLAG_CHIRP_Z if chirp_z else 0)
        check(N.lib().crdfft_bluestein_plan_create(ctypes.byref(self.handle), self.n, self.batch, idist, odist, flags),
              f"crdfft_bluestein_plan_create(n={n}, batch={batch})")

    def conv_size(self):
        return self.info()["n1"]


class NdFftPlan:
    """Row-major N-D C2C of `batch` volumes in place (nd_fft.hpp:28-196): inverse scaled by 1/prod(dims)."""

    def __init__(self, dims, batch=1, transpose_all=False):
        self.dims = [int(d) for d in dims]
        arr = (ctypes.c_uint64 * len(self.dims))(*self.dims)
        self.handle = ctypes.c_void_p()
        flags = N.ND_TRANSPOSE_ALL if transpose_all else 0
        check(N.lib().crdfft_nd_plan_create(ctypes.byref(self.handle), arr, len(self.dims), int(batch), flags),
              f"crdfft_nd_plan_create({self.dims})")

    def execute(self, data, direction=N.FORWARD, stream=None):
        check(N.lib().crdfft_nd_execute(self.handle, _ptr(data), int(direction), _stream(stream)),
              "crdfft_nd_execute")

    def __del__(self):
        try:
            if self.handle:
                N.lib().crdfft_nd_plan_destroy(self.handle)
        except Exception:
            pass


class RealFftPlan:
    """real[n] <-> complex[n/2+1] for `batch` transforms (real_fft.hpp:18-101)."""

    def __init__(self, n, batch=1, self_check=True):
        self.n = int(n)
        self.batch = int(batch)
        self._r2c = FftPlan(n, batch, "r2c", self_check=self_check)
        self._c2r = FftPlan(n, batch, "c2r", self_check=self_check)

    def size(self):
        return self.n

    def rfft(self, real_in, out, stream=None):
        self._r2c.execute(real_in, out, N.FORWARD, stream)

    def irfft(self, data_in, real_out, stream=None):
        self._c2r.execute(data_in, real_out, N.INVERSE, stream)


def fft(data, direction=N.FORWARD):
    """One-shot transform of host complex64 rows (shape (..., n)), unnormalized (fft.hpp:1005-1012)."""
    x = np.ascontiguousarray(data, dtype=np.complex64)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    plan = FftPlan(n, batch)
    buf = DeviceBuffer.from_array(x)
    plan.execute(buf, buf, direction)
    synchronize()
    return buf.download(x.shape, np.complex64)


def ifft_normalized(data):
    """Inverse scaled by 1/n (fft.hpp:1015-1029)."""
    y = fft(data, N.INVERSE)
    n = y.shape[-1]
    return (y * np.float32(1.0 / n)).astype(np.complex64) if n else y
