"""crd_amd: MI355X-native (gfx950) batched FFT path of crd-hesap-fft, behind its FftPlan plan/execute interface.

The transforms run in hand-written HIP kernels in libcrdfft.so (crd_amd/csrc), reached through the C-ABI of
include/crdfft.h. This package is the Python side of that ABI: ctypes bindings and a mirror of the reference's
FftPlan / RealFftPlan / BluesteinPlan / NdFftPlan / DctPlan / fft / ifft_normalized interface.
"""
from ._native import (C2C, C2R, FORWARD, INVERSE, R2C, CrdfftError, LIB_PATH, lib)  # noqa: F401
from .fft import (BluesteinPlan, DeviceBuffer, NdFftPlan, Event, FftDirection, FftPlan, RealFftPlan, Stream, describe, device_count,  # noqa
                  device_props, fft, ifft_normalized, synchronize)
from .dct import DctPlan  # noqa: F401

__all__ = [
    "FftPlan", "RealFftPlan", "BluesteinPlan", "NdFftPlan", "DctPlan", "DeviceBuffer", "Stream", "Event", "FftDirection", "describe", "device_count",
    "device_props", "fft", "ifft_normalized", "synchronize", "CrdfftError", "lib", "LIB_PATH", "FORWARD",
    "INVERSE", "C2C", "R2C", "C2R",
]
