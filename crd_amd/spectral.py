"""The spectral front end over the GPU real FFT: the reference's hesap-dsp stft / spectrogram / welch_psd /
periodogram (engine/numerics/hesap-dsp/include/crd/hesap/dsp/spectral.hpp:43-222, :334-386), each a fused launch of
libcrdfft (frame load, mean removal, window, rfft, |X|^2 / averaging epilogue).

SpectralPlan works on device buffers; the functions below take and return host arrays (float32 signals) like the
reference's, at FP32.
"""
import ctypes

import numpy as np

from . import _native as N
from ._native import check, lib
from .fft import DeviceBuffer, _ptr, _stream


class SpectralPlan:
    """Frames of `nperseg` samples (a power of two, 4..32768) at hop nperseg - noverlap; window "hann" (periodic) or
    "boxcar"; detrend subtracts each frame's mean."""

    def __init__(self, nperseg, noverlap, window="hann", detrend=False):
        self.nperseg = int(nperseg)
        self.noverlap = int(noverlap)
        self.nfreq = self.nperseg // 2 + 1
        self.handle = ctypes.c_void_p()
        win = {"hann": N.WINDOW_HANN, "boxcar": N.WINDOW_BOXCAR}[window]
        flags = N.SPECTRAL_DETREND if detrend else 0
        check(lib().crdfft_spectral_plan_create(ctypes.byref(self.handle), self.nperseg, self.noverlap, win, flags),
              f"crdfft_spectral_plan_create({nperseg}, {noverlap})")

    def frames(self, length):
        return int(lib().crdfft_spectral_frames(self.handle, int(length)))

    def stft(self, x, length, out, stream=None):
        check(lib().crdfft_stft(self.handle, _ptr(x), int(length), _ptr(out), _stream(stream)), "crdfft_stft")

    def spectrogram(self, x, length, fs, out, stream=None):
        check(lib().crdfft_spectrogram(self.handle, _ptr(x), int(length), float(fs), _ptr(out), _stream(stream)),
              "crdfft_spectrogram")

    def welch(self, x, length, fs, psd, stream=None):
        check(lib().crdfft_welch(self.handle, _ptr(x), int(length), float(fs), _ptr(psd), _stream(stream)),
              "crdfft_welch")

    def __del__(self):
        try:
            if self.handle:
                lib().crdfft_spectral_plan_destroy(self.handle)
        except Exception:
            pass


def _upload(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return x, DeviceBuffer.from_array(x)


def stft(x, nperseg, noverlap):
    """Complex [nframes x nperseg/2+1] rfft of each hann-windowed frame (no detrend, no scaling)."""
    from .fft import synchronize

    x, d = _upload(x)
    plan = SpectralPlan(nperseg, noverlap, "hann", detrend=False)
    nf = plan.frames(len(x))
    out = DeviceBuffer(max(nf, 1) * plan.nfreq * 8)
    plan.stft(d, len(x), out)
    synchronize()
    return out.download((nf, plan.nfreq), np.complex64)


def spectrogram(x, fs, nperseg, noverlap):
    """Density periodogram rows [nframes x nperseg/2+1] (hann, detrend constant)."""
    from .fft import synchronize

    x, d = _upload(x)
    plan = SpectralPlan(nperseg, noverlap, "hann", detrend=True)
    nf = plan.frames(len(x))
    out = DeviceBuffer(max(nf, 1) * plan.nfreq * 4)
    plan.spectrogram(d, len(x), fs, out)
    synchronize()
    return out.download((nf, plan.nfreq), np.float32)


def welch_psd(x, fs, nperseg, noverlap=0):
    """Welch PSD (hann, density, detrend constant); noverlap 0 means nperseg/2, as the reference's default."""
    from .fft import synchronize

    noverlap = nperseg // 2 if noverlap == 0 else noverlap
    x, d = _upload(x)
    plan = SpectralPlan(nperseg, noverlap, "hann", detrend=True)
    psd = DeviceBuffer(plan.nfreq * 4)
    plan.welch(d, len(x), fs, psd)
    synchronize()
    return psd.download((plan.nfreq,), np.float32)


def periodogram(x, fs):
    """One boxcar frame of the whole signal (a power of two long), detrend constant, density."""
    from .fft import synchronize

    x, d = _upload(x)
    plan = SpectralPlan(len(x), 0, "boxcar", detrend=True)
    out = DeviceBuffer(plan.nfreq * 4)
    plan.spectrogram(d, len(x), fs, out)
    synchronize()
    return out.download((plan.nfreq,), np.float32)
