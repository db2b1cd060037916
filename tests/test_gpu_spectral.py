"""The fused spectral front end (STFT, spectrogram, Welch, periodogram over the GPU real FFT) against the oracle's FP64
restatement of the reference's hesap-dsp spectral.hpp and against the reference's own scipy vectors.

Bars: FP32 output against FP64 truth, rel-L2 <= 2e-6 * log2(nperseg) over all values, and per value 50x that bound
relative, or 5e-6 of the largest value absolute for the bins near the noise floor; the reference's fixture
(spectral_refs.inc) within 2e-5 relative at the bins above 1e-3 of the peak and 2e-6 of the peak absolute below, the
FP32 counterpart of its 1e-9..1e-11 FP64 gates. Welch's fixed-order reduction
gives bit-identical reruns (the reference's "bit-identical across thread counts" property, test_spectral.cpp:43-62).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "spectral_golden.npz")


def bound(n):
    return 2e-6 * max(1.0, np.log2(n))


def close(y, ref, n, floor=5e-6):
    """rel-L2 over all values, and per value: relative where |ref| is above floor * max|ref|, absolute below."""
    y = np.asarray(y, dtype=np.float64 if not np.iscomplexobj(y) else np.complex128)
    ref = np.asarray(ref)
    scale = np.max(np.abs(ref)) if ref.size else 1.0
    err = np.linalg.norm((y - ref).ravel()) / max(np.linalg.norm(ref.ravel()), 1e-300)
    per = np.abs(y - ref) <= np.maximum(np.abs(ref) * 50 * bound(n), floor * scale)
    return err <= bound(n) and bool(np.all(per)), err


def signal(length, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(length)
    # a tone, a chirp-like term, noise and an offset (so the mean removal matters)
    return (np.sin(2 * np.pi * 0.013 * t) + 0.5 * np.cos(2 * np.pi * 1e-6 * t * t) + 0.3 * rng.standard_normal(length)
            + 1.7).astype(np.float32)


CASES = [(4, 1, 40), (8, 4, 100), (64, 32, 1000), (256, 0, 4096), (256, 201, 5000), (1024, 512, 20000),
         (4096, 1000, 50000), (8192, 4096, 70000), (16384, 8192, 100000), (32768, 16384, 200000)]


@pytest.mark.parametrize("nperseg,noverlap,length", CASES)
def test_stft_matches_fp64(gpu, po, nperseg, noverlap, length):
    from crd_amd import spectral

    x = signal(length, nperseg)
    z = spectral.stft(x, nperseg, noverlap)
    ref = po.stft64(x.astype(np.float64), nperseg, noverlap)
    assert z.shape == ref.shape
    ok, err = close(z, ref, nperseg)
    assert ok, f"nperseg={nperseg} noverlap={noverlap}: rel-L2 {err:.3e}"


@pytest.mark.parametrize("nperseg,noverlap,length", CASES)
def test_spectrogram_and_welch_match_fp64(gpu, po, nperseg, noverlap, length):
    from crd_amd import spectral

    x = signal(length, nperseg + 1)
    s = spectral.spectrogram(x, 250.0, nperseg, noverlap)
    ref = po.periodograms64(x.astype(np.float64), 250.0, nperseg, noverlap, window="hann", detrend=True)
    assert s.shape == ref.shape
    ok, err = close(s, ref, nperseg)
    assert ok, f"spectrogram nperseg={nperseg}: rel-L2 {err:.3e}"
    w = spectral.welch_psd(x, 250.0, nperseg, noverlap if noverlap else nperseg // 2)
    wref = po.welch64(x.astype(np.float64), 250.0, nperseg, noverlap if noverlap else nperseg // 2)
    ok, err = close(w, wref, nperseg)
    assert ok, f"welch nperseg={nperseg}: rel-L2 {err:.3e}"


def test_reference_vectors(gpu):
    """The reference's scipy vectors: welch_psd(x, 100, 64), spectrogram(x, 100, 64, 32), periodogram(x, 100)."""
    from crd_amd import spectral

    with np.load(GOLDEN, allow_pickle=False) as g:
        x, pxx, spec, pgram = g["x"], g["welch_pxx"], g["spec"], g["periodogram"]
    for name, got, want, n in (("welch", spectral.welch_psd(x, 100.0, 64), pxx, 64),
                               ("spectrogram", spectral.spectrogram(x, 100.0, 64, 32), spec, 64),
                               ("periodogram", spectral.periodogram(x, 100.0), pgram, 512)):
        assert got.shape == want.shape
        top = np.max(np.abs(want))
        big = np.abs(want) > 1e-3 * top
        rel = np.max(np.abs(got[big] - want[big]) / np.abs(want[big]))
        small = np.max(np.abs(got[~big] - want[~big])) / top if np.any(~big) else 0.0
        l2 = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert rel <= 2e-5 and small <= 2e-6 and l2 <= bound(n), (name, rel, small, l2)


def test_welch_is_bit_identical_and_edge_cases(gpu, po):
    from crd_amd import spectral

    x = signal(300000, 7)
    a = spectral.welch_psd(x, 1.0, 1024)
    b = spectral.welch_psd(x, 1.0, 1024)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # fewer samples than one frame: no frames, zero PSD; exactly one frame
    assert spectral.stft(x[:1000], 1024, 512).shape == (0, 513)
    assert np.all(spectral.welch_psd(x[:1000], 1.0, 1024) == 0)
    one = spectral.spectrogram(x[:1024], 1.0, 1024, 512)
    assert one.shape == (1, 513)
    ok, _ = close(one, po.periodograms64(x[:1024].astype(np.float64), 1.0, 1024, 512), 1024)
    assert ok
    # a NaN poisons the frames that contain it and no other
    y = x[:10240].copy()
    y[3000] = np.nan
    s = spectral.spectrogram(y, 1.0, 1024, 512)
    hop = 512
    bad = [f for f in range(s.shape[0]) if f * hop <= 3000 < f * hop + 1024]
    for f in range(s.shape[0]):
        assert np.all(np.isnan(s[f])) == (f in bad)


def test_spectral_rejects_bad_plans(gpu):
    import ctypes

    from crd_amd import _native as N

    h = ctypes.c_void_p()
    for nperseg, noverlap, win, flags, want in ((48, 0, 1, 0, N.INVALID_SIZE), (2, 0, 1, 0, N.INVALID_SIZE),
                                                (64, 64, 1, 0, N.INVALID_SIZE), (64, 0, 7, 0, N.INVALID_ARGUMENT),
                                                (64, 0, 1, 8, N.INVALID_ARGUMENT), (65536, 0, 1, 0, N.UNSUPPORTED)):
        assert N.lib().crdfft_spectral_plan_create(ctypes.byref(h), nperseg, noverlap, win, flags) == want
