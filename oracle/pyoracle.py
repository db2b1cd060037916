"""Test infrastructure only: numpy-facing ctypes binding of the CPU oracle (oracle.h, liboracle.so).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the crd_amd package.
Every array here is host memory; complex arrays are numpy complex64/complex128 (interleaved {re, im}, the layout
of the reference's Complex<T>, engine/numerics/hesap/include/crd/hesap/complex.hpp:27-35).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile liboracle.so with oracle/Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, sz, i, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64
        sig = {
            "oracle_lcg_fill64": (None, [vp, sz, u64]),
            "oracle_lcg_fill32": (None, [vp, sz, u64]),
            "oracle_normal_fill32": (None, [vp, sz, u64]),
            "oracle_dft64": (None, [vp, vp, sz, i]),
            "oracle_fft64": (None, [vp, sz, i]),
            "oracle_fft64_batch": (None, [vp, sz, sz, sz, i, i]),
            "oracle_rfft64": (None, [vp, vp, sz]),
            "oracle_irfft64": (None, [vp, vp, sz]),
            "oracle_c2c_f32_to_f64": (None, [vp, vp, sz, sz, i, i]),
            "oracle_plan32_create": (vp, [sz]),
            "oracle_plan32_create_planted": (vp, [sz]),
            "oracle_plan32_destroy": (None, [vp]),
            "oracle_plan32_execute": (None, [vp, vp, i]),
            "oracle_plan32_execute_batched": (None, [vp, vp, sz, i]),
            "oracle_plan32_execute_reference": (None, [vp, vp, i]),
            "oracle_c2c32_batch": (None, [vp, sz, sz, sz, i, i]),
            "oracle_rfft32_batch": (None, [vp, vp, sz, sz, i]),
            "oracle_irfft32_batch": (None, [vp, vp, sz, sz, i]),
            "oracle_bluestein32": (None, [vp, sz, i]),
            "oracle_hardware_threads": (i, []),
            "oracle_hann64": (None, [vp, sz, i]),
            "oracle_periodograms64": (sz, [vp, sz, ctypes.c_double, sz, sz, i, i, vp]),
            "oracle_welch64": (sz, [vp, sz, ctypes.c_double, sz, sz, vp]),
            "oracle_stft64": (sz, [vp, sz, sz, sz, vp]),
            "oracle_dct_direct64": (None, [vp, vp, sz, i]),
            "oracle_dct64": (None, [vp, vp, sz, sz, i]),
            "oracle_mlp_forward": (None, [vp, ctypes.c_uint32, ctypes.c_uint32, vp, sz, vp]),
            "oracle_mlp_loss_grad": (ctypes.c_double, [vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, sz, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lcg_complex(count, seed, dtype=np.complex64):
    """`count` complex values from the reference tests' LCG (test_fft.cpp:31-44): re, im alternate."""
    out = np.empty(2 * count, dtype=np.float64)
    lib().oracle_lcg_fill64(_p(out), out.size, seed & 0xFFFFFFFFFFFFFFFF)
    c = out.view(np.complex128)
    return c.astype(dtype) if dtype != np.complex128 else c


def lcg_real(count, seed, dtype=np.float64):
    """`count` reals from the same LCG stream (the reference's DCT tests fill rows this way, test_dct.cpp:23-31)."""
    out = np.empty(count, dtype=np.float64)
    lib().oracle_lcg_fill64(_p(out), out.size, seed & 0xFFFFFFFFFFFFFFFF)
    return out.astype(dtype) if dtype != np.float64 else out


def normal_real(count, seed):
    """`count` float32 standard normals (Box-Muller over the same LCG)."""
    out = np.empty(count, dtype=np.float32)
    lib().oracle_normal_fill32(_p(out), count, seed & 0xFFFFFFFFFFFFFFFF)
    return out


def dft64(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().oracle_dft64(_p(x), _p(out), x.size, int(inverse))
    return out


def fft64(x, inverse=False, threads=0):
    """FP64 truth of every row of x (shape (..., n)), unnormalized both ways."""
    y = np.array(x, dtype=np.complex128, order="C", copy=True)
    n = y.shape[-1]
    batch = y.size // n if n else 0
    if n and batch:
        lib().oracle_fft64_batch(_p(y), n, batch, n, int(inverse), threads)
    return y


def fft64_of_f32(x, inverse=False, threads=0):
    """FP64 truth of complex64 rows, widened inside the oracle."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    n = x.shape[-1]
    batch = x.size // n
    out = np.empty(x.shape, dtype=np.complex128)
    lib().oracle_c2c_f32_to_f64(_p(x), _p(out), n, batch, int(inverse), threads)
    return out


def rfft64(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    rows = x.reshape(-1, n)
    out = np.empty((rows.shape[0], n // 2 + 1), dtype=np.complex128)
    for r in range(rows.shape[0]):
        lib().oracle_rfft64(_p(rows[r]), _p(out[r]), n)
    return out.reshape(x.shape[:-1] + (n // 2 + 1,))


def frames(length, nperseg, noverlap):
    hop = nperseg - noverlap
    return 0 if length < nperseg else (length - noverlap) // hop


def hann64(m, sym=False):
    out = np.empty(m, dtype=np.float64)
    lib().oracle_hann64(_p(out), m, 1 if sym else 0)
    return out


def periodograms64(x, fs, nperseg, noverlap, window="hann", detrend=True):
    """Density periodogram rows of every frame (spectrogram), FP64."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nseg = frames(len(x), nperseg, noverlap)
    out = np.empty((max(nseg, 1), nperseg // 2 + 1), dtype=np.float64)
    got = lib().oracle_periodograms64(_p(x), len(x), float(fs), nperseg, noverlap, 1 if window == "hann" else 0,
                                      1 if detrend else 0, _p(out))
    return out[:got]


def welch64(x, fs, nperseg, noverlap=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    noverlap = nperseg // 2 if noverlap is None else noverlap
    out = np.empty(nperseg // 2 + 1, dtype=np.float64)
    lib().oracle_welch64(_p(x), len(x), float(fs), nperseg, noverlap, _p(out))
    return out


def stft64(x, nperseg, noverlap):
    x = np.ascontiguousarray(x, dtype=np.float64)
    nf = frames(len(x), nperseg, noverlap)
    out = np.empty((max(nf, 1), nperseg // 2 + 1), dtype=np.complex128)
    got = lib().oracle_stft64(_p(x), len(x), nperseg, noverlap, _p(out))
    return out[:got]


DCT2, DCT3, DST2, DST3 = 0, 1, 2, 3


def dct_direct64(x, kind):
    """The direct O(n^2) sum of one real sequence (FP64)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().oracle_dct_direct64(_p(x), _p(y), len(x), int(kind))
    return y


def dct64(x, kind):
    """Rows of a pow2 length through the 2n-point FP64 FFT route."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    y = np.empty_like(x)
    lib().oracle_dct64(_p(x), _p(y), n, x.size // n, int(kind))
    return y


def irfft64(X, n):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    rows = X.reshape(-1, n // 2 + 1)
    out = np.empty((rows.shape[0], n), dtype=np.float64)
    for r in range(rows.shape[0]):
        lib().oracle_irfft64(_p(rows[r]), _p(out[r]), n)
    return out.reshape(X.shape[:-1] + (n,))


class Plan32:
    """FftPlan<f32>-shaped restatement: execute / execute_batched / execute_reference, in place."""

    def __init__(self, n, planted=False):
        f = lib().oracle_plan32_create_planted if planted else lib().oracle_plan32_create
        self._h = f(n)
        if not self._h:
            raise ValueError(f"oracle Plan32 needs a power of two, got {n}")
        self.n = n

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_plan32_destroy(self._h)
            self._h = None

    def execute(self, x, inverse=False):
        y = np.array(x, dtype=np.complex64, copy=True)
        lib().oracle_plan32_execute(self._h, _p(y), int(inverse))
        return y

    def execute_batched(self, x, batch, inverse=False):
        y = np.array(x, dtype=np.complex64, copy=True)
        lib().oracle_plan32_execute_batched(self._h, _p(y), batch, int(inverse))
        return y

    def execute_reference(self, x, inverse=False):
        y = np.array(x, dtype=np.complex64, copy=True)
        lib().oracle_plan32_execute_reference(self._h, _p(y), int(inverse))
        return y


def c2c32(x, inverse=False, threads=0):
    """The CPU restatement over every row of x (complex64, shape (batch, n)); run_batch job split."""
    y = np.array(x, dtype=np.complex64, order="C", copy=True)
    n = y.shape[-1]
    batch = y.size // n
    lib().oracle_c2c32_batch(_p(y), n, batch, n, int(inverse), threads)
    return y


def c2c32_inplace(y, inverse=False, threads=0):
    n = y.shape[-1]
    lib().oracle_c2c32_batch(_p(y), n, y.size // n, n, int(inverse), threads)


def rfft32(x, threads=0):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[-1]
    batch = x.size // n
    out = np.empty(x.shape[:-1] + (n // 2 + 1,), dtype=np.complex64)
    lib().oracle_rfft32_batch(_p(x), _p(out), n, batch, threads)
    return out


def irfft32(X, n, threads=0):
    X = np.ascontiguousarray(X, dtype=np.complex64)
    batch = X.size // (n // 2 + 1)
    out = np.empty(X.shape[:-1] + (n,), dtype=np.float32)
    lib().oracle_irfft32_batch(_p(X), _p(out), n, batch, threads)
    return out


def bluestein32(x, inverse=False):
    y = np.array(x, dtype=np.complex64, copy=True)
    lib().oracle_bluestein32(_p(y), y.size, int(inverse))
    return y


def hardware_threads():
    return lib().oracle_hardware_threads()


def rel_l2(y, ref):
    """||y - ref||_2 / ||ref||_2 over the whole array, in FP64."""
    y = np.asarray(y, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((y - ref).ravel()) / (den if den > 0 else 1.0))


def max_rel(y, ref):
    """The reference tests' metric: max|y - ref| / (1 + max|ref|) (test_fft.cpp:68-99)."""
    y = np.asarray(y, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    return float(np.max(np.abs(y - ref)) / (1.0 + np.max(np.abs(ref)))) if ref.size else 0.0


def mlp_forward(params, inputs, hidden, x):
    """FP64 forward of the cost model (oracle/mlp.cpp)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    y = np.empty(n, dtype=np.float64)
    lib().oracle_mlp_forward(_p(p), inputs, hidden, _p(x), n, _p(y))
    return y


def mlp_loss_grad(params, inputs, hidden, x, t):
    """FP64 mean squared error and its gradient (flat, the parameters' layout)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.empty_like(p)
    loss = lib().oracle_mlp_loss_grad(_p(p), inputs, hidden, _p(x), _p(t), x.shape[0], _p(g))
    return loss, g
