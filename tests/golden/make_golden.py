"""Writes tests/golden/fft_golden.npz: input/output vectors that pin the oracle.

Outputs come from numpy.fft (numpy 2.2's pocketfft, FP64), a third-party library independent of both the oracle
and the reference. Inputs come from the reference tests' LCG (tests/numerics/hesap-fft/test_fft.cpp:31-44:
s = s*6364136223846793005 + 1442695040888963407, value = (s >> 11)/2^53*2 - 1, seed 0x9E3779B97F4A7C15 ^ n as in
test_fft.cpp:71), restated here in plain Python so the fixture also pins the oracle's LCG, and from the reference
self-check's closed-form integer pattern (src/fft_self_check.cpp:43-48: values {1,-2,2,-1,-2,1,2}[(k*(salt%5+2) +
salt*3 + k/7) % 7]).

Run: python tests/golden/make_golden.py  (a few seconds; numpy only).
"""
import os

import numpy as np

MASK = (1 << 64) - 1
C2C_SIZES = [1, 2, 3, 4, 5, 7, 8, 12, 16, 32, 60, 64, 97, 100, 128, 243, 256, 1000, 1024, 4096]
R2C_SIZES = [4, 8, 16, 64, 256, 1024, 4096]
K_VALUES = [1, -2, 2, -1, -2, 1, 2]


def lcg(count, seed):
    s = seed & MASK
    out = np.empty(count, dtype=np.float64)
    for i in range(count):
        s = (s * 6364136223846793005 + 1442695040888963407) & MASK
        out[i] = (float(s >> 11) / float(1 << 53)) * 2.0 - 1.0
    return out


def pattern(n, salt):
    return np.array([K_VALUES[(k * (salt % 5 + 2) + salt * 3 + k // 7) % 7] for k in range(n)], dtype=np.float64)


def main():
    out = {}
    for n in C2C_SIZES:
        seed = 0x9E3779B97F4A7C15 ^ n
        v = lcg(2 * n, seed)
        x = v[0::2] + 1j * v[1::2]
        out[f"c2c_in_{n}"] = x
        out[f"c2c_fwd_{n}"] = np.fft.fft(x)
        out[f"c2c_inv_{n}"] = np.fft.ifft(x) * n
    for n in R2C_SIZES:
        seed = 0x5DEECE66D ^ n
        x = lcg(n, seed)
        out[f"r2c_in_{n}"] = x
        out[f"r2c_out_{n}"] = np.fft.rfft(x)
    # the self-check's integer inputs: sizes 2..128, salt pair per size
    for i, n in enumerate([2, 4, 8, 16, 32, 64, 128]):
        x = pattern(n, 2 * i) + 1j * pattern(n, 2 * i + 1)
        out[f"kat_in_{n}"] = x
        out[f"kat_fwd_{n}"] = np.fft.fft(x)
        out[f"kat_inv_{n}"] = np.fft.ifft(x) * n
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
