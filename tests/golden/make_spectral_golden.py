"""Writes tests/golden/spectral_golden.npz: the reference's own Welch / spectrogram / periodogram vectors.

The values are the data arrays of the reference's test fixture tests/numerics/hesap-dsp/spectral_refs.inc (generated
there by scipy.signal.welch / spectrogram / periodogram, FP64), read as numbers from that file and stored unchanged:
  x            ref_welch_x      (512 samples; the signal of test_spectral.cpp:22-41)
  welch_pxx    ref_welch_pxx    welch_psd(x, fs=100, nperseg=64), noverlap 32, hann, density, detrend constant
  spec         ref_spec         spectrogram(x, fs=100, nperseg=64, noverlap=32): 15 x 33 rows (test_spectral.cpp:79)
  periodogram  ref_periodogram  periodogram(x, fs=100): boxcar, one segment of 512 (test_spectral.cpp:226)
Tolerances of the reference's own gates: welch 1e-11 relative; spectrogram / periodogram 1e-9 relative or 1e-11 /
1e-12 absolute near zero (test_spectral.cpp:41, :88, :231).

Run once in a checkout that holds the reference at /root/reference: python tests/golden/make_spectral_golden.py
"""
import os
import re

import numpy as np

SRC = "/root/reference/tests/numerics/hesap-dsp/spectral_refs.inc"
NAMES = {"ref_welch_x": "x", "ref_welch_pxx": "welch_pxx", "ref_spec": "spec", "ref_periodogram": "periodogram"}


def main():
    text = open(SRC).read()
    out = {}
    for m in re.finditer(r"inline const double (\w+)\[\] = \{([^}]*)\}", text):
        if m.group(1) in NAMES:
            vals = [float(v) for v in m.group(2).replace("\n", " ").split(",") if v.strip()]
            out[NAMES[m.group(1)]] = np.array(vals, dtype=np.float64)
    ints = dict((k, int(v)) for k, v in re.findall(r"inline const int (\w+) = (-?\d+);", text))
    out["spec"] = out["spec"].reshape(ints["ref_spec_nseg"], ints["ref_spec_nfreq"])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spectral_golden.npz")
    np.savez_compressed(path, **out)
    print(path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
