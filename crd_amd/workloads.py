"""Synthetic workloads of the batched FFT path (seeded, host-side generators only)."""
import numpy as np


def smooth_sizes(lo=64, hi=16384, primes=(2, 3, 5, 7)):
    """Every integer in [lo, hi] whose prime factors are all in `primes` (365 for 7-smooth in [64, 16384])."""
    out = []
    for n in range(lo, hi + 1):
        m = n
        for p in primes:
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


def mixed_batch(total=20000, mean_batch=8, lo=64, hi=16384, seed=20000):
    """The mixed-size batch: lengths drawn log-uniformly over the 7-smooth sizes in [lo, hi] (a uniform draw in
    log(n), snapped to the nearest size in log), each with a Poisson(mean_batch) count of transforms (a zero count
    draws again), until `total` transforms; the last group is cut to fit. Returns the per-transform lengths in draw
    order (each group contiguous)."""
    sizes = np.array(smooth_sizes(lo, hi), dtype=np.int64)
    logs = np.log(sizes.astype(np.float64))
    rng = np.random.default_rng(seed)
    lengths = []
    while len(lengths) < total:
        u = rng.uniform(np.log(lo), np.log(hi))
        n = int(sizes[int(np.argmin(np.abs(logs - u)))])
        b = int(rng.poisson(mean_batch))
        if b == 0:
            continue
        lengths.extend([n] * min(b, total - len(lengths)))
    return np.array(lengths, dtype=np.int64)


def uniform_complex(count, seed):
    """count complex64 values with re, im i.i.d. U[-1, 1) (seeded)."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1.0, 1.0, 2 * count).astype(np.float32)
    return v.view(np.complex64)
