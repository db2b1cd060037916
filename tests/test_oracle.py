"""Pins the CPU oracle (oracle/): FP64 truth against numpy's pocketfft fixtures, closed forms and invariants; the FP32
restatement against FP64 at the reference's tolerance classes; a planted mistake that must be caught."""
import os

import numpy as np
import pytest

C2C_SIZES = [1, 2, 3, 4, 5, 7, 8, 12, 16, 32, 60, 64, 97, 100, 128, 243, 256, 1000, 1024, 4096]
R2C_SIZES = [4, 8, 16, 64, 256, 1024, 4096]


def rel(y, ref):
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((y - ref).ravel()) / (den if den > 0 else 1.0))


def test_lcg_matches_fixture(po, golden):
    for n in (1, 7, 64, 4096):
        x = po.lcg_complex(n, 0x9E3779B97F4A7C15 ^ n, np.complex128)
        assert np.array_equal(x, golden[f"c2c_in_{n}"])


@pytest.mark.parametrize("n", C2C_SIZES)
def test_fft64_matches_pocketfft(po, golden, n):
    x = golden[f"c2c_in_{n}"]
    assert rel(po.fft64(x), golden[f"c2c_fwd_{n}"]) < 1e-14
    assert rel(po.fft64(x, inverse=True), golden[f"c2c_inv_{n}"]) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 8, 12, 60, 97, 128, 243, 256])
def test_dft64_matches_pocketfft(po, golden, n):
    x = golden[f"c2c_in_{n}"]
    assert rel(po.dft64(x), golden[f"c2c_fwd_{n}"]) < 1e-13
    assert rel(po.dft64(x, inverse=True), golden[f"c2c_inv_{n}"]) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 6, 16, 97, 1000, 1024, 2187, 4096, 10007, 65536])
def test_fft64_closed_forms(po, n):
    # impulse at 0 -> all ones; constant -> n at bin 0; tone at f -> n at bin f; impulse at 1 -> W_n^k
    imp = np.zeros(n, complex)
    imp[0] = 1
    assert np.max(np.abs(po.fft64(imp) - 1)) < 1e-12
    const = np.ones(n, complex)
    want = np.zeros(n, complex)
    want[0] = n
    assert np.max(np.abs(po.fft64(const) - want)) < 1e-12 * n
    f = 3 % n
    j = np.arange(n)
    tone = np.exp(2j * np.pi * ((f * j) % n) / n)
    want = np.zeros(n, complex)
    want[f] = n
    assert np.max(np.abs(po.fft64(tone) - want)) < 1e-11 * n
    if n > 1:
        imp1 = np.zeros(n, complex)
        imp1[1] = 1
        assert np.max(np.abs(po.fft64(imp1) - np.exp(-2j * np.pi * j / n))) < 1e-13 * np.log2(n) + 1e-15


@pytest.mark.parametrize("n", [8, 60, 97, 1024, 3125, 16384])
def test_fft64_invariants(po, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X, Y = po.fft64(x), po.fft64(y)
    # Parseval
    assert abs(np.sum(np.abs(X) ** 2) - n * np.sum(np.abs(x) ** 2)) < 1e-10 * n * np.sum(np.abs(x) ** 2)
    # linearity
    assert rel(po.fft64(2.5 * x - 1j * y), 2.5 * X - 1j * Y) < 1e-13
    # shift theorem: x[j - s] -> X[k] W^{s k}
    s = 5 % n
    k = np.arange(n)
    assert rel(po.fft64(np.roll(x, s)), X * np.exp(-2j * np.pi * s * k / n)) < 1e-12
    # inverse of forward is n * x
    assert rel(po.fft64(X, inverse=True), n * x) < 1e-13


@pytest.mark.parametrize("n", R2C_SIZES)
def test_rfft64_matches_pocketfft(po, golden, n):
    x = golden[f"r2c_in_{n}"]
    X = po.rfft64(x)
    assert rel(X, golden[f"r2c_out_{n}"]) < 1e-14
    assert abs(X[0].imag) == 0 and abs(X[-1].imag) < 1e-12
    assert np.max(np.abs(po.irfft64(X, n) - x)) < 1e-13


TOL_CLASS = {False: 2e-4, True: 1e-3}  # max-rel, n < 2^19 and n >= 2^19 (test_fft_kernels.cpp:151-163)


@pytest.mark.parametrize("lg", list(range(0, 18)) + [19, 20])
def test_plan32_against_fp64(po, lg):
    n = 1 << lg
    x = po.lcg_complex(n, 0x12345 ^ n)
    p = po.Plan32(n)
    for inv in (False, True):
        y = p.execute(x, inv)
        ref = po.fft64(x.astype(np.complex128), inverse=inv)
        assert po.max_rel(y, ref) < TOL_CLASS[n >= (1 << 19)]
        assert po.rel_l2(y, ref) <= 2e-6 * max(lg, 1)
        yr = p.execute_reference(x, inv)
        assert po.rel_l2(yr, ref) <= 2e-6 * max(lg, 1)


def test_plan32_known_answers(po, golden):
    """The reference self-check's integer inputs: FP32 outputs within 1.0e-7 of max|X| (fft_self_check.cpp:4-6)."""
    for n in (2, 4, 8, 16, 32, 64, 128):
        x = golden[f"kat_in_{n}"]
        p = po.Plan32(n)
        for inv, key in ((False, "fwd"), (True, "inv")):
            ref = golden[f"kat_{key}_{n}"]
            y = p.execute(x.astype(np.complex64), inv)
            assert np.max(np.abs(y - ref)) <= 1.0e-7 * np.max(np.abs(ref)) * max(1, np.log2(n) / 2)


def test_planted_twiddle_sign_is_caught(po):
    """A conjugated twiddle table (the classic sign slip) must fail the parity bar by orders of magnitude."""
    for n in (64, 4096, 1 << 17):
        x = po.lcg_complex(n, 99)
        ref = po.fft64(x.astype(np.complex128))
        good = po.Plan32(n).execute(x)
        bad = po.Plan32(n, planted=True).execute(x)
        assert po.rel_l2(good, ref) <= 2e-6 * np.log2(n)
        assert po.rel_l2(bad, ref) > 0.1


def test_plan32_element_major_batch(po):
    for n, batch in ((8, 7), (64, 19), (1024, 5)):
        x = po.lcg_complex(n * batch, n + batch)
        y = po.Plan32(n).execute_batched(x, batch)
        ref = np.fft.fft(x.reshape(n, batch).astype(np.complex128), axis=0)
        assert po.rel_l2(y.reshape(n, batch), ref) <= 2e-6 * np.log2(n)


def test_batch_split_is_bit_identical_across_jobs(po):
    """run_batch: every job count gives the serial bits (batch_arms.hpp:11-14)."""
    for n in (64, 4096, 1 << 17):
        x = po.lcg_complex(n * 9, 5).reshape(9, n)
        ys = [po.c2c32(x, threads=t) for t in (1, 2, 3, 8)]
        for y in ys[1:]:
            assert np.array_equal(y.view(np.uint32), ys[0].view(np.uint32))


@pytest.mark.parametrize("n", [3, 5, 7, 12, 97, 100, 1000, 4096 + 1, 7 * 11 * 13])
def test_bluestein32_against_fp64(po, n):
    x = po.lcg_complex(n, 17 * n)
    for inv in (False, True):
        y = po.bluestein32(x, inv)
        ref = po.fft64(x.astype(np.complex128), inverse=inv)
        if inv:
            ref = ref / n  # the reference's Bluestein inverse is normalized (bluestein.hpp:86)
        assert po.rel_l2(y, ref) <= 2e-6 * np.log2(n) * 2


@pytest.mark.parametrize("n", [4, 8, 64, 1024, 1 << 16, 1 << 20])
def test_real32_against_fp64(po, n):
    x = po.normal_real(3 * n, 11).reshape(3, n)
    X = po.rfft32(x)
    ref = po.rfft64(x.astype(np.float64))
    assert po.rel_l2(X, ref) <= 2e-6 * np.log2(n)
    back = po.irfft32(X, n)
    assert np.max(np.abs(back - x)) <= 2e-6 * np.log2(n) * np.max(np.abs(x))


def _mlp_params(inputs, hidden):
    """Zero parameters and views of each tensor of the flat layout: W1[hidden][inputs], b1[hidden],
    W2[hidden][hidden] (row = output unit), b2[hidden], w3[hidden], b3."""
    sizes = [hidden * inputs, hidden, hidden * hidden, hidden, hidden, 1]
    p = np.zeros(sum(sizes))
    parts = np.split(p, np.cumsum(sizes)[:-1])
    return p, (parts[0].reshape(hidden, inputs), parts[1], parts[2].reshape(hidden, hidden), parts[3], parts[4],
               parts[5])


def test_mlp_forward_closed_forms(po):
    """The oracle's forward pass on parameters whose output is known in closed form: all weights zero gives b3 for
    every row; one active path x_i -> unit j -> unit k -> output gives e * relu(d * relu(a x_i + c) + q) + f, and the
    same path with a negative pre-activation gives f (the ReLU clips). Inputs differ from hidden units in count and
    the path's indices are distinct, so a transposed W1 or W2, a dropped bias or a missing ReLU fails."""
    inputs, hidden = 7, 5
    rng = np.random.default_rng(3)
    x = rng.uniform(-2, 2, (13, inputs))
    p, (w1, b1, w2, b2, w3, b3) = _mlp_params(inputs, hidden)
    b3[0] = 0.625
    assert np.all(po.mlp_forward(p, inputs, hidden, x) == 0.625)
    i, j, k = 6, 1, 3
    a, c, d, q, e, f = 1.5, 0.25, -2.0, 0.5, 0.75, -0.125
    w1[j, i], b1[j], w2[k, j], b2[k], w3[k], b3[0] = a, c, d, q, e, f
    relu = lambda v: np.maximum(v, 0.0)
    want = e * relu(d * relu(a * x[:, i] + c) + q) + f
    got = po.mlp_forward(p, inputs, hidden, x)
    assert np.max(np.abs(got - want)) <= 1e-14
    assert np.any(want != f) and np.any(want == f)  # both sides of the second ReLU occur in the sample
    # inputs other than x_i have no path to the output
    x2 = x.copy()
    x2[:, [m for m in range(inputs) if m != i]] = rng.uniform(-2, 2, (13, inputs - 1))
    assert np.array_equal(po.mlp_forward(p, inputs, hidden, x2), got)


def test_mlp_gradient_matches_finite_differences(po):
    """The oracle's backpropagation against central differences of its own forward (FP64)."""
    inputs, hidden, n = 7, 5, 11
    rng = np.random.default_rng(1)
    count = hidden * inputs + hidden + hidden * hidden + hidden + hidden + 1
    p = rng.standard_normal(count) * 0.5
    x = rng.standard_normal((n, inputs))
    t = rng.standard_normal(n)
    loss, g = po.mlp_loss_grad(p, inputs, hidden, x, t)
    y = po.mlp_forward(p, inputs, hidden, x)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-12
    h = 1e-6
    num = np.empty_like(p)
    for i in range(count):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num[i] = (np.mean((po.mlp_forward(pp, inputs, hidden, x) - t) ** 2)
                  - np.mean((po.mlp_forward(pm, inputs, hidden, x) - t) ** 2)) / (2 * h)
    assert np.max(np.abs(num - g)) <= 1e-6 * max(1.0, np.max(np.abs(g)))


# ---- the spectral front end (hesap-dsp spectral.hpp), pinned by the reference's own scipy vectors ----

SPECTRAL_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "spectral_golden.npz")


@pytest.fixture(scope="module")
def spectral_golden():
    with np.load(SPECTRAL_GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_welch_oracle_matches_reference_vectors(po, spectral_golden):
    """welch_psd(x, fs=100, nperseg=64) against the reference fixture at its own gate, 1e-11 relative
    (test_spectral.cpp:33-41)."""
    g = spectral_golden
    psd = po.welch64(g["x"], 100.0, 64)
    assert psd.shape == g["welch_pxx"].shape
    assert np.all(np.abs(psd - g["welch_pxx"]) <= 1e-11 * np.abs(g["welch_pxx"]))


def test_spectrogram_oracle_matches_reference_vectors(po, spectral_golden):
    """spectrogram(x, 100, 64, 32): 15 x 33 rows, 1e-9 relative or 1e-11 absolute (test_spectral.cpp:79-88)."""
    g = spectral_golden
    s = po.periodograms64(g["x"], 100.0, 64, 32, window="hann", detrend=True)
    assert s.shape == g["spec"].shape
    d = np.abs(s - g["spec"])
    assert np.all((d <= 1e-9 * np.abs(g["spec"])) | (d <= 1e-11))


def test_periodogram_oracle_matches_reference_vectors(po, spectral_golden):
    """periodogram(x, 100): one boxcar frame of all 512 samples (test_spectral.cpp:226-231)."""
    g = spectral_golden
    p = po.periodograms64(g["x"], 100.0, len(g["x"]), 0, window="boxcar", detrend=True)[0]
    d = np.abs(p - g["periodogram"])
    assert np.all((d <= 1e-9 * np.abs(g["periodogram"])) | (d <= 1e-12))


def test_hann_and_stft_oracle(po):
    """The periodic hann of windows.hpp is 0.5 - 0.5 cos(2 pi i / m); the STFT frames are numpy's rfft of the
    windowed frames (numpy's pocketfft, FP64); frame count (len - noverlap) / hop."""
    for m in (4, 64, 1000):
        i = np.arange(m)
        assert np.allclose(po.hann64(m), 0.5 - 0.5 * np.cos(2 * np.pi * i / m), atol=1e-15, rtol=0)
        assert np.allclose(po.hann64(m, sym=True), 0.5 - 0.5 * np.cos(2 * np.pi * i / (m - 1)), atol=1e-15, rtol=0)
    x = np.random.default_rng(3).standard_normal(1000)
    for nperseg, noverlap in ((64, 32), (128, 0), (256, 201)):
        z = po.stft64(x, nperseg, noverlap)
        hop = nperseg - noverlap
        assert z.shape[0] == (len(x) - noverlap) // hop
        w = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(nperseg) / nperseg)
        ref = np.array([np.fft.rfft(x[f * hop:f * hop + nperseg] * w) for f in range(z.shape[0])])
        assert np.max(np.abs(z - ref)) <= 1e-12 * np.max(np.abs(ref))
    # a signal shorter than one frame has no frames, and Welch of it is all zeros
    assert po.stft64(x[:63], 64, 32).shape[0] == 0
    assert np.all(po.welch64(x[:63], 1.0, 64) == 0)


KINDS = ("dct2", "dct3", "dst2", "dst3")


def scipy_r2r(x, kind):
    import scipy.fft as sf

    return {0: lambda v: sf.dct(v, 2), 1: lambda v: sf.dct(v, 3), 2: lambda v: sf.dst(v, 2),
            3: lambda v: sf.dst(v, 3)}[kind](x)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 16, 64, 100, 256, 1024])
def test_dct_oracle_matches_scipy(po, n):
    """dct.hpp states its conventions as scipy norm=None (dct.hpp:5-17, verified there against scipy): the direct sums
    and the 2N-point FFT route both equal scipy.fft.dct/dst types 2 and 3 at the reference's f64 gate (1e-12 of
    1 + max|y|, test_dct.cpp:83-94), on the reference tests' LCG rows."""
    for kind in range(4):
        x = po.lcg_real(n, 0xC7C7 ^ (n * 7 + kind))
        ref = scipy_r2r(x, kind)
        d = po.dct_direct64(x, kind)
        assert np.max(np.abs(d - ref)) / (1 + np.max(np.abs(ref))) < 1e-12, (KINDS[kind], n)
        if n >= 2 and (n & (n - 1)) == 0:
            f = po.dct64(x[None, :], kind)[0]
            assert np.max(np.abs(f - ref)) / (1 + np.max(np.abs(ref))) < 1e-12, (KINDS[kind], n)


def test_dct_oracle_closed_forms_and_inverse_pairs(po):
    """Closed forms: DCT-II of a constant row c is (2Nc, 0, ...); DCT-II of cos(pi (2n+1) k0 / 2N) is N at k0 and 0
    elsewhere; DST-II of sin(pi (2n+1)(k0+1) / 2N) is N at k0 (2N at k0 = N-1); DCT-III of a unit impulse is all ones. Inverse pairs:
    dct3(dct2(x)) = dst3(dst2(x)) = 2N x (test_dct.cpp:97-131). A transposed pair (dct3 where dct2 belongs) or a
    dropped half-sample phase fails these."""
    for n in (4, 64, 1024):
        i = np.arange(n)
        y = po.dct64(np.full((1, n), 0.75), 0)[0]
        assert abs(y[0] - 2 * n * 0.75) < 1e-12 * n and np.max(np.abs(y[1:])) < 1e-12 * n
        for k0 in (1, n // 3, n - 1):
            y = po.dct64(np.cos(np.pi * (2 * i + 1) * k0 / (2 * n))[None, :], 0)[0]
            want = np.zeros(n)
            want[k0] = n
            assert np.max(np.abs(y - want)) < 1e-10 * n
            y = po.dct64(np.sin(np.pi * (2 * i + 1) * (k0 + 1) / (2 * n))[None, :], 2)[0]
            want = np.zeros(n)
            want[k0] = 2 * n if k0 == n - 1 else n  # k0 + 1 = N: the row is (-1)^n, sum of squares N, not N/2
            assert np.max(np.abs(y - want)) < 1e-10 * n
        imp = np.zeros((1, n))
        imp[0, 0] = 1.0
        assert np.allclose(po.dct64(imp, 1)[0], 1.0, atol=1e-14, rtol=0)
        x = po.lcg_real(n, 0x1234 ^ n)[None, :]
        for fwd, inv in ((0, 1), (2, 3)):
            r = po.dct64(po.dct64(x, fwd), inv)
            assert np.max(np.abs(r - 2 * n * x)) / (1 + 2 * n) < 1e-12
        # the wrong inverse does not round-trip
        assert np.max(np.abs(po.dct64(po.dct64(x, 0), 3) - 2 * n * x)) > 1e-3
