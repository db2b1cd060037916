"""Parity of the plain FP32 GPU path with the oracle, through the C-ABI.

Bars: rel-L2 <= 2e-6 * log2(n) against the oracle's FP64 truth and against the oracle's FP32 restatement of the
reference's CPU algorithm; the reference's own max-norm classes (2e-4 below 2^19, 1e-3 from 2^19,
test_fft_kernels.cpp:151-163); bit-identical reruns; untouched guard words; NaN/Inf propagation and signed zero /
subnormal handling as the reference tests them (test_fft_kernels.cpp:105-135, 341-408).
"""
import os
import subprocess
import shutil

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bound(n):
    return 2e-6 * max(1.0, np.log2(n))


def run_c2c(gpu, x, inverse=False, in_place=False, flags_plan=None):
    batch, n = x.shape
    plan = flags_plan or gpu.FftPlan(n, batch)
    src = gpu.DeviceBuffer.from_array(x)
    dst = src if in_place else gpu.DeviceBuffer(x.nbytes)
    plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
    gpu.synchronize()
    return dst.download(x.shape, np.complex64)


@pytest.mark.parametrize("lg", list(range(0, 25)))
def test_c2c_matches_fp64_and_restatement(gpu, po, lg):
    n = 1 << lg
    batch = max(1, min(37, (1 << 21) // n))
    x = po.lcg_complex(n * batch, 0x12345 ^ n).reshape(batch, n)
    for inverse in (False, True):
        y = run_c2c(gpu, x, inverse)
        ref = po.fft64_of_f32(x, inverse=inverse)
        assert po.rel_l2(y, ref) <= bound(n), f"n={n} inverse={inverse}"
        assert po.max_rel(y, ref) < (1e-3 if n >= (1 << 19) else 2e-4)
        if lg <= 20:
            r32 = po.c2c32(x[: min(batch, 4)], inverse=inverse)
            assert po.rel_l2(y[: min(batch, 4)], r32) <= 2 * bound(n)


@pytest.mark.parametrize("n,batch", [(2, 1), (8, 255), (64, 7), (64, 65), (1024, 3), (4096, 5), (16384, 2),
                                     (1 << 15, 3), (1 << 16, 5), (1 << 20, 3)])
def test_odd_batches_and_in_place(gpu, po, n, batch):
    x = po.lcg_complex(n * batch, n * 31 + batch).reshape(batch, n)
    y = run_c2c(gpu, x)
    yi = run_c2c(gpu, x, in_place=True)
    assert np.array_equal(y.view(np.uint32), yi.view(np.uint32)), "in place differs from out of place"
    assert po.rel_l2(y, po.fft64_of_f32(x)) <= bound(n)


def test_rerun_is_bit_identical(gpu, po):
    for n in (4096, 1 << 16):
        x = po.lcg_complex(n * 6, 3).reshape(6, n)
        a = run_c2c(gpu, x)
        b = run_c2c(gpu, x)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_chunked_four_step_equals_one_chunk(gpu, po):
    """The four-step's chunking of the batch is scheduling: any chunk gives the same bits."""
    n, batch = 1 << 16, 9
    x = po.lcg_complex(n * batch, 8).reshape(batch, n)
    ys = []
    for chunk in (1, 2, 4, 9):
        plan = gpu.FftPlan(n, batch, tuning={"chunk": chunk})
        ys.append(run_c2c(gpu, x, flags_plan=plan))
    for y in ys[1:]:
        assert np.array_equal(y.view(np.uint32), ys[0].view(np.uint32))


def test_padded_distances_and_guard_words(gpu, po):
    """idist/odist > n: the gaps and a guard tail keep their bits (test_fft_kernels.cpp:105-135)."""
    for n, batch in ((256, 5), (1 << 16, 3)):
        pad = 24
        dist = n + pad
        x = po.lcg_complex(n * batch, 77).reshape(batch, n)
        host_in = np.full((batch, dist), np.nan + 1j * np.nan, dtype=np.complex64)
        host_in[:, :n] = x
        guard = np.frombuffer(np.uint32(0x7FC0DEAD).tobytes() * 2, dtype=np.complex64)[0]
        host_out = np.full(batch * dist + 64, guard, dtype=np.complex64)
        src = gpu.DeviceBuffer.from_array(host_in)
        dst = gpu.DeviceBuffer.from_array(host_out)
        plan = gpu.FftPlan(n, batch, idist=dist, odist=dist)
        plan.execute(src, dst, gpu.FORWARD)
        gpu.synchronize()
        out = dst.download(host_out.shape, np.complex64)
        body = out[: batch * dist].reshape(batch, dist)
        assert po.rel_l2(body[:, :n], po.fft64_of_f32(x)) <= bound(n)
        assert np.array_equal(body[:, n:].view(np.uint32), np.full((batch, pad), guard).view(np.uint32))
        assert np.array_equal(out[batch * dist:].view(np.uint32), np.full(64, guard).view(np.uint32))


def test_specials_propagate(gpu, po):
    """One NaN or Inf poisons every output of its transform and no other (test_fft_kernels.cpp:388-408);
    subnormal inputs are not flushed (test_fft_kernels.cpp:341-387)."""
    for n in (64, 4096, 1 << 16):
        x = po.lcg_complex(n * 3, 5).reshape(3, n)
        x[0, n // 3] = np.complex64(np.nan)
        x[1, 1] = np.complex64(complex(np.inf, 0.0))
        y = run_c2c(gpu, x)
        assert np.all(~np.isfinite(y[0])) and np.all(~np.isfinite(y[1]))
        assert np.all(np.isfinite(y[2]))
        tiny = np.zeros((1, n), np.complex64)
        tiny[0, 0] = np.complex64(complex(1e-40, -1e-40))  # subnormal in FP32
        yt = run_c2c(gpu, tiny)
        assert np.all(yt.real == np.float32(1e-40)) and np.all(yt.imag == np.float32(-1e-40))
        z = np.zeros((1, n), np.complex64)
        z[0, :] = np.complex64(complex(-0.0, -0.0))
        yz = run_c2c(gpu, z)
        assert np.all(yz == 0)


@pytest.mark.parametrize("n,batch", [(2, 3), (16, 19), (64, 17), (256, 16), (1024, 7), (4096, 5), (8192, 5),
                                     (8192, 17), (1 << 14, 100), (1 << 15, 3), (1 << 17, 2), (1 << 20, 3),
                                     (1 << 22, 1), (1 << 24, 1)])
def test_element_major_batched(gpu, po, n, batch):
    """execute_batched, element-major data[i*batch + t] (fft.hpp:462-470): one strided-column kernel up to n = 4096,
    the two-pass strided four-step above (the reference's engine takes every power of two, fft_kernels.cpp:593-784),
    against FP64 and the oracle's FP32 Plan32::execute_batched, both directions, odd batches included."""
    x = po.lcg_complex(n * batch, n ^ batch)
    plan = gpu.FftPlan(n, 1)
    buf = gpu.DeviceBuffer.from_array(x)
    for inverse in (False, True):
        buf.upload(x)
        plan.execute_batched(buf, batch, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        y = buf.download((n, batch), np.complex64)
        ref = po.fft64(x.reshape(n, batch).T.astype(np.complex128), inverse=inverse).T
        assert po.rel_l2(y, ref) <= bound(n)
        r32 = po.Plan32(n).execute_batched(x, batch, inverse).reshape(n, batch)
        assert po.rel_l2(y, r32) <= 2 * bound(n)


@pytest.mark.parametrize("lg", [1, 4, 7, 10, 12, 15, 17])
def test_reference_radix2_on_device(gpu, po, lg):
    n = 1 << lg
    batch = 3
    x = po.lcg_complex(n * batch, lg).reshape(batch, n)
    plan = gpu.FftPlan(n, batch)
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    for inverse in (False, True):
        plan.execute_reference(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        y = dst.download(x.shape, np.complex64)
        assert po.rel_l2(y, po.fft64_of_f32(x, inverse=inverse)) <= bound(n)
        r = po.Plan32(n).execute_reference(x[0], inverse)
        assert po.rel_l2(y[0], r) <= 2 * bound(n)
        # in place (fft.hpp:417-460 works on the caller's span): the pair-swap bit reversal, then the same stages,
        # so the bits equal the out-of-place form's
        inplace = gpu.DeviceBuffer.from_array(x)
        plan.execute_reference(inplace, None, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        assert np.array_equal(inplace.download(x.shape, np.complex64).view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("passes", [False, True])
@pytest.mark.parametrize("lg", [2, 3, 5, 8, 9, 10, 11, 13, 15, 16, 17, 20, 21])
def test_r2c_c2r(gpu, po, lg, passes):
    """R2C/C2R against FP64 and the FP32 restatement: the one-kernel forms (n <= 32768) and the separate passes."""
    n = 1 << lg
    batch = max(1, min(9, (1 << 21) // n))
    x = po.normal_real(n * batch, lg).reshape(batch, n)
    r2c = gpu.FftPlan(n, batch, "r2c", passes=passes)
    c2r = gpu.FftPlan(n, batch, "c2r", passes=passes)
    xin = gpu.DeviceBuffer.from_array(x)
    spec = gpu.DeviceBuffer(8 * (n // 2 + 1) * batch)
    back = gpu.DeviceBuffer(4 * n * batch)
    r2c.execute(xin, spec, gpu.FORWARD)
    gpu.synchronize()
    X = spec.download((batch, n // 2 + 1), np.complex64)
    ref = po.rfft64(x.astype(np.float64))
    assert po.rel_l2(X, ref) <= bound(n)
    assert po.rel_l2(X[: min(batch, 3)], po.rfft32(x[: min(batch, 3)])) <= 2 * bound(n)
    assert np.all(X[:, 0].imag == 0) and np.all(X[:, -1].imag == 0)
    c2r.execute(spec, back, gpu.INVERSE)
    gpu.synchronize()
    xb = back.download((batch, n), np.float32)
    assert np.max(np.abs(xb - x)) <= bound(n) * np.max(np.abs(x))
    # C2R of a given half spectrum against FP64 (not only the round trip)
    Y = (po.lcg_complex(batch * (n // 2 + 1), 3 * lg)).reshape(batch, n // 2 + 1)
    Y[:, 0] = Y[:, 0].real
    Y[:, -1] = Y[:, -1].real
    spec.upload(Y)
    c2r.execute(spec, back, gpu.INVERSE)
    gpu.synchronize()
    yb = back.download((batch, n), np.float32)
    refb = po.irfft64(Y.astype(np.complex128), n)
    assert po.rel_l2(yb, refb) <= bound(n)


@pytest.mark.parametrize("n,batch,pad_in,pad_out", [(4, 300, 2, 3), (512, 1001, 6, 1), (1024, 301, 4, 2),
                                                    (2048, 77, 10, 5), (32768, 9, 2, 7), (65536, 5, 2, 3),
                                                    (1 << 18, 3, 4, 1)])
def test_r2c_c2r_padded_rows(gpu, po, n, batch, pad_in, pad_out):
    """Rows idist / odist apart with gaps (many rows per workgroup, a partial last workgroup): the one-kernel forms
    write nothing outside their rows and equal the separate passes within FP32 rounding."""
    h = n // 2
    ri, co = n + pad_in, h + 1 + pad_out  # R2C: reals in, complex out
    x = po.normal_real(batch * ri, n + batch).reshape(batch, ri).astype(np.float32)
    outs = []
    for passes in (False, True):
        plan = gpu.FftPlan(n, batch, "r2c", idist=ri, odist=co, passes=passes)
        guard = np.full((batch, co), 3.0 - 1.0j, dtype=np.complex64)
        dst = gpu.DeviceBuffer.from_array(guard)
        plan.execute(gpu.DeviceBuffer.from_array(x), dst, gpu.FORWARD)
        gpu.synchronize()
        X = dst.download((batch, co), np.complex64)
        assert np.all(X[:, h + 1:] == guard[:, h + 1:])
        assert po.rel_l2(X[:, : h + 1], po.rfft64(x[:, :n].astype(np.float64))) <= bound(n)
        outs.append(X[:, : h + 1])
    assert po.rel_l2(outs[0], outs[1]) <= bound(n)
    # C2R: complex rows in (odd pad), reals out with an even row pitch
    ci, ro = h + 1 + pad_in, n + 2 * pad_out
    Y = po.lcg_complex(batch * ci, n).reshape(batch, ci)
    Y[:, 0] = Y[:, 0].real
    Y[:, h] = Y[:, h].real
    for passes in (False, True):
        plan = gpu.FftPlan(n, batch, "c2r", idist=ci, odist=ro, passes=passes)
        guard = np.full((batch, ro), 9.5, dtype=np.float32)
        dst = gpu.DeviceBuffer.from_array(guard)
        plan.execute(gpu.DeviceBuffer.from_array(Y), dst, gpu.INVERSE)
        gpu.synchronize()
        yb = dst.download((batch, ro), np.float32)
        assert np.all(yb[:, n:] == 9.5)
        assert po.rel_l2(yb[:, :n], po.irfft64(Y[:, : h + 1].astype(np.complex128), n)) <= bound(n)


def test_real_four_step_over_several_chunks(gpu, po):
    """n = 2^16 real (half length 32768 on the four-step, 512 transforms per 128 MiB chunk in both directions) over
    1030 rows: the fused row pass (R2C) and fused column pass (C2R) across two chunk boundaries and a partial last
    chunk, checked on rows either side of each boundary."""
    n, batch = 1 << 16, 1030
    h = n // 2
    x = po.normal_real(n * batch, 77).reshape(batch, n).astype(np.float32)
    r2c = gpu.FftPlan(n, batch, "r2c")
    c2r = gpu.FftPlan(n, batch, "c2r")
    assert r2c.info()["algorithm"] == 2 and r2c.info()["chunk"] < batch
    xin = gpu.DeviceBuffer.from_array(x)
    spec = gpu.DeviceBuffer(8 * (h + 1) * batch)
    back = gpu.DeviceBuffer(4 * n * batch)
    r2c.execute(xin, spec, gpu.FORWARD)
    c2r.execute(spec, back, gpu.INVERSE)
    gpu.synchronize()
    rows = [0, 1, 511, 512, 1022, 1023, 1024, 1025, batch - 1]
    X = spec.download((batch, h + 1), np.complex64)[rows]
    assert po.rel_l2(X, po.rfft64(x[rows].astype(np.float64))) <= bound(n)
    xb = back.download((batch, n), np.float32)[rows]
    assert np.max(np.abs(xb - x[rows])) <= bound(n) * np.max(np.abs(x[rows]))


def test_normalized_inverse_flag(gpu, po):
    from crd_amd import _native as N
    import ctypes

    for n in (1024, 1 << 18):
        x = po.lcg_complex(n * 2, 4).reshape(2, n)
        h = ctypes.c_void_p()
        N.check(N.lib().crdfft_plan_create(ctypes.byref(h), n, 2, 0, 0, N.C2C, N.FP32, 2))
        try:
            src = gpu.DeviceBuffer.from_array(x)
            N.check(N.lib().crdfft_execute(h, src.ptr, src.ptr, N.FORWARD, None))
            N.check(N.lib().crdfft_execute(h, src.ptr, src.ptr, N.INVERSE, None))
            gpu.synchronize()
            y = src.download(x.shape, np.complex64)
        finally:
            N.lib().crdfft_plan_destroy(h)
        assert np.max(np.abs(y - x)) <= bound(n) * 4


def test_python_mirror_fft_and_ifft_normalized(gpu, po):
    x = po.lcg_complex(3 * 512, 12).reshape(3, 512)
    y = gpu.fft(x)
    assert po.rel_l2(y, po.fft64_of_f32(x)) <= bound(512)
    back = gpu.ifft_normalized(y)
    assert np.max(np.abs(back - x)) < 1e-5


def test_error_paths(gpu):
    from crd_amd import _native as N

    with pytest.raises(N.CrdfftError) as e:
        gpu.FftPlan(12, 1)
    assert e.value.status == N.INVALID_SIZE
    with pytest.raises(N.CrdfftError) as e:
        gpu.FftPlan(64, 4, idist=32)
    assert e.value.status == N.INVALID_ARGUMENT
    r2c = gpu.FftPlan(64, 1, "r2c")
    buf = gpu.DeviceBuffer(1024)
    with pytest.raises(N.CrdfftError) as e:
        r2c.execute(buf, buf, gpu.FORWARD)  # R2C is never in place
    assert e.value.status == N.INVALID_ARGUMENT
    with pytest.raises(N.CrdfftError) as e:
        r2c.execute(buf, gpu.DeviceBuffer(1024), gpu.INVERSE)  # rfft is forward only
    assert e.value.status == N.INVALID_ARGUMENT
    c2c = gpu.FftPlan(64, 1)
    with pytest.raises(N.CrdfftError):
        c2c.execute(buf, buf, 5)
    with pytest.raises(N.CrdfftError) as e:
        c2c.execute_reference(buf, buf, 7)  # no such direction
    assert e.value.status == N.INVALID_ARGUMENT


def test_empty_batches_are_no_ops(gpu):
    """Zero transforms (the reference's run_batch over an empty range, batch_arms.hpp:46-77; a signal shorter than
    one frame, spectral.hpp:43-126): every execute returns OK and leaves the buffers as they were, except Welch,
    whose PSD of no frames is all zeros."""
    from crd_amd import _native as N
    from crd_amd import spectral
    from crd_amd.dct import DctPlan

    guard = np.full(1024, 0x7FC00001, dtype=np.uint32)  # a quiet-NaN pattern
    a, b = gpu.DeviceBuffer.from_array(guard), gpu.DeviceBuffer.from_array(guard)
    plans = [gpu.FftPlan(4096, 0), gpu.FftPlan(1 << 16, 0), gpu.FftPlan(1 << 20, 0), gpu.FftPlan(4096, 0, "r2c"),
             gpu.FftPlan(1 << 18, 0, "r2c"), gpu.FftPlan(4096, 0, "c2r"), gpu.FftPlan(1 << 18, 0, "c2r"),
             gpu.BluesteinPlan(1000, 0), gpu.BluesteinPlan(1009, 0), gpu.BluesteinPlan(1000, 0, passes=True)]
    for p in plans:
        t = p.info()["type"]
        dirs = (gpu.FORWARD, gpu.INVERSE) if t == N.C2C else ((gpu.FORWARD,) if t == N.R2C else (gpu.INVERSE,))
        for d in dirs:
            p.execute(a, b, d)
    gpu.FftPlan(4096, 8).execute_batched(a, 0)
    for passes in (False, True):
        dp = DctPlan(4096, 0, passes=passes)
        for kind in range(4):
            dp.execute(kind, a, b)
    sp = spectral.SpectralPlan(256, 128)
    assert sp.frames(100) == 0
    sp.stft(a, 100, b)
    sp.spectrogram(a, 100, 1.0, b)
    gpu.synchronize()
    assert np.array_equal(a.download(guard.shape, np.uint32), guard)
    assert np.array_equal(b.download(guard.shape, np.uint32), guard)
    sp.welch(a, 100, 1.0, b)
    gpu.synchronize()
    psd = b.download(guard.shape, np.uint32)
    assert not psd[:129].any() and np.array_equal(psd[129:], guard[129:])


@pytest.mark.parametrize("lg,batch,chunk", [(18, 40, 7), (20, 8, 2), (20, 5, 2), (22, 3, 1)])
def test_two_pipelines_match_one(gpu, lg, batch, chunk):
    """Long-column four-step plans run alternate chunks on a second stream, forked from and joined back to the
    caller's: the output is bit-identical to the same kernels on one stream (odd chunk counts, a partial last chunk,
    both directions, a second caller stream and the null stream), and matches FP64."""
    n = 1 << lg
    two = gpu.FftPlan(n, batch, tuning={"chunk": chunk, "hints": 3 | 32})
    one = gpu.FftPlan(n, batch, tuning={"chunk": chunk, "hints": 3})
    assert two.info()["hints"] & 32 and not one.info()["hints"] & 32 and two.info()["chunk"] == chunk
    x = (np.random.default_rng(lg).standard_normal((batch, 2 * n), dtype=np.float32)).view(np.complex64)
    src = gpu.DeviceBuffer.from_array(x)
    a, b = gpu.DeviceBuffer(x.nbytes), gpu.DeviceBuffer(x.nbytes)
    other = gpu.Stream()
    for direction in (gpu.FORWARD, gpu.INVERSE):
        one.execute(src, a, direction)
        for stream in (None, other):
            b.fill(0x7F, stream)
            two.execute(src, b, direction, stream)
            if stream is not None:
                stream.synchronize()
            gpu.synchronize()
            assert np.array_equal(a.download(x.shape, np.uint64), b.download(x.shape, np.uint64)), (direction, stream)
    # in place: chunk c + 1's column pass (reading its own transforms) runs beside chunk c's row pass (writing its own)
    one.execute(src, a, gpu.FORWARD)
    b.copy_from(src)
    two.execute(b, b, gpu.FORWARD)
    gpu.synchronize()
    assert np.array_equal(a.download(x.shape, np.uint64), b.download(x.shape, np.uint64)), "in place"
    gpu.FftPlan(n, batch).execute(src, b, gpu.FORWARD)  # the planner's own choice
    gpu.synchronize()
    y = b.download(x.shape, np.complex64)
    rows = min(batch, 2)
    want = np.fft.fft(x[:rows].astype(np.complex128), axis=1)
    assert np.linalg.norm(y[:rows] - want) / np.linalg.norm(want) <= 2e-6 * lg


def test_allocation_failures_are_reported(gpu):
    """CRDFFT_FAULT_ALLOC=k fails the k-th device allocation the library makes (the reference's refusing allocator,
    fft.hpp:52, exercised by its fatal tests, tests/numerics/hesap-fft/CMakeLists.txt:26-36): each plan kind, failed
    at every allocation in turn, returns OUT_OF_MEMORY and never crashes, or (a 2/3/5/7-smooth Bluestein plan, whose
    mixed-radix launch falls back to chirp-z) still builds a plan that computes the right transform."""
    from crd_amd import _native as N
    from crd_amd import spectral
    from crd_amd.dct import DctPlan

    rng = np.random.default_rng(3)
    x = (rng.standard_normal((4, 2000), dtype=np.float32)).view(np.complex64)
    makers = {
        "c2c 4096": lambda: gpu.FftPlan(4096, 8),
        "c2c 2^16": lambda: gpu.FftPlan(1 << 16, 8),
        "c2c 2^20": lambda: gpu.FftPlan(1 << 20, 2),
        "r2c 2^16": lambda: gpu.FftPlan(1 << 16, 4, "r2c"),
        "c2r 2^18": lambda: gpu.FftPlan(1 << 18, 4, "c2r"),
        "bluestein 1000": lambda: gpu.BluesteinPlan(1000, 4),
        "bluestein 1009": lambda: gpu.BluesteinPlan(1009, 4),
        "bluestein 10007 passes": lambda: gpu.BluesteinPlan(10007, 2),
        "dct 4096": lambda: DctPlan(4096, 4),
        "dct 2^16": lambda: DctPlan(1 << 16, 2),
        "spectral": lambda: spectral.SpectralPlan(1024, 512),
        "nd": lambda: gpu.NdFftPlan((16, 32, 64), 2),
    }
    try:
        for name, make in makers.items():
            failures = 0
            for k in range(1, 25):
                os.environ["CRDFFT_FAULT_ALLOC"] = str(k)
                try:
                    p = make()
                except N.CrdfftError as e:
                    assert e.status == N.OUT_OF_MEMORY, (name, k, e.status)
                    failures += 1
                    continue
                finally:
                    os.environ["CRDFFT_FAULT_ALLOC"] = "0"
                if name == "bluestein 1000":
                    buf = gpu.DeviceBuffer.from_array(x[:, :1000])
                    p.execute(buf, buf, gpu.FORWARD)
                    gpu.synchronize()
                    y = buf.download((4, 1000), np.complex64)
                    want = np.fft.fft(x[:, :1000].astype(np.complex128), axis=1)
                    assert np.linalg.norm(y - want) / np.linalg.norm(want) <= 2e-6 * 10, (name, k)
                del p
            assert failures >= 1 or name == "bluestein 1000", name
    finally:
        del os.environ["CRDFFT_FAULT_ALLOC"]
    # without the fault the same plans build and run
    p = gpu.FftPlan(1 << 16, 2)
    xin = (rng.standard_normal((2, 2 << 16), dtype=np.float32)).view(np.complex64)
    buf = gpu.DeviceBuffer.from_array(xin)
    p.execute(buf, buf, gpu.FORWARD)
    gpu.synchronize()
    y = buf.download(xin.shape, np.complex64)
    want = np.fft.fft(xin.astype(np.complex128), axis=1)
    assert np.linalg.norm(y - want) / np.linalg.norm(want) <= 2e-6 * 16


def test_self_check_fault_seam(gpu):
    """The demotion path (select_checked_kernel_level, kernel_dispatch.hpp:105-145), forced by the host fault seam
    as the reference forces its own (fault_seam.hpp:139): CRDFFT_FAULT_SELF_CHECK=1 fails the selected kernels'
    check, so the plan is created on its floor (the radix-2 reference kernels) and reports DEMOTED; =2 fails the
    floor's check too, and creation refuses the plan."""
    from crd_amd import _native as N

    for n, t in ((4096, "c2c"), (1 << 18, "c2c"), (1024, "r2c"), (1024, "c2r")):
        os.environ["CRDFFT_FAULT_SELF_CHECK"] = "2"
        try:
            with pytest.raises(N.CrdfftError) as e:
                gpu.FftPlan(n, 2, t)
            assert e.value.status == N.SELF_CHECK_FAILED
        finally:
            del os.environ["CRDFFT_FAULT_SELF_CHECK"]
        os.environ["CRDFFT_FAULT_SELF_CHECK"] = "1"
        try:
            demoted = gpu.FftPlan(n, 2, t)
        finally:
            del os.environ["CRDFFT_FAULT_SELF_CHECK"]
        assert demoted.info()["check"] == N.CHECK_DEMOTED
        assert demoted.self_check() == N.OK  # the floor, checked again without the seam
        p = gpu.FftPlan(n, 2, t)
        assert p.info()["check"] == N.CHECK_PASSED
        assert p.self_check() == N.OK


@pytest.mark.parametrize("n,t,normalize", [(64, "c2c", False), (4096, "c2c", True), (1 << 17, "c2c", False),
                                           (1 << 18, "c2c", True), (1024, "r2c", False), (1 << 16, "r2c", False),
                                           (1024, "c2r", False), (1 << 16, "c2r", False)])
def test_demoted_plan_matches_fp64(gpu, po, n, t, normalize):
    """A plan demoted to its floor computes the same transforms, against FP64: C2C both directions (in place, out
    of place, the normalized inverse with padded rows whose padding must stay untouched), R2C and C2R through the
    separate recombine passes around the floor's complex transform."""
    from crd_amd import _native as N

    batch = 3
    os.environ["CRDFFT_FAULT_SELF_CHECK"] = "1"
    try:
        if t == "c2c":
            dist = n + 8
            flags = N.FLAG_NORMALIZE_INVERSE if normalize else 0
            plan = gpu.FftPlan(n, batch, t, idist=dist, odist=dist, flags=flags)
        else:
            plan = gpu.FftPlan(n, batch, t)
    finally:
        del os.environ["CRDFFT_FAULT_SELF_CHECK"]
    assert plan.info()["check"] == N.CHECK_DEMOTED
    if t == "c2c":
        x = np.zeros((batch, dist), np.complex64)
        x[:, :n] = po.lcg_complex(n * batch, n).reshape(batch, n)
        x[:, n:] = np.complex64(7 + 7j)  # guard words
        for inverse in (False, True):
            src = gpu.DeviceBuffer.from_array(x)
            dst = gpu.DeviceBuffer.from_array(x)
            d = gpu.INVERSE if inverse else gpu.FORWARD
            plan.execute(src, dst, d)
            plan.execute(src, None, d)  # in place
            gpu.synchronize()
            ref = po.fft64_of_f32(x[:, :n], inverse=inverse)
            if inverse and normalize:
                ref = ref / n
            for buf in (dst, src):
                y = buf.download(x.shape, np.complex64)
                assert po.rel_l2(y[:, :n], ref) <= bound(n)
                assert np.array_equal(y[:, n:], x[:, n:])
    elif t == "r2c":
        x = po.normal_real(n * batch, n).reshape(batch, n)
        out = gpu.DeviceBuffer(8 * (n // 2 + 1) * batch)
        plan.execute(gpu.DeviceBuffer.from_array(x), out, gpu.FORWARD)
        gpu.synchronize()
        y = out.download((batch, n // 2 + 1), np.complex64)
        assert po.rel_l2(y, np.fft.rfft(x.astype(np.float64), axis=1)) <= bound(n)
    else:
        spec = np.fft.rfft(po.normal_real(n * batch, n).reshape(batch, n).astype(np.float64), axis=1)
        spec = spec.astype(np.complex64)
        out = gpu.DeviceBuffer(4 * n * batch)
        plan.execute(gpu.DeviceBuffer.from_array(spec), out, gpu.INVERSE)
        gpu.synchronize()
        y = out.download((batch, n), np.float32)
        assert po.rel_l2(y, np.fft.irfft(spec.astype(np.complex128), n=n, axis=1)) <= bound(n)


@pytest.mark.parametrize("name", ["wrapper_use", "reference_call_sites"])
def test_cpp_interface_runs(gpu, tmp_path, name):
    """The C++20 interface (include/crd_amd/fft.hpp) drives the library from host programs: the device-native forms,
    and call sites written exactly as the reference's tests write them over host spans (FftPlan(alloc, n),
    FftAlgorithm hints, fft / ifft_normalized with the allocator, in-place execute_reference, element-major
    execute_batched up to 8192, RealFftPlan / BluesteinPlan / NdFftPlan / DctPlan), checked against an FP64 DFT at
    the reference's f32 gate."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / name)
    src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
    libdir = os.path.join(ROOT, "crd_amd")
    subprocess.run([gxx, "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"), src, "-L", libdir, "-lcrdfft",
                    f"-Wl,-rpath,{libdir}", "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


def smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def test_bluestein_large_prime_over_two_pipelines(gpu):
    """A prime n = 1000003: the chirp-z passes over a 2^21-point plan whose long columns run two chunk pipelines
    (16 transforms per pass chunk, 8 per pipeline chunk; a last pass chunk of 4 runs on one), against FP64."""
    n, batch = 1000003, 20
    plan = gpu.BluesteinPlan(n, batch)
    assert plan.conv_size() == 1 << 21
    x = (np.random.default_rng(11).standard_normal((batch, 2 * n), dtype=np.float32)).view(np.complex64)
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    for inverse in (False, True):
        plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        y = dst.download(x.shape, np.complex64)
        for r in (0, 17, 19):
            xr = x[r].astype(np.complex128)
            ref = np.fft.ifft(xr) if inverse else np.fft.fft(xr)
            assert np.linalg.norm(y[r] - ref) / np.linalg.norm(ref) <= 2 * bound(n), (r, inverse)


@pytest.mark.parametrize("mode", ["default", "chirp_z", "passes"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 97, 100, 1000, 4097, 8191, 8192, 10007, 7 * 11 * 13 * 17,
                               3 * 5 * 7 * 11 * 13])
def test_bluestein_any_length(gpu, po, n, mode):
    """BluesteinPlan (bluestein.hpp): forward unnormalized, inverse normalized, vs FP64 and the FP32 restatement;
    2/3/5/7-smooth n as one mixed-radix launch, the chirp-z one-kernel form (m <= 16384) and its five passes."""
    batch = 3
    x = po.lcg_complex(n * batch, 7 * n + 1).reshape(batch, n)
    plan = gpu.BluesteinPlan(n, batch, passes=mode == "passes", chirp_z=mode == "chirp_z")
    direct = mode == "default" and 2 <= n <= 16384 and smooth7(n)
    info = plan.info()
    if direct:
        assert plan.conv_size() == 0 and info["workspace_bytes"] == 0
    else:
        assert plan.conv_size() >= 2 * n - 1
        one_kernel = mode != "passes" and 4 <= plan.conv_size() <= 16384
        assert (info["e"] > 0) == one_kernel and (info["workspace_bytes"] == 0) == one_kernel
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    for inverse in (False, True):
        plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        y = dst.download(x.shape, np.complex64)
        ref = po.fft64_of_f32(x, inverse=inverse)
        if inverse:
            ref = ref / n
        assert po.rel_l2(y, ref) <= 2 * bound(max(n, 2)), f"n={n} inverse={inverse}"
        r32 = np.stack([po.bluestein32(row, inverse) for row in x])
        assert po.rel_l2(y, r32) <= 4 * bound(max(n, 2))
    # in place
    plan.execute(src, src, gpu.FORWARD)
    gpu.synchronize()
    assert po.rel_l2(src.download(x.shape, np.complex64), po.fft64_of_f32(x)) <= 2 * bound(max(n, 2))


@pytest.mark.parametrize("n,batch,idist,odist", [(5, 1000, 0, 0), (3, 777, 4, 6), (100, 301, 128, 100),
                                                  (1000, 65, 1000, 1024), (6000, 9, 6001, 6000)])
def test_bluestein_one_kernel_batches(gpu, po, n, batch, idist, odist):
    """The one-kernel form over many sequences per workgroup with a partial last workgroup, and padded layouts:
    equal to the five-pass form within FP32 rounding, and nothing outside [0, n) of an output row is written."""
    idist, odist = idist or n, odist or n
    rng = np.random.default_rng(n + batch)
    xin = (rng.standard_normal((batch, idist)) + 1j * rng.standard_normal((batch, idist))).astype(np.complex64)
    src = gpu.DeviceBuffer.from_array(xin)
    guard = np.full((batch, odist), 7.0 + 3.0j, dtype=np.complex64)
    outs = []
    for passes in (False, True):
        plan = gpu.BluesteinPlan(n, batch, idist, odist, passes=passes, chirp_z=True)
        dst = gpu.DeviceBuffer.from_array(guard)
        for inverse in (False, True):
            plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
            gpu.synchronize()
            y = dst.download(guard.shape, np.complex64)
            assert np.all(y[:, n:] == guard[:, n:])
            ref = po.fft64_of_f32(xin[:, :n], inverse=inverse)
            if inverse:
                ref = ref / n
            assert po.rel_l2(y[:, :n], ref) <= 2 * bound(n), (passes, inverse)
            outs.append(y[:, :n])
    assert po.rel_l2(outs[0], outs[2]) <= 2 * bound(n) and po.rel_l2(outs[1], outs[3]) <= 2 * bound(n)


@pytest.mark.parametrize("dims", [(8, 16), (64, 64, 64), (3, 5, 7), (16, 1, 32), (5, 64), (64, 5), (2, 8192, 4),
                                  (256, 256), (1, 1), (12,)])
def test_nd_plan(gpu, po, dims):
    """NdFftPlan (nd_fft.hpp): forward = N-D DFT, inverse = 1/prod(dims) IDFT, against numpy's FP64 N-D DFT."""
    batch = 2
    total = int(np.prod(dims))
    x = po.lcg_complex(total * batch, total).reshape((batch,) + tuple(dims))
    plan = gpu.NdFftPlan(dims, batch)
    buf = gpu.DeviceBuffer.from_array(x)
    axes = tuple(range(1, len(dims) + 1))
    for inverse in (False, True):
        buf.upload(x)
        plan.execute(buf, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        y = buf.download(x.shape, np.complex64)
        xd = x.astype(np.complex128)
        ref = np.fft.ifftn(xd, axes=axes) if inverse else np.fft.fftn(xd, axes=axes)
        assert po.rel_l2(y, ref) <= 3 * bound(max(total, 2)), f"dims={dims} inverse={inverse}"


@pytest.mark.parametrize("kind", ["four_step", "fused", "bluestein", "nd"])
def test_one_plan_on_several_streams(gpu, po, kind):
    """One plan executed on three streams at once, interleaved without synchronisation: each stream has its own
    scratch, so every result equals the plan's result on one stream (the reference's plans are per thread,
    fft.hpp:17-18; here one plan may be shared)."""
    streams = [gpu.Stream() for _ in range(3)]
    if kind == "nd":
        dims, batch = (16, 64, 64), 2
        n = int(np.prod(dims))
        plan = gpu.NdFftPlan(dims, batch, transpose_all=True)
    else:
        n, batch = {"four_step": (1 << 16, 24), "fused": (1 << 16, 24), "bluestein": (3 * 5 * 7 * 11, 40)}[kind]
        if kind == "bluestein":
            plan = gpu.BluesteinPlan(n, batch)
        else:
            plan = gpu.FftPlan(n, batch, tuning={"algorithm": 5} if kind == "fused" else {"algorithm": 2, "chunk": 8})
    xs = [po.lcg_complex(n * batch, 1000 + i).reshape(batch, n) for i in range(3)]
    bufs = [gpu.DeviceBuffer.from_array(x) for x in xs]
    outs = [gpu.DeviceBuffer(x.nbytes) for x in xs]

    def run(i, s):
        if kind == "nd":
            plan.execute(outs[i], gpu.FORWARD, s)
        else:
            plan.execute(bufs[i], outs[i], gpu.FORWARD, s)

    # reference: one stream, one at a time
    want = []
    for i in range(3):
        if kind == "nd":
            outs[i].copy_from(bufs[i])
            gpu.synchronize()  # the copy runs on the null stream, which the library's streams do not wait for
        run(i, streams[0])
        streams[0].synchronize()
        want.append(outs[i].download(xs[i].shape, np.complex64))
    # three streams interleaved, several rounds
    for _ in range(4):
        if kind == "nd":
            gpu.synchronize()
            for i in range(3):
                outs[i].copy_from(bufs[i])
            gpu.synchronize()
        for i in range(3):
            run(i, streams[i])
        for s in streams:
            s.synchronize()
        for i in range(3):
            got = outs[i].download(xs[i].shape, np.complex64)
            assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), f"{kind}: stream {i} differs"
