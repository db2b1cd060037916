"""crdfft_execute_noalloc (include/crdfft.h): the no-allocation form of crdfft_execute, through the C-ABI, against
crdfft_execute on the same plan to the library's bit-identity bar (the same per-transform arithmetic in chunks that
fit the static scratch), on two calls in a row with the second smaller, over several chunks of the scratch, the
real transforms' four-step, and the refusals."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _exec(N, plan, src, dst, direction, noalloc):
    fn = N.lib().crdfft_execute_noalloc if noalloc else N.lib().crdfft_execute
    return fn(plan.handle, ctypes.c_void_p(src.address), ctypes.c_void_p(dst.address), int(direction), None)


@pytest.mark.parametrize("shapes", [
    [(1 << 22, 20, "c2c"), (1 << 16, 7, "c2c")],    # ten 64 MiB chunks, then a smaller call
    [(1 << 23, 2, "c2c"), (4096, 33, "c2c")],       # one transform per chunk, then a one-kernel plan
    [(1 << 20, 6, "r2c"), (1 << 17, 3, "c2r")],     # the real transforms' four-step
])
def test_noalloc_matches_execute(gpu, shapes):
    from crd_amd import _native as N

    assert N.lib().crdfft_noalloc_scratch_bytes() == 64 << 20
    rng = np.random.default_rng(5)
    for n, batch, t in shapes:
        print(f"noalloc n={n} batch={batch} {t}", flush=True)
        plan = gpu.FftPlan(n, batch, t)
        if t == "c2c":
            x = rng.standard_normal((batch, 2 * n), dtype=np.float32).view(np.complex64)
            out_shape, out_dtype, d = x.shape, np.complex64, gpu.FORWARD
        elif t == "r2c":
            x = rng.standard_normal((batch, n), dtype=np.float32)
            out_shape, out_dtype, d = (batch, n // 2 + 1), np.complex64, gpu.FORWARD
        else:
            x = np.fft.rfft(rng.standard_normal((batch, n)), axis=1).astype(np.complex64)
            out_shape, out_dtype, d = (batch, n), np.float32, gpu.INVERSE
        src = gpu.DeviceBuffer.from_array(x)
        ref = gpu.DeviceBuffer(np.empty(out_shape, out_dtype).nbytes)
        got = gpu.DeviceBuffer(ref.nbytes)
        got.fill(0x7F)  # not the answer
        assert _exec(N, plan, src, ref, d, False) == N.OK
        assert _exec(N, plan, src, got, d, True) == N.OK
        gpu.synchronize()
        a = ref.download(out_shape, out_dtype)
        b = got.download(out_shape, out_dtype)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        if t == "c2c":
            # in place too
            inplace = gpu.DeviceBuffer.from_array(x)
            assert _exec(N, plan, inplace, inplace, d, True) == N.OK
            gpu.synchronize()
            assert np.array_equal(inplace.download(out_shape, out_dtype).view(np.uint32), a.view(np.uint32))


def test_noalloc_refusals(gpu):
    from crd_amd import _native as N

    buf = gpu.DeviceBuffer(8 << 20)
    # one 2^24 transform's intermediate (128 MiB) exceeds the 64 MiB scratch
    big = gpu.FftPlan(1 << 24, 1)
    big_buf = gpu.DeviceBuffer(8 << 24)
    assert _exec(N, big, big_buf, big_buf, gpu.FORWARD, True) == N.INVALID_SIZE
    # forms with structured workspaces of their own
    split = gpu.FftPlan(1 << 16, 2, precision="split_fp16")
    assert _exec(N, split, buf, buf, gpu.FORWARD, True) == N.UNSUPPORTED
    assert N.lib().crdfft_execute_noalloc(None, None, None, 0, None) == N.INVALID_ARGUMENT
