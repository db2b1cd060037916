"""The GPU real-to-real transforms (crdfft_dct_*, the reference's DctPlan, dct.hpp:34-170) against the FP64 oracle.

Bars: the reference's own f32 gate, max|y - ref| / (1 + max|ref|) < 1e-4 (test_dct.cpp:88-95), at every length, and
rel-L2 <= 2e-6 * log2(n) against FP64 truth (the direct O(n^2) sums up to n = 1024, the oracle's 2n-point FFT route
above); the one-kernel form and the separate passes agree to that bound; dct3(dct2(x)) = 2n x and
dst3(dst2(x)) = 2n x to the same bound at the full sizes; a rerun is bit-identical (test_dct.cpp:134-146).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = ("dct2", "dct3", "dst2", "dst3")


def bound(n):
    return 2e-6 * max(1.0, np.log2(n))


def truth(po, x, kind):
    n = x.shape[-1]
    if n <= 1024:
        return np.stack([po.dct_direct64(row, kind) for row in x.astype(np.float64)])
    return po.dct64(x.astype(np.float64), kind)


def run(gpu, x, kind, passes=False, idist=0, odist=0, in_place=False):
    from crd_amd.dct import DctPlan

    n = x.shape[-1]
    plan = DctPlan(n, x.shape[0], idist, odist, passes=passes)
    src = gpu.DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32))
    dst = src if in_place else gpu.DeviceBuffer(x.shape[0] * (odist or n) * 4)
    plan.execute(kind, src, dst)
    gpu.synchronize()
    return plan, dst.download((x.shape[0], odist or n), np.float32)


def check(y, ref, n, what):
    gate = np.max(np.abs(y - ref)) / (1 + np.max(np.abs(ref)))
    l2 = np.linalg.norm((y - ref).ravel()) / max(np.linalg.norm(ref.ravel()), 1e-300)
    assert gate < 1e-4 and l2 <= bound(n), f"{what}: gate {gate:.3e} rel-L2 {l2:.3e}"


@pytest.mark.parametrize("passes", [False, True])
@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 256, 512, 1024, 4096, 32768, 65536])
def test_r2r_matches_fp64(gpu, po, n, passes):
    batch = 3 if n <= 4096 else 2
    for kind in range(4):
        x = po.lcg_real(n * batch, 0xC7C7 ^ (n * 7 + kind), np.float32).reshape(batch, n)
        plan, y = run(gpu, x, kind, passes)
        info = plan.info()
        one_kernel = not passes and 4 <= n <= 32768
        assert (info["e"] > 0) == one_kernel and (info["workspace_bytes"] == 0) == (one_kernel or n == 2)
        check(y, truth(po, x, kind), n, f"{KINDS[kind]} n={n} passes={passes}")


@pytest.mark.parametrize("n", [8, 1024, 32768, 1 << 20])
def test_round_trips_and_forms_agree(gpu, po, n):
    """dct3(dct2(x)) = 2n x and dst3(dst2(x)) = 2n x (test_dct.cpp:97-131) on the GPU alone; at 2^20 the separate
    passes over the four-step real FFT."""
    batch = 4 if n <= 32768 else 2
    x = po.lcg_real(n * batch, 0x1234 ^ n, np.float32).reshape(batch, n)
    for fwd, inv in ((0, 1), (2, 3)):
        _, t = run(gpu, x, fwd)
        _, r = run(gpu, t, inv)
        l2 = np.linalg.norm(r - 2 * n * x.astype(np.float64)) / np.linalg.norm(2 * n * x.astype(np.float64))
        assert l2 <= 2 * bound(n), (KINDS[fwd], n, l2)
        if n <= 32768:
            _, tp = run(gpu, x, fwd, passes=True)
            assert np.linalg.norm(t - tp) / np.linalg.norm(tp) <= bound(n)


def test_many_rows_padded_layouts_in_place_and_rerun(gpu, po):
    """Many rows per workgroup with a partial last workgroup; padded rows (nothing past n of an output row is
    written); in place; a rerun is bit-identical."""
    from crd_amd.dct import DctPlan

    for n, batch, idist, odist in ((8, 1001, 8, 8), (64, 333, 70, 66), (2048, 17, 2048, 2050), (2, 77, 3, 5)):
        x = po.lcg_real(batch * idist, n + batch, np.float32).reshape(batch, idist)
        for kind in range(4):
            plan = DctPlan(n, batch, idist, odist)
            src = gpu.DeviceBuffer.from_array(x)
            guard = np.full((batch, odist), 5.5, dtype=np.float32)
            dst = gpu.DeviceBuffer.from_array(guard)
            plan.execute(kind, src, dst)
            gpu.synchronize()
            y = dst.download(guard.shape, np.float32)
            assert np.all(y[:, n:] == 5.5)
            check(y[:, :n], truth(po, x[:, :n], kind), n, f"{KINDS[kind]} n={n} padded")
            plan.execute(kind, src, dst)
            gpu.synchronize()
            assert np.array_equal(dst.download(guard.shape, np.float32).view(np.uint32), y.view(np.uint32))
    x = po.lcg_real(16 * 4096, 99, np.float32).reshape(16, 4096)
    for kind in range(4):
        _, y = run(gpu, x, kind, in_place=True)
        check(y, truth(po, x, kind), 4096, f"{KINDS[kind]} in place")


def test_dct_rejects_bad_arguments(gpu):
    from crd_amd import _native as N

    lib = N.lib()
    h = ctypes.c_void_p()
    for n, batch, idist, odist, flags in ((0, 1, 0, 0, 0), (3, 1, 0, 0, 0), (48, 1, 0, 0, 0), (1, 1, 0, 0, 0),
                                          (1 << 26, 1, 0, 0, 0)):
        assert lib.crdfft_dct_plan_create(ctypes.byref(h), n, batch, idist, odist, flags) == N.INVALID_SIZE
    assert lib.crdfft_dct_plan_create(ctypes.byref(h), 64, 2, 32, 0, 0) == N.INVALID_ARGUMENT
    assert lib.crdfft_dct_plan_create(ctypes.byref(h), 64, 1, 0, 0, 1 << 9) == N.INVALID_ARGUMENT
    assert lib.crdfft_dct_plan_create(ctypes.byref(h), 64, 2, 64, 80, 0) == N.OK
    buf = gpu.DeviceBuffer(2 * 80 * 4)
    assert lib.crdfft_dct_execute(h, 4, buf.address, buf.address, None) == N.INVALID_ARGUMENT
    assert lib.crdfft_dct_execute(h, 0, buf.address, buf.address, None) == N.INVALID_ARGUMENT  # in place, 64 != 80
    assert lib.crdfft_dct_plan_destroy(h) == N.OK
