"""The fused four-step (both passes of a transform in one workgroup of a persistent launch) against the oracle.

Bars as for the two-kernel path: rel-L2 <= 2e-6 * log2(n) against the oracle's FP64 truth and the reference's
max-norm class (2e-4 below 2^19, test_fft_kernels.cpp:151-163). The fused kernel runs the same column/row
arithmetic as the two-kernel four-step of the same split, so the two agree to FP32 rounding; batches are sized past
the persistent grid so every workgroup runs several transforms through its slot.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (log2 n, columns per phase-A tile) of every instantiated fused kernel; 0 = the planner's default size
CASES = [(15, 0), (15, 128), (16, 0), (16, 64), (17, 64)]


def bound(n):
    return 2e-6 * max(1.0, np.log2(n))


def fused_plan(gpu, n, batch, tile1, **kw):
    tuning = {"algorithm": 5, "tile1": tile1}
    plan = gpu.FftPlan(n, batch, tuning=tuning, **kw)
    info = plan.info()
    assert info["algorithm"] == 5, info
    return plan, info


def run(gpu, plan, x, inverse=False, in_place=False):
    src = gpu.DeviceBuffer.from_array(x)
    dst = src if in_place else gpu.DeviceBuffer(x.nbytes)
    plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
    gpu.synchronize()
    return dst.download(x.shape, np.complex64)


def sample_rows(batch):
    rows = sorted(set([0, 1, 2, batch // 3, batch // 2, batch - 3, batch - 2, batch - 1]))
    return [r for r in rows if 0 <= r < batch]


@pytest.mark.parametrize("lg,tile1", CASES)
def test_fused_matches_fp64_past_the_grid(gpu, po, lg, tile1):
    n = 1 << lg
    probe, info = fused_plan(gpu, n, 1 << 20, tile1, self_check=False)
    slots = info["chunk"]
    del probe
    batch = slots + slots // 2 + 3  # every workgroup runs one or two transforms, the last ones partly idle
    print(f"fused n={n} tile1={tile1} slots={slots} batch={batch}", flush=True)
    plan, _ = fused_plan(gpu, n, batch, tile1)
    x = po.lcg_complex(n * batch, 0xF05E ^ n).reshape(batch, n)
    rows = sample_rows(batch)
    for inverse in (False, True):
        y = run(gpu, plan, x, inverse)
        ref = po.fft64_of_f32(x[rows], inverse=inverse)
        assert po.rel_l2(y[rows], ref) <= bound(n), f"n={n} inverse={inverse}"
        assert po.max_rel(y[rows], ref) < 2e-4
        # every transform of the batch agrees with the two-kernel four-step of the same split
        two = gpu.FftPlan(n, batch, tuning={"algorithm": 2, "n1": info["n1"]})
        z = run(gpu, two, x, inverse)
        assert po.rel_l2(y, z) <= 1e-6, f"n={n} inverse={inverse} fused vs two-kernel"


@pytest.mark.parametrize("lg,tile1", CASES)
def test_fused_small_batches_in_place_and_rerun(gpu, po, lg, tile1):
    n = 1 << lg
    for batch in (1, 3):
        plan, _ = fused_plan(gpu, n, batch, tile1)
        x = po.lcg_complex(n * batch, n + batch).reshape(batch, n)
        y = run(gpu, plan, x)
        yi = run(gpu, plan, x, in_place=True)
        y2 = run(gpu, plan, x)
        assert np.array_equal(y.view(np.uint32), yi.view(np.uint32)), "in place differs from out of place"
        assert np.array_equal(y.view(np.uint32), y2.view(np.uint32)), "rerun differs"
        assert po.rel_l2(y, po.fft64_of_f32(x)) <= bound(n)


def test_fused_padded_distances_guard_words_and_normalized_inverse(gpu, po):
    from crd_amd import _native as N

    n, batch, pad = 1 << 16, 5, 40
    dist = n + pad
    x = po.lcg_complex(n * batch, 91).reshape(batch, n)
    host_in = np.full((batch, dist), np.nan + 1j * np.nan, dtype=np.complex64)
    host_in[:, :n] = x
    guard = np.frombuffer(np.uint32(0x7FC0DEAD).tobytes() * 2, dtype=np.complex64)[0]
    host_out = np.full(batch * dist + 64, guard, dtype=np.complex64)
    src = gpu.DeviceBuffer.from_array(host_in)
    dst = gpu.DeviceBuffer.from_array(host_out)
    t = N.Tuning(algorithm=N.ALGO_FUSED)
    h = ctypes.c_void_p()
    N.check(N.lib().crdfft_plan_create_tuned(ctypes.byref(h), n, batch, dist, dist, N.C2C, N.FP32,
                                             N.FLAG_NORMALIZE_INVERSE, ctypes.byref(t)))
    try:
        N.check(N.lib().crdfft_execute(h, src.ptr, dst.ptr, N.FORWARD, None))
        gpu.synchronize()
        out = dst.download(host_out.shape, np.complex64)
        body = out[: batch * dist].reshape(batch, dist)
        assert po.rel_l2(body[:, :n], po.fft64_of_f32(x)) <= bound(n)
        assert np.array_equal(body[:, n:].view(np.uint32), np.full((batch, pad), guard).view(np.uint32))
        assert np.array_equal(out[batch * dist:].view(np.uint32), np.full(64, guard).view(np.uint32))
        # inverse of the forward output, normalized: back to x
        N.check(N.lib().crdfft_execute(h, dst.ptr, dst.ptr, N.INVERSE, None))
        gpu.synchronize()
        back = dst.download(host_out.shape, np.complex64)[: batch * dist].reshape(batch, dist)[:, :n]
        assert np.max(np.abs(back - x)) <= bound(n) * 4
    finally:
        N.lib().crdfft_plan_destroy(h)


def test_fused_specials_stay_in_their_transform(gpu, po):
    n = 1 << 16
    plan, _ = fused_plan(gpu, n, 3, 0)
    x = po.lcg_complex(n * 3, 5).reshape(3, n)
    x[0, n // 3] = np.complex64(np.nan)
    x[1, 1] = np.complex64(complex(np.inf, 0.0))
    y = run(gpu, plan, x)
    assert np.all(~np.isfinite(y[0])) and np.all(~np.isfinite(y[1]))
    assert np.all(np.isfinite(y[2]))
    assert po.rel_l2(y[2:], po.fft64_of_f32(x[2:])) <= bound(n)


def test_fused_real_transforms(gpu, po):
    """R2C / C2R of n = 2^17 run their half-length complex transform (2^16) through the fused kernel."""
    n, batch = 1 << 17, 3
    xr = po.normal_real(n * batch, 17).reshape(batch, n)
    plan = gpu.FftPlan(n, batch, type="r2c", tuning={"algorithm": 5})
    assert plan.info()["algorithm"] == 5
    src = gpu.DeviceBuffer.from_array(xr)
    dst = gpu.DeviceBuffer(batch * (n // 2 + 1) * 8)
    plan.execute(src, dst, gpu.FORWARD)
    gpu.synchronize()
    X = dst.download((batch, n // 2 + 1), np.complex64)
    assert po.rel_l2(X, po.rfft64(xr.astype(np.float64))) <= bound(n)
    inv = gpu.FftPlan(n, batch, type="c2r", tuning={"algorithm": 5})
    back = gpu.DeviceBuffer(xr.nbytes)
    inv.execute(dst, back, gpu.INVERSE)
    gpu.synchronize()
    y = back.download(xr.shape, np.float32)
    assert np.max(np.abs(y - xr)) <= bound(n) * np.max(np.abs(xr))


@pytest.mark.parametrize("n,batch,n1,tile1", [(1 << 17, 5, 256, 16), (1 << 18, 3, 512, 16),
                                             (1 << 20, 3, 1024, 8), (1 << 20, 3, 2048, 8),
                                             (1 << 20, 2, 1024, 16), (1 << 21, 2, 2048, 8)])
def test_persistent_column_pass(gpu, po, n, batch, n1, tile1):
    """The column pass as a persistent prefetching launch (plan hint bit 3) gives the results of the one-tile-per-
    workgroup column pass, for full-table (n <= 2^18) and factored twiddles."""
    x = po.lcg_complex(n * batch, n ^ 0x5EED).reshape(batch, n)
    base = {"algorithm": 2, "n1": n1, "tile1": tile1}
    pf = gpu.FftPlan(n, batch, tuning=dict(base, hints=11))
    assert pf.info()["hints"] & 8
    plain = gpu.FftPlan(n, batch, tuning=dict(base, hints=3))
    for inverse in (False, True):
        y = run(gpu, pf, x, inverse)
        z = run(gpu, plain, x, inverse)
        ref = po.fft64_of_f32(x, inverse=inverse)
        assert po.rel_l2(y, ref) <= bound(n), f"n={n} inverse={inverse}"
        assert po.rel_l2(y, z) <= 1e-7, f"n={n} inverse={inverse} persistent vs one tile per workgroup"
