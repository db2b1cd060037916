"""The cooperative single-pass four-step (CRDFFT_ALGO_COOP, k_coop): one square transform per group of workgroups of
one XCD, against FP64 truth and against the two-kernel four-step, in group mode and with every workgroup alone.

Bars: rel-L2 <= 2e-6 * log2(n) against FP64 (numpy's pocketfft through the oracle) both directions, and the plan's
last launch reports no wait that gave up. The groups form from the XCC_ID register: mode 1 (groups) is expected on an
idle device, mode 2 (every workgroup alone) is also a valid result and is forced by the solo hint.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALGO_COOP = 6
SOLO = 16


def bound(n):
    return 2e-6 * max(1.0, np.log2(n))


def status(gpu, plan, stream=None):
    from crd_amd import _native as N

    gave_up, mode = ctypes.c_uint32(), ctypes.c_uint32()
    s = None if stream is None else stream.handle
    assert N.lib().crdfft_plan_coop_status(plan.handle, s, ctypes.byref(gave_up), ctypes.byref(mode)) == 0
    return gave_up.value, mode.value


@pytest.mark.parametrize("solo", [False, True])
@pytest.mark.parametrize("n,batch", [(4096, 1), (4096, 37), (4096, 3000), (65536, 1), (65536, 7), (65536, 33),
                                     (65536, 300)])
def test_coop_matches_fp64(gpu, po, n, batch, solo):
    plan = gpu.FftPlan(n, batch, tuning={"algorithm": ALGO_COOP, "hints": 3 | (SOLO if solo else 0)})
    info = plan.info()
    assert info["algorithm"] == ALGO_COOP and info["n1"] * info["n2"] == n
    x = po.lcg_complex(n * batch, n + batch + solo).reshape(batch, n)
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    for inverse in (False, True):
        plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
        gave_up, mode = status(gpu, plan)
        assert gave_up == 0, f"{gave_up} waits gave up"
        assert mode == (2 if solo else mode) and mode in (1, 2)
        y = dst.download(x.shape, np.complex64)
        pick = slice(0, min(batch, 16))
        ref = po.fft64_of_f32(x[pick], inverse=inverse)
        assert po.rel_l2(y[pick], ref) <= bound(n), f"n={n} batch={batch} inverse={inverse} mode={mode}"
        if batch > 16:
            ref = po.fft64_of_f32(x[-8:], inverse=inverse)
            assert po.rel_l2(y[-8:], ref) <= bound(n)


def test_coop_equals_two_pass_in_place_and_reruns(gpu, po):
    """Against the two-kernel four-step (the same arithmetic up to the twiddle factorisation), in place, and a rerun
    gives the same bits (each transform's work does not depend on which group ran it)."""
    n, batch = 65536, 64
    x = po.lcg_complex(n * batch, 99).reshape(batch, n)
    coop = gpu.FftPlan(n, batch, tuning={"algorithm": ALGO_COOP})
    plain = gpu.FftPlan(n, batch, tuning={"algorithm": 2})
    a = gpu.DeviceBuffer.from_array(x)
    b = gpu.DeviceBuffer(x.nbytes)
    coop.execute(a, b, gpu.FORWARD)
    gpu.synchronize()
    y1 = b.download(x.shape, np.complex64)
    coop.execute(a, b, gpu.FORWARD)
    gpu.synchronize()
    assert np.array_equal(b.download(x.shape, np.complex64).view(np.uint32), y1.view(np.uint32))
    plain.execute(a, b, gpu.FORWARD)
    gpu.synchronize()
    assert po.rel_l2(y1, b.download(x.shape, np.complex64)) <= bound(n)
    coop.execute(a, a, gpu.FORWARD)
    assert status(gpu, coop) == (0, status(gpu, coop)[1])
    assert np.array_equal(a.download(x.shape, np.complex64).view(np.uint32), y1.view(np.uint32))


def test_coop_on_several_streams(gpu, po):
    """Three streams share one plan, each with its own slots and control words; launches that cannot be resident at
    the same time fall back to every workgroup alone, so every result is still exact."""
    n, batch = 65536, 24
    plan = gpu.FftPlan(n, batch, tuning={"algorithm": ALGO_COOP})
    streams = [gpu.Stream() for _ in range(3)]
    xs = [po.lcg_complex(n * batch, 500 + i).reshape(batch, n) for i in range(3)]
    bufs = [gpu.DeviceBuffer.from_array(x) for x in xs]
    outs = [gpu.DeviceBuffer(x.nbytes) for x in xs]
    for _ in range(2):
        for i, s in enumerate(streams):
            plan.execute(bufs[i], outs[i], gpu.FORWARD, s)
    for i, s in enumerate(streams):
        gave_up, _ = status(gpu, plan, s)
        assert gave_up == 0
        y = outs[i].download(xs[i].shape, np.complex64)
        assert po.rel_l2(y[:6], po.fft64_of_f32(xs[i][:6])) <= bound(n)
