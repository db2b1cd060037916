"""Split-precision MFMA four-step (n = 2^16): radix-16 stages as DFT-matrix products on v_mfma_f32_16x16x32 in
split BF16 (hi+mid+lo) and FP16 (hi+lo with per-sequence power-of-two scaling), FP32 accumulation, against the
oracle's FP64 truth. The error bound is the tuned form's target, rel-L2 <= 1e-6 against FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1 << 16


def wide_dynamic_range(batch, seed):
    """|x| log-uniform in [1e-4, 1e4], uniform phase (seeded)."""
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-4, 4, (batch, N))
    ph = rng.uniform(0, 2 * np.pi, (batch, N))
    return (mag * np.exp(1j * ph)).astype(np.complex64)


@pytest.mark.parametrize("precision", ["split_bf16", "split_fp16"])
def test_split_precision_against_fp64(gpu, po, precision):
    batch = 6
    x = wide_dynamic_range(batch, 44)
    plan = gpu.FftPlan(N, batch, precision=precision)
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    errs = {}
    for inverse in (False, True):
        plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        y = dst.download(x.shape, np.complex64)
        ref = po.fft64_of_f32(x, inverse=inverse)
        errs[inverse] = po.rel_l2(y, ref)
        print(f"{precision} inverse={inverse} rel_l2={errs[inverse]:.3e}")
        assert errs[inverse] <= 1e-6
    # the same inputs through the plain FP32 path, for scale
    plain = gpu.FftPlan(N, batch)
    plain.execute(src, dst, gpu.FORWARD)
    gpu.synchronize()
    e32 = po.rel_l2(dst.download(x.shape, np.complex64), po.fft64_of_f32(x))
    print(f"plain fp32 rel_l2={e32:.3e}")


@pytest.mark.parametrize("precision", ["split_bf16", "split_fp16"])
def test_split_precision_specials_and_batches(gpu, po, precision):
    # odd batch, in place, a zero transform and one with a single huge value
    batch = 3
    x = wide_dynamic_range(batch, 7)
    x[1] = 0
    x[2] = 0
    x[2, 123] = np.complex64(3e4 - 2e4j)
    plan = gpu.FftPlan(N, batch, precision=precision)
    buf = gpu.DeviceBuffer.from_array(x)
    plan.execute(buf, buf, gpu.FORWARD)
    gpu.synchronize()
    y = buf.download(x.shape, np.complex64)
    assert np.all(y[1] == 0)
    ref = po.fft64_of_f32(x)
    assert po.rel_l2(y[0], ref[0]) <= 1e-6
    assert po.rel_l2(y[2], ref[2]) <= 1e-6


def test_split_precision_rejects_other_shapes(gpu):
    from crd_amd import _native as N_

    for n, t in ((4096, "c2c"), (1 << 20, "c2c"), (N, "r2c")):
        with pytest.raises(N_.CrdfftError) as e:
            gpu.FftPlan(n, 2, t, precision="split_fp16")
        assert e.value.status == N_.UNSUPPORTED
