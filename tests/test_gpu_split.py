"""Split-precision MFMA transforms: radix-16 (256-point sequences, v_mfma_f32_16x16x32) and radix-32 (1024-point
sequences, v_mfma_f32_32x32x16) stages as complex DFT-matrix products in split BF16 (hi+mid+lo) and FP16 (hi+lo
with per-sequence power-of-two scaling), FP32 accumulation, the exchange between a sequence's two stages done with
cross-lane swaps. Against the oracle's FP64 truth at the tuned form's target, rel-L2 <= 1e-6, for every length the
planner takes: 256 and 1024 (one sequence per transform), 2^16 (256 x 256), 2^18 (256 x 1024), 2^20 (1024 x 1024)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1 << 16


def wide_dynamic_range(batch, seed, n=N):
    """|x| log-uniform in [1e-4, 1e4], uniform phase (seeded)."""
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-4, 4, (batch, n))
    ph = rng.uniform(0, 2 * np.pi, (batch, n))
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
@pytest.mark.parametrize("n,batch", [(256, 37), (1024, 19), (1 << 18, 3), (1 << 20, 2)])
def test_split_radix16_and_radix32_lengths(gpu, po, precision, n, batch):
    """The radix-32 stages (1024-point sequences) and the mixed four-step against FP64, both directions, with odd
    batches (a partial last workgroup of whole transforms at n = 256, 1024) and in place."""
    x = wide_dynamic_range(batch, n ^ batch, n)
    plan = gpu.FftPlan(n, batch, precision=precision)
    info = plan.info()
    assert info["precision"] != 0
    for inverse in (False, True):
        src = gpu.DeviceBuffer.from_array(x)
        dst = gpu.DeviceBuffer(x.nbytes)
        d = gpu.INVERSE if inverse else gpu.FORWARD
        plan.execute(src, dst, d)
        plan.execute(src, None, d)  # in place
        gpu.synchronize()
        ref = po.fft64_of_f32(x, inverse=inverse)
        for buf in (dst, src):
            y = buf.download(x.shape, np.complex64)
            err = po.rel_l2(y, ref)
            print(f"{precision} n={n} inverse={inverse} rel_l2={err:.3e}")
            assert err <= 1e-6


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


@pytest.mark.parametrize("n", [1024, 1 << 16])
def test_split_normalized_inverse(gpu, po, n):
    """CRDFFT_FLAG_NORMALIZE_INVERSE on a split plan: ifft_normalized's 1/n rides on the last store."""
    from crd_amd import _native as N_

    x = wide_dynamic_range(2, 3, n)
    plan = gpu.FftPlan(n, 2, precision="split_fp16", flags=N_.FLAG_NORMALIZE_INVERSE)
    buf = gpu.DeviceBuffer.from_array(x)
    plan.execute(buf, None, gpu.INVERSE)
    gpu.synchronize()
    y = buf.download(x.shape, np.complex64)
    assert po.rel_l2(y, po.fft64_of_f32(x, inverse=True) / n) <= 1e-6


def test_split_precision_rejects_other_shapes(gpu):
    from crd_amd import _native as N_

    for n, t in ((4096, "c2c"), (1 << 17, "c2c"), (1 << 22, "c2c"), (N, "r2c")):
        with pytest.raises(N_.CrdfftError) as e:
            gpu.FftPlan(n, 2, t, precision="split_fp16")
        assert e.value.status == N_.UNSUPPORTED


@pytest.mark.parametrize("precision", ["split_bf16", "split_fp16"])
def test_split_row_pass_walks_several_tiles(gpu, po, precision):
    """A batch with more row-pass tiles (16 rows of one transform each) than the resident grid holds, so workgroups
    walk several tiles and prefetch the next: every transform against FP64, out of place and in place."""
    batch = 80  # 1280 row-pass tiles per chunk, more than the resident workgroups of one launch
    x = wide_dynamic_range(batch, 91)
    plan = gpu.FftPlan(N, batch, precision=precision)
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    plan.execute(src, dst, gpu.FORWARD)
    plan.execute(src, None, gpu.FORWARD)
    gpu.synchronize()
    ref = po.fft64_of_f32(x)
    for buf in (dst, src):
        y = buf.download(x.shape, np.complex64)
        per = [po.rel_l2(y[i], ref[i]) for i in range(batch)]
        assert max(per) <= 1e-6, (precision, int(np.argmax(per)), max(per))


@pytest.mark.parametrize("precision", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("n,batch,chunk", [(1 << 16, 7, 3), (1 << 18, 5, 2), (1 << 20, 3, 2)])
def test_split_chunks(gpu, po, precision, n, batch, chunk):
    """Several chunks through one workspace, the last one partial (the default chunks of 2^16 and 2^18 hold 256 and
    64 transforms, more than the other tests' batches): every transform against FP64, both directions."""
    x = wide_dynamic_range(batch, 7 * n + batch, n)
    plan = gpu.FftPlan(n, batch, precision=precision, tuning={"chunk": chunk})
    assert plan.info()["chunk"] == chunk
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    for inverse in (False, True):
        plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
        gpu.synchronize()
        y = dst.download(x.shape, np.complex64)
        ref = po.fft64_of_f32(x, inverse=inverse)
        per = [po.rel_l2(y[i], ref[i]) for i in range(batch)]
        print(f"{precision} n={n} chunk={chunk} inverse={inverse} max rel_l2={max(per):.3e}")
        assert max(per) <= 1e-6, (precision, n, inverse, int(np.argmax(per)), max(per))
