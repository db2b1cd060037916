"""hipGraph capture of the plans' executes (crdfft_graph_*): one graph holding a C2C single-kernel and a four-step
transform, the cooperative single pass, a two-pipeline four-step (a second stream forked and joined inside the
capture), R2C/C2R in one kernel and in two fused passes, a DCT, an any-length C2C
(mixed radix and chirp-z) and an STFT, replayed several times with eager work and synchronisations between the
replays; every replay must reproduce the eager results bit for bit (the same kernels on the same inputs).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(gpu, po):
    from crd_amd import _native as N
    from crd_amd import spectral
    from crd_amd.dct import DctPlan

    lib = N.lib()
    s = gpu.Stream()
    rng = np.random.default_rng(5)
    work = []  # (name, run(stream), output buffer, output bytes)

    def c2c(name, n, batch, tuning=None):
        plan = gpu.FftPlan(n, batch, tuning=tuning)
        x = (rng.standard_normal((batch, 2 * n), dtype=np.float32)).view(np.complex64)
        src, dst = gpu.DeviceBuffer.from_array(x), gpu.DeviceBuffer(x.nbytes)
        work.append((name, lambda st, p=plan, a=src, b=dst: p.execute(a, b, gpu.FORWARD, st), dst, x.nbytes,
                     (plan, src)))

    c2c("c2c 4096", 4096, 64)
    c2c("c2c 2^16 four-step", 1 << 16, 8)
    c2c("c2c 2^16 cooperative", 1 << 16, 4, {"algorithm": 6})
    c2c("c2c 2^20 two pipelines", 1 << 20, 5, {"chunk": 2, "hints": 3 | 32})  # fork/join of a second stream
    for n in (4096, 1 << 16):
        x = rng.standard_normal((4, n), dtype=np.float32)
        r2c = gpu.FftPlan(n, 4, "r2c")
        c2r = gpu.FftPlan(n, 4, "c2r")
        src = gpu.DeviceBuffer.from_array(x)
        spec = gpu.DeviceBuffer(8 * (n // 2 + 1) * 4)
        back = gpu.DeviceBuffer(x.nbytes)
        work.append((f"r2c {n}", lambda st, p=r2c, a=src, b=spec: p.execute(a, b, gpu.FORWARD, st), spec,
                     8 * (n // 2 + 1) * 4, (r2c, src)))
        work.append((f"c2r {n}", lambda st, p=c2r, a=spec, b=back: p.execute(a, b, gpu.INVERSE, st), back, x.nbytes,
                     (c2r,)))
    for n, kw in ((1000, {}), (1009, {})):
        plan = gpu.BluesteinPlan(n, 8, **kw)
        x = (rng.standard_normal((8, 2 * n), dtype=np.float32)).view(np.complex64)
        src, dst = gpu.DeviceBuffer.from_array(x), gpu.DeviceBuffer(x.nbytes)
        work.append((f"any-length {n}", lambda st, p=plan, a=src, b=dst: p.execute(a, b, gpu.FORWARD, st), dst,
                     x.nbytes, (plan, src)))
    dplan = DctPlan(4096, 16)
    xd = rng.standard_normal((16, 4096), dtype=np.float32)
    dsrc, ddst = gpu.DeviceBuffer.from_array(xd), gpu.DeviceBuffer(xd.nbytes)
    work.append(("dct2 4096", lambda st: dplan.dct2(dsrc, ddst, st), ddst, xd.nbytes, ()))
    splan = spectral.SpectralPlan(1024, 512)
    sig = rng.standard_normal(1 << 18, dtype=np.float32)
    ssrc = gpu.DeviceBuffer.from_array(sig)
    frames = splan.frames(sig.size)
    sdst = gpu.DeviceBuffer(frames * 513 * 8)
    work.append(("stft 1024", lambda st: splan.stft(ssrc, sig.size, sdst, st), sdst, frames * 513 * 8, ()))

    # eager: also creates each plan's scratch for this stream before any capture
    for _, run, _, _, _ in work:
        run(s)
    s.synchronize()
    eager = [out.download((nbytes // 4,), np.uint32) for _, _, out, nbytes, _ in work]

    assert lib.crdfft_graph_begin(s.handle) == N.OK
    for _, run, _, _, _ in work:
        run(s)
    graph = ctypes.c_void_p()
    assert lib.crdfft_graph_end(s.handle, ctypes.byref(graph)) == N.OK
    try:
        other = gpu.FftPlan(2048, 32)
        junk = gpu.DeviceBuffer(8 * 2048 * 32)
        for replay in range(3):
            for _, _, out, nbytes, _ in work:
                assert lib.crdfft_fill_bytes(out.address, 0x5A, nbytes, s.handle) == N.OK  # eager, outside the graph
            other.execute(junk, junk, gpu.FORWARD, s)  # unrelated eager work between replays
            assert lib.crdfft_graph_launch(graph, s.handle) == N.OK
            s.synchronize()
            for (name, _, out, nbytes, _), want in zip(work, eager):
                got = out.download((nbytes // 4,), np.uint32)
                assert np.array_equal(got, want), f"replay {replay}: {name}"
            gpu.synchronize()
    finally:
        assert lib.crdfft_graph_destroy(graph) == N.OK
    assert lib.crdfft_graph_begin(None) == N.INVALID_ARGUMENT


def test_two_pipeline_plan_first_run_inside_a_capture(gpu):
    """A two-pipeline plan executed for the first time inside a capture cannot create the captured stream's second
    pipeline (that allocates): it records both halves on the captured stream, and its replays match an eager run of
    the same plan shape with the two pipelines, bit for bit."""
    from crd_amd import _native as N

    lib = N.lib()
    n, batch = 1 << 20, 5
    tuning = {"chunk": 2, "hints": 3 | 32}
    eager_plan = gpu.FftPlan(n, batch, tuning=tuning)
    captured_plan = gpu.FftPlan(n, batch, tuning=tuning)
    x = (np.random.default_rng(9).standard_normal((batch, 2 * n), dtype=np.float32)).view(np.complex64)
    src = gpu.DeviceBuffer.from_array(x)
    want, got = gpu.DeviceBuffer(x.nbytes), gpu.DeviceBuffer(x.nbytes)
    s = gpu.Stream()
    eager_plan.execute(src, want, gpu.FORWARD, s)
    s.synchronize()
    assert lib.crdfft_graph_begin(s.handle) == N.OK
    captured_plan.execute(src, got, gpu.FORWARD, s)  # its first execute: the plan's own workspace, one pipeline
    graph = ctypes.c_void_p()
    assert lib.crdfft_graph_end(s.handle, ctypes.byref(graph)) == N.OK
    try:
        for _ in range(2):
            assert lib.crdfft_fill_bytes(got.address, 0x5A, x.nbytes, s.handle) == N.OK
            assert lib.crdfft_graph_launch(graph, s.handle) == N.OK
            s.synchronize()
            assert np.array_equal(got.download(x.shape, np.uint64), want.download(x.shape, np.uint64))
    finally:
        assert lib.crdfft_graph_destroy(graph) == N.OK
