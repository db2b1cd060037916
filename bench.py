#!/usr/bin/env python3
"""Batched FP32 C2C FFT throughput on one MI355X (gfx950), through the crdfft C-ABI.

Workload: the bandwidth-roofline configuration of the reference's batched FFT path, plain FP32 1D complex-to-complex
forward transforms, unnormalized, interleaved complex, contiguous: n = 2^16 x 4096 transforms (4 GiB of algorithmic
traffic) and n = 2^20 x 64 transforms (1 GiB). Inputs are i.i.d. standard normal (seeded), resident in HBM before the
timed region. One step = one forward transform of each shape. value = algorithmic bytes (16 per complex point:
one read, one write) of a step / step time, in GB/s; the percentage of the 8.0 TB/s HBM peak is reported beside it.

Also reported (not the value): the R2C -> C2R round trip of both shapes, n = 4096 x 1024 (rotating over 4 buffers so
the 64 MiB set cannot sit in the 256 MB Infinity Cache), the accuracy of a sample against FP64, a device-copy
ceiling, the per-kernel roofline of the dominant kernel timed live with hipEvents around each launch, and the CPU
baseline (the oracle's FP32 restatement of the reference's CPU algorithm, one plan per thread over contiguous ranges,
timed on a bounded sample on this host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "batched C2C GB/s (% HBM peak); tuned vs plain FP32 FFTs/s at rel-L2≤1e-6"
HBM_PEAK_GBS = 8000.0
SHAPES = [(1 << 16, 4096), (1 << 20, 64)]
EXTRAS = ("copy", "roundtrip", "c2c4096", "mixed", "split", "spectral", "r2r")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def fill_normal(buf, count_floats, seed, chunk=1 << 24):
    """Upload count_floats i.i.d. N(0,1) float32 values (seeded) to buf; returns the first chunk (host copy)."""
    rng = np.random.default_rng(seed)
    first = None
    done = 0
    while done < count_floats:
        m = min(chunk, count_floats - done)
        a = rng.standard_normal(m, dtype=np.float32)
        if first is None:
            first = a.copy()
        buf.upload(a, offset=done * 4)
        done += m
    return first


def time_loop(fn, stream, steps, warmup, Event):
    for _ in range(warmup):
        fn()
    stream.synchronize()
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / steps


def fnv1a64(data):
    """FNV-1a, 64-bit, over the bytes of `data` (the digest the reference's self-check records, used here to show
    that two runs or builds produced identical output bits)."""
    h = 0xCBF29CE484222325
    for byte in bytes(data):
        h = ((h ^ byte) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def step_times(fn, stream, reps, Event):
    """Per-step times (ms) of `reps` steps, each bracketed by its own events."""
    out = []
    for _ in range(reps):
        e0, e1 = Event(), Event()
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        out.append(e0.elapsed_ms(e1))
    return out


def rel_l2(y, ref):
    return float(np.linalg.norm((y - ref).ravel()) / np.linalg.norm(ref.ravel()))


def step_bytes_per_core_estimate(cores):
    """GB/s of the headline step mix at the reference's published AVX-512 single-thread times, scaled by cores."""
    published_ms = {1 << 16: 0.08764, 1 << 20: 2.66472}
    total_bytes = sum(16.0 * n * b for n, b in SHAPES)
    total_s = sum(published_ms[n] * 1e-3 * b for n, b in SHAPES) / cores
    return total_bytes / total_s / 1e9


def cpu_baseline(seconds_per_shape, threads):
    """The oracle's FP32 restatement on a bounded sample of the same shapes, on `threads` host threads."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    from oracle import pyoracle as po

    # single-core calibration against the reference's published AVX-512 codelet (0.00292 ms per 4096-point f32
    # transform, level_plans_gen.hpp:83): 256 transforms of n = 4096 on one thread, best of 5
    x1 = np.random.default_rng(8).standard_normal(2 * 4096 * 256, dtype=np.float32).view(np.complex64).reshape(256, 4096)
    best1 = float("inf")
    for _ in range(5):
        y1 = x1.copy()
        t0 = time.perf_counter()
        po.c2c32_inplace(y1, inverse=False, threads=1)
        best1 = min(best1, time.perf_counter() - t0)
    total_bytes = 0.0
    total_s = 0.0
    parts = []
    for n, batch in SHAPES:
        # a sample of whole transforms: grow until one pass takes >= seconds_per_shape
        count = max(threads, 1)
        x = np.random.default_rng(7).standard_normal(2 * n * count, dtype=np.float32).view(np.complex64)
        x = x.reshape(count, n)
        while True:
            y = x.copy()
            t0 = time.perf_counter()
            po.c2c32_inplace(y, inverse=False, threads=threads)
            dt = time.perf_counter() - t0
            if dt >= seconds_per_shape or count >= batch:
                break
            grow = min(batch, max(count * 2, int(count * seconds_per_shape / max(dt, 1e-3))))
            if grow == count:
                break
            count = grow
            x = np.random.default_rng(7).standard_normal(2 * n * count, dtype=np.float32).view(np.complex64)
            x = x.reshape(count, n)
        b = 16.0 * n * count
        parts.append(f"{count} x n={n} in {dt:.2f} s")
        # scale each shape to its share of one step (batch transforms), so the CPU figure is the same step mix
        total_bytes += 16.0 * n * batch
        total_s += dt * batch / count
    return {
        "value": total_bytes / total_s / 1e9,
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": "FP32 Stockham/four-step restatement (oracle/), one plan per thread over contiguous ranges: "
                  + "; ".join(parts) + "; scaled to one step of the GPU workload",
        "single_core_ms_per_n4096": round(1000.0 * best1 / 256, 5),
        # the reference's own generated AVX-512 codelets run this transform in 0.00292 ms on one core
        # (level_plans_gen.hpp:83); the portable restatement is this many times slower per core
        "single_core_vs_reference_published": round(1000.0 * best1 / 256 / 0.00292, 2),
        "reference_published_estimate": {
            "value": round(step_bytes_per_core_estimate(threads), 2), "unit": "GB/s",
            "what": "the reference's published single-core rates (0.08764 ms per 2^16 and 2.66472 ms per 2^20 "
                    "transform, level_plans_gen.hpp:87, :91) times this host's cores, for the same step mix; an "
                    "estimate, not a measurement"},
    }


def bench_c2c4096(stream, Event):
    """n = 4096 x 1024 (the reference's config-1 shape, 64 MiB) over 4 rotating buffer pairs so the set cannot sit
    in the 256 MB Infinity Cache; 16 B per point algorithmic; one single-pass kernel per execute."""
    import crd_amd
    from crd_amd import DeviceBuffer, FftPlan
    from crd_amd import _native as N

    n, batch, nb = 4096, 1024, 4
    p = FftPlan(n, batch)
    ins = [DeviceBuffer(8 * n * batch) for _ in range(nb)]
    outs = [DeviceBuffer(8 * n * batch) for _ in range(nb)]
    for i, b in enumerate(ins):
        fill_normal(b, 2 * n * batch, seed=3000 + i)
    state = {"i": 0}

    def rot():
        i = state["i"] % nb
        state["i"] += 1
        p.execute(ins[i], outs[i], crd_amd.FORWARD, stream)

    ms = time_loop(rot, stream, 40, 10, Event)

    # the ceiling for this size: the streaming copy (16-byte elements, one read and one write per byte) over the same
    # rotating buffer pairs, 64 MiB per launch
    def rot_copy():
        i = state["i"] % nb
        state["i"] += 1
        N.check(N.lib().crdfft_stream_copy(outs[i].ptr, ins[i].ptr, 8 * n * batch, stream.handle),
                "crdfft_stream_copy")

    copy_ms = time_loop(rot_copy, stream, 40, 10, Event)
    return {"ms": round(ms, 5), "gbs": round(16.0 * n * batch / (ms * 1e-3) / 1e9, 1),
            "frac_of_peak": round(16.0 * n * batch / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "ffts_per_s": round(batch / (ms * 1e-3), 1), "rotating_buffers": nb,
            "copy_same_bytes": {"ms": round(copy_ms, 5), "gbs": round(16.0 * n * batch / (copy_ms * 1e-3) / 1e9, 1)},
            "vs_copy": round(copy_ms / ms, 3)}


def bench_mixed(stream, Event):
    """20,000 C2C transforms of 7-smooth lengths in [64, 16384] (log-uniform lengths, Poisson(8) batches, seeded):
    the autotuned persistent launch replayed from a hipGraph, against the reference's method (one plan per length
    and batch group, Bluestein for every length that is not a power of two)."""
    import crd_amd
    from crd_amd import many, workloads

    lengths = workloads.mixed_batch()
    total = int(lengths.sum())
    x = workloads.uniform_complex(total, 20000)
    src = crd_amd.DeviceBuffer.from_array(x)
    dst = crd_amd.DeviceBuffer(x.nbytes)
    t0 = time.time()
    choices, records = many.autotune(lengths, reps=3)
    tune_s = time.time() - t0
    out = {"transforms": int(len(lengths)), "lengths": int(len(set(lengths.tolist()))), "points": total,
           "algorithmic_bytes": 16 * total, "autotune_seconds": round(tune_s, 2), "autotune_timings": len(records)}
    for name, ch in (("tuned", choices), ("default_radix_plans", None)):
        plan = many.ManyPlan(lengths, choices=ch)
        plan.capture_graph(src, dst, crd_amd.FORWARD, stream)
        ms = time_loop(lambda: plan.launch_graph(stream), stream, 20, 10, Event)
        out[name] = {"ms": round(ms, 4), "ffts_per_s": round(len(lengths) / (ms * 1e-3), 1),
                     "gbs": round(16.0 * total / (ms * 1e-3) / 1e9, 1)}
        if name == "tuned":
            info = plan.info()
            out["tiles"] = int(info["tiles"])
            out["imbalance"] = round(info["imbalance"], 3)
            y = dst.download(x.shape, np.complex64)
            errs_num, errs_den, pos = 0.0, 0.0, 0
            rng = np.random.default_rng(1)
            pick = set(rng.choice(len(lengths), 200, replace=False).tolist())
            for i, n in enumerate(lengths.tolist()):
                if i in pick:
                    ref = np.fft.fft(x[pos:pos + n].astype(np.complex128))
                    errs_num += float(np.sum(np.abs(y[pos:pos + n] - ref) ** 2))
                    errs_den += float(np.sum(np.abs(ref) ** 2))
                pos += n
            out["tuned"]["rel_l2_vs_fp64_sample"] = (errs_num / errs_den) ** 0.5
        del plan
    # the reference's method: a plan per (length, group) in draw order, Bluestein off the powers of two
    groups, start = [], 0
    for i in range(1, len(lengths) + 1):
        if i == len(lengths) or lengths[i] != lengths[start]:
            groups.append((int(lengths[start]), start, i - start))
            start = i
    offs = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    plans = {}
    calls = []
    for n, first, cnt in groups:
        key = (n, cnt)
        if key not in plans:
            plans[key] = crd_amd.FftPlan(n, cnt, self_check=False) if n & (n - 1) == 0 else \
                crd_amd.BluesteinPlan(n, cnt, chirp_z=True)
        calls.append((plans[key], src.address + 8 * int(offs[first]), dst.address + 8 * int(offs[first])))

    def plain():
        for p, a, b in calls:
            p.execute(a, b, crd_amd.FORWARD, stream)

    ms = time_loop(plain, stream, 3, 1, Event)
    out["plain_per_length_bluestein"] = {"ms": round(ms, 4), "ffts_per_s": round(len(lengths) / (ms * 1e-3), 1),
                                         "gbs": round(16.0 * total / (ms * 1e-3) / 1e9, 1), "plans": len(plans),
                                         "launch_groups": len(calls)}
    del plans, calls
    # the same mixed-radix kernel and default radix plans, one launch per (length, group) in draw order: the
    # algorithm without the single persistent launch
    mplans = {}
    mcalls = []
    for n, first, cnt in groups:
        key = (n, cnt)
        if key not in mplans:
            mplans[key] = many.ManyPlan(np.full(cnt, n, dtype=np.uint64))
        mcalls.append((mplans[key], src.address + 8 * int(offs[first]), dst.address + 8 * int(offs[first])))

    def per_length():
        for p, a, b in mcalls:
            p.execute(a, b, crd_amd.FORWARD, stream)

    ms = time_loop(per_length, stream, 3, 1, Event)
    out["per_length_launches_mixed_radix"] = {"ms": round(ms, 4), "ffts_per_s": round(len(lengths) / (ms * 1e-3), 1),
                                              "plans": len(mplans), "launch_groups": len(mcalls)}
    del mplans, mcalls
    out["tuned_vs_plain_ffts_per_s"] = round(out["plain_per_length_bluestein"]["ms"] / out["tuned"]["ms"], 2)
    # where the factor comes from: the algorithm (mixed radix instead of chirp-z, both one launch per group), the
    # single persistent launch (default radix plans), and the autotuner's radix plans
    out["gain_breakdown"] = {
        "algorithm": round(out["plain_per_length_bluestein"]["ms"] / out["per_length_launches_mixed_radix"]["ms"], 2),
        "persistent_launch": round(out["per_length_launches_mixed_radix"]["ms"] / out["default_radix_plans"]["ms"], 2),
        "tuning": round(out["default_radix_plans"]["ms"] / out["tuned"]["ms"], 3),
    }
    return out


def bench_spectral(stream, Event):
    """The fused spectral front end on a 2^26-sample FP32 signal (256 MiB), frames of 4096 at 50% overlap (32767
    frames): Welch PSD, spectrogram rows and STFT, beside a plain R2C plan over the same frames already cut and
    windowed (the FFT alone, on a framed copy the fused path never writes)."""
    import crd_amd
    from crd_amd import DeviceBuffer, FftPlan, spectral

    length, nperseg, noverlap = 1 << 26, 4096, 2048
    rng = np.random.default_rng(11)
    x = rng.standard_normal(length, dtype=np.float32)
    dx = DeviceBuffer.from_array(x)
    wplan = spectral.SpectralPlan(nperseg, noverlap, "hann", detrend=True)
    splan = spectral.SpectralPlan(nperseg, noverlap, "hann", detrend=False)
    frames = wplan.frames(length)
    nfreq = nperseg // 2 + 1
    psd = DeviceBuffer(nfreq * 4)
    rows = DeviceBuffer(frames * nfreq * 8)
    read_bytes = 4.0 * frames * nperseg  # each frame's samples once (the overlap reads the signal twice)
    out = {"signal_samples": length, "nperseg": nperseg, "noverlap": noverlap, "frames": frames}
    ms = time_loop(lambda: wplan.welch(dx, length, 1.0, psd, stream), stream, 10, 10, Event)
    out["welch"] = {"ms": round(ms, 4), "frames_per_s": round(frames / (ms * 1e-3), 1),
                    "gbs_frame_reads": round(read_bytes / (ms * 1e-3) / 1e9, 1)}
    ms = time_loop(lambda: wplan.spectrogram(dx, length, 1.0, rows, stream), stream, 10, 10, Event)
    out["spectrogram"] = {"ms": round(ms, 4),
                          "gbs": round((read_bytes + 4.0 * frames * nfreq) / (ms * 1e-3) / 1e9, 1)}
    ms = time_loop(lambda: splan.stft(dx, length, rows, stream), stream, 10, 10, Event)
    out["stft"] = {"ms": round(ms, 4), "gbs": round((read_bytes + 8.0 * frames * nfreq) / (ms * 1e-3) / 1e9, 1)}
    # the FFT alone on a framed, windowed copy (what a frame-by-frame port would run after its own framing pass)
    framed = DeviceBuffer(frames * nperseg * 4)
    r2c = FftPlan(nperseg, frames, "r2c", self_check=False)
    ms = time_loop(lambda: r2c.execute(framed, rows, crd_amd.FORWARD, stream), stream, 10, 10, Event)
    out["r2c_of_framed_copy_ms"] = round(ms, 4)
    out["welch_vs_r2c_alone"] = round(out["r2c_of_framed_copy_ms"] / out["welch"]["ms"], 2)
    return out


def bench_r2r(stream, Event):
    """DCT-II/III and DST-II/III of 4096 rows of 4096 FP32 samples (64 MiB in, 64 MiB out): the one-kernel form and
    the separate passes around the R2C / C2R plans; GB/s counts one read and one write of every sample. Then one
    any-length C2C, n = 1000 over 16777 rows: the default (one mixed-radix launch, 1000 = 2^3 5^3), the chirp-z one-kernel
    form (m = 2048) and its five passes."""
    import crd_amd
    from crd_amd import DeviceBuffer
    from crd_amd.dct import DctPlan

    n, rows = 4096, 4096
    x = np.random.default_rng(12).standard_normal((rows, n), dtype=np.float32)
    src, dst = DeviceBuffer.from_array(x), DeviceBuffer(x.nbytes)
    out = {"n": n, "rows": rows}
    names = ("dct2", "dct3", "dst2", "dst3")
    for passes in (False, True):
        plan = DctPlan(n, rows, passes=passes)
        form = {}
        for kind in range(4):
            ms = time_loop(lambda: plan.execute(kind, src, dst, stream), stream, 20, 10, Event)
            form[names[kind]] = {"ms": round(ms, 4), "gbs": round(8.0 * n * rows / (ms * 1e-3) / 1e9, 1)}
        out["separate_passes" if passes else "one_kernel"] = form
    out["one_kernel_vs_passes_dct2"] = round(out["separate_passes"]["dct2"]["ms"] / out["one_kernel"]["dct2"]["ms"], 2)
    bn, bb = 1000, 16777
    xb = np.random.default_rng(13).standard_normal((bb, 2 * bn), dtype=np.float32).view(np.complex64)
    bsrc, bdst = DeviceBuffer.from_array(xb), DeviceBuffer(xb.nbytes)
    blue = {"n": bn, "batch": bb}
    for name, kw in (("mixed_radix", {}), ("one_kernel", {"chirp_z": True}), ("separate_passes", {"passes": True})):
        plan = crd_amd.BluesteinPlan(bn, bb, **kw)
        ms = time_loop(lambda: plan.execute(bsrc, bdst, crd_amd.FORWARD, stream), stream, 20, 10, Event)
        blue[name] = {"ms": round(ms, 4), "gbs": round(16.0 * bn * bb / (ms * 1e-3) / 1e9, 1)}
    blue["one_kernel_vs_passes"] = round(blue["separate_passes"]["ms"] / blue["one_kernel"]["ms"], 2)
    out["bluestein"] = blue
    return out


def bench_split(stream, Event):
    """n = 2^16, batch 2048, |x| log-uniform in [1e-4, 1e4] with uniform phase: plain FP32 against the split-BF16
    and split-FP16 MFMA forms, time and rel-L2 against FP64 (numpy, float64) on a sample. The same for every other
    length the split forms take (256 and 1024: one radix-16 / radix-32 sequence per transform; 2^18 = 256 x 1024 and
    2^20 = 1024 x 1024: radix-32 stages in one or both passes), each at 2^27 points per batch."""
    import crd_amd
    from crd_amd import _native as N

    precs = ("fp32", "split_bf16", "split_fp16")
    points = 1 << 27
    src = crd_amd.DeviceBuffer(8 * points)
    dst = crd_amd.DeviceBuffer(8 * points)
    seed_pts = 1 << 22
    rng = np.random.default_rng(4)
    mag = 10.0 ** rng.uniform(-4, 4, seed_pts)
    x = (mag * np.exp(1j * rng.uniform(0, 2 * np.pi, seed_pts))).astype(np.complex64)
    src.upload(x)
    for off in range(seed_pts, points, seed_pts):  # the seeded 2^22 points repeated over the buffer
        N.check(N.lib().crdfft_memcpy_d2d(src.address + 8 * off, src.ptr, 8 * seed_pts, stream.handle),
                "crdfft_memcpy_d2d")
    stream.synchronize()

    def one_length(n, sample):
        batch = points // n
        ref = np.fft.fft(x[:sample * n].reshape(sample, n).astype(np.complex128), axis=1)
        plans = {prec: crd_amd.FftPlan(n, batch, precision=prec) for prec in precs}
        # the three forms timed in alternation after ~0.2 s of warm executes each, so none of them is measured while
        # the clocks recover from host-side work; the median of 5 rounds of 8 executes
        for prec in precs:
            time_loop(lambda: plans[prec].execute(src, dst, crd_amd.FORWARD, stream), stream, 200, 0, Event)
        times = {prec: [] for prec in precs}
        for _ in range(5):
            for prec in precs:
                times[prec].append(time_loop(lambda: plans[prec].execute(src, dst, crd_amd.FORWARD, stream), stream,
                                             8, 1, Event))
        res = {"n": n, "batch": batch}
        for prec in precs:
            plans[prec].execute(src, dst, crd_amd.FORWARD, stream)
            stream.synchronize()
            y = dst.download((sample, n), np.complex64)
            ms = float(np.median(times[prec]))
            res[prec] = {"ms": round(ms, 4), "ffts_per_s": round(batch / (ms * 1e-3), 1),
                         "gbs": round(16.0 * n * batch / (ms * 1e-3) / 1e9, 1),
                         "rel_l2_vs_fp64": rel_l2(y.astype(np.complex128), ref)}
        del plans
        res["split_bf16_vs_fp32_ffts_per_s"] = round(res["fp32"]["ms"] / res["split_bf16"]["ms"], 3)
        res["split_fp16_vs_fp32_ffts_per_s"] = round(res["fp32"]["ms"] / res["split_fp16"]["ms"], 3)
        return res

    out = one_length(1 << 16, 4)
    out["other_lengths"] = [one_length(n, s) for n, s in ((256, 64), (1024, 64), (1 << 18, 2), (1 << 20, 2))]
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="only the value (for profiler runs)")
    ap.add_argument("--cpu-seconds", type=float, default=6.0)
    ap.add_argument("--inputs", choices=["normal", "device"], default="normal",
                    help="normal: seeded N(0,1) uploaded from the host; device: a byte fill on the device (profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB streaming copy (the HBM-counter calibration kernel)")
    ap.add_argument("--extras", default=",".join(EXTRAS),
                    help="comma list of the extra configurations to run (" + ", ".join(EXTRAS) + ")")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the batched FFT shards by splitting "
              "the batch, with no exchange, so the one-GPU figure is the per-GPU share", file=sys.stderr)
        sys.exit(2)

    import crd_amd
    from crd_amd import DeviceBuffer, Event, FftPlan, Stream

    if crd_amd.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    props = crd_amd.device_props(0)
    log(f"device {props['arch']} CUs={props['compute_units']}")
    stream = Stream()

    plans, bufs, firsts = [], [], []
    for idx, (n, batch) in enumerate(SHAPES):
        log(f"plan c2c n={n} batch={batch}")
        p = FftPlan(n, batch)
        nbytes = 8 * n * batch
        xin = DeviceBuffer(nbytes)
        xout = DeviceBuffer(nbytes)
        if args.inputs == "normal":
            firsts.append(fill_normal(xin, 2 * n * batch, seed=1000 + idx))
        else:
            xin.fill(0x3E)  # every float 0.18575... : finite, normal, no work skipped
            firsts.append(None)
        plans.append(p)
        bufs.append((xin, xout))

    def step():
        for p, (xin, xout) in zip(plans, bufs):
            p.execute(xin, xout, crd_amd.FORWARD, stream)

    log(f"timing {args.steps} steps after {args.warmup} warmup")
    ms_step = time_loop(step, stream, args.steps, args.warmup, Event)
    per_step = sorted(step_times(step, stream, max(20, args.steps), Event))
    step_bytes = sum(16.0 * n * b for n, b in SHAPES)
    value = step_bytes / (ms_step * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (i.i.d. N(0,1) from a fixed seed, resident in HBM)",
        "config": {
            "workload": "plain FP32 1D C2C forward, interleaved, contiguous, unnormalized: "
                        "n=65536 x batch 4096 + n=1048576 x batch 64 per step",
            "shapes": [{"n": n, "batch": b} for n, b in SHAPES],
            "parallelism": "single GPU",
        },
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "step_ms_best": round(per_step[0], 4),
        "step_ms_median": round(per_step[len(per_step) // 2], 4),
    }

    # per-shape figures and the live per-kernel roofline
    per_shape = []
    best = None
    for (n, batch), p, (xin, xout) in zip(SHAPES, plans, bufs):
        ms = time_loop(lambda: p.execute(xin, xout, crd_amd.FORWARD, stream), stream, max(3, args.steps),
                       1, Event)
        gbs = 16.0 * n * batch / (ms * 1e-3) / 1e9
        acc = None
        for _ in range(3):
            acc = p.execute_profiled(xin, xout, crd_amd.FORWARD, stream, acc)
        info = p.info()
        first_phase = {1: "single", 5: "fused four-step (both passes)"}.get(info["algorithm"], "column pass")
        names = [first_phase, "row pass (transposed store)", "real", "copy"]
        # algorithmic bytes per complex point of one launch (one read + one write of the transform's input and
        # output, intermediates not counted): 16 for a single-pass kernel, 8 for each pass of the four-step (the
        # column pass reads the input, the row pass writes the output; the intermediate between them is excluded)
        two_pass = info["algorithm"] == 2
        kernels = []
        for ph in range(4):
            if acc.launches[ph] == 0:
                continue
            avg_ms = acc.ms[ph] / acc.launches[ph]
            pts = acc.points[ph] / acc.launches[ph]
            bpp = 8.0 if (two_pass and ph in (0, 1)) else 16.0
            ach = bpp * pts / (avg_ms * 1e-3) / 1e9
            k = {"phase": names[ph], "launches_per_call": acc.launches[ph] // 3, "avg_ms": round(avg_ms, 5),
                 "points_per_launch": int(pts), "algorithmic_bytes_per_point": bpp,
                 "achieved_gbs": round(ach, 1), "l2_miss_bytes_per_point_moved": 16.0,
                 "total_ms": acc.ms[ph] / 3}
            kernels.append(k)
            if best is None or k["total_ms"] > best[0]["total_ms"]:
                best = (k, n, batch, info)
        if two_pass and len(kernels) >= 2:
            # the transform as a whole: 16 B per point over the sum of its two passes' launch times
            pts = kernels[0]["points_per_launch"]
            both = kernels[0]["avg_ms"] + kernels[1]["avg_ms"]
            kernels_transform = {"passes": 2, "ms_per_chunk": round(both, 5), "points_per_chunk": pts,
                                 "achieved_gbs": round(16.0 * pts / (both * 1e-3) / 1e9, 1)}
        else:
            kernels_transform = None
        stream.synchronize()
        head = xout.download((4096,), np.complex64)  # the first 32 KiB of the output of the last execute
        per_shape.append({"n": n, "batch": batch, "ms": round(ms, 4), "gbs": round(gbs, 1),
                          "ffts_per_s": round(batch / (ms * 1e-3), 1), "plan": {k: info[k] for k in (
                              "algorithm", "n1", "n2", "e", "e1", "tile1", "e2", "tile2", "chunk")},
                          "kernels": kernels, "transform_level": kernels_transform,
                          "output_head_fnv1a64": fnv1a64(head.tobytes())})
    result["per_shape"] = per_shape
    k, n, batch, info = best
    bpp = k["algorithmic_bytes_per_point"]
    result["roofline"] = {
        "bound": "hbm",
        "achieved": k["achieved_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(k["achieved_gbs"] / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": f"{k['phase']} of n={n}: algorithmic {bpp:g} B per complex point per launch (one read or write "
                  f"of the transform's own input/output per pass, intermediates excluded); "
                  f"{k['points_per_launch']} points per launch, avg {k['avg_ms']} ms, hipEvents on the launch stream",
    }
    tl = next((s_["transform_level"] for s_ in per_shape if s_["n"] == n), None)
    if tl is not None:
        result["roofline"]["transform_level"] = {
            "achieved": tl["achieved_gbs"], "frac": round(tl["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "what": f"both passes of n={n} per chunk: 16 B per point over the sum of the two launches' average time"}
    # HBM traffic per launch: not measurable inside this process (it needs two rocprofv3 --pmc runs of this same
    # command); the figure comes from those runs' summary, and the line says so
    traffic_file = os.path.join(HERE, "profiles", "hbm_traffic.json")
    if os.path.exists(traffic_file):
        try:
            tr = json.load(open(traffic_file))
            key = f"n={n} {k['phase']}"
            if key in tr:
                result["roofline"]["traffic"] = tr[key]["bytes_per_launch"]
                result["roofline"]["traffic_source"] = (
                    "stored: profiles/hbm_traffic.json, rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this "
                    "command (calibrated on a copy kernel), L2-miss bytes of every pass including the "
                    "intermediate; not measured in this run")
        except (ValueError, KeyError):
            pass

    if args.calibrate_copy:
        from crd_amd import _native as N

        cal = DeviceBuffer(1 << 30)
        cal2 = DeviceBuffer(1 << 30)
        N.check(N.lib().crdfft_stream_copy(cal2.ptr, cal.ptr, 1 << 30, stream.handle), "crdfft_stream_copy")
        stream.synchronize()
        del cal, cal2

    # accuracy of a sample vs FP64 (numpy's pocketfft in float64)
    acc_report = []
    for (n, batch), (xin, xout), first in zip(SHAPES, bufs, firsts):
        if first is None:
            continue
        m = min(len(first) // (2 * n), 4)
        x = first[: 2 * n * m].view(np.complex64).reshape(m, n)
        y = xout.download((m, n), np.complex64)
        ref = np.fft.fft(x.astype(np.complex128), axis=1)
        acc_report.append({"n": n, "transforms": m, "rel_l2_vs_fp64": rel_l2(y.astype(np.complex128), ref)})
    result["accuracy"] = acc_report

    want = set(x for x in args.extras.split(",") if x)
    unknown = want - set(EXTRAS)
    if unknown:
        print(f"unknown --extras {sorted(unknown)}", file=sys.stderr)
        sys.exit(2)
    if not args.no_extras:
        extras = {}
        # device copy ceiling (read + write bytes): the library's streaming copy kernel (one 16-byte element per
        # work-item) and, beside it, hipMemcpyAsync device to device
        from crd_amd import _native as N

        xin, xout = bufs[0]
        if "copy" in want:
            ms = time_loop(lambda: N.check(N.lib().crdfft_stream_copy(xout.ptr, xin.ptr, xin.nbytes, stream.handle),
                                           "crdfft_stream_copy"), stream, 5, 1, Event)
            extras["device_copy_gbs"] = round(2.0 * xin.nbytes / (ms * 1e-3) / 1e9, 1)
            ms = time_loop(lambda: xout.copy_from(xin, stream=stream), stream, 5, 1, Event)
            extras["hip_memcpy_d2d_gbs"] = round(2.0 * xin.nbytes / (ms * 1e-3) / 1e9, 1)
            # the same copy over working sets that fit the L2 (32 MiB), the 256 MB Infinity Cache (64-256 MiB) or
            # neither: what a byte that leaves L2 costs when the Infinity Cache serves it instead of DRAM
            sweep = []
            for mib in (16, 32, 64, 128, 512, 2048):
                nb = mib << 20
                ms = time_loop(lambda: N.check(N.lib().crdfft_stream_copy(xout.ptr, xin.ptr, nb, stream.handle),
                                               "crdfft_stream_copy"), stream, max(10, min(200, (4 << 30) // nb)),
                               3, Event)
                sweep.append({"set_mib": 2 * mib, "gbs": round(2.0 * nb / (ms * 1e-3) / 1e9, 1)})
            extras["copy_by_working_set"] = sweep
        # R2C -> C2R round trip of both shapes
        rts = []
        for idx, (n, batch) in enumerate(SHAPES if "roundtrip" in want else []):
            r2c = FftPlan(n, batch, "r2c")
            c2r = FftPlan(n, batch, "c2r")
            real_in = DeviceBuffer(4 * n * batch)
            spec = DeviceBuffer(8 * (n // 2 + 1) * batch)
            back = DeviceBuffer(4 * n * batch)
            first = fill_normal(real_in, n * batch, seed=2000 + idx)

            def rt():
                r2c.execute(real_in, spec, crd_amd.FORWARD, stream)
                c2r.execute(spec, back, crd_amd.INVERSE, stream)

            ms = time_loop(rt, stream, max(3, args.steps // 2), args.warmup, Event)
            bytes_rt = float(batch) * (4 * n + 8 * (n // 2 + 1)) * 2
            m = min(4, batch)
            xb = back.download((m, n), np.float32)
            err = float(np.max(np.abs(xb - first[: m * n].reshape(m, n))))
            rts.append({"n": n, "batch": batch, "ms": round(ms, 4), "gbs": round(bytes_rt / (ms * 1e-3) / 1e9, 1),
                        "max_abs_roundtrip_err": err})
            del r2c, c2r, real_in, spec, back
        if rts:
            extras["r2c_c2r_roundtrip"] = rts
        if "c2c4096" in want:
            extras["c2c_n4096_b1024"] = bench_c2c4096(stream, Event)
        if "mixed" in want:
            extras["mixed_size"] = bench_mixed(stream, Event)
        if "split" in want:
            extras["split_precision"] = bench_split(stream, Event)
        if "spectral" in want:
            extras["spectral"] = bench_spectral(stream, Event)
        if "r2r" in want:
            extras["r2r_and_bluestein"] = bench_r2r(stream, Event)
        result["extras"] = extras

    if not args.no_cpu:
        threads = os.cpu_count() or 1
        log(f"cpu baseline on {threads} threads")
        result["cpu_baseline"] = cpu_baseline(args.cpu_seconds, threads)
    else:
        result["cpu_baseline"] = None

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
