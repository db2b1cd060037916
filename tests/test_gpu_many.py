"""Mixed-size persistent launch (crdfft_many_*) and its autotuner, through the C-ABI, against the oracle's FP64 truth
(mixed-radix DFT for 2/3/5/7-smooth lengths) and the FP32 restatement's Bluestein (the reference's route for these
lengths, bluestein.hpp)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bound(n):
    return 2e-6 * max(1.0, np.log2(n))


def run_many(gpu, lengths, x, choices=None, inverse=False, offsets=None, total=None):
    from crd_amd import many

    plan = many.ManyPlan(lengths, offsets=offsets, choices=choices)
    src = gpu.DeviceBuffer.from_array(x)
    dst = gpu.DeviceBuffer(x.nbytes)
    dst.upload(np.zeros_like(x))
    plan.execute(src, dst, gpu.INVERSE if inverse else gpu.FORWARD)
    gpu.synchronize()
    return plan, dst.download(x.shape, np.complex64)


def split(y, lengths, offsets=None):
    out, pos = [], 0
    for i, n in enumerate(lengths):
        o = pos if offsets is None else offsets[i]
        out.append(y[o:o + n])
        pos += n
    return out


def check_all(po, xs, ys, inverse=False):
    for x, y in zip(xs, ys):
        n = len(x)
        ref = po.fft64_of_f32(x[None, :], inverse=inverse)[0]
        assert po.rel_l2(y, ref) <= bound(n), f"n={n}"


def test_mixed_lengths_forward_inverse(gpu, po):
    from crd_amd import workloads

    rng = np.random.default_rng(3)
    sizes = workloads.smooth_sizes(1, 16384)
    lengths = [int(s) for s in rng.choice(sizes, 120)] + [1, 2, 3, 5, 7, 16384, 15876, 2187, 3125, 2401]
    lengths = [n for n in lengths for _ in range(int(rng.integers(1, 4)))]
    x = workloads.uniform_complex(sum(lengths), 11)
    for inverse in (False, True):
        plan, y = run_many(gpu, lengths, x, inverse=inverse)
        check_all(po, split(x, lengths), split(y, lengths), inverse)
    info = plan.info()
    assert info["transforms"] == len(lengths) and info["points"] == sum(lengths)
    assert 1 <= info["workgroups"] <= 256


@pytest.mark.parametrize("n", [12, 60, 64, 96, 640, 1000, 2187, 4096, 6561, 7 * 7 * 7 * 8, 15876, 16384])
def test_every_candidate_is_exact(gpu, po, n):
    from crd_amd import many, workloads

    cands = many.radix_candidates(n)
    assert cands and all(int(np.prod(c.radices())) == n for c in cands)
    count = 5
    lengths = [n] * count
    x = workloads.uniform_complex(n * count, n)
    ref = po.fft64_of_f32(x.reshape(count, n))
    for c in cands:
        _, y = run_many(gpu, lengths, x, choices=[c])
        assert po.rel_l2(y.reshape(count, n), ref) <= bound(n), c.as_dict()


def test_offsets_gaps_and_in_place(gpu, po):
    from crd_amd import many, workloads

    lengths = [64, 1000, 7, 4096, 64, 1000]
    offsets, pos = [], 5
    for n in lengths:
        offsets.append(pos)
        pos += n + 3
    total = pos + 5
    guard = np.complex64(complex(1234.5, -6.75))
    x = np.full(total, guard, dtype=np.complex64)
    payload = workloads.uniform_complex(sum(lengths), 4)
    p = 0
    for o, n in zip(offsets, lengths):
        x[o:o + n] = payload[p:p + n]
        p += n
    plan = many.ManyPlan(lengths, offsets=offsets)
    buf = gpu.DeviceBuffer.from_array(x)
    plan.execute(buf, buf, gpu.FORWARD)  # in place
    gpu.synchronize()
    y = buf.download(x.shape, np.complex64)
    mask = np.ones(total, bool)
    for o, n in zip(offsets, lengths):
        mask[o:o + n] = False
    assert np.all(y[mask] == guard)
    check_all(po, split(x, lengths, offsets), split(y, lengths, offsets))


def test_graph_replays_match_eager(gpu, po):
    """hipGraph capture of the persistent launch: every replay, with eager work and synchronisations between
    replays, gives the eager launch's bits."""
    from crd_amd import many, workloads

    lengths = workloads.mixed_batch(total=600, seed=5)
    x = workloads.uniform_complex(int(lengths.sum()), 9)
    plan = many.ManyPlan(lengths)
    src = gpu.DeviceBuffer.from_array(x)
    eager = gpu.DeviceBuffer(x.nbytes)
    plan.execute(src, eager, gpu.FORWARD)
    gpu.synchronize()
    want = eager.download(x.shape, np.complex64)
    out = gpu.DeviceBuffer(x.nbytes)
    s = gpu.Stream()
    plan.capture_graph(src, out, gpu.FORWARD, s)
    for rep in range(4):
        out.fill(0)
        gpu.synchronize()
        plan.launch_graph(s)
        s.synchronize()
        got = out.download(x.shape, np.complex64)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"replay {rep}"
        plan.execute(src, eager, gpu.INVERSE)  # eager work between replays
        gpu.synchronize()
    check_all(po, split(x, list(lengths))[:50], split(want, list(lengths))[:50])


def test_autotune_picks_valid_fastest_plans(gpu):
    """Every length gets its fastest timed candidate, unless the planner's defaults ran the whole batch faster, in
    which case every length gets its default (candidate 0)."""
    from crd_amd import many

    lengths = [64, 1000, 1000, 4096, 15876]
    choices, records = many.autotune(lengths, reps=3)
    assert sorted(int(c.n) for c in choices) == sorted(set(lengths))
    fastest, default = [], []
    for c in choices:
        assert int(np.prod(c.radices())) == c.n
        mine = [r for r in records if r.plan.n == c.n]
        assert len(mine) == len(many.radix_candidates(int(c.n)))
        assert all(r.us_per_transform > 0 for r in mine)
        best = min(mine, key=lambda r: r.us_per_transform)
        first = many.radix_candidates(int(c.n))[0]
        fastest.append(best.plan.radices() == c.radices() and best.plan.group == c.group)
        default.append(first.radices() == c.radices() and first.group == c.group)
    assert all(fastest) or all(default), (fastest, default)


def test_rejects_non_smooth_lengths(gpu):
    from crd_amd import many
    from crd_amd import _native as N

    with pytest.raises(N.CrdfftError) as e:
        many.ManyPlan([64, 11])
    assert e.value.status == N.INVALID_SIZE
    with pytest.raises(N.CrdfftError):
        many.ManyPlan([64, 32768])


def test_full_mixed_workload_sample(gpu, po):
    """The 20,000-transform mixed-size batch (945 MB algorithmic): a seeded sample of 300 transforms against FP64
    and, over every transform, Parseval's identity (sum |X|^2 = n sum |x|^2)."""
    from crd_amd import workloads

    lengths = workloads.mixed_batch()
    x = workloads.uniform_complex(int(lengths.sum()), 20000)
    plan, y = run_many(gpu, lengths, x)
    xs, ys = split(x, list(lengths)), split(y, list(lengths))
    idx = np.random.default_rng(0).choice(len(lengths), 300, replace=False)
    check_all(po, [xs[i] for i in idx], [ys[i] for i in idx])
    ex = np.array([np.sum(np.abs(v.astype(np.complex128)) ** 2) for v in xs])
    ey = np.array([np.sum(np.abs(v.astype(np.complex128)) ** 2) for v in ys])
    rel = np.abs(ey - lengths * ex) / (lengths * ex)
    assert np.max(rel) < 1e-5
