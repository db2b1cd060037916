"""The C-ABI boundary on the CPU: the library loads, exports every declared entry point, the planner answers without
a device, arguments are validated, and the C++20 interface compiles."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from crd_amd import _native as N

    lib = N.lib()
    names = N.header_symbols()
    assert len(names) >= 30
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_python_binding_covers_every_symbol():
    from crd_amd import _native as N

    unbound = [s for s in N.header_symbols() if s not in N._SIGS]
    assert not unbound, f"entry points without a ctypes signature: {unbound}"


def test_version_and_status_strings():
    from crd_amd import _native as N

    assert N.lib().crdfft_version().startswith(b"crdfft")
    for st in range(8):
        assert N.status_string(st) != "unknown status"
    assert N.status_string(99) == "unknown status"


@pytest.mark.parametrize("lg", range(0, 25))
def test_planner_covers_every_pow2(lg):
    import crd_amd

    n = 1 << lg
    info = crd_amd.describe(n, 4)
    assert info["complex_n"] == n
    if n == 1:
        assert info["algorithm"] == 0
    elif n <= 16384:
        assert info["algorithm"] == 1 and info["e"] >= 1
    else:
        assert info["algorithm"] == 2
        assert info["n1"] * info["n2"] == n
        assert info["workspace_bytes"] == info["chunk"] * n * 8
        # the intermediate of one pass pair stays within the 256 MiB budget
        assert info["workspace_bytes"] <= (256 << 20) or info["chunk"] == 1


def test_planner_real_transforms_use_half_length():
    import crd_amd

    for n in (4, 64, 1 << 16, 1 << 20):
        for t in ("r2c", "c2r"):
            assert crd_amd.describe(n, 2, t)["complex_n"] == n // 2


def test_planner_defaults_follow_the_measured_tables():
    """Values per thread and cache hints of the single kernel (C2C: the per-length sweep; R2C/C2R: the one-kernel
    real forms' instantiations, so the plan's stage table serves them), and the 512-point row pass of the four-step."""
    import crd_amd

    want_c2c = {2: 2, 4: 4, 8: 8, 16: 4, 64: 4, 128: 8, 256: 16, 512: 16, 1024: 8, 4096: 8, 8192: 32, 16384: 16}
    for n, e in want_c2c.items():
        info = crd_amd.describe(n, 4)
        assert info["e"] == e, n
        assert info["hints"] == (3 if n >= 256 else 0), n
    want_half = {2: 2, 4: 4, 8: 8, 16: 4, 256: 4, 512: 8, 8192: 8, 16384: 16}
    for h, e in want_half.items():
        for t in ("r2c", "c2r"):
            assert crd_amd.describe(2 * h, 4, t)["e"] == e, (t, h)
    for lg, e2 in ((17, 8), (19, 8), (20, 8), (16, 16), (21, 16)):
        assert crd_amd.describe(1 << lg, 4)["e2"] == e2, lg


def test_planner_four_step_schedules():
    """The measured four-step schedules: 256 MiB chunks in one pipeline up to 1024-point columns; two pipelines of
    128 MiB chunks from 2048-point columns when the batch needs at least two of them; C2R at half-size chunks; R2C
    without cache hints; the n = 2^16 real splits."""
    import crd_amd
    from crd_amd import _native as N

    two = N.HINT_TWO_PIPELINES
    d = crd_amd.describe(1 << 16, 4096)
    assert (d["n1"], d["chunk"], d["hints"]) == (256, 512, 3)
    d = crd_amd.describe(1 << 18, 1024)
    assert (d["n1"], d["chunk"], d["hints"]) == (1024, 128, 3)
    d = crd_amd.describe(1 << 20, 64)
    assert (d["n1"], d["chunk"], d["hints"]) == (2048, 16, 3 | two)
    d = crd_amd.describe(1 << 20, 16)  # one chunk: one pipeline
    assert (d["chunk"], d["hints"]) == (16, 3)
    d = crd_amd.describe(1 << 24, 4)
    assert (d["n1"], d["chunk"], d["hints"]) == (4096, 1, 3 | two)
    assert crd_amd.describe(1 << 20, 64, "r2c")["hints"] & two == 0
    d = crd_amd.describe(1 << 16, 4096, "c2r")
    assert (d["n1"], d["n2"], d["e2"], d["chunk"], d["hints"]) == (256, 128, 8, 512, 3)
    d = crd_amd.describe(1 << 20, 64, "c2r")
    assert (d["chunk"], d["hints"]) == (32, 3)
    d = crd_amd.describe(1 << 16, 4096, "r2c")  # 256 x 128 like C2R, 32 rows of 8 values, half-size chunks
    assert (d["n1"], d["n2"], d["e2"], d["tile2"], d["chunk"], d["hints"]) == (256, 128, 8, 32, 512, 0)
    d = crd_amd.describe(1 << 20, 64, "r2c")  # 1024-point columns: half-size chunks
    assert (d["n1"], d["chunk"], d["hints"]) == (1024, 32, 0)
    # a tuned plan can switch the pipelines either way
    info = N.PlanInfo()
    t = N.Tuning(hints=3)
    assert N.lib().crdfft_plan_describe_tuned(1 << 20, 64, 0, 0, ctypes.byref(t), ctypes.byref(info)) == N.OK
    assert info.hints == 3 and info.chunk == 32
    t = N.Tuning(hints=3 | two, chunk=4)
    assert N.lib().crdfft_plan_describe_tuned(1 << 18, 64, 0, 0, ctypes.byref(t), ctypes.byref(info)) == N.OK
    assert info.hints == 3 | two and info.chunk == 4


def test_tuning_hints_none_is_explicit():
    """CRDFFT_HINTS_NONE asks for plain loads and stores; 0 keeps the planner's default."""
    import crd_amd
    from crd_amd import _native as N

    info = N.PlanInfo()
    t = N.Tuning(hints=N.HINTS_NONE)
    assert N.lib().crdfft_plan_describe_tuned(4096, 4, 0, 0, ctypes.byref(t), ctypes.byref(info)) == N.OK
    assert info.hints == 0
    t = N.Tuning()
    assert N.lib().crdfft_plan_describe_tuned(4096, 4, 0, 0, ctypes.byref(t), ctypes.byref(info)) == N.OK
    assert info.hints == crd_amd.describe(4096, 4)["hints"] == 3


def test_planner_rejects_bad_shapes():
    import crd_amd
    from crd_amd import _native as N

    for n in (0, 3, 12, 1000, 1 << 25):
        with pytest.raises(N.CrdfftError) as e:
            crd_amd.describe(n, 1)
        assert e.value.status == N.INVALID_SIZE
    for n in (1, 2):
        with pytest.raises(N.CrdfftError) as e:
            crd_amd.describe(n, 1, "r2c")
        assert e.value.status == N.INVALID_SIZE
    info = N.PlanInfo()
    assert N.lib().crdfft_plan_describe(64, 1, 7, 0, ctypes.byref(info)) == N.INVALID_ARGUMENT
    assert N.lib().crdfft_plan_describe(64, 1, 0, 9, ctypes.byref(info)) == N.INVALID_ARGUMENT
    assert N.lib().crdfft_plan_describe(64, 1, 0, 0, None) == N.INVALID_ARGUMENT


def test_null_handles_are_rejected_without_a_device():
    from crd_amd import _native as N

    L = N.lib()
    assert L.crdfft_plan_destroy(None) == N.INVALID_ARGUMENT
    assert L.crdfft_execute(None, None, None, 0, None) == N.INVALID_ARGUMENT
    assert L.crdfft_execute_batched(None, None, 1, 0, None) == N.INVALID_ARGUMENT
    assert L.crdfft_self_check(None) == N.INVALID_ARGUMENT
    assert L.crdfft_plan_create(None, 64, 1, 0, 0, 0, 0, 0) == N.INVALID_ARGUMENT
    # graph capture: the null stream cannot be captured, and null graphs are refused before any HIP call
    g = ctypes.c_void_p()
    assert L.crdfft_graph_begin(None) == N.INVALID_ARGUMENT
    assert L.crdfft_graph_end(None, ctypes.byref(g)) == N.INVALID_ARGUMENT
    assert L.crdfft_graph_launch(None, None) == N.INVALID_ARGUMENT
    assert L.crdfft_graph_destroy(None) == N.INVALID_ARGUMENT


@pytest.mark.parametrize("name", ["wrapper_use.cpp", "reference_call_sites.cpp"])
def test_cpp_interface_compiles(name):
    """Both host programs compile against the C++ interface: the device-native forms (wrapper_use.cpp) and call
    sites written exactly as the reference's tests write them, over host spans (reference_call_sites.cpp)."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = os.path.join(ROOT, "tests", "cpp", name)
    r = subprocess.run([gxx, "-std=c++20", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_algorithm_hints_map_onto_device_plans(tmp_path):
    """FftPlan(alloc, n, FftAlgorithm) (fft.hpp:232): every one of the nine hints yields a plan the planner accepts,
    FourStep and SixStep select the four-step above n = 16384, and SixStep takes the square split there (the
    reference's six-step changes only the split, fft.hpp:233-236). Host-only: the mapping goes through
    crdfft_plan_describe_tuned, which touches no device."""
    from crd_amd import _native as N

    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    code = r"""
#include <crd_amd/fft.hpp>
#include <cstdio>
using namespace crd_amd::hesap::fft;
int main()
{
    for (int a = 0; a < 9; ++a)
    {
        for (unsigned lg : {4u, 12u, 16u, 17u, 20u})
        {
            const crdfft_tuning t = detail::algorithm_tuning(1ull << lg, static_cast<FftAlgorithm>(a));
            crdfft_plan_info info{};
            const int st = crdfft_plan_describe_tuned(1ull << lg, 1, CRDFFT_C2C, CRDFFT_FP32, &t, &info);
            std::printf("%s %u %d %d %llu\n", fft_algorithm_name(static_cast<FftAlgorithm>(a)), lg, st,
                        info.algorithm, static_cast<unsigned long long>(info.n1));
        }
    }
    return 0;
}
"""
    exe = str(tmp_path / "algo_hints")
    libdir = os.path.join(ROOT, "crd_amd")
    r = subprocess.run([gxx, "-std=c++20", "-x", "c++", "-", "-I", os.path.join(ROOT, "include"), "-L", libdir,
                        "-lcrdfft", f"-Wl,-rpath,{libdir}", "-o", exe], input=code, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    rows = [ln.split() for ln in out.stdout.strip().splitlines()]
    assert len(rows) == 45
    for name, lg, st, algo, n1 in rows:
        lg, st, algo, n1 = int(lg), int(st), int(algo), int(n1)
        assert st == N.OK, (name, lg)
        if lg <= 14:
            assert algo == 1, (name, lg)  # one LDS-resident kernel
        else:
            assert algo == 2, (name, lg)  # the four-step
        if name == "SixStep" and lg > 14:
            assert n1 == 1 << (lg // 2), (name, lg, n1)
    names = sorted({r[0] for r in rows})
    assert names == sorted(["Stockham", "StandaloneHier", "Ip4", "DeepSplit3", "DeepSplit4", "FourStep", "SixStep",
                            "TiledFourStep", "InPlace3"])


def test_c_header_compiles_as_c():
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    code = '#include "crdfft.h"\n#include "crdfft_runtime.h"\nint main(void){crdfft_plan p = 0; (void)p; return 0;}\n'
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-"], input=code, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
