// The known-answer self-check of a plan (crdfft_self_check), the GPU counterpart of the reference's HW-RESILIENCE
// check (engine/numerics/hesap-fft/src/fft_self_check.cpp:1-211): run inputs whose transforms are known in closed
// form through the plan's own kernels and compare on the host.
//
// The reference pins its kernels to FNV-1a digests of its exact output bits; a GPU build cannot reproduce another
// implementation's bits, so the answers here are analytic: an impulse at j = 1 transforms to W_n^k, a tone
// exp(+2 pi i 3 j / n) to n at bin 3 and zero elsewhere, a constant to n at bin 0; for real plans a cosine of
// frequency 1 to n/2 at bin 1 and back. Tolerance: 2e-6 * log2(n) * (1 + max |X|), the plain path's accuracy class.
// A failing check returns CRDFFT_SELF_CHECK_FAILED; plan creation then demotes the plan to its floor (the radix-2
// reference kernels) and checks again (crdfft_check_status). CRDFFT_FAULT_SELF_CHECK in the environment corrupts one
// output value before the comparison, the host fault seam (reference crd/core/fault_seam.hpp:139) that lets a test
// force either path: "1" for the selected kernels only, "2" for every check.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "../../include/crdfft_runtime.h"
#include "planner.hpp"
#include "scratch.hpp"

namespace
{

constexpr double kTwoPi = 6.283185307179586476925286766559;

struct DevBuf
{
    void* p = nullptr;
    ~DevBuf()
    {
        if (p != nullptr)
        {
            (void)hipFree(p);
        }
    }
};

// The fault seam: "1" fails the check of a plan's selected kernels (so it demotes to its floor), "2" fails every
// check (the floor's too).
bool fault_forced(crdfft_plan plan)
{
    const char* v = std::getenv("CRDFFT_FAULT_SELF_CHECK");
    if (v == nullptr)
    {
        return false;
    }
    return v[0] == '2' || (v[0] == '1' && !crdfft::plan_on_floor(plan));
}

// The single-transform plan the check runs: the same kernels as `plan`, on its floor when the plan is.
crdfft_status check_plan(crdfft_plan plan, const crdfft_plan_info& info, crdfft_plan* one);

// Indices checked for an n-point output: all of them up to 2^16, else a stride that visits 2^16 of them plus the
// first and last 64.
bool checked(uint64_t k, uint64_t n)
{
    if (n <= (uint64_t{1} << 16))
    {
        return true;
    }
    const uint64_t stride = n >> 16;
    return k < 64 || k + 64 >= n || (k % stride) == (k / stride) % stride;
}

crdfft_tuning same_kernels(const crdfft_plan_info& info)
{
    crdfft_tuning t{};
    t.n1 = info.n1;
    t.e = info.e;
    t.e1 = info.e1;
    t.tile1 = info.tile1;
    t.e2 = info.e2;
    t.tile2 = info.tile2;
    t.chunk = 1;
    t.algorithm = info.algorithm;
    return t;
}

crdfft_status check_plan(crdfft_plan plan, const crdfft_plan_info& info, crdfft_plan* one)
{
    const crdfft_tuning same = same_kernels(info);
    crdfft_status st = crdfft_plan_create_tuned(one, info.n, 1, 0, 0, info.type, CRDFFT_FP32,
                                                CRDFFT_FLAG_NO_SELF_CHECK, &same);
    if (st == CRDFFT_OK && crdfft::plan_on_floor(plan))
    {
        st = crdfft::plan_use_floor(*one);
        if (st != CRDFFT_OK)
        {
            crdfft_plan_destroy(*one);
            *one = nullptr;
        }
    }
    return st;
}

} // namespace

extern "C" crdfft_status crdfft_self_check(crdfft_plan plan)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    crdfft_plan_info info{};
    crdfft_status st = crdfft_plan_get_info(plan, &info);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    if (info.precision != CRDFFT_FP32)
    {
        return CRDFFT_OK; // the split-precision kernels check themselves against FP64 in their own tests
    }
    const uint64_t n = info.n;
    const double lg = std::max(1.0, std::log2(static_cast<double>(n)));
    const double tol = 2e-6 * lg;
    const bool fault = fault_forced(plan);

    if (info.type == CRDFFT_C2C)
    {
        // the plan's distances are its own; the check builds a fresh plan-shaped call on packed buffers of at most
        // the plan's batch (three cases run as three calls of one transform each when the batch is 1)
        const uint64_t cases = 3;
        std::vector<std::complex<float>> host(n), out(n);
        DevBuf din, dout;
        if (crdfft::device_malloc(&din.p, n * 8) != hipSuccess || crdfft::device_malloc(&dout.p, n * 8) != hipSuccess)
        {
            return CRDFFT_OUT_OF_MEMORY;
        }
        // a single-transform plan with the same kernels as `plan` (same split, values per thread and tiles; only
        // the chunk differs, which does not change the arithmetic)
        crdfft_plan one = nullptr;
        st = check_plan(plan, info, &one);
        if (st != CRDFFT_OK)
        {
            return st;
        }
        crdfft_status result = CRDFFT_OK;
        for (int dir = 0; dir < 2 && result == CRDFFT_OK; ++dir)
        {
            for (uint64_t c = 0; c < cases && result == CRDFFT_OK; ++c)
            {
                for (uint64_t j = 0; j < n; ++j)
                {
                    if (c == 0)
                    {
                        host[j] = (j == 1 % n) ? std::complex<float>(1.0f, 0.0f) : std::complex<float>(0.0f, 0.0f);
                    }
                    else if (c == 1)
                    {
                        const double a = (dir == 0 ? 1.0 : -1.0) * kTwoPi * static_cast<double>((3 * j) % n) /
                                         static_cast<double>(n);
                        host[j] = std::complex<float>(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
                    }
                    else
                    {
                        host[j] = std::complex<float>(1.0f, 0.0f);
                    }
                }
                if (hipMemcpy(din.p, host.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess)
                {
                    result = CRDFFT_HIP_ERROR;
                    break;
                }
                st = crdfft_execute(one, din.p, dout.p, dir, nullptr);
                if (st != CRDFFT_OK)
                {
                    result = st;
                    break;
                }
                if (hipMemcpy(out.data(), dout.p, n * 8, hipMemcpyDeviceToHost) != hipSuccess)
                {
                    result = CRDFFT_HIP_ERROR;
                    break;
                }
                if (fault && c == 0)
                {
                    uint32_t bits;
                    std::memcpy(&bits, &out[n - 1], 4);
                    bits ^= 0x00400000u;
                    std::memcpy(&out[n - 1], &bits, 4);
                }
                const double maxref = (c == 0) ? 1.0 : static_cast<double>(n);
                double maxerr = 0.0;
                for (uint64_t k = 0; k < n; ++k)
                {
                    if (!checked(k, n))
                    {
                        continue;
                    }
                    std::complex<double> want(0.0, 0.0);
                    if (c == 0)
                    {
                        const double a = -(dir == 0 ? 1.0 : -1.0) * kTwoPi * static_cast<double>(k % n) /
                                         static_cast<double>(n);
                        want = std::complex<double>(std::cos(a), std::sin(a));
                    }
                    else if (c == 1)
                    {
                        want = (k == 3 % n) ? std::complex<double>(static_cast<double>(n), 0.0) : want;
                    }
                    else
                    {
                        want = (k == 0) ? std::complex<double>(static_cast<double>(n), 0.0) : want;
                    }
                    const std::complex<double> got(out[k].real(), out[k].imag());
                    const double err = std::abs(got - want);
                    if (!(err <= maxerr))
                    {
                        maxerr = err; // NaN also lands here and fails below
                    }
                }
                if (!(maxerr <= tol * (1.0 + maxref)))
                {
                    result = CRDFFT_SELF_CHECK_FAILED;
                }
            }
        }
        crdfft_plan_destroy(one);
        return result;
    }

    // R2C / C2R: a cosine of frequency 1 and its half spectrum
    const uint64_t h = n / 2;
    DevBuf dr, dc;
    if (crdfft::device_malloc(&dr.p, n * 4) != hipSuccess || crdfft::device_malloc(&dc.p, (h + 1) * 8) != hipSuccess)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    crdfft_plan one = nullptr;
    st = check_plan(plan, info, &one);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    crdfft_status result = CRDFFT_OK;
    if (info.type == CRDFFT_R2C)
    {
        std::vector<float> x(n);
        for (uint64_t j = 0; j < n; ++j)
        {
            x[j] = static_cast<float>(std::cos(kTwoPi * static_cast<double>(j) / static_cast<double>(n)));
        }
        std::vector<std::complex<float>> X(h + 1);
        if (hipMemcpy(dr.p, x.data(), n * 4, hipMemcpyHostToDevice) != hipSuccess)
        {
            result = CRDFFT_HIP_ERROR;
        }
        if (result == CRDFFT_OK)
        {
            result = crdfft_execute(one, dr.p, dc.p, CRDFFT_FORWARD, nullptr);
        }
        if (result == CRDFFT_OK && hipMemcpy(X.data(), dc.p, (h + 1) * 8, hipMemcpyDeviceToHost) != hipSuccess)
        {
            result = CRDFFT_HIP_ERROR;
        }
        if (result == CRDFFT_OK)
        {
            if (fault)
            {
                X[1] = std::complex<float>(X[1].real() * 1.5f, X[1].imag());
            }
            double maxerr = 0.0;
            for (uint64_t k = 0; k <= h; ++k)
            {
                if (!checked(k, h + 1))
                {
                    continue;
                }
                const double want = (k == 1) ? static_cast<double>(h) : 0.0;
                const double err = std::abs(std::complex<double>(X[k].real() - want, X[k].imag()));
                if (!(err <= maxerr))
                {
                    maxerr = err;
                }
            }
            if (!(maxerr <= tol * (1.0 + static_cast<double>(h))))
            {
                result = CRDFFT_SELF_CHECK_FAILED;
            }
        }
    }
    else
    {
        std::vector<std::complex<float>> X(h + 1, std::complex<float>(0.0f, 0.0f));
        X[1] = std::complex<float>(static_cast<float>(h), 0.0f);
        std::vector<float> x(n);
        if (hipMemcpy(dc.p, X.data(), (h + 1) * 8, hipMemcpyHostToDevice) != hipSuccess)
        {
            result = CRDFFT_HIP_ERROR;
        }
        if (result == CRDFFT_OK)
        {
            result = crdfft_execute(one, dc.p, dr.p, CRDFFT_INVERSE, nullptr);
        }
        if (result == CRDFFT_OK && hipMemcpy(x.data(), dr.p, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        {
            result = CRDFFT_HIP_ERROR;
        }
        if (result == CRDFFT_OK)
        {
            if (fault)
            {
                x[0] += 0.5f;
            }
            double maxerr = 0.0;
            for (uint64_t j = 0; j < n; ++j)
            {
                if (!checked(j, n))
                {
                    continue;
                }
                const double want = std::cos(kTwoPi * static_cast<double>(j) / static_cast<double>(n));
                const double err = std::abs(static_cast<double>(x[j]) - want);
                if (!(err <= maxerr))
                {
                    maxerr = err;
                }
            }
            if (!(maxerr <= tol * 2.0))
            {
                result = CRDFFT_SELF_CHECK_FAILED;
            }
        }
    }
    crdfft_plan_destroy(one);
    return result;
}
