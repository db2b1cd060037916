// The spectral front end's C-ABI (crdfft_spectral_*, crdfft_stft, crdfft_spectrogram, crdfft_welch): the GPU form of
// the reference's hesap-dsp stft / spectrogram / welch_psd / periodogram (engine/numerics/hesap-dsp/include/crd/
// hesap/dsp/spectral.hpp:43-222, :334-386) over its RealFftPlan, each one fused launch (spectral_kernels.hip).
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "kernels.hpp"
#include "planner.hpp"
#include "scratch.hpp"
#include "real.hpp"
#include "spectral.hpp"

namespace
{

// Welch: the lane-group kernel runs one resident round (CUs x workgroups per CU, at most this many workgroups); the
// grid is fixed per device and frame count, so the PSD is bit-identical run to run. The chunked path for the
// largest frames writes this many rows per chunk.
constexpr uint64_t kWelchWorkgroups = 2048;

crdfft_status st_of(hipError_t e)
{
    if (e == hipSuccess)
    {
        return CRDFFT_OK;
    }
    return e == hipErrorOutOfMemory ? CRDFFT_OUT_OF_MEMORY : CRDFFT_HIP_ERROR;
}

#define SPEC_TRY(expr)                                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        const hipError_t spec_e_ = (expr);                                                                             \
        if (spec_e_ != hipSuccess)                                                                                     \
            return st_of(spec_e_);                                                                                     \
    } while (0)

} // namespace

struct crdfft_spectral_plan_s
{
    uint64_t nperseg = 0, noverlap = 0, hop = 0, h = 0;
    uint64_t welch_wgs = 0; // the resident round of the Welch kernel
    int32_t window = CRDFFT_WINDOW_HANN;
    uint32_t flags = 0;
    int e = 0;
    unsigned fpg = 1; // frames per workgroup
    double winsum2 = 0.0;
    crdfft::DeviceBuffer win, wtab, tw, partial;
    crdfft::PerStream partial_of;
};

namespace
{

uint64_t frames_of(const crdfft_spectral_plan_s* p, uint64_t len)
{
    return len < p->nperseg ? 0 : (len - p->noverlap) / p->hop;
}

crdfft_status run(crdfft_spectral_plan p, const float* x, uint64_t len, int mode, double fs, void* out, void* stream)
{
    if (p == nullptr || x == nullptr || out == nullptr || (mode != 0 && !(fs > 0.0)))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t frames = frames_of(p, len);
    const uint64_t nfreq = p->h + 1;
    crdfft::SpecArgs a{x,
                       p->window == CRDFFT_WINDOW_HANN ? p->win.as<float>() : nullptr,
                       p->wtab.as<crdfft::cfloat>(),
                       p->tw.as<crdfft::cfloat>(),
                       out,
                       frames,
                       p->hop,
                       mode == 0 ? 1.0f : static_cast<float>(1.0 / (fs * p->winsum2)),
                       (p->flags & CRDFFT_SPECTRAL_DETREND) != 0 ? 1 : 0,
                       nfreq};
    if (mode != 2)
    {
        if (frames == 0)
        {
            return CRDFFT_OK;
        }
        SPEC_TRY(crdfft::launch_spec(p->h, mode, a, (frames + p->fpg - 1) / p->fpg, s));
        return CRDFFT_OK;
    }
    float* psd = static_cast<float*>(out);
    if (frames == 0)
    {
        SPEC_TRY(crdfft::launch_fill_bytes(psd, 0, nfreq * sizeof(float), s)); // a kernel: safe in a captured graph
        return CRDFFT_OK;
    }
    float* partial = static_cast<float*>(p->partial_of.get(p->partial, s));
    if (partial == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    const float inv = static_cast<float>(1.0 / static_cast<double>(frames));
    if (p->h <= static_cast<uint64_t>(crdfft::kSpecWelchMaxH))
    {
        const uint64_t wgs = std::min<uint64_t>((frames + p->fpg - 1) / p->fpg, p->welch_wgs);
        a.out = partial;
        SPEC_TRY(crdfft::launch_spec(p->h, 2, a, wgs, s));
        SPEC_TRY(crdfft::launch_spec_sum(partial, psd, nfreq, wgs * p->fpg, true, inv, s));
        return CRDFFT_OK;
    }
    // the largest frames: density rows for chunks of frames into the scratch, each chunk added in order
    const uint64_t chunk = kWelchWorkgroups * p->fpg;
    for (uint64_t f0 = 0; f0 < frames; f0 += chunk)
    {
        const uint64_t cnt = std::min(chunk, frames - f0);
        crdfft::SpecArgs c = a;
        c.x = x + f0 * p->hop;
        c.frames = cnt;
        c.out = partial;
        SPEC_TRY(crdfft::launch_spec(p->h, 1, c, (cnt + p->fpg - 1) / p->fpg, s));
        SPEC_TRY(crdfft::launch_spec_sum(partial, psd, nfreq, cnt, f0 == 0, f0 + cnt == frames ? inv : 1.0f, s));
    }
    return CRDFFT_OK;
}

} // namespace

extern "C" {

crdfft_status crdfft_spectral_plan_create(crdfft_spectral_plan* plan, uint64_t nperseg, uint64_t noverlap,
                                          int32_t window, uint32_t flags)
{
    if (plan == nullptr || (window != CRDFFT_WINDOW_BOXCAR && window != CRDFFT_WINDOW_HANN) ||
        (flags & ~CRDFFT_SPECTRAL_DETREND) != 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *plan = nullptr;
    if (nperseg < 4 || (nperseg & (nperseg - 1)) != 0 || noverlap >= nperseg)
    {
        return CRDFFT_INVALID_SIZE; // a RealFftPlan: a power of two >= 4 (real_fft.hpp:24)
    }
    int e = 0;
    unsigned fpg = 0;
    if (!crdfft::spec_shape(nperseg / 2, &e, &fpg))
    {
        return CRDFFT_UNSUPPORTED;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    {
        return CRDFFT_NO_DEVICE;
    }
    auto* p = new (std::nothrow) crdfft_spectral_plan_s;
    if (p == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    p->nperseg = nperseg;
    p->noverlap = noverlap;
    p->hop = nperseg - noverlap;
    p->h = nperseg / 2;
    p->window = window;
    p->flags = flags;
    p->e = e;
    p->fpg = fpg;
    // periodic hann (windows.hpp:63-124, sym = false): 0.5 + 0.5 cos(-pi + 2 pi i / nperseg), from FP64
    std::vector<float> w(nperseg, 1.0f);
    if (window == CRDFFT_WINDOW_HANN)
    {
        for (uint64_t i = 0; i < nperseg; ++i)
        {
            const double fac = -M_PI + 2.0 * M_PI * static_cast<double>(i) / static_cast<double>(nperseg);
            w[i] = static_cast<float>(0.5 + 0.5 * std::cos(fac));
        }
    }
    for (float v : w)
    {
        p->winsum2 += static_cast<double>(v) * static_cast<double>(v);
    }
    if (p->h <= static_cast<uint64_t>(crdfft::kSpecWelchMaxH))
    {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t re = hipGetDevice(&dev);
        if (re == hipSuccess)
            re = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (re == hipSuccess)
            re = crdfft::spec_welch_residency(p->h, &per_cu);
        if (re != hipSuccess)
        {
            delete p;
            return st_of(re);
        }
        p->welch_wgs = std::max<uint64_t>(1, std::min<uint64_t>(static_cast<uint64_t>(cus) * std::max(per_cu, 1),
                                                                kWelchWorkgroups));
    }
    const uint64_t entries = crdfft::stage_twiddle_entries(p->h, e);
    hipError_t err = p->win.alloc(nperseg * sizeof(float));
    if (err == hipSuccess)
        err = hipMemcpy(p->win.ptr, w.data(), nperseg * sizeof(float), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = p->wtab.alloc((p->h + 1) * sizeof(crdfft::cfloat));
    if (err == hipSuccess)
        err = crdfft::build_roots(p->wtab.as<crdfft::cfloat>(), p->h + 1, 1, nperseg, nullptr);
    if (err == hipSuccess)
        err = p->tw.alloc(std::max<uint64_t>(entries, 1) * sizeof(crdfft::cfloat));
    if (err == hipSuccess && entries > 0)
        err = crdfft::build_stage_twiddles(p->tw.as<crdfft::cfloat>(), p->h, e, nullptr);
    if (err == hipSuccess)
        err = p->partial.alloc(kWelchWorkgroups * fpg * (p->h + 1) * sizeof(float));
    if (err == hipSuccess)
        err = hipStreamSynchronize(nullptr);
    if (err != hipSuccess)
    {
        delete p;
        return st_of(err);
    }
    *plan = p;
    return CRDFFT_OK;
}

crdfft_status crdfft_spectral_plan_destroy(crdfft_spectral_plan plan)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    delete plan;
    return CRDFFT_OK;
}

uint64_t crdfft_spectral_frames(crdfft_spectral_plan plan, uint64_t len)
{
    return plan == nullptr ? 0 : frames_of(plan, len);
}

crdfft_status crdfft_stft(crdfft_spectral_plan plan, const float* x, uint64_t len, void* out, void* stream)
{
    return run(plan, x, len, 0, 0.0, out, stream);
}

crdfft_status crdfft_spectrogram(crdfft_spectral_plan plan, const float* x, uint64_t len, double fs, float* out,
                                 void* stream)
{
    return run(plan, x, len, 1, fs, out, stream);
}

crdfft_status crdfft_welch(crdfft_spectral_plan plan, const float* x, uint64_t len, double fs, float* psd,
                           void* stream)
{
    return run(plan, x, len, 2, fs, psd, stream);
}

} // extern "C"
