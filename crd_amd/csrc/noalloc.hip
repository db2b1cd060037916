// crdfft_execute_noalloc: crdfft_execute without any device allocation, for callers that cannot allocate (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The four-step passes use this
// file's static device array as their intermediate instead of the plan's per-stream workspace (which a first execute
// on a new stream allocates), in chunks that fit it, on the caller's stream alone (no second chunk pipeline, whose
// stream and workspace a first use would allocate).
//
// The array is kScratchBytes per device, loaded with the library's code object and held until the process ends.
// Every call on one device shares it: calls that overlap on one device (two streams, two host threads) race on it,
// so a caller serializes them. Its contents are the last call's intermediate; nothing in it has to start at zero
// (the column pass writes every value the row pass reads).
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/crdfft.h"
#include "planner.hpp"
#include "scratch.hpp"

namespace
{

constexpr uint64_t kScratchBytes = uint64_t{64} << 20; // 64 MiB: one n = 2^23 transform, or 512 of 2^14

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mutex;
std::vector<void*> g_ptr; // per device: the array's device address, from hipGetSymbolAddress

// The array's address on the current device (looked up once per device, with that device current).
crdfft_status scratch_address(void** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
    {
        return CRDFFT_HIP_ERROR;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ptr.size() <= static_cast<size_t>(dev))
    {
        g_ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
    }
    if (g_ptr[dev] == nullptr)
    {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
        {
            return CRDFFT_HIP_ERROR;
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return CRDFFT_OK;
}

struct OverrideGuard
{
    explicit OverrideGuard(const crdfft::ScratchOverride* o) { crdfft::set_scratch_override(o); }
    ~OverrideGuard() { crdfft::set_scratch_override(nullptr); }
};

} // namespace

extern "C" crdfft_status crdfft_execute_noalloc(crdfft_plan plan, const void* in, void* out, int32_t direction,
                                                void* stream)
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
    // plan kinds whose only scratch is the four-step intermediate (or none); the cooperative, fused, split-precision
    // and chirp-z forms keep structured workspaces of their own
    const bool kind_ok = info.precision == CRDFFT_FP32 &&
                         (info.algorithm == CRDFFT_ALGO_COPY || info.algorithm == CRDFFT_ALGO_SINGLE ||
                          info.algorithm == CRDFFT_ALGO_FOUR_STEP);
    if (!kind_ok)
    {
        return CRDFFT_UNSUPPORTED;
    }
    if (info.algorithm == CRDFFT_ALGO_FOUR_STEP && info.complex_n * 8 > kScratchBytes)
    {
        return CRDFFT_INVALID_SIZE; // one transform's intermediate does not fit the static scratch
    }
    void* ptr = nullptr;
    st = scratch_address(&ptr);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    const crdfft::ScratchOverride ov{ptr, kScratchBytes};
    const OverrideGuard guard(&ov);
    return crdfft_execute(plan, in, out, direction, stream);
}

extern "C" uint64_t crdfft_noalloc_scratch_bytes(void) { return kScratchBytes; }
