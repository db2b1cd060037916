// Device scratch owned by a plan: a buffer, and a per-stream set of copies so that one plan can be shared by host
// threads that execute it on distinct streams.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdlib>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace crdfft
{

// Fault seam for the allocation-failure tests (the reference's refusing allocator, fft.hpp:52): with
// CRDFFT_FAULT_ALLOC=k in the environment, the k-th library-owned device allocation counted from the moment the
// variable took that value fails as out of memory, so every error path of plan creation can be driven from the host
// without exhausting the device. Unset, empty or 0: no effect.
inline bool alloc_fault_forced()
{
    static std::mutex mutex;
    static std::string last;
    static uint64_t count = 0;
    const char* v = std::getenv("CRDFFT_FAULT_ALLOC");
    const std::string cur = v != nullptr ? v : "";
    std::lock_guard<std::mutex> lock(mutex);
    if (cur != last)
    {
        last = cur;
        count = 0;
    }
    const long k = cur.empty() ? 0 : std::atol(cur.c_str());
    return k > 0 && ++count == static_cast<uint64_t>(k);
}

// hipMalloc behind the fault seam: every device allocation a plan owns goes through here.
inline hipError_t device_malloc(void** p, size_t bytes)
{
    if (alloc_fault_forced())
    {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

struct DeviceBuffer
{
    void* ptr = nullptr;
    uint64_t bytes = 0;
    hipError_t alloc(uint64_t b)
    {
        bytes = b;
        if (b == 0)
        {
            return hipSuccess;
        }
        return device_malloc(&ptr, b);
    }
    ~DeviceBuffer()
    {
        if (ptr != nullptr)
        {
            (void)hipFree(ptr);
        }
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

// A scratch buffer per stream: the buffer allocated with the plan serves the first stream that executes, every other
// stream gets its own on its first execute, so one plan can be shared by host threads on distinct streams.
struct PerStream
{
    std::mutex mutex;
    bool claimed = false;
    hipStream_t first = nullptr;
    std::vector<std::pair<hipStream_t, std::unique_ptr<DeviceBuffer>>> more;
    static constexpr size_t kMaxStreams = 64;

    // The buffer of stream s (base serves the first stream); nullptr when no more can be allocated.
    void* get(DeviceBuffer& base, hipStream_t s)
    {
        if (base.bytes == 0)
        {
            return base.ptr;
        }
        std::lock_guard<std::mutex> lock(mutex);
        if (!claimed)
        {
            claimed = true;
            first = s;
        }
        if (first == s)
        {
            return base.ptr;
        }
        for (auto& e : more)
        {
            if (e.first == s)
            {
                return e.second->ptr;
            }
        }
        if (more.size() >= kMaxStreams)
        {
            return nullptr;
        }
        auto buf = std::make_unique<DeviceBuffer>();
        if (buf->alloc(base.bytes) != hipSuccess)
        {
            (void)hipGetLastError();
            return nullptr;
        }
        more.emplace_back(s, std::move(buf));
        return more.back().second->ptr;
    }
};

// Caller-free scratch for one execute on this host thread (crdfft_execute_noalloc): while set, the four-step
// passes use `ptr` (bytes long) as their intermediate, in chunks that fit it, on the caller's stream alone, and
// every path that would allocate per-stream scratch refuses instead.
struct ScratchOverride
{
    void* ptr;
    uint64_t bytes;
};
const ScratchOverride* scratch_override();
void set_scratch_override(const ScratchOverride* o);

} // namespace crdfft
