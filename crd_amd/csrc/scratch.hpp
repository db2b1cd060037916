// Device scratch owned by a plan: a buffer, and a per-stream set of copies so that one plan can be shared by host
// threads that execute it on distinct streams.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace crdfft
{

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
        return hipMalloc(&ptr, b);
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

} // namespace crdfft
