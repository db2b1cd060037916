// C++20 host interface of the MI355X batched FFT path: the reference's FftPlan / RealFftPlan / BluesteinPlan /
// NdFftPlan / DctPlan / fft / ifft_normalized (engine/numerics/hesap-fft/include/crd/hesap/fft/fft.hpp:220-1029,
// real_fft.hpp:18-101, bluestein.hpp, nd_fft.hpp, dct.hpp) and hesap-dsp's spectral front end over the crdfft C-ABI,
// header-only.
//
// Same names, argument meaning and conventions as the reference: forward and inverse unnormalized, irfft scaled by
// 1/n, Complex<T> interleaved {re, im}, execute in place on n values, execute_batched element-major
// (data[i*batch + t]). Two ways in:
//   - the reference's own call shapes, unchanged: FftPlan<f32>(alloc, n), FftPlan<f32>(alloc, n, FftAlgorithm),
//     RealFftPlan / BluesteinPlan / DctPlan (alloc, n), NdFftPlan(alloc, dims), fft(alloc, span, dir),
//     ifft_normalized(alloc, span), execute / execute_batched / execute_reference on any span type with data() and
//     size() (crd::containers::Span included). A span over host memory is staged through device memory and the
//     call returns with the result in place, like the reference's synchronous CPU call; a span over device memory
//     runs asynchronously on the plan's stream. The allocator is accepted and not used: a device plan's tables and
//     workspace are device memory (the reference allocates its host tables from it, fft.hpp:478-491);
//   - the device-native forms: DeviceSpan views, a plan of `batch` contiguous transforms and a HIP stream.
// Differences, all forced by the device boundary:
//   - a contract violation throws FftError carrying the crdfft status (the reference asserts, fft.hpp:237);
//   - the FftAlgorithm hint names CPU engines; algorithm_tuning() maps each onto a device plan kind.
// Only T = float exists: the GPU path is FP32.
#pragma once

#include <concepts>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "../crdfft.h"
#include "../crdfft_runtime.h"

namespace crd_amd::hesap::fft
{

enum class FftDirection : std::uint8_t
{
    Forward = 0,
    Inverse = 1,
};

// FftAlgorithm (fft_kernels.hpp:28-39): the engine a reference plan runs forward. The device plans have their own
// kinds (crdfft_algorithm); algorithm_tuning() below records how each hint maps onto them.
enum class FftAlgorithm : std::uint8_t
{
    Stockham,
    StandaloneHier,
    Ip4,
    DeepSplit3,
    DeepSplit4,
    FourStep,
    SixStep,
    TiledFourStep,
    InPlace3,
};

// fft_algorithm_name (fft_kernels.hpp:41-64): the enumerator's name.
[[nodiscard]] constexpr const char* fft_algorithm_name(FftAlgorithm algorithm) noexcept
{
    switch (algorithm)
    {
        case FftAlgorithm::Stockham: return "Stockham";
        case FftAlgorithm::StandaloneHier: return "StandaloneHier";
        case FftAlgorithm::Ip4: return "Ip4";
        case FftAlgorithm::DeepSplit3: return "DeepSplit3";
        case FftAlgorithm::DeepSplit4: return "DeepSplit4";
        case FftAlgorithm::FourStep: return "FourStep";
        case FftAlgorithm::SixStep: return "SixStep";
        case FftAlgorithm::TiledFourStep: return "TiledFourStep";
        case FftAlgorithm::InPlace3: return "InPlace3";
    }
    return "?";
}

// Layout of crd::hesap::Complex<T> (engine/numerics/hesap/include/crd/hesap/complex.hpp:27-35).
template <typename T> struct Complex
{
    T re;
    T im;
};
static_assert(sizeof(Complex<float>) == 8, "interleaved FP32 complex");

class FftError : public std::runtime_error
{
public:
    FftError(crdfft_status status, const char* what)
        : std::runtime_error(std::string(what) + ": " + crdfft_status_string(status)), m_status(status)
    {
    }
    [[nodiscard]] crdfft_status status() const noexcept { return m_status; }

private:
    crdfft_status m_status;
};

inline void check(crdfft_status st, const char* what)
{
    if (st != CRDFFT_OK)
    {
        throw FftError(st, what);
    }
}

// A view of device memory: `size` values of T at `data`.
template <typename T> class DeviceSpan
{
public:
    constexpr DeviceSpan() = default;
    constexpr DeviceSpan(T* data, std::size_t size) : m_data(data), m_size(size) {}
    [[nodiscard]] constexpr T* data() const noexcept { return m_data; }
    [[nodiscard]] constexpr std::size_t size() const noexcept { return m_size; }

private:
    T* m_data = nullptr;
    std::size_t m_size = 0;
};

// An owning device allocation of `count` values of T.
template <typename T> class DeviceArray
{
public:
    explicit DeviceArray(std::size_t count) : m_size(count)
    {
        void* p = nullptr;
        check(crdfft_malloc(&p, count * sizeof(T)), "crdfft_malloc");
        m_data = static_cast<T*>(p);
    }
    ~DeviceArray()
    {
        if (m_data != nullptr)
        {
            (void)crdfft_free(m_data);
        }
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept : m_data(std::exchange(o.m_data, nullptr)), m_size(o.m_size) {}
    DeviceArray& operator=(DeviceArray&&) = delete;

    [[nodiscard]] DeviceSpan<T> span() const noexcept { return {m_data, m_size}; }
    [[nodiscard]] T* data() const noexcept { return m_data; }
    [[nodiscard]] std::size_t size() const noexcept { return m_size; }
    void upload(const T* host) { check(crdfft_memcpy_h2d(m_data, host, m_size * sizeof(T)), "crdfft_memcpy_h2d"); }
    void download(T* host) const
    {
        check(crdfft_memcpy_d2h(host, m_data, m_size * sizeof(T)), "crdfft_memcpy_d2h");
    }

private:
    T* m_data = nullptr;
    std::size_t m_size = 0;
};

// Any contiguous view with data() and size() (crd::containers::Span / ConstSpan, std::span, DeviceSpan) whose
// elements are `Bytes` bytes: 8 for complex FP32, 4 for real FP32.
template <typename S, std::size_t Bytes>
concept SpanOf = requires(const S& s) {
    { s.data() };
    { s.size() } -> std::convertible_to<std::size_t>;
} && sizeof(*std::declval<const S&>().data()) == Bytes;

namespace detail
{
inline bool on_device(const void* p)
{
    std::int32_t dev = 0;
    check(crdfft_pointer_is_device(p, &dev), "crdfft_pointer_is_device");
    return dev != 0;
}

// The memory of a caller's span as the kernels see it: the span itself when it is device memory; else a device copy
// (filled from the host when `in`), which finish() synchronizes the stream for and writes back (when `out`). The
// host path is synchronous, as the reference's CPU call is; the device path stays asynchronous on the stream.
class Staged
{
public:
    Staged(const void* p, std::size_t bytes, bool in, bool out, void* stream)
        : m_user(const_cast<void*>(p)), m_bytes(bytes), m_out(out), m_stream(stream)
    {
        if (p == nullptr || bytes == 0 || on_device(p))
        {
            m_dev = m_user;
            return;
        }
        check(crdfft_malloc(&m_dev, bytes), "crdfft_malloc(staging)");
        m_owned = true;
        if (in)
        {
            const crdfft_status st = crdfft_memcpy_h2d(m_dev, p, bytes);
            if (st != CRDFFT_OK)
            {
                (void)crdfft_free(m_dev);
                throw FftError(st, "crdfft_memcpy_h2d(staging)");
            }
        }
    }
    ~Staged()
    {
        if (m_owned)
        {
            (void)crdfft_free(m_dev);
        }
    }
    Staged(const Staged&) = delete;
    Staged& operator=(const Staged&) = delete;
    [[nodiscard]] void* device() const noexcept { return m_dev; }
    [[nodiscard]] bool staged() const noexcept { return m_owned; }
    void finish()
    {
        if (!m_owned)
        {
            return;
        }
        check(crdfft_stream_synchronize(m_stream), "crdfft_stream_synchronize");
        if (m_out)
        {
            check(crdfft_memcpy_d2h(m_user, m_dev, m_bytes), "crdfft_memcpy_d2h(staging)");
        }
    }

private:
    void* m_user;
    void* m_dev = nullptr;
    std::size_t m_bytes;
    bool m_out;
    void* m_stream;
    bool m_owned = false;
};

// How a reference FftAlgorithm hint maps onto a device plan of n (the hint is a CPU engine; every device plan kind
// computes the same transform):
//   Stockham, StandaloneHier, Ip4, DeepSplit3, DeepSplit4, InPlace3, TiledFourStep -> the planner's plan (one
//     LDS-resident kernel up to n = 16384, the two-pass four-step above);
//   FourStep -> the four-step (CRDFFT_ALGO_FOUR_STEP) where n > 16384, else the planner's;
//   SixStep  -> the four-step with the square split n1 = 2^floor(log2(n)/2) (the reference's six-step changes only
//     the split, fft.hpp:233-236) where n > 16384 and that split has kernels, else the planner's.
inline crdfft_tuning algorithm_tuning(std::uint64_t n, FftAlgorithm algorithm)
{
    crdfft_tuning t{};
    if (n <= 16384 || (algorithm != FftAlgorithm::FourStep && algorithm != FftAlgorithm::SixStep))
    {
        return t;
    }
    t.algorithm = CRDFFT_ALGO_FOUR_STEP;
    if (algorithm == FftAlgorithm::SixStep)
    {
        unsigned lg = 0;
        while ((std::uint64_t{1} << lg) < n)
        {
            ++lg;
        }
        crdfft_tuning sq = t;
        sq.n1 = std::uint64_t{1} << (lg / 2);
        crdfft_plan_info info{};
        if (crdfft_plan_describe_tuned(n, 1, CRDFFT_C2C, CRDFFT_FP32, &sq, &info) == CRDFFT_OK)
        {
            return sq;
        }
    }
    return t;
}
} // namespace detail

template <typename T> class FftPlan
{
    static_assert(std::is_same_v<T, float>, "the MI355X FFT path is FP32");

public:
    // FftPlan(IAllocator*, n) (fft.hpp:224-227): one transform of n, the planner's plan; any allocator type.
    template <typename Alloc> FftPlan(Alloc* /*alloc*/, std::size_t n) : FftPlan(n, 1, nullptr) {}
    // FftPlan(IAllocator*, n, FftAlgorithm) (fft.hpp:232): the hint mapped by detail::algorithm_tuning.
    template <typename Alloc>
    FftPlan(Alloc* /*alloc*/, std::size_t n, FftAlgorithm forward) : m_n(n), m_batch(1), m_stream(nullptr)
    {
        const crdfft_tuning t = detail::algorithm_tuning(n, forward);
        check(crdfft_plan_create_tuned(&m_plan, n, 1, 0, 0, CRDFFT_C2C, CRDFFT_FP32, CRDFFT_FLAG_NONE, &t),
              "crdfft_plan_create_tuned");
    }

    // A plan of `batch` contiguous transforms of n (fft.hpp:224-227); `stream` is a hipStream_t or null.
    explicit FftPlan(std::size_t n, std::size_t batch = 1, void* stream = nullptr) : m_n(n), m_batch(batch), m_stream(stream)
    {
        check(crdfft_plan_create(&m_plan, n, batch, 0, 0, CRDFFT_C2C, CRDFFT_FP32, CRDFFT_FLAG_NONE),
              "crdfft_plan_create");
    }
    ~FftPlan()
    {
        if (m_plan != nullptr)
        {
            (void)crdfft_plan_destroy(m_plan);
        }
    }
    // pinned like the reference's plan (fft.hpp:348-351)
    FftPlan(const FftPlan&) = delete;
    FftPlan& operator=(const FftPlan&) = delete;
    FftPlan(FftPlan&&) = delete;
    FftPlan& operator=(FftPlan&&) = delete;

    [[nodiscard]] std::size_t size() const noexcept { return m_n; }
    [[nodiscard]] std::size_t batch() const noexcept { return m_batch; }

    // In place over the plan's batch (fft.hpp:409-413): data.size() == n * batch.
    void execute(DeviceSpan<Complex<T>> data, FftDirection dir) const
    {
        require(data.size() == m_n * m_batch, "FftPlan::execute: size mismatch");
        check(crdfft_execute(m_plan, data.data(), data.data(), static_cast<int32_t>(dir), m_stream), "crdfft_execute");
    }
    // Out of place.
    void execute(DeviceSpan<const Complex<T>> in, DeviceSpan<Complex<T>> out, FftDirection dir) const
    {
        require(in.size() == m_n * m_batch && out.size() == m_n * m_batch, "FftPlan::execute: size mismatch");
        check(crdfft_execute(m_plan, in.data(), out.data(), static_cast<int32_t>(dir), m_stream), "crdfft_execute");
    }
    // Element-major batch in place (fft.hpp:462-470): data.size() == n * batch.
    void execute_batched(DeviceSpan<Complex<T>> data, std::size_t batch, FftDirection dir) const
    {
        require(data.size() == m_n * batch, "FftPlan::execute_batched: size mismatch");
        check(crdfft_execute_batched(m_plan, data.data(), batch, static_cast<int32_t>(dir), m_stream),
              "crdfft_execute_batched");
    }
    // The radix-2 reference transform (fft.hpp:415-460), out of place.
    void execute_reference(DeviceSpan<const Complex<T>> in, DeviceSpan<Complex<T>> out, FftDirection dir) const
    {
        require(in.size() == m_n * m_batch && out.size() == m_n * m_batch, "FftPlan::execute_reference: size mismatch");
        check(crdfft_execute_reference(m_plan, in.data(), out.data(), static_cast<int32_t>(dir), m_stream),
              "crdfft_execute_reference");
    }
    // The reference's call shapes on any span (fft.hpp:409-413, :462-470, :417-460): in place, host or device memory.
    template <SpanOf<8> S> void execute(S&& data, FftDirection dir) const
    {
        require(data.size() == m_n * m_batch, "FftPlan::execute: size mismatch");
        detail::Staged d(data.data(), data.size() * 8, true, true, m_stream);
        check(crdfft_execute(m_plan, d.device(), d.device(), static_cast<int32_t>(dir), m_stream), "crdfft_execute");
        d.finish();
    }
    template <SpanOf<8> S> void execute_batched(S&& data, std::size_t batch, FftDirection dir) const
    {
        require(data.size() == m_n * batch, "FftPlan::execute_batched: size mismatch");
        detail::Staged d(data.data(), data.size() * 8, true, true, m_stream);
        check(crdfft_execute_batched(m_plan, d.device(), batch, static_cast<int32_t>(dir), m_stream),
              "crdfft_execute_batched");
        d.finish();
    }
    template <SpanOf<8> S> void execute_reference(S&& data, FftDirection dir) const
    {
        require(data.size() == m_n * m_batch, "FftPlan::execute_reference: size mismatch");
        detail::Staged d(data.data(), data.size() * 8, true, true, m_stream);
        check(crdfft_execute_reference(m_plan, d.device(), d.device(), static_cast<int32_t>(dir), m_stream),
              "crdfft_execute_reference");
        d.finish();
    }
    [[nodiscard]] crdfft_plan_info state() const
    {
        crdfft_plan_info info{};
        check(crdfft_plan_get_info(m_plan, &info), "crdfft_plan_get_info");
        return info;
    }
    [[nodiscard]] crdfft_plan handle() const noexcept { return m_plan; }

private:
    static void require(bool ok, const char* what)
    {
        if (!ok)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, what);
        }
    }
    std::size_t m_n;
    std::size_t m_batch;
    void* m_stream;
    crdfft_plan m_plan = nullptr;
};

// real[n] <-> complex[n/2+1] over `batch` transforms (real_fft.hpp:18-101).
template <typename T> class RealFftPlan
{
    static_assert(std::is_same_v<T, float>, "the MI355X FFT path is FP32");

public:
    // RealFftPlan(IAllocator*, n) (real_fft.hpp:21-36): one transform.
    template <typename Alloc> RealFftPlan(Alloc* /*alloc*/, std::size_t n) : RealFftPlan(n, 1, nullptr) {}
    explicit RealFftPlan(std::size_t n, std::size_t batch = 1, void* stream = nullptr) : m_n(n), m_batch(batch), m_stream(stream)
    {
        check(crdfft_plan_create(&m_r2c, n, batch, 0, 0, CRDFFT_R2C, CRDFFT_FP32, CRDFFT_FLAG_NONE),
              "crdfft_plan_create(R2C)");
        const crdfft_status st = crdfft_plan_create(&m_c2r, n, batch, 0, 0, CRDFFT_C2R, CRDFFT_FP32, CRDFFT_FLAG_NONE);
        if (st != CRDFFT_OK)
        {
            (void)crdfft_plan_destroy(m_r2c);
            throw FftError(st, "crdfft_plan_create(C2R)");
        }
    }
    ~RealFftPlan()
    {
        (void)crdfft_plan_destroy(m_r2c);
        (void)crdfft_plan_destroy(m_c2r);
    }
    RealFftPlan(const RealFftPlan&) = delete;
    RealFftPlan& operator=(const RealFftPlan&) = delete;

    [[nodiscard]] std::size_t size() const noexcept { return m_n; }

    // real_fft.hpp:41-64: out[k], k <= n/2, per transform
    void rfft(DeviceSpan<const T> real_in, DeviceSpan<Complex<T>> out) const
    {
        if (real_in.size() != m_n * m_batch || out.size() != (m_n / 2 + 1) * m_batch)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "RealFftPlan::rfft: size mismatch");
        }
        check(crdfft_execute(m_r2c, real_in.data(), out.data(), CRDFFT_FORWARD, m_stream), "crdfft_execute(R2C)");
    }
    // real_fft.hpp:67-93: normalized, irfft(rfft(x)) == x
    void irfft(DeviceSpan<const Complex<T>> in, DeviceSpan<T> real_out) const
    {
        if (in.size() != (m_n / 2 + 1) * m_batch || real_out.size() != m_n * m_batch)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "RealFftPlan::irfft: size mismatch");
        }
        check(crdfft_execute(m_c2r, in.data(), real_out.data(), CRDFFT_INVERSE, m_stream), "crdfft_execute(C2R)");
    }
    // The reference's call shapes on any span (real_fft.hpp:41, :67), host or device memory.
    template <SpanOf<4> In, SpanOf<8> Out> void rfft(In&& real_in, Out&& out) const
    {
        if (real_in.size() != m_n * m_batch || out.size() != (m_n / 2 + 1) * m_batch)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "RealFftPlan::rfft: size mismatch");
        }
        detail::Staged x(real_in.data(), real_in.size() * 4, true, false, m_stream);
        detail::Staged y(out.data(), out.size() * 8, false, true, m_stream);
        check(crdfft_execute(m_r2c, x.device(), y.device(), CRDFFT_FORWARD, m_stream), "crdfft_execute(R2C)");
        y.finish();
        x.finish();
    }
    template <SpanOf<8> In, SpanOf<4> Out> void irfft(In&& in, Out&& real_out) const
    {
        if (in.size() != (m_n / 2 + 1) * m_batch || real_out.size() != m_n * m_batch)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "RealFftPlan::irfft: size mismatch");
        }
        detail::Staged x(in.data(), in.size() * 8, true, false, m_stream);
        detail::Staged y(real_out.data(), real_out.size() * 4, false, true, m_stream);
        check(crdfft_execute(m_c2r, x.device(), y.device(), CRDFFT_INVERSE, m_stream), "crdfft_execute(C2R)");
        y.finish();
        x.finish();
    }

private:
    std::size_t m_n;
    std::size_t m_batch;
    void* m_stream;
    crdfft_plan m_r2c = nullptr;
    crdfft_plan m_c2r = nullptr;
};

namespace detail
{
// An owned crdfft handle of one of the plan kinds (the reference's plans are pinned, so are these).
template <typename H, crdfft_status (*Destroy)(H)> class Handle
{
public:
    Handle() = default;
    ~Handle()
    {
        if (m_h != nullptr)
        {
            (void)Destroy(m_h);
        }
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    H* out() noexcept { return &m_h; }
    [[nodiscard]] H get() const noexcept { return m_h; }

private:
    H m_h = nullptr;
};
} // namespace detail

// Any-length C2C (bluestein.hpp:29-136): forward unnormalized, inverse scaled by 1/n; `batch` contiguous transforms.
template <typename T> class BluesteinPlan
{
    static_assert(std::is_same_v<T, float>, "the MI355X FFT path is FP32");

public:
    // BluesteinPlan(IAllocator*, n) (bluestein.hpp:32): one transform.
    template <typename Alloc> BluesteinPlan(Alloc* /*alloc*/, std::size_t n) : BluesteinPlan(n, 1, nullptr) {}
    explicit BluesteinPlan(std::size_t n, std::size_t batch = 1, void* stream = nullptr)
        : m_n(n), m_batch(batch), m_stream(stream)
    {
        check(crdfft_bluestein_plan_create(m_plan.out(), n, batch, 0, 0, CRDFFT_FLAG_NONE),
              "crdfft_bluestein_plan_create");
    }
    [[nodiscard]] std::size_t size() const noexcept { return m_n; }
    // bluestein.hpp:87, in place on any span, host or device memory
    template <SpanOf<8> S> void execute(S&& data, FftDirection dir) const
    {
        if (data.size() != m_n * m_batch)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "BluesteinPlan::execute: size mismatch");
        }
        detail::Staged d(data.data(), data.size() * 8, true, true, m_stream);
        check(crdfft_execute(m_plan.get(), d.device(), d.device(), static_cast<int32_t>(dir), m_stream),
              "crdfft_execute(Bluestein)");
        d.finish();
    }

private:
    std::size_t m_n;
    std::size_t m_batch;
    void* m_stream;
    detail::Handle<crdfft_plan, crdfft_plan_destroy> m_plan;
};

// Row-major N-D C2C of `batch` volumes in place (nd_fft.hpp:28-196): inverse scaled by 1/prod(dims).
template <typename T> class NdFftPlan
{
    static_assert(std::is_same_v<T, float>, "the MI355X FFT path is FP32");

public:
    NdFftPlan(const std::size_t* dims, std::uint32_t ndim, std::size_t batch = 1, void* stream = nullptr)
        : m_total(batch), m_stream(stream)
    {
        std::uint64_t d[8] = {};
        if (ndim == 0 || ndim > 8)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "NdFftPlan: 1..8 dimensions");
        }
        for (std::uint32_t i = 0; i < ndim; ++i)
        {
            d[i] = dims[i];
            m_total *= dims[i];
        }
        check(crdfft_nd_plan_create(m_plan.out(), d, ndim, batch, 0), "crdfft_nd_plan_create");
    }
    // NdFftPlan(IAllocator*, ConstSpan<usize> dims) (nd_fft.hpp:31): one volume; the reference's job_plans count
    // (host threads) has no device meaning and is accepted and ignored.
    template <typename Alloc, typename Dims>
        requires requires(const Dims& d) {
            { d.data() };
            { d.size() } -> std::convertible_to<std::size_t>;
        }
    NdFftPlan(Alloc* /*alloc*/, const Dims& dims, std::uint32_t /*job_plans*/ = 1)
        : NdFftPlan(reinterpret_cast<const std::size_t*>(dims.data()), static_cast<std::uint32_t>(dims.size()), 1,
                    nullptr)
    {
        static_assert(sizeof(*dims.data()) == sizeof(std::size_t), "dims are usize");
    }
    // nd_fft.hpp:105, in place on any span, host or device memory
    template <SpanOf<8> S> void execute(S&& data, FftDirection dir) const
    {
        if (data.size() != m_total)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "NdFftPlan::execute: size mismatch");
        }
        detail::Staged d(data.data(), data.size() * 8, true, true, m_stream);
        check(crdfft_nd_execute(m_plan.get(), d.device(), static_cast<int32_t>(dir), m_stream), "crdfft_nd_execute");
        d.finish();
    }

private:
    std::size_t m_total;
    void* m_stream;
    detail::Handle<crdfft_nd_plan, crdfft_nd_plan_destroy> m_plan;
};

// DCT-II/III and DST-II/III of `batch` rows of n reals (dct.hpp:34-170, scipy norm=None conventions).
template <typename T> class DctPlan
{
    static_assert(std::is_same_v<T, float>, "the MI355X FFT path is FP32");

public:
    // DctPlan(IAllocator*, n) (dct.hpp:38): one row.
    template <typename Alloc> DctPlan(Alloc* /*alloc*/, std::size_t n) : DctPlan(n, 1, nullptr) {}
    explicit DctPlan(std::size_t n, std::size_t batch = 1, void* stream = nullptr)
        : m_n(n), m_batch(batch), m_stream(stream)
    {
        check(crdfft_dct_plan_create(m_plan.out(), n, batch, 0, 0, CRDFFT_FLAG_NONE), "crdfft_dct_plan_create");
    }
    [[nodiscard]] std::size_t size() const noexcept { return m_n; }
    // dct.hpp:62, :99, :120, :164 on any spans (ConstSpan in, Span out), host or device memory
    template <SpanOf<4> X, SpanOf<4> Y> void dct2(X&& x, Y&& y) const { run(CRDFFT_DCT2, x, y); }
    template <SpanOf<4> X, SpanOf<4> Y> void dct3(X&& x, Y&& y) const { run(CRDFFT_DCT3, x, y); }
    template <SpanOf<4> X, SpanOf<4> Y> void dst2(X&& x, Y&& y) const { run(CRDFFT_DST2, x, y); }
    template <SpanOf<4> X, SpanOf<4> Y> void dst3(X&& x, Y&& y) const { run(CRDFFT_DST3, x, y); }

private:
    template <typename X, typename Y> void run(crdfft_r2r_kind kind, const X& x, const Y& y) const
    {
        if (x.size() != m_n * m_batch || y.size() != m_n * m_batch)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, "DctPlan: size mismatch");
        }
        detail::Staged xs(x.data(), x.size() * 4, true, false, m_stream);
        detail::Staged ys(y.data(), y.size() * 4, false, true, m_stream);
        check(crdfft_dct_execute(m_plan.get(), kind, static_cast<const float*>(xs.device()),
                                 static_cast<float*>(ys.device()), m_stream),
              "crdfft_dct_execute");
        ys.finish();
        xs.finish();
    }
    std::size_t m_n;
    std::size_t m_batch;
    void* m_stream;
    detail::Handle<crdfft_dct_plan, crdfft_dct_plan_destroy> m_plan;
};

// hesap-dsp's STFT / spectrogram / Welch over frames of nperseg samples at hop nperseg - noverlap
// (spectral.hpp:43-222, :334-386), each one fused launch.
class SpectralPlan
{
public:
    SpectralPlan(std::size_t nperseg, std::size_t noverlap, crdfft_window window, bool detrend,
                 void* stream = nullptr)
        : m_nfreq(nperseg / 2 + 1), m_stream(stream)
    {
        check(crdfft_spectral_plan_create(m_plan.out(), nperseg, noverlap, window,
                                          detrend ? CRDFFT_SPECTRAL_DETREND : 0u),
              "crdfft_spectral_plan_create");
    }
    [[nodiscard]] std::size_t frames(std::size_t len) const { return crdfft_spectral_frames(m_plan.get(), len); }
    void stft(DeviceSpan<const float> x, DeviceSpan<Complex<float>> out) const
    {
        need(out.size() == frames(x.size()) * m_nfreq, "SpectralPlan::stft: size mismatch");
        check(crdfft_stft(m_plan.get(), x.data(), x.size(), out.data(), m_stream), "crdfft_stft");
    }
    void spectrogram(DeviceSpan<const float> x, double fs, DeviceSpan<float> out) const
    {
        need(out.size() == frames(x.size()) * m_nfreq, "SpectralPlan::spectrogram: size mismatch");
        check(crdfft_spectrogram(m_plan.get(), x.data(), x.size(), fs, out.data(), m_stream), "crdfft_spectrogram");
    }
    void welch(DeviceSpan<const float> x, double fs, DeviceSpan<float> psd) const
    {
        need(psd.size() == m_nfreq, "SpectralPlan::welch: size mismatch");
        check(crdfft_welch(m_plan.get(), x.data(), x.size(), fs, psd.data(), m_stream), "crdfft_welch");
    }

private:
    static void need(bool ok, const char* what)
    {
        if (!ok)
        {
            throw FftError(CRDFFT_INVALID_ARGUMENT, what);
        }
    }
    std::size_t m_nfreq;
    void* m_stream;
    detail::Handle<crdfft_spectral_plan, crdfft_spectral_plan_destroy> m_plan;
};

// One-shot transient plan, in place, unnormalized (fft.hpp:1005-1012). Synchronous.
template <typename T> void fft(DeviceSpan<Complex<T>> data, FftDirection dir = FftDirection::Forward)
{
    const FftPlan<T> plan(data.size());
    plan.execute(data, dir);
    check(crdfft_device_synchronize(), "crdfft_device_synchronize");
}

// Normalized inverse, ifft_normalized(fft(x)) == x (fft.hpp:1015-1029): the 1/n rides on the last store.
template <typename T> void ifft_normalized(DeviceSpan<Complex<T>> data)
{
    static_assert(std::is_same_v<T, float>, "the MI355X FFT path is FP32");
    crdfft_plan plan = nullptr;
    check(crdfft_plan_create(&plan, data.size(), 1, 0, 0, CRDFFT_C2C, CRDFFT_FP32, CRDFFT_FLAG_NORMALIZE_INVERSE),
          "crdfft_plan_create");
    const crdfft_status st = crdfft_execute(plan, data.data(), data.data(), CRDFFT_INVERSE, nullptr);
    const crdfft_status sy = crdfft_device_synchronize();
    (void)crdfft_plan_destroy(plan);
    check(st, "crdfft_execute");
    check(sy, "crdfft_device_synchronize");
}

// The reference's one-shot calls (fft.hpp:1005-1029), with its allocator argument and any span, host or device
// memory; synchronous like the reference.
template <typename T, typename Alloc, SpanOf<8> S>
void fft(Alloc* alloc, S&& data, FftDirection dir = FftDirection::Forward)
{
    const FftPlan<T> plan(alloc, data.size());
    plan.execute(data, dir);
    check(crdfft_device_synchronize(), "crdfft_device_synchronize");
}
template <typename T, typename Alloc, SpanOf<8> S> void ifft_normalized(Alloc* /*alloc*/, S&& data)
{
    static_assert(std::is_same_v<T, float>, "the MI355X FFT path is FP32");
    detail::Staged d(data.data(), data.size() * 8, true, true, nullptr);
    ifft_normalized<T>(DeviceSpan<Complex<T>>(static_cast<Complex<T>*>(d.device()), data.size()));
    d.finish();
}

} // namespace crd_amd::hesap::fft
