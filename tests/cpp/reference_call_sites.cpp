// Call sites written the way the reference's own tests write them (tests/numerics/hesap-fft/test_fft.cpp:83-84,
// :138-139, :164-165, :235-237, :412-431, :523; bluestein / nd_fft / dct tests), over host memory, compiled against
// include/crd_amd/fft.hpp with only the one-line namespace mapping a maintainer adds at the call site. The `crd`
// namespace below stands in for the reference's own foundation types (allocator, spans, Complex, integer aliases):
// just enough of their interface for these calls, no behaviour.
//
// Exit status 0 when every transform matches an FP64 naive DFT within the reference's f32 gate
// (max|y - ref| / (1 + max|ref|) < 2e-4 below 2^19, test_fft_kernels.cpp:151-163) and the round trips return.
// Run by tests/test_gpu_fft.py (needs a device); compiled by tests/test_abi.py on the CPU.
#include <crd_amd/fft.hpp>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace crd
{
using usize = std::size_t;
using u64 = std::uint64_t;
using f32 = float;
using f64 = double;
namespace memory
{
class IAllocator
{
public:
    virtual ~IAllocator() = default;
};
class TlsfAllocator final : public IAllocator
{
};
} // namespace memory
namespace containers
{
template <typename T> class Span
{
public:
    Span(T* p, usize n) : m_p(p), m_n(n) {}
    [[nodiscard]] T* data() const { return m_p; }
    [[nodiscard]] usize size() const { return m_n; }

private:
    T* m_p;
    usize m_n;
};
template <typename T> using ConstSpan = Span<const T>;
} // namespace containers
namespace hesap
{
template <typename T> struct Complex
{
    T re;
    T im;
};
} // namespace hesap
} // namespace crd

// The integration line: the reference's namespace resolves to the device implementation.
namespace crd::hesap::fft
{
using namespace crd_amd::hesap::fft;
}

using crd::f32;
using crd::usize;
namespace fft = crd::hesap::fft;
namespace cont = crd::containers;
using crd::hesap::Complex;

namespace
{
// the LCG of test_fft.cpp:31-44
void fill_lcg(std::vector<Complex<f32>>& x, usize n, crd::u64 seed)
{
    x.resize(n);
    crd::u64 s = seed;
    auto next = [&s]() -> double
    {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (static_cast<double>(s >> 11) / static_cast<double>(1ULL << 53)) * 2.0 - 1.0;
    };
    for (usize i = 0; i < n; ++i)
    {
        x[i] = Complex<f32>{static_cast<f32>(next()), static_cast<f32>(next())};
    }
}

// FP64 DFT of an FP32 input (direct sum below 4096 points, else the same sum over a radix-2 recursion in FP64)
void dft64(const Complex<f32>* x, usize n, usize stride, bool inverse, std::vector<double>& re, std::vector<double>& im)
{
    re.assign(n, 0.0);
    im.assign(n, 0.0);
    if (n == 1)
    {
        re[0] = x[0].re;
        im[0] = x[0].im;
        return;
    }
    if (n <= 64)
    {
        const double sgn = inverse ? 1.0 : -1.0;
        for (usize k = 0; k < n; ++k)
        {
            for (usize j = 0; j < n; ++j)
            {
                const double ang = sgn * 6.283185307179586476925286766559 * static_cast<double>((j * k) % n) /
                                   static_cast<double>(n);
                const double c = std::cos(ang), s = std::sin(ang);
                re[k] += x[j * stride].re * c - x[j * stride].im * s;
                im[k] += x[j * stride].re * s + x[j * stride].im * c;
            }
        }
        return;
    }
    std::vector<double> er, ei, orr, oi;
    dft64(x, n / 2, stride * 2, inverse, er, ei);
    dft64(x + stride, n / 2, stride * 2, inverse, orr, oi);
    const double sgn = inverse ? 1.0 : -1.0;
    for (usize k = 0; k < n / 2; ++k)
    {
        const double ang = sgn * 6.283185307179586476925286766559 * static_cast<double>(k) / static_cast<double>(n);
        const double c = std::cos(ang), s = std::sin(ang);
        const double tr = orr[k] * c - oi[k] * s, ti = orr[k] * s + oi[k] * c;
        re[k] = er[k] + tr;
        im[k] = ei[k] + ti;
        re[k + n / 2] = er[k] - tr;
        im[k + n / 2] = ei[k] - ti;
    }
}

// the reference's gate metric: max|y - ref| / (1 + max|ref|)
double gate(const std::vector<Complex<f32>>& y, const std::vector<double>& re, const std::vector<double>& im,
            usize off = 0, usize stride = 1)
{
    double err = 0.0, mag = 0.0;
    for (usize k = 0; k < re.size(); ++k)
    {
        const Complex<f32>& v = y[off + k * stride];
        err = std::fmax(err, std::hypot(v.re - re[k], v.im - im[k]));
        mag = std::fmax(mag, std::hypot(re[k], im[k]));
    }
    return err / (1.0 + mag);
}

int failures = 0;
void expect(bool ok, const char* what, double v)
{
    std::printf("%-58s %.3e %s\n", what, v, ok ? "ok" : "FAIL");
    failures += ok ? 0 : 1;
}
} // namespace

int main()
{
    crd::memory::TlsfAllocator alloc;
    std::vector<double> re, im;

    // FftPlan(alloc, n) + execute on a host span, and execute_reference in place (test_fft.cpp:83-84, :235-237)
    for (const usize n : {usize{8}, usize{4096}, usize{1} << 16})
    {
        std::vector<Complex<f32>> a, b;
        fill_lcg(a, n, 0x9E3779B97F4A7C15ULL ^ n);
        b = a;
        dft64(a.data(), n, 1, false, re, im);
        const fft::FftPlan<f32> plan(&alloc, n);
        plan.execute(cont::Span<Complex<f32>>(a.data(), n), fft::FftDirection::Forward);
        plan.execute_reference(cont::Span<Complex<f32>>(b.data(), n), fft::FftDirection::Forward);
        char what[96];
        std::snprintf(what, sizeof what, "FftPlan(alloc, %zu).execute(Span) vs FP64", n);
        expect(gate(a, re, im) < 2e-4, what, gate(a, re, im));
        std::snprintf(what, sizeof what, "FftPlan(alloc, %zu).execute_reference(Span) vs FP64", n);
        expect(gate(b, re, im) < 2e-4, what, gate(b, re, im));
    }

    // the FftAlgorithm overload (fft.hpp:232): each hint computes the same transform
    for (const fft::FftAlgorithm algo : {fft::FftAlgorithm::Stockham, fft::FftAlgorithm::FourStep,
                                         fft::FftAlgorithm::SixStep, fft::FftAlgorithm::TiledFourStep})
    {
        const usize n = usize{1} << 17;
        std::vector<Complex<f32>> a;
        fill_lcg(a, n, 7 + static_cast<unsigned>(algo));
        dft64(a.data(), n, 1, true, re, im);
        const fft::FftPlan<f32> plan(&alloc, n, algo);
        plan.execute(cont::Span<Complex<f32>>(a.data(), n), fft::FftDirection::Inverse);
        char what[96];
        std::snprintf(what, sizeof what, "FftPlan(alloc, 2^17, %s) inverse vs FP64", fft::fft_algorithm_name(algo));
        expect(gate(a, re, im) < 2e-4, what, gate(a, re, im));
    }

    // fft(alloc, span, dir) then ifft_normalized(alloc, span) (test_fft.cpp:164-165)
    {
        const usize n = 2048;
        std::vector<Complex<f32>> x, orig;
        fill_lcg(x, n, 3);
        orig = x;
        fft::fft<f32>(&alloc, cont::Span<Complex<f32>>(x.data(), n), fft::FftDirection::Forward);
        fft::ifft_normalized<f32>(&alloc, cont::Span<Complex<f32>>(x.data(), n));
        double err = 0.0;
        for (usize i = 0; i < n; ++i)
        {
            err = std::fmax(err, std::hypot(x[i].re - orig[i].re, x[i].im - orig[i].im));
        }
        expect(err < 1e-5, "ifft_normalized(alloc, fft(alloc, x)) - x", err);
    }

    // execute_batched, element-major data[i*batch + t] (test_fft.cpp:412-431)
    for (const usize m : {usize{64}, usize{8192}})
    {
        const usize batch = 5;
        std::vector<Complex<f32>> work, single;
        fill_lcg(work, m * batch, 11 + m);
        const std::vector<Complex<f32>> in = work;
        const fft::FftPlan<f32> plan(&alloc, m);
        plan.execute_batched(cont::Span<Complex<f32>>(work.data(), m * batch), batch, fft::FftDirection::Forward);
        double worst = 0.0;
        for (usize t = 0; t < batch; ++t)
        {
            dft64(in.data() + t, m, batch, false, re, im);
            worst = std::fmax(worst, gate(work, re, im, t, batch));
        }
        char what[96];
        std::snprintf(what, sizeof what, "FftPlan(alloc, %zu).execute_batched(Span, 5) vs FP64", m);
        expect(worst < 2e-4, what, worst);
    }

    // RealFftPlan(alloc, n): rfft(ConstSpan, Span), irfft(ConstSpan, Span) (real_fft.hpp:21, :41, :67)
    {
        const usize n = 1024;
        std::vector<f32> x(n), back(n);
        for (usize i = 0; i < n; ++i)
        {
            x[i] = std::sin(0.05f * static_cast<f32>(i)) + 0.25f * std::cos(0.9f * static_cast<f32>(i));
        }
        std::vector<Complex<f32>> spec(n / 2 + 1);
        const fft::RealFftPlan<f32> plan(&alloc, n);
        plan.rfft(cont::ConstSpan<f32>(x.data(), n), cont::Span<Complex<f32>>(spec.data(), n / 2 + 1));
        plan.irfft(cont::ConstSpan<Complex<f32>>(spec.data(), n / 2 + 1), cont::Span<f32>(back.data(), n));
        std::vector<Complex<f32>> xc(n);
        for (usize i = 0; i < n; ++i)
        {
            xc[i] = {x[i], 0.0f};
        }
        dft64(xc.data(), n, 1, false, re, im);
        re.resize(n / 2 + 1);
        im.resize(n / 2 + 1);
        expect(gate(spec, re, im) < 2e-4, "RealFftPlan(alloc, 1024).rfft(ConstSpan, Span) vs FP64",
               gate(spec, re, im));
        double err = 0.0;
        for (usize i = 0; i < n; ++i)
        {
            err = std::fmax(err, std::fabs(back[i] - x[i]));
        }
        expect(err < 1e-5, "irfft(rfft(x)) - x", err);
    }

    // BluesteinPlan(alloc, n) (bluestein.hpp:32, :87): forward unnormalized, inverse normalized
    {
        const usize n = 1000;
        std::vector<Complex<f32>> x, orig;
        fill_lcg(x, n, 5);
        orig = x;
        const fft::BluesteinPlan<f32> plan(&alloc, n);
        plan.execute(cont::Span<Complex<f32>>(x.data(), n), fft::FftDirection::Forward);
        std::vector<double> dre(n, 0.0), dim(n, 0.0);
        for (usize k = 0; k < n; ++k)
        {
            for (usize j = 0; j < n; ++j)
            {
                const double ang = -6.283185307179586476925286766559 * static_cast<double>((j * k) % n) / n;
                dre[k] += orig[j].re * std::cos(ang) - orig[j].im * std::sin(ang);
                dim[k] += orig[j].re * std::sin(ang) + orig[j].im * std::cos(ang);
            }
        }
        expect(gate(x, dre, dim) < 2e-4, "BluesteinPlan(alloc, 1000).execute(Span) vs FP64 DFT", gate(x, dre, dim));
        plan.execute(cont::Span<Complex<f32>>(x.data(), n), fft::FftDirection::Inverse);
        double err = 0.0;
        for (usize i = 0; i < n; ++i)
        {
            err = std::fmax(err, std::hypot(x[i].re - orig[i].re, x[i].im - orig[i].im));
        }
        expect(err < 1e-5, "Bluestein inverse(forward(x)) - x", err);
    }

    // NdFftPlan(alloc, ConstSpan<usize> dims) (nd_fft.hpp:31, :105): inverse normalized
    {
        const usize dims[2] = {32, 48};
        const usize total = dims[0] * dims[1];
        std::vector<Complex<f32>> x, orig;
        fill_lcg(x, total, 9);
        orig = x;
        const fft::NdFftPlan<f32> plan(&alloc, cont::ConstSpan<usize>(dims, 2));
        plan.execute(cont::Span<Complex<f32>>(x.data(), total), fft::FftDirection::Forward);
        // X[0][0] is the sum of every value
        double sr = 0.0, si = 0.0;
        for (const auto& v : orig)
        {
            sr += v.re;
            si += v.im;
        }
        const double dc = std::hypot(x[0].re - sr, x[0].im - si) / (1.0 + std::hypot(sr, si));
        expect(dc < 1e-5, "NdFftPlan(alloc, {32,48}) X[0,0] = sum x", dc);
        plan.execute(cont::Span<Complex<f32>>(x.data(), total), fft::FftDirection::Inverse);
        double err = 0.0;
        for (usize i = 0; i < total; ++i)
        {
            err = std::fmax(err, std::hypot(x[i].re - orig[i].re, x[i].im - orig[i].im));
        }
        expect(err < 1e-5, "NdFftPlan inverse(forward(x)) - x", err);
    }

    // DctPlan(alloc, n): dct3(dct2(x)) = 2n x (dct.hpp:38, :62, :99)
    {
        const usize n = 512;
        std::vector<f32> x(n), mid(n), back(n);
        for (usize i = 0; i < n; ++i)
        {
            x[i] = std::cos(0.3f * static_cast<f32>(i * i % 97));
        }
        const fft::DctPlan<f32> plan(&alloc, n);
        plan.dct2(cont::ConstSpan<f32>(x.data(), n), cont::Span<f32>(mid.data(), n));
        plan.dct3(cont::ConstSpan<f32>(mid.data(), n), cont::Span<f32>(back.data(), n));
        double err = 0.0;
        for (usize i = 0; i < n; ++i)
        {
            err = std::fmax(err, std::fabs(back[i] / (2.0 * n) - x[i]));
        }
        expect(err < 1e-5, "DctPlan(alloc, 512): dct3(dct2(x)) / 2n - x", err);
    }

    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
