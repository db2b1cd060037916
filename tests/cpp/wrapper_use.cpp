// Host program over the C++20 interface (include/crd_amd/fft.hpp): forward + normalized inverse round trip of a
// batch and an R2C/C2R round trip, exit status 0 when both come back within the FP32 bound. Built and run by
// tests/test_gpu_fft.py (needs a device); compiled for syntax by tests/test_abi.py on the CPU.
#include <crd_amd/fft.hpp>

#include <cmath>
#include <cstdio>
#include <vector>

using namespace crd_amd::hesap::fft;

int main()
{
    constexpr std::size_t n = 4096, batch = 3;
    std::vector<Complex<float>> host(n * batch), back(n * batch);
    for (std::size_t i = 0; i < host.size(); ++i)
    {
        host[i] = {std::sin(0.37f * static_cast<float>(i)), std::cos(0.11f * static_cast<float>(i * i % 977))};
    }
    DeviceArray<Complex<float>> dev(n * batch);
    dev.upload(host.data());
    FftPlan<float> plan(n, batch);
    plan.execute(dev.span(), FftDirection::Forward);
    plan.execute(dev.span(), FftDirection::Inverse);
    check(crdfft_device_synchronize(), "sync");
    dev.download(back.data());
    double err = 0.0;
    for (std::size_t i = 0; i < host.size(); ++i)
    {
        err = std::fmax(err, std::fabs(back[i].re / n - host[i].re) + std::fabs(back[i].im / n - host[i].im));
    }
    std::vector<float> real(n * batch), real_back(n * batch);
    for (std::size_t i = 0; i < real.size(); ++i)
    {
        real[i] = std::sin(0.01f * static_cast<float>(i));
    }
    DeviceArray<float> rdev(n * batch), rback(n * batch);
    DeviceArray<Complex<float>> spec((n / 2 + 1) * batch);
    rdev.upload(real.data());
    RealFftPlan<float> rplan(n, batch);
    rplan.rfft(DeviceSpan<const float>(rdev.data(), rdev.size()), spec.span());
    rplan.irfft(DeviceSpan<const Complex<float>>(spec.data(), spec.size()), rback.span());
    check(crdfft_device_synchronize(), "sync");
    rback.download(real_back.data());
    double rerr = 0.0;
    for (std::size_t i = 0; i < real.size(); ++i)
    {
        rerr = std::fmax(rerr, std::fabs(real_back[i] - real[i]));
    }
    DeviceArray<Complex<float>> one(n);
    one.upload(host.data());
    fft<float>(one.span(), FftDirection::Forward);
    ifft_normalized<float>(one.span());
    std::vector<Complex<float>> once(n);
    one.download(once.data());
    double nerr = 0.0;
    for (std::size_t i = 0; i < n; ++i)
    {
        nerr = std::fmax(nerr, std::fabs(once[i].re - host[i].re) + std::fabs(once[i].im - host[i].im));
    }
    bool threw = false;
    try
    {
        FftPlan<float> bad(3);
    }
    catch (const FftError& e)
    {
        threw = e.status() == CRDFFT_INVALID_SIZE;
    }
    std::printf("c2c roundtrip %.3e  r2c roundtrip %.3e  ifft_normalized %.3e  invalid-size throws %d\n", err, rerr,
                nerr, threw ? 1 : 0);
    return (err < 1e-5 && rerr < 1e-5 && nerr < 1e-5 && threw) ? 0 : 1;
}
