// Host program over the C++20 interface (include/crd_amd/fft.hpp): round trips through FftPlan, RealFftPlan,
// ifft_normalized, BluesteinPlan, NdFftPlan and DctPlan, and the Welch PSD of a tone; exit status 0 when every round
// trip comes back within the FP32 bound and the PSD peaks at the tone's bin. Built and run by
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
    // any length (inverse normalized), N-D (inverse normalized), DCT-III(DCT-II(x)) = 2n x, Welch of a tone
    constexpr std::size_t bn = 1000, nd = 64 * 64;
    DeviceArray<Complex<float>> bdev(bn * batch), ndev(nd);
    bdev.upload(host.data());
    ndev.upload(host.data());
    BluesteinPlan<float> bplan(bn, batch);
    bplan.execute(bdev.span(), FftDirection::Forward);
    bplan.execute(bdev.span(), FftDirection::Inverse);
    const std::size_t dims[2] = {64, 64};
    NdFftPlan<float> ndplan(dims, 2);
    ndplan.execute(ndev.span(), FftDirection::Forward);
    ndplan.execute(ndev.span(), FftDirection::Inverse);
    DctPlan<float> dplan(n, batch);
    DeviceArray<float> dmid(n * batch), dback(n * batch);
    dplan.dct2(DeviceSpan<const float>(rdev.data(), rdev.size()), dmid.span());
    dplan.dct3(DeviceSpan<const float>(dmid.data(), dmid.size()), dback.span());
    constexpr std::size_t len = 1 << 16, nperseg = 1024;
    std::vector<float> tone(len);
    for (std::size_t i = 0; i < len; ++i)
    {
        tone[i] = std::sin(2.0f * 3.14159265f * 100.0f * static_cast<float>(i) / static_cast<float>(nperseg));
    }
    DeviceArray<float> tdev(len), psd(nperseg / 2 + 1);
    tdev.upload(tone.data());
    SpectralPlan splan(nperseg, nperseg / 2, CRDFFT_WINDOW_HANN, true);
    splan.welch(DeviceSpan<const float>(tdev.data(), len), 1.0, psd.span());
    check(crdfft_device_synchronize(), "sync");
    std::vector<Complex<float>> bback(bn * batch), ndback(nd);
    bdev.download(bback.data());
    ndev.download(ndback.data());
    std::vector<float> dct_back(n * batch), p(nperseg / 2 + 1);
    dback.download(dct_back.data());
    psd.download(p.data());
    double berr = 0.0, nderr = 0.0, derr = 0.0;
    for (std::size_t i = 0; i < bn * batch; ++i)
    {
        berr = std::fmax(berr, std::fabs(bback[i].re - host[i].re) + std::fabs(bback[i].im - host[i].im));
    }
    for (std::size_t i = 0; i < nd; ++i)
    {
        nderr = std::fmax(nderr, std::fabs(ndback[i].re - host[i].re) + std::fabs(ndback[i].im - host[i].im));
    }
    for (std::size_t i = 0; i < n * batch; ++i)
    {
        derr = std::fmax(derr, std::fabs(dct_back[i] / (2.0 * n) - real[i]));
    }
    std::size_t peak = 0;
    for (std::size_t k = 1; k < p.size(); ++k)
    {
        peak = p[k] > p[peak] ? k : peak;
    }
    std::printf("c2c roundtrip %.3e  r2c roundtrip %.3e  ifft_normalized %.3e  invalid-size throws %d\n", err, rerr,
                nerr, threw ? 1 : 0);
    std::printf("bluestein roundtrip %.3e  nd roundtrip %.3e  dct3(dct2)/2n %.3e  welch peak bin %zu\n", berr, nderr,
                derr, peak);
    const bool ok = err < 1e-5 && rerr < 1e-5 && nerr < 1e-5 && threw && berr < 1e-5 && nderr < 1e-5 && derr < 1e-5 &&
                    peak == 100;
    return ok ? 0 : 1;
}
