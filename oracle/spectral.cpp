// Test infrastructure only: the FP64 oracle of the spectral front end over the real FFT (the reference's
// engine/numerics/hesap-dsp/include/crd/hesap/dsp/spectral.hpp and windows.hpp), pinned by the scipy vectors of the
// reference's own fixture (tests/numerics/hesap-dsp/spectral_refs.inc, via tests/golden/spectral_golden.npz) and by
// numpy's rfft of explicitly windowed frames.
#include <cmath>
#include <vector>

#include "oracle.h"

namespace
{

constexpr double kPi = 3.14159265358979323846;

// general_cosine hann (windows.hpp:63-124): w[i] = 0.5 + 0.5 cos(fac_i), fac = linspace(-pi, pi, N), N = m (sym) or
// m + 1 (periodic, the last sample dropped).
void hann(size_t m, bool sym, std::vector<double>& w)
{
    w.assign(m, 1.0);
    const size_t n = sym ? m : m + 1;
    if (n <= 1)
    {
        return;
    }
    for (size_t i = 0; i < m; ++i)
    {
        const double fac = -kPi + 2.0 * kPi * static_cast<double>(i) / static_cast<double>(n - 1);
        w[i] = 0.5 + 0.5 * std::cos(fac);
    }
}

size_t frame_count(size_t len, size_t nperseg, size_t noverlap)
{
    const size_t hop = nperseg - noverlap;
    return len < nperseg ? 0 : (len - noverlap) / hop;
}

} // namespace

extern "C" {

void oracle_hann64(double* out, size_t m, int sym)
{
    std::vector<double> w;
    hann(m, sym != 0, w);
    for (size_t i = 0; i < m; ++i)
    {
        out[i] = w[i];
    }
}

// segment_periodograms_arm (spectral.hpp:43-126): frame s = x[s*hop, s*hop + nperseg), minus its mean when
// detrend != 0, times the window (hann periodic, or boxcar when window == 0), rfft, scale/(fs*sum w^2) |X|^2 doubled
// except DC and Nyquist. out: nseg x (nperseg/2 + 1) doubles; returns nseg.
size_t oracle_periodograms64(const double* x, size_t len, double fs, size_t nperseg, size_t noverlap, int window,
                             int detrend, double* out)
{
    const size_t hop = nperseg - noverlap;
    const size_t nfreq = nperseg / 2 + 1;
    const size_t nseg = frame_count(len, nperseg, noverlap);
    std::vector<double> w;
    if (window != 0)
    {
        hann(nperseg, false, w);
    }
    else
    {
        w.assign(nperseg, 1.0);
    }
    double winsum2 = 0.0;
    for (double v : w)
    {
        winsum2 += v * v;
    }
    const double scale = 1.0 / (fs * winsum2);
    std::vector<double> seg(nperseg), spec(2 * nfreq);
    for (size_t s = 0; s < nseg; ++s)
    {
        const double* xs = x + s * hop;
        double mean = 0.0;
        if (detrend != 0)
        {
            for (size_t i = 0; i < nperseg; ++i)
            {
                mean += xs[i];
            }
            mean /= static_cast<double>(nperseg);
        }
        for (size_t i = 0; i < nperseg; ++i)
        {
            seg[i] = (xs[i] - mean) * w[i];
        }
        oracle_rfft64(seg.data(), spec.data(), nperseg);
        for (size_t k = 0; k < nfreq; ++k)
        {
            double p = scale * (spec[2 * k] * spec[2 * k] + spec[2 * k + 1] * spec[2 * k + 1]);
            if (k != 0 && k != nfreq - 1)
            {
                p *= 2.0;
            }
            out[s * nfreq + k] = p;
        }
    }
    return nseg;
}

// welch_psd (spectral.hpp:140-172): the per-segment periodograms (hann, detrend constant) averaged in segment order.
size_t oracle_welch64(const double* x, size_t len, double fs, size_t nperseg, size_t noverlap, double* psd)
{
    const size_t nfreq = nperseg / 2 + 1;
    const size_t nseg = frame_count(len, nperseg, noverlap);
    std::vector<double> rows(nseg * nfreq);
    oracle_periodograms64(x, len, fs, nperseg, noverlap, 1, 1, rows.data());
    for (size_t k = 0; k < nfreq; ++k)
    {
        psd[k] = 0.0;
    }
    for (size_t s = 0; s < nseg; ++s)
    {
        for (size_t k = 0; k < nfreq; ++k)
        {
            psd[k] += rows[s * nfreq + k];
        }
    }
    for (size_t k = 0; nseg > 0 && k < nfreq; ++k)
    {
        psd[k] /= static_cast<double>(nseg);
    }
    return nseg;
}

// stft_arm (spectral.hpp:334-386): frame s = rfft(x[s*hop, s*hop + nperseg) * hann periodic), no detrend, no
// scaling. out: nframes x (nperseg/2 + 1) complex (interleaved doubles); returns nframes.
size_t oracle_stft64(const double* x, size_t len, size_t nperseg, size_t noverlap, double* out)
{
    const size_t hop = nperseg - noverlap;
    const size_t nfreq = nperseg / 2 + 1;
    const size_t nframes = frame_count(len, nperseg, noverlap);
    std::vector<double> w, seg(nperseg);
    hann(nperseg, false, w);
    for (size_t f = 0; f < nframes; ++f)
    {
        for (size_t i = 0; i < nperseg; ++i)
        {
            seg[i] = x[f * hop + i] * w[i];
        }
        oracle_rfft64(seg.data(), out + 2 * f * nfreq, nperseg);
    }
    return nframes;
}

} // extern "C"
