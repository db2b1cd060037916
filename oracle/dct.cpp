// Test infrastructure only: the FP64 oracle of the real-to-real transforms (the reference's DctPlan,
// engine/numerics/hesap-fft/include/crd/hesap/fft/dct.hpp:10-17 definitions, scipy norm=None):
//   kind 0 DCT-II : y[k] = 2 sum_n x[n] cos(pi (2n+1) k / 2N)
//   kind 1 DCT-III: y[k] = x[0] + 2 sum_{n>=1} x[n] cos(pi n (2k+1) / 2N)
//   kind 2 DST-II : y[k] = 2 sum_n x[n] sin(pi (2n+1) (k+1) / 2N)
//   kind 3 DST-III: y[k] = (-1)^k x[N-1] + 2 sum_{n<=N-2} x[n] sin(pi (n+1) (2k+1) / 2N)
// Two independent routes: the direct O(N^2) sums (the reference's own gate, dct.hpp:171-239) and an O(N log N) route
// through a 2N-point FP64 FFT of the zero-padded, phase-rotated input (not Makhoul's shuffle, which the GPU path
// uses), so an error in the shuffle or in the half-angle phases cannot cancel. Pinned against scipy.fft.dct/dst and
// against each other in tests/test_oracle.py.
#include <cmath>
#include <complex>
#include <vector>

#include "oracle.h"

namespace
{

constexpr double kPi = 3.14159265358979323846;
using cd = std::complex<double>;

// cos / sin of pi * num / den with num reduced mod 2 den, so large arguments stay exact in units of pi
double cospi_frac(uint64_t num, uint64_t den) { return std::cos(kPi * static_cast<double>(num % (2 * den)) / static_cast<double>(den)); }
double sinpi_frac(uint64_t num, uint64_t den) { return std::sin(kPi * static_cast<double>(num % (2 * den)) / static_cast<double>(den)); }

void direct(const double* x, double* y, size_t n, int kind)
{
    const uint64_t den = 2 * n;
    for (size_t k = 0; k < n; ++k)
    {
        double acc = 0.0;
        switch (kind)
        {
            case 0:
                for (size_t j = 0; j < n; ++j)
                    acc += x[j] * cospi_frac((2 * j + 1) * k, den);
                y[k] = 2.0 * acc;
                break;
            case 1:
                for (size_t j = 1; j < n; ++j)
                    acc += x[j] * cospi_frac(j * (2 * k + 1), den);
                y[k] = x[0] + 2.0 * acc;
                break;
            case 2:
                for (size_t j = 0; j < n; ++j)
                    acc += x[j] * sinpi_frac((2 * j + 1) * (k + 1), den);
                y[k] = 2.0 * acc;
                break;
            default:
                for (size_t j = 0; j + 1 < n; ++j)
                    acc += x[j] * sinpi_frac((j + 1) * (2 * k + 1), den);
                y[k] = ((k % 2 == 0) ? 1.0 : -1.0) * x[n - 1] + 2.0 * acc;
                break;
        }
    }
}

// e^{i pi m / 2N}
cd half_phase(size_t m, size_t n) { return {cospi_frac(m, 2 * n), sinpi_frac(m, 2 * n)}; }

// One transform through the 2N-point FFT (oracle_fft64, unnormalized both ways):
//   DCT-II : y[k] = 2 Re(e^{-i pi k/2N} X[k]),        X = DFT_2N(x padded)
//   DST-II : y[k] = -2 Im(e^{-i pi (k+1)/2N} X[k+1])
//   DCT-III: y[k] = 2 Re(IDFT_2N(u)[k]) - x[0],        u[m] = x[m] e^{i pi m/2N}, m < N
//   DST-III: y[k] = Im(IDFT_2N(u)[k]),                 u[m] = c_m x[m-1] e^{i pi m/2N}, 1 <= m <= N, c_N = 1, else 2
void via_fft(const double* x, double* y, size_t n, int kind, std::vector<cd>& buf)
{
    const size_t m2 = 2 * n;
    buf.assign(m2, cd{0.0, 0.0});
    if (kind == 0 || kind == 2)
    {
        for (size_t j = 0; j < n; ++j)
            buf[j] = x[j];
        oracle_fft64(reinterpret_cast<double*>(buf.data()), m2, 0);
        for (size_t k = 0; k < n; ++k)
        {
            if (kind == 0)
                y[k] = 2.0 * (std::conj(half_phase(k, n)) * buf[k]).real();
            else
                y[k] = -2.0 * (std::conj(half_phase(k + 1, n)) * buf[k + 1]).imag();
        }
        return;
    }
    if (kind == 1)
    {
        for (size_t j = 0; j < n; ++j)
            buf[j] = x[j] * half_phase(j, n);
    }
    else
    {
        for (size_t j = 1; j <= n; ++j)
            buf[j] = (j == n ? 1.0 : 2.0) * x[j - 1] * half_phase(j, n);
    }
    oracle_fft64(reinterpret_cast<double*>(buf.data()), m2, 1);
    for (size_t k = 0; k < n; ++k)
        y[k] = kind == 1 ? 2.0 * buf[k].real() - x[0] : buf[k].imag();
}

} // namespace

extern "C" {

void oracle_dct_direct64(const double* x, double* y, size_t n, int kind) { direct(x, y, n, kind); }

void oracle_dct64(const double* x, double* y, size_t n, size_t batch, int kind)
{
    std::vector<cd> buf;
    for (size_t t = 0; t < batch; ++t)
        via_fft(x + t * n, y + t * n, n, kind, buf);
}

} // extern "C"
