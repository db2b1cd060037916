/* Test infrastructure only: the CPU oracle of the batched FFT path.
 *
 * Two halves, both plain C++ compiled by oracle/Makefile into oracle/liboracle.so:
 *   - FP64 truth (oracle_fft64*, oracle_dft64, oracle_rfft64/irfft64): the ground truth every GPU result is gated on.
 *   - An FP32 restatement of the reference's CPU algorithm (oracle_plan32*, oracle_c2c32_batch, oracle_rfft32_batch,
 *     oracle_irfft32_batch, oracle_bluestein32*): the "reference FP32" the GPU path is compared with, and the CPU
 *     baseline that bench.py times.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and only as the checker
 * (or, for cpu_baseline, as the thing the GPU is compared with). The product library libcrdfft.so never links it.
 *
 * Conventions (reference engine/numerics/hesap-fft/include/crd/hesap/fft/fft.hpp:16, :251-260):
 *   complex values are interleaved {re, im}; forward X[k] = sum_j x[j] e^{-2 pi i jk/n}; inverse uses +i;
 *   neither direction is scaled; irfft scales by 1/n (real_fft.hpp:66-93).
 *
 * Pinning (what each function is checked against, tests/test_oracle.py):
 *   oracle_dft64     closed forms (impulse, constant, single tone) and numpy.fft (pocketfft, FP64) fixtures.
 *   oracle_fft64     numpy.fft fixtures for pow2 and 2/3/5/7-smooth n and primes (Bluestein leg), oracle_dft64,
 *                    Parseval, linearity, the shift theorem.
 *   oracle_rfft64 / oracle_irfft64   numpy.fft.rfft/irfft fixtures, Hermitian symmetry, round trip.
 *   oracle_plan32*   FP64 truth at the reference's tolerance classes (test_fft_kernels.cpp:151-163), the reference's
 *                    known-answer integer inputs (fft_self_check.cpp:43-48) to its stated 1.0e-7 of max|X| bound,
 *                    and a planted wrong twiddle sign that the tests must catch.
 *   oracle_*32_batch bit-identity across thread counts (batch_arms.hpp:11-14) and FP64 truth.
 *   oracle_bluestein32   FP64 truth for primes and 7-smooth n.
 *   oracle_dct_direct64 / oracle_dct64   scipy.fft.dct/dst types 2 and 3 (norm=None, the reference's stated
 *                    convention), each other, closed forms (constant input, a single cosine), and the inverse pairs
 *                    dct3(dct2(x)) = dst3(dst2(x)) = 2N x.
 *   oracle_welch64 / oracle_periodograms64   the reference's scipy vectors (spectral_refs.inc): Welch, spectrogram,
 *                    periodogram; oracle_stft64 numpy rfft of explicitly windowed frames; oracle_hann64 the closed form
 *                    and scipy's hann.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- shared helpers ---- */
/* The reference tests' LCG (tests/numerics/hesap-fft/test_fft.cpp:31-44): s = s*6364136223846793005 +
 * 1442695040888963407, value = (s >> 11) / 2^53 * 2 - 1. Writes `count` doubles (re, im alternate). */
void oracle_lcg_fill64(double* out, size_t count, uint64_t seed);
/* Same stream narrowed to float. */
void oracle_lcg_fill32(float* out, size_t count, uint64_t seed);
/* Standard normal pairs by Box-Muller over the same LCG (u in (0,1] from (s>>11)+1). */
void oracle_normal_fill32(float* out, size_t count, uint64_t seed);

/* ---- FP64 truth ---- */
void oracle_dft64(const double* in, double* out, size_t n, int inverse);
void oracle_fft64(double* data, size_t n, int inverse);
void oracle_fft64_batch(double* data, size_t n, size_t batch, size_t dist, int inverse, int threads);
/* real[n] -> complex[n/2+1] (any n >= 1), and back (scaled by 1/n). */
void oracle_rfft64(const double* x, double* out, size_t n);
void oracle_irfft64(const double* in, double* x, size_t n);
/* FP64 truth over FP32 storage: widen, transform in FP64, keep FP64 output (out has 2*n*batch doubles). */
void oracle_c2c_f32_to_f64(const float* in, double* out, size_t n, size_t batch, int inverse, int threads);

/* ---- FP32 restatement of the reference's CPU path ---- */
typedef struct oracle_plan32 oracle_plan32;
/* A FftPlan<f32>-shaped plan of pow2 n: Stockham below 2^17, four-step from there (fft.hpp:63-66, :112-145). */
oracle_plan32* oracle_plan32_create(size_t n);
/* Same, with every twiddle's imaginary part negated: a planted mistake the parity tests must catch. */
oracle_plan32* oracle_plan32_create_planted(size_t n);
void oracle_plan32_destroy(oracle_plan32* plan);
/* In place, n contiguous complex (fft.hpp:409-413). */
void oracle_plan32_execute(oracle_plan32* plan, float* data, int inverse);
/* Element-major batch: element i of transform t at data[i*batch + t] (fft.hpp:462-470). */
void oracle_plan32_execute_batched(oracle_plan32* plan, float* data, size_t batch, int inverse);
/* The radix-2 reference transform (fft.hpp:417-460). */
void oracle_plan32_execute_reference(oracle_plan32* plan, float* data, int inverse);
/* `batch` transforms at distance `dist` complex, split over `threads` jobs of contiguous ranges, one plan per job
 * (batch_arms.hpp:46-77). threads <= 0 means every hardware thread. Non-pow2 n runs Bluestein (bluestein.hpp). */
void oracle_c2c32_batch(float* data, size_t n, size_t batch, size_t dist, int inverse, int threads);
/* Real transforms of pow2 n >= 4 (real_fft.hpp:41-93); in: batch*n floats, out: batch*(n/2+1) complex. */
void oracle_rfft32_batch(const float* in, float* out, size_t n, size_t batch, int threads);
void oracle_irfft32_batch(const float* in, float* out, size_t n, size_t batch, int threads);
/* Bluestein of any n >= 1 over a pow2 plan (bluestein.hpp:32-136); inverse is scaled by 1/n like the reference. */
void oracle_bluestein32(float* data, size_t n, int inverse);
/* Hardware threads the batch functions use for threads <= 0. */
int oracle_hardware_threads(void);

/* ---- FP64 spectral front end over the real FFT (hesap-dsp spectral.hpp / windows.hpp; spectral.cpp) ----
 * Frames s = x[s*hop, s*hop + nperseg), hop = nperseg - noverlap, count (len - noverlap)/hop (0 if len < nperseg). */
/* hann window of m points, symmetric (sym != 0) or periodic (windows.hpp:63-124). */
void oracle_hann64(double* out, size_t m, int sym);
/* Density periodogram of every frame: window 1 = hann periodic, 0 = boxcar; detrend 1 subtracts each frame's mean;
 * out nseg x (nperseg/2 + 1); returns nseg (segment_periodograms_arm, spectral.hpp:43-126; periodogram = one boxcar
 * frame of the whole signal, :186-222). */
size_t oracle_periodograms64(const double* x, size_t len, double fs, size_t nperseg, size_t noverlap, int window,
                             int detrend, double* out);
/* Welch PSD: the hann, detrend-constant periodograms averaged in frame order (spectral.hpp:140-172). */
size_t oracle_welch64(const double* x, size_t len, double fs, size_t nperseg, size_t noverlap, double* psd);
/* STFT: complex rfft of every hann-windowed frame, no detrend, no scaling (spectral.hpp:334-386). */
size_t oracle_stft64(const double* x, size_t len, size_t nperseg, size_t noverlap, double* out);

/* ---- FP64 real-to-real transforms (dct.hpp definitions, scipy norm=None; dct.cpp). kind: 0 DCT-II, 1 DCT-III,
 * 2 DST-II, 3 DST-III. ---- */
/* The direct O(n^2) sums (dct.hpp:171-239), any n >= 1. */
void oracle_dct_direct64(const double* x, double* y, size_t n, int kind);
/* The same transforms through a 2n-point FP64 FFT of the padded, phase-rotated input; n a power of two; rows of n. */
void oracle_dct64(const double* x, double* y, size_t n, size_t batch, int kind);

/* ---- FP64 restatement of the plan cost model (mlp.cpp): 3-layer ReLU perceptron, flat parameters
 * W1[h][in], b1[h], W2[h][h], b2[h], w3[h], b3; pinned by central finite differences. ---- */
void oracle_mlp_forward(const double* params, uint32_t inputs, uint32_t hidden, const double* x, uint64_t n, double* y);
double oracle_mlp_loss_grad(const double* params, uint32_t inputs, uint32_t hidden, const double* x, const double* t,
                            uint64_t n, double* grad);

#ifdef __cplusplus
}
#endif
