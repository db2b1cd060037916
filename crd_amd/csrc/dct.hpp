// Shared layout of the real-to-real transforms (dct_kernels.hip) and their host plan (dct.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace crdfft
{

// The transform kinds, numbered as crdfft_r2r_kind.
enum DctKind : int
{
    kDct2 = 0,
    kDct3 = 1,
    kDst2 = 2,
    kDst3 = 3
};

// Largest length of the one-kernel form: the half-length complex FFT (h = n/2 <= 16384) stays in one workgroup.
constexpr uint64_t kDctOneKernelMax = 32768;

// `count` rows of n reals, row t at in + t * idist and out + t * odist.
//   cs[k]  = (cos theta_k, sin theta_k), theta_k = pi k / 2n, k <= n/2
//   rec[k] = W_n^k = e^{-2 pi i k / n}, k < n/2 (the real-FFT recombination twiddle)
//   tw     = stage table of the (n/2)-point core at e values per thread
struct DctArgs
{
    const float* in;
    float* out;
    const cfloat* cs;
    const cfloat* rec;
    const cfloat* tw;
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
};

// Values per thread of the one-kernel form at length n, or 0 where it has none (n < 4 or n > kDctOneKernelMax).
int dct_one_kernel_e(uint64_t n);
hipError_t launch_dct_one_kernel(uint64_t n, int e, int kind, const DctArgs& a, hipStream_t s);
// n = 2: the four transforms as their 2-term sums.
hipError_t launch_dct_n2(int kind, const DctArgs& a, hipStream_t s);

// FP64-built tables: cs (h + 1 entries) and rec (h entries) for length n = 2h.
hipError_t build_dct_tables(cfloat* cs, cfloat* rec, uint64_t n, hipStream_t s);

// The separate-pass form over the library's real FFT plans (any pow2 n >= 4):
//   forward kinds: w = shuffle(x) (odd samples reversed, negated for DST-II) -> R2C -> spec -> post
//   inverse kinds: spec = Makhoul pre-rotation of x (h + 1 bins) -> C2R (scaled 1/n) -> v -> unshuffle * n
hipError_t launch_dct_shuffle(const float* x, float* w, uint64_t n, uint64_t count, uint64_t idist, bool sine,
                              hipStream_t s);
hipError_t launch_dct_fwd_post(const cfloat* spec, float* y, const cfloat* cs, uint64_t n, uint64_t count,
                               uint64_t odist, bool sine, hipStream_t s);
hipError_t launch_dct_inv_pre(const float* x, cfloat* spec, const cfloat* cs, uint64_t n, uint64_t count,
                              uint64_t idist, bool sine, hipStream_t s);
hipError_t launch_dct_inv_post(const float* v, float* y, uint64_t n, uint64_t count, uint64_t odist, bool sine,
                               hipStream_t s);

} // namespace crdfft
