// Host view of the fused spectral kernels (spectral_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace crdfft
{

struct SpecArgs
{
    const float* x;      // the signal
    const float* window; // N floats, or nullptr for the boxcar
    const cfloat* wtab;  // W_N^k, k <= H (FP64-built)
    const cfloat* tw;    // stage twiddles of the H-point core
    void* out;           // complex rows, density rows, or partial rows (by mode)
    uint64_t frames;
    uint64_t hop;
    float scale;         // 1 / (fs * sum w^2) of the density modes
    int32_t detrend;     // subtract each frame's mean
    uint64_t ostride;    // output row pitch in elements (complex for mode 0, float for mode 1), >= H + 1
};

// The largest H with the fused Welch lane-group mode; larger frames run density rows in chunks.
constexpr int kSpecWelchMaxH = 8192;

// Values per thread and frames per workgroup of the kernel for H = nperseg / 2 (false: no instantiation).
bool spec_shape(uint64_t h, int* e, unsigned* frames_per_group);
// mode 0: complex rows (STFT); 1: density rows; 2: Welch partial rows, one per lane group. groups = workgroups.
hipError_t launch_spec(uint64_t h, int mode, const SpecArgs& a, uint64_t groups, hipStream_t s);
// Resident workgroups per CU of the Welch lane-group kernel (the persistent grid is this times the CU count).
hipError_t spec_welch_residency(uint64_t h, int* per_cu);
// acc[k] = ((first ? 0 : acc[k]) + rows[0][k] + ... + rows[nrows-1][k]) * scale, summed in a fixed order.
hipError_t launch_spec_sum(const float* rows, float* acc, uint64_t nfreq, uint64_t nrows, bool first, float scale,
                           hipStream_t s);

} // namespace crdfft
