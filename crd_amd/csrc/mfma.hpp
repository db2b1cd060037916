// Host view of the split-precision MFMA passes (mfma_kernels.hip): the four-step of n = 256 x 256.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace crdfft
{

constexpr int kSplitBf16 = 1; // CRDFFT_SPLIT_BF16
constexpr int kSplitFp16 = 2; // CRDFFT_SPLIT_FP16

struct SplitArgs
{
    const cfloat* in;
    cfloat* out;
    const float* tab;     // operand splits of the DFT matrices (build_split_tables), this direction's
    const cfloat* tw256;  // W_256^e, e < 256
    const cfloat* thi;    // four-step twiddle W_65536^{j << hshift}
    const cfloat* tlo;    // W_65536^{j}, j < 2^hshift
    uint32_t hshift;
    uint64_t count;       // 65536-point transforms
    uint64_t idist;
    uint64_t odist;
};

uint64_t split_table_floats();
hipError_t build_split_tables(float* tab, int mode, bool inverse, hipStream_t s);
// pass 0: column pass with the four-step twiddle (in: the transforms, out: the intermediate, same layout);
// pass 1: row pass with the transposed store (in: the intermediate, out: the transforms)
hipError_t launch_split_pass(int pass, int mode, bool inverse, const SplitArgs& a, hipStream_t s);

} // namespace crdfft
