// Host view of the split-precision MFMA passes (mx_kernels.hip): radix-16 (256-point) and radix-32 (1024-point)
// sequences as DFT-matrix products, one whole transform of 256 or 1024 points, or the two passes of a four-step
// n = n1 * n2 with n1, n2 in {256, 1024}.
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
    const uint16_t* tab; // operand splits of the radix's DFT matrices (build_split_tables), this direction's
    const cfloat* tws;   // W_S^e, e < S: the stage twiddles of the sequence length S
    const cfloat* thi;   // pass 0: four-step twiddle W_n^{j << hshift}
    const cfloat* tlo;   // pass 0: W_n^{j}, j < 2^hshift
    const cfloat* tfull; // pass 0, when set: W_n^{c k} at [k * n2 + c] (one coalesced load per value)
    uint32_t hshift;
    uint64_t count;      // transforms (pass 2) or four-step matrices (passes 0, 1)
    uint64_t idist;
    uint64_t odist;
    uint64_t n_seq;      // pass 0: columns (n2); pass 1: rows (n1); pass 2: 1
    float scale;         // pass 1 and 2: applied to every output
    const cfloat* tcm;   // pass 0, when set and tfull is not: W_n^{c m} at [m * n2 + c] for m < R
};

// The radix of a sequence length (256 -> 16, 1024 -> 32), false for other lengths.
bool split_radix_of(uint64_t len, int* radix);
// 16-bit entries of one direction's operand table (f16 or bf16 values by mode)
uint64_t split_table_entries(int radix);
hipError_t build_split_tables(uint16_t* tab, int radix, int mode, bool inverse, hipStream_t s);
// pass 0: column pass with the four-step twiddle (in: the transforms, out: the intermediate, same layout);
// pass 1: row pass with the transposed store (in: the intermediate, out: the transforms);
// pass 2: whole transforms of one sequence each
hipError_t launch_split_pass(int pass, int radix, int mode, bool inverse, const SplitArgs& a, hipStream_t s);

} // namespace crdfft
