// Host-side launchers of the real-transform passes and the byte fill (real_kernels.hip).
#pragma once

#include "kernels.hpp"

namespace crdfft
{

// In place on `count` transforms at `dist` complex: Z[0..h) from the complex plan -> X[0..h] of the real input.
hipError_t launch_r2c_post(cfloat* data, const cfloat* w, uint64_t h, uint64_t count, uint64_t dist, hipStream_t s);
// X[0..h] (in) -> scale * Z[0..h) (out) for the inverse complex plan.
hipError_t launch_c2r_pre(const cfloat* in, cfloat* out, const cfloat* w, uint64_t h, uint64_t count,
                          uint64_t idist, uint64_t odist, float scale, hipStream_t s);
hipError_t launch_fill_bytes(void* dst, int value, uint64_t bytes, hipStream_t s);

// The one-kernel C2R of n = 2h <= 32768: `count` rows of h + 1 bins (idist complex apart) to rows of n reals (odist
// complex apart), every output times `scale`; w = W_n^k (k <= h), tw = the stage table of (h, e).
struct C2rArgs
{
    const cfloat* in;
    cfloat* out;
    const cfloat* w;
    const cfloat* tw;
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
    float scale;
};
// Values per thread of the one-kernel C2R at half length h, or 0 where it has none.
int c2r_one_kernel_e(uint64_t h);
hipError_t launch_c2r_one_kernel(uint64_t h, int e, const C2rArgs& a, hipStream_t s);

} // namespace crdfft
