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

} // namespace crdfft
