// Layout and device entry points of the plan cost model (mlp_kernels.hip, mlp.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crdfft
{

struct MlpDims
{
    uint32_t inputs;
    uint32_t hidden;
};

// Offsets of each tensor in the flat parameter (and gradient) vector.
struct MlpOffsets
{
    uint64_t w1, b1, w2, b2, w3, b3, total;
};

inline MlpOffsets mlp_offsets(const MlpDims& d)
{
    MlpOffsets o{};
    o.w1 = 0;
    o.b1 = o.w1 + static_cast<uint64_t>(d.hidden) * d.inputs;
    o.w2 = o.b1 + d.hidden;
    o.b2 = o.w2 + static_cast<uint64_t>(d.hidden) * d.hidden;
    o.w3 = o.b2 + d.hidden;
    o.b3 = o.w3 + d.hidden;
    o.total = o.b3 + 1;
    return o;
}

// h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = w3 h2 + b3 for `rows` rows.
hipError_t mlp_forward(const MlpDims& d, const float* params, const float* x, uint32_t rows, float* h1, float* h2,
                       float* y, hipStream_t s);
// Gradient of mean((y - t)^2) into grad (flat); g, dh2, dh1 are row scratch.
hipError_t mlp_backward(const MlpDims& d, const float* params, const float* x, const float* t, uint32_t rows,
                        const float* h1, const float* h2, const float* y, float* g, float* dh2, float* dh1,
                        float* grad, hipStream_t s);
hipError_t mlp_adam(float* p, const float* g, float* m1, float* m2, uint64_t count, float lr, uint32_t step,
                    hipStream_t s);
hipError_t mlp_gather(const float* src, const uint32_t* idx, float* dst, uint32_t rows, uint32_t width, hipStream_t s);

} // namespace crdfft
