// The learned plan cost model's device kernels: a 3-layer perceptron (inputs -> H -> H -> 1, ReLU) in FP32, forward,
// backward and Adam. One thread per output value; every reduction over the minibatch is a loop inside one thread, so
// training is deterministic (no atomics) and the same inputs give the same bits on every run.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "mlp.hpp"

namespace crdfft
{

namespace
{

// out[b][m] = act(bias[m] + sum_k w[m][k] in[b][k])
__global__ void k_dense(const float* in, const float* w, const float* bias, float* out, uint32_t rows, uint32_t k_in,
                        uint32_t m_out, int relu)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<uint64_t>(rows) * m_out)
    {
        return;
    }
    const uint32_t b = static_cast<uint32_t>(idx / m_out);
    const uint32_t m = static_cast<uint32_t>(idx - static_cast<uint64_t>(b) * m_out);
    const float* x = in + static_cast<uint64_t>(b) * k_in;
    const float* wr = w + static_cast<uint64_t>(m) * k_in;
    float acc = bias[m];
    for (uint32_t k = 0; k < k_in; ++k)
    {
        acc = __fmaf_rn(wr[k], x[k], acc);
    }
    out[idx] = relu ? fmaxf(acc, 0.0f) : acc;
}

// g[b] = 2 (y[b] - t[b]) / rows   (d mean-squared-error / d y)
__global__ void k_dloss(const float* y, const float* t, float* g, uint32_t rows)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < rows)
    {
        g[b] = 2.0f * (y[b] - t[b]) / static_cast<float>(rows);
    }
}

// dW[m][k] = sum_b dout[b][m] in[b][k];  dbias[m] = sum_b dout[b][m]  (thread m*(k_in+1) + k, k == k_in: bias)
__global__ void k_grad_w(const float* dout, const float* in, float* dw, float* dbias, uint32_t rows, uint32_t k_in,
                         uint32_t m_out)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<uint64_t>(m_out) * (k_in + 1))
    {
        return;
    }
    const uint32_t m = static_cast<uint32_t>(idx / (k_in + 1));
    const uint32_t k = static_cast<uint32_t>(idx - static_cast<uint64_t>(m) * (k_in + 1));
    float acc = 0.0f;
    if (k < k_in)
    {
        for (uint32_t b = 0; b < rows; ++b)
        {
            acc = __fmaf_rn(dout[static_cast<uint64_t>(b) * m_out + m], in[static_cast<uint64_t>(b) * k_in + k], acc);
        }
        dw[static_cast<uint64_t>(m) * k_in + k] = acc;
    }
    else
    {
        for (uint32_t b = 0; b < rows; ++b)
        {
            acc += dout[static_cast<uint64_t>(b) * m_out + m];
        }
        dbias[m] = acc;
    }
}

// din[b][k] = [act[b][k] > 0] * sum_m dout[b][m] w[m][k]   (the ReLU of the layer below)
__global__ void k_grad_in(const float* dout, const float* w, const float* act, float* din, uint32_t rows,
                          uint32_t k_in, uint32_t m_out)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<uint64_t>(rows) * k_in)
    {
        return;
    }
    const uint32_t b = static_cast<uint32_t>(idx / k_in);
    const uint32_t k = static_cast<uint32_t>(idx - static_cast<uint64_t>(b) * k_in);
    float acc = 0.0f;
    for (uint32_t m = 0; m < m_out; ++m)
    {
        acc = __fmaf_rn(dout[static_cast<uint64_t>(b) * m_out + m], w[static_cast<uint64_t>(m) * k_in + k], acc);
    }
    din[idx] = act[idx] > 0.0f ? acc : 0.0f;
}

__global__ void k_adam(float* p, const float* g, float* m1, float* m2, uint64_t count, float lr, float beta1,
                       float beta2, float eps, float bc1, float bc2)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count)
    {
        return;
    }
    const float gi = g[i];
    const float a = beta1 * m1[i] + (1.0f - beta1) * gi;
    const float b = beta2 * m2[i] + (1.0f - beta2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

// dst[r][:] = src[idx[r]][:]
__global__ void k_gather_rows(const float* src, const uint32_t* idx, float* dst, uint32_t rows, uint32_t width)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<uint64_t>(rows) * width)
    {
        return;
    }
    const uint32_t r = static_cast<uint32_t>(i / width);
    const uint32_t c = static_cast<uint32_t>(i - static_cast<uint64_t>(r) * width);
    dst[i] = src[static_cast<uint64_t>(idx[r]) * width + c];
}

unsigned grid_for(uint64_t items) { return static_cast<unsigned>((items + 255) / 256); }

} // namespace

hipError_t mlp_forward(const MlpDims& d, const float* params, const float* x, uint32_t rows, float* h1, float* h2,
                       float* y, hipStream_t s)
{
    const MlpOffsets o = mlp_offsets(d);
    hipLaunchKernelGGL(k_dense, dim3(grid_for(static_cast<uint64_t>(rows) * d.hidden)), dim3(256), 0, s, x,
                       params + o.w1, params + o.b1, h1, rows, d.inputs, d.hidden, 1);
    hipLaunchKernelGGL(k_dense, dim3(grid_for(static_cast<uint64_t>(rows) * d.hidden)), dim3(256), 0, s, h1,
                       params + o.w2, params + o.b2, h2, rows, d.hidden, d.hidden, 1);
    hipLaunchKernelGGL(k_dense, dim3(grid_for(rows)), dim3(256), 0, s, h2, params + o.w3, params + o.b3, y, rows,
                       d.hidden, 1u, 0);
    return hipGetLastError();
}

hipError_t mlp_backward(const MlpDims& d, const float* params, const float* x, const float* t, uint32_t rows,
                        const float* h1, const float* h2, const float* y, float* g, float* dh2, float* dh1,
                        float* grad, hipStream_t s)
{
    const MlpOffsets o = mlp_offsets(d);
    hipLaunchKernelGGL(k_dloss, dim3(grid_for(rows)), dim3(256), 0, s, y, t, g, rows);
    // output layer: dw3, db3 from g and h2; dh2 = relu'(h2) * g w3
    hipLaunchKernelGGL(k_grad_w, dim3(grid_for(d.hidden + 1)), dim3(256), 0, s, g, h2, grad + o.w3, grad + o.b3, rows,
                       d.hidden, 1u);
    hipLaunchKernelGGL(k_grad_in, dim3(grid_for(static_cast<uint64_t>(rows) * d.hidden)), dim3(256), 0, s, g,
                       params + o.w3, h2, dh2, rows, d.hidden, 1u);
    // hidden layer 2
    hipLaunchKernelGGL(k_grad_w, dim3(grid_for(static_cast<uint64_t>(d.hidden) * (d.hidden + 1))), dim3(256), 0, s,
                       dh2, h1, grad + o.w2, grad + o.b2, rows, d.hidden, d.hidden);
    hipLaunchKernelGGL(k_grad_in, dim3(grid_for(static_cast<uint64_t>(rows) * d.hidden)), dim3(256), 0, s, dh2,
                       params + o.w2, h1, dh1, rows, d.hidden, d.hidden);
    // hidden layer 1
    hipLaunchKernelGGL(k_grad_w, dim3(grid_for(static_cast<uint64_t>(d.hidden) * (d.inputs + 1))), dim3(256), 0, s,
                       dh1, x, grad + o.w1, grad + o.b1, rows, d.inputs, d.hidden);
    return hipGetLastError();
}

hipError_t mlp_adam(float* p, const float* g, float* m1, float* m2, uint64_t count, float lr, uint32_t step,
                    hipStream_t s)
{
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - static_cast<float>(std::pow(0.9, static_cast<double>(step)));
    const float bc2 = 1.0f - static_cast<float>(std::pow(0.999, static_cast<double>(step)));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(count)), dim3(256), 0, s, p, g, m1, m2, count, lr, beta1, beta2, eps,
                       bc1, bc2);
    return hipGetLastError();
}

hipError_t mlp_gather(const float* src, const uint32_t* idx, float* dst, uint32_t rows, uint32_t width, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(static_cast<uint64_t>(rows) * width)), dim3(256), 0, s, src, idx,
                       dst, rows, width);
    return hipGetLastError();
}

} // namespace crdfft
