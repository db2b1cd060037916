// Auxiliary passes of the N-dimensional plan (nd.cpp): a batched tiled transpose (for axes the strided column kernel
// does not cover) and a scale pass (the inverse's 1/prod(d), nd_fft.hpp:104).
#include "dft.hpp"
#include "kernels.hpp"

namespace crdfft
{
using namespace dev;

namespace
{

constexpr int kTile = 32;

// out[t][c][r] = in[t][r][c] for r < rows, c < cols: 32x32 tiles through LDS (pitch 33: conflict-free columns),
// 32x8 threads, each moving four values.
__global__ __launch_bounds__(256) void k_transpose(const cf* in, cf* out, uint64_t rows, uint64_t cols,
                                                   uint64_t count, uint64_t idist, uint64_t odist)
{
    __shared__ float tre[kTile][kTile + 1];
    __shared__ float tim[kTile][kTile + 1];
    const uint64_t tiles_c = (cols + kTile - 1) / kTile;
    const uint64_t tiles_r = (rows + kTile - 1) / kTile;
    const uint64_t per = tiles_c * tiles_r;
    const uint64_t b = blockIdx.x;
    const uint64_t t = b / per;
    if (t >= count)
    {
        return;
    }
    const uint64_t rem = b - t * per;
    const uint64_t r0 = (rem / tiles_c) * kTile;
    const uint64_t c0 = (rem % tiles_c) * kTile;
    const int tx = threadIdx.x % kTile;
    const int ty = threadIdx.x / kTile;
    for (int k = ty; k < kTile; k += 8)
    {
        const uint64_t r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols)
        {
            const cf v = in[t * idist + r * cols + c];
            tre[k][tx] = v.x;
            tim[k][tx] = v.y;
        }
    }
    __syncthreads();
    for (int k = ty; k < kTile; k += 8)
    {
        const uint64_t c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols)
        {
            out[t * odist + c * rows + r] = cf{tre[tx][k], tim[tx][k]};
        }
    }
}

__global__ void k_scale(cf* data, uint64_t count, float s)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    {
        const cf v = data[i];
        data[i] = cf{v.x * s, v.y * s};
    }
}

// data[t*dist + i] *= s for i < n, t < count (the rows of a padded batch; the padding is left alone)
__global__ void k_scale_rows(cf* data, uint64_t n, uint64_t count, uint64_t dist, float s)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < n * count; idx += stride)
    {
        const uint64_t t = idx / n;
        cf* p = data + t * dist + (idx - t * n);
        const cf v = *p;
        *p = cf{v.x * s, v.y * s};
    }
}

} // namespace

hipError_t launch_scale_rows(cfloat* data, uint64_t n, uint64_t count, uint64_t dist, float scale, hipStream_t s)
{
    if (n * count == 0)
    {
        return hipSuccess;
    }
    const uint64_t blocks = std::min<uint64_t>((n * count + 255) / 256, 65536);
    hipLaunchKernelGGL(k_scale_rows, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       reinterpret_cast<cf*>(data), n, count, dist, scale);
    return hipGetLastError();
}

hipError_t launch_transpose(const cfloat* in, cfloat* out, uint64_t rows, uint64_t cols, uint64_t count,
                            uint64_t idist, uint64_t odist, hipStream_t s)
{
    const uint64_t blocks = ((rows + kTile - 1) / kTile) * ((cols + kTile - 1) / kTile) * count;
    if (blocks == 0)
    {
        return hipSuccess;
    }
    if (blocks > 0x7FFFFFFFull)
    {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_transpose, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       reinterpret_cast<const cf*>(in), reinterpret_cast<cf*>(out), rows, cols, count, idist, odist);
    return hipGetLastError();
}

hipError_t launch_scale(cfloat* data, uint64_t count, float scale, hipStream_t s)
{
    if (count == 0)
    {
        return hipSuccess;
    }
    const uint64_t blocks = std::min<uint64_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL(k_scale, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, reinterpret_cast<cf*>(data),
                       count, scale);
    return hipGetLastError();
}

} // namespace crdfft
