// Real-transform passes (RealFftPlan, engine/numerics/hesap-fft/include/crd/hesap/fft/real_fft.hpp:41-93) and a
// byte-fill kernel.
//
// n reals are read as h = n/2 complex z[j] = x[2j] + i x[2j+1] (no pack pass: that is the memory layout), the
// complex plan runs on z, and one thread per pair {k, h-k} does the Hermitian recombine, so each output slot has a
// single writer and the pass can run in place:
//   E[k] = (Z[k] + conj Z[h-k]) / 2,  O[k] = (Z[k] - conj Z[h-k]) / (2i),  X[k] = E[k] + W_n^k O[k],
//   X[h-k] = conj E[k] + W_n^{h-k} conj O[k],  X[0] = Re Z0 + Im Z0,  X[h] = Re Z0 - Im Z0.
// C2R inverts it, Z[k] = E[k] + i O[k] with E = (X[k] + conj X[h-k])/2, O = W_n^{-k} (X[k] - conj X[h-k])/2,
// and folds irfft's 1/h into Z before the inverse complex transform.
#include "dft.hpp"
#include "kernels.hpp"
#include "real.hpp"

namespace crdfft
{
using namespace dev;

namespace
{

__global__ void k_r2c_post(cf* data, const cf* w, uint64_t h, uint64_t count, uint64_t dist)
{
    const uint64_t per = h / 2 + 1;
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= per * count)
    {
        return;
    }
    const uint64_t t = idx / per;
    const uint64_t k = idx - t * per;
    cf* z = data + t * dist;
    if (k == 0)
    {
        const cf z0 = z[0];
        z[0] = cf{z0.x + z0.y, 0.0f};
        z[h] = cf{z0.x - z0.y, 0.0f};
        return;
    }
    const cf a = z[k];
    const cf b = z[h - k];
    // E = (a + conj b)/2, O = (a - conj b)/(2i) = ((a.y + b.y), (b.x - a.x)) / 2
    const cf e = {(a.x + b.x) * 0.5f, (a.y - b.y) * 0.5f};
    const cf o = {(a.y + b.y) * 0.5f, (b.x - a.x) * 0.5f};
    z[k] = cadd(e, cmul(o, w[k]));
    if (2 * k != h)
    {
        const cf ec = {e.x, -e.y};
        const cf oc = {o.x, -o.y};
        z[h - k] = cadd(ec, cmul(oc, w[h - k]));
    }
}

__device__ __forceinline__ cf c2r_z(cf xk, cf xh, cf wk, float scale)
{
    const cf e = {(xk.x + xh.x) * 0.5f, (xk.y - xh.y) * 0.5f};
    const cf d = {(xk.x - xh.x) * 0.5f, (xk.y + xh.y) * 0.5f};
    const cf o = cmulc(d, wk); // W^{-k} D
    // Z = E + i O
    return {(e.x - o.y) * scale, (e.y + o.x) * scale};
}

__global__ void k_c2r_pre(const cf* in, cf* out, const cf* w, uint64_t h, uint64_t count, uint64_t idist,
                          uint64_t odist, float scale)
{
    const uint64_t per = h / 2 + 1;
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= per * count)
    {
        return;
    }
    const uint64_t t = idx / per;
    const uint64_t k = idx - t * per;
    const cf* x = in + t * idist;
    cf* z = out + t * odist;
    const cf xk = x[k];
    const cf xh = x[h - k];
    z[k] = c2r_z(xk, xh, w[k], scale);
    if (k != 0 && 2 * k != h)
    {
        z[h - k] = c2r_z(xh, xk, w[h - k], scale);
    }
}

__global__ void k_fill_bytes(uint8_t* dst, uint8_t value, uint64_t bytes)
{
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < bytes; i += stride)
    {
        dst[i] = value;
    }
}

} // namespace

hipError_t launch_r2c_post(cfloat* data, const cfloat* w, uint64_t h, uint64_t count, uint64_t dist, hipStream_t s)
{
    const uint64_t items = (h / 2 + 1) * count;
    if (items == 0)
    {
        return hipSuccess;
    }
    if (items > 0xFFFFFFFFull)
    {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_r2c_post, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<cf*>(data), reinterpret_cast<const cf*>(w), h, count, dist);
    return hipGetLastError();
}

hipError_t launch_c2r_pre(const cfloat* in, cfloat* out, const cfloat* w, uint64_t h, uint64_t count,
                          uint64_t idist, uint64_t odist, float scale, hipStream_t s)
{
    const uint64_t items = (h / 2 + 1) * count;
    if (items == 0)
    {
        return hipSuccess;
    }
    if (items > 0xFFFFFFFFull)
    {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_c2r_pre, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const cf*>(in), reinterpret_cast<cf*>(out), reinterpret_cast<const cf*>(w), h,
                       count, idist, odist, scale);
    return hipGetLastError();
}

hipError_t launch_fill_bytes(void* dst, int value, uint64_t bytes, hipStream_t s)
{
    if (bytes == 0)
    {
        return hipSuccess;
    }
    const uint64_t blocks = std::min<uint64_t>((bytes + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_bytes, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       static_cast<uint8_t*>(dst), static_cast<uint8_t>(value), bytes);
    return hipGetLastError();
}

} // namespace crdfft
