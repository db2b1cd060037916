// Pointwise passes of the Bluestein (chirp-z) transform for any length n (the reference's BluesteinPlan,
// engine/numerics/hesap-fft/include/crd/hesap/fft/bluestein.hpp:1-136):
//   X[k] = conj(c[k]) * ((x conj(c)) (*) b)[k],  c[k] = exp(+i pi k^2 / n),  b = c symmetrised to length M >= 2n-1,
// the convolution as IFFT_M(FFT_M(a) * FFT_M(b)) over the library's power-of-two plan. The inverse conjugates in and
// out and scales by 1/n (bluestein.hpp:86, :127-133). One thread per element; every kernel guards its range.
#include "dft.hpp"
#include "kernels.hpp"

namespace crdfft
{
using namespace dev;

namespace
{

// w[t][k] = (inverse ? conj x : x)[t][k] * conj(c[k]) for k < n, 0 for n <= k < m
__global__ void k_blue_pre(const cf* x, cf* w, const cf* chirp, uint64_t n, uint64_t m, uint64_t count,
                           uint64_t idist, int inverse)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= m * count)
    {
        return;
    }
    const uint64_t t = idx / m;
    const uint64_t k = idx - t * m;
    cf v = {0.0f, 0.0f};
    if (k < n)
    {
        cf xv = x[t * idist + k];
        if (inverse)
        {
            xv.y = -xv.y;
        }
        v = cmulc(xv, chirp[k]);
    }
    w[idx] = v;
}

// w[t][k] *= bhat[k]
__global__ void k_blue_mul(cf* w, const cf* bhat, uint64_t m, uint64_t count)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= m * count)
    {
        return;
    }
    const uint64_t k = idx % m;
    w[idx] = cmul(w[idx], bhat[k]);
}

// y[t][k] = w[t][k] * conj(c[k]) / m, then for the inverse conj(.) / n
__global__ void k_blue_post(const cf* w, cf* y, const cf* chirp, uint64_t n, uint64_t m, uint64_t count,
                            uint64_t odist, int inverse)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n * count)
    {
        return;
    }
    const uint64_t t = idx / n;
    const uint64_t k = idx - t * n;
    const float cm = 1.0f / static_cast<float>(m);
    cf v = cmulc(w[t * m + k], chirp[k]);
    v.x *= cm;
    v.y *= cm;
    if (inverse)
    {
        const float s = 1.0f / static_cast<float>(n);
        v.x *= s;
        v.y = -v.y * s;
    }
    y[t * odist + k] = v;
}

// c[k] = exp(+i pi (k^2 mod 2n) / n) in FP64, narrowed
__global__ void k_blue_chirp(cf* chirp, uint64_t n)
{
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n)
    {
        return;
    }
    const uint64_t two_n = 2 * n;
    const uint64_t kk = static_cast<uint64_t>((static_cast<unsigned __int128>(k % two_n) * (k % two_n)) % two_n);
    // angle pi * kk / n reduced to (-pi, pi]
    int64_t s = static_cast<int64_t>(kk);
    if (kk > n)
    {
        s -= static_cast<int64_t>(two_n);
    }
    double sn, cs;
    sincospi(static_cast<double>(s) / static_cast<double>(n), &sn, &cs);
    chirp[k] = cf{static_cast<float>(cs), static_cast<float>(sn)};
}

// b[0] = c[0], b[k] = b[m-k] = c[k] for 0 < k < n, else 0
__global__ void k_blue_b(const cf* chirp, cf* b, uint64_t n, uint64_t m)
{
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= m)
    {
        return;
    }
    cf v = {0.0f, 0.0f};
    if (k < n)
    {
        v = chirp[k];
    }
    else if (k > m - n)
    {
        v = chirp[m - k];
    }
    b[k] = v;
}

unsigned blocks_for(uint64_t items) { return static_cast<unsigned>((items + 255) / 256); }

} // namespace

hipError_t launch_blue_pre(const cfloat* x, cfloat* w, const cfloat* chirp, uint64_t n, uint64_t m, uint64_t count,
                           uint64_t idist, bool inverse, hipStream_t s)
{
    const uint64_t items = m * count;
    if (items == 0)
        return hipSuccess;
    if (items > 0xFFFFFFFFull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_blue_pre, dim3(blocks_for(items)), dim3(256), 0, s, reinterpret_cast<const cf*>(x),
                       reinterpret_cast<cf*>(w), reinterpret_cast<const cf*>(chirp), n, m, count, idist,
                       inverse ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_blue_mul(cfloat* w, const cfloat* bhat, uint64_t m, uint64_t count, hipStream_t s)
{
    const uint64_t items = m * count;
    if (items == 0)
        return hipSuccess;
    if (items > 0xFFFFFFFFull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_blue_mul, dim3(blocks_for(items)), dim3(256), 0, s, reinterpret_cast<cf*>(w),
                       reinterpret_cast<const cf*>(bhat), m, count);
    return hipGetLastError();
}

hipError_t launch_blue_post(const cfloat* w, cfloat* y, const cfloat* chirp, uint64_t n, uint64_t m, uint64_t count,
                            uint64_t odist, bool inverse, hipStream_t s)
{
    const uint64_t items = n * count;
    if (items == 0)
        return hipSuccess;
    if (items > 0xFFFFFFFFull)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_blue_post, dim3(blocks_for(items)), dim3(256), 0, s, reinterpret_cast<const cf*>(w),
                       reinterpret_cast<cf*>(y), reinterpret_cast<const cf*>(chirp), n, m, count, odist,
                       inverse ? 1 : 0);
    return hipGetLastError();
}

hipError_t build_blue_tables(cfloat* chirp, cfloat* b, uint64_t n, uint64_t m, hipStream_t s)
{
    hipLaunchKernelGGL(k_blue_chirp, dim3(blocks_for(n)), dim3(256), 0, s, reinterpret_cast<cf*>(chirp), n);
    hipLaunchKernelGGL(k_blue_b, dim3(blocks_for(m)), dim3(256), 0, s, reinterpret_cast<const cf*>(chirp),
                       reinterpret_cast<cf*>(b), n, m);
    return hipGetLastError();
}

} // namespace crdfft
