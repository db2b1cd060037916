// Real-transform passes (RealFftPlan, engine/numerics/hesap-fft/include/crd/hesap/fft/real_fft.hpp:41-93), the
// one-kernel C2R of n <= 32768 (k_c2r1; the one-kernel R2C is the spectral kernel's plain mode) and a byte-fill
// kernel.
//
// n reals are read as h = n/2 complex z[j] = x[2j] + i x[2j+1] (no pack pass: that is the memory layout), the
// complex plan runs on z, and one thread per pair {k, h-k} does the Hermitian recombine, so each output slot has a
// single writer and the pass can run in place:
//   E[k] = (Z[k] + conj Z[h-k]) / 2,  O[k] = (Z[k] - conj Z[h-k]) / (2i),  X[k] = E[k] + W_n^k O[k],
//   X[h-k] = conj E[k] + W_n^{h-k} conj O[k],  X[0] = Re Z0 + Im Z0,  X[h] = Re Z0 - Im Z0.
// C2R inverts it, Z[k] = E[k] + i O[k] with E = (X[k] + conj X[h-k])/2, O = W_n^{-k} (X[k] - conj X[h-k])/2,
// and folds irfft's 1/h into Z before the inverse complex transform.
#include "buffer.hpp"
#include "dft.hpp"
#include "fft_core.hpp"
#include "kernels.hpp"
#include "real.hpp"

#include <algorithm>

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

// k_c2r1<H, E>: the whole C2R of one row of n = 2H reals per slot of T = H/E threads (the k_rows layout of
// fft_core): slot k loads X[k] and X[H-k] (buffer loads; an inactive slot has no records and reads zeros), forms
// scale * Z[k] as k_c2r_pre does, runs the inverse H-point core and stores z[j] = (x[2j], x[2j+1]). One HBM read of
// the half spectrum and one write of the reals, where the separate form reads and writes the row twice.
constexpr int round_up_c(int v, int m) { return (v + m - 1) / m * m; }

template <int H, int E> struct C2rCfg
{
    static constexpr int T = H / E;
    static constexpr int M = T >= 256 ? 1 : 256 / T;
    static constexpr int threads = T * M;
    static constexpr int LF = lds_floats(H);
    static constexpr int RS = (H == E) ? 0 : round_up_c(2 * LF, 64) + (T < 64 ? T : 0);
    static constexpr int lds_bytes = RS * M * 4;
};

template <int H, int E>
__global__ __launch_bounds__((C2rCfg<H, E>::threads)) void k_c2r1(C2rArgs a)
{
    using C = C2rCfg<H, E>;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int local = tid / C::T;
    const int i = tid - local * C::T;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * C::M + local;
    const bool active = t < a.count;
    const __amdgpu_buffer_rsrc_t rx =
        make_rsrc(reinterpret_cast<const cf*>(a.in) + (active ? t * a.idist : 0), active ? (H + 1) * 8u : 0u);
    const __amdgpu_buffer_rsrc_t rw = make_rsrc(a.w, (H + 1) * 8u);
    cf v[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const uint32_t k = static_cast<uint32_t>(i + C::T * m);
        v[m] = c2r_z(bld<0>(rx, 8u * k, 0), bld<0>(rx, 8u * (H - k), 0), bld<0>(rw, 8u * k, 0), a.scale);
    }
    const LdsView<1> view{lds + local * C::RS, C::LF};
    fft_core<H, E, true, 1>(v, i, view, reinterpret_cast<const cf*>(a.tw));
    const __amdgpu_buffer_rsrc_t ry =
        make_rsrc(reinterpret_cast<cf*>(a.out) + (active ? t * a.odist : 0), active ? H * 8u : 0u);
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        bst<0>(ry, 8u * static_cast<uint32_t>(i + C::T * m), 0, v[m]);
    }
}

#define CRDFFT_C2R1_LIST(X)                                                                                            \
    X(2, 2) X(4, 4) X(8, 8) X(16, 4) X(32, 4) X(64, 4) X(128, 4) X(256, 4) X(512, 8) X(1024, 8) X(2048, 8) X(4096, 8)  \
        X(8192, 8) X(16384, 16)

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

int c2r_one_kernel_e(uint64_t h)
{
#define X(H, E)                                                                                                        \
    if (h == H)                                                                                                        \
        return E;
    CRDFFT_C2R1_LIST(X)
#undef X
    return 0;
}

hipError_t launch_c2r_one_kernel(uint64_t h, int e, const C2rArgs& a, hipStream_t s)
{
#define X(H, E)                                                                                                        \
    if (h == H && e == E)                                                                                              \
    {                                                                                                                  \
        using C = C2rCfg<H, E>;                                                                                        \
        auto kern = k_c2r1<H, E>;                                                                                      \
        if (C::lds_bytes > 64 * 1024)                                                                                  \
        {                                                                                                              \
            hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                                  \
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, C::lds_bytes);            \
            if (err != hipSuccess)                                                                                     \
                return err;                                                                                            \
        }                                                                                                              \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            C2rArgs part = a;                                                                                          \
            const uint64_t left = a.count - done;                                                                      \
            const uint64_t groups = std::min<uint64_t>((left + C::M - 1) / C::M, 0xFFFFFFFFull / C::threads);         \
            part.count = std::min<uint64_t>(left, groups * C::M);                                                      \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups)), dim3(C::threads), C::lds_bytes, s, part);   \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_C2R1_LIST(X)
#undef X
    return hipErrorInvalidValue;
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
