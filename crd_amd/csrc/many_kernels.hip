// The mixed-size persistent kernel: thousands of C2C transforms of 2/3/5/7-smooth lengths (up to 16384) in one
// launch.
//
// The host (many.cpp) groups the transforms by length into tiles of at most 16384 points (several transforms of one
// length) and balances the tiles over a grid of one 1024-thread workgroup per CU (longest-processing-time first), so
// the kernel has no queue, no atomics and no inter-workgroup wait: workgroup w walks tiles [begin[w], begin[w+1]).
// Each tile runs the length's radix plan (chosen per length by the autotuner) as a Stockham autosort whose radix is a
// runtime value dispatched to compile-time butterflies: stage s reads butterfly b of transform j at slots
// b + m*N/R, twiddles by W_{PR}^{m (b mod P)}, and writes slot (b - b mod P)*R + (b mod P) + m*P. The first stage reads
// global memory and the last writes it (both unit-stride across lanes); the stages between exchange through LDS (split
// re/im, one pad float per 16). Radices 3, 5 and 7 are new relative to the reference, which sends every length that
// is not a power of two to Bluestein (engine/numerics/hesap-fft/include/crd/hesap/fft/bluestein.hpp).
#include "dft.hpp"
#include "fft_core.hpp"
#include "many.hpp"

namespace crdfft
{
using namespace dev;

namespace
{

constexpr int kThreads = kManyThreads;
constexpr int kTilePoints = kManyTilePoints;
constexpr int kLdsFloats = kTilePoints + (kTilePoints >> 4) + 16;

struct TileCtx
{
    const cf* in;
    cf* out;
    const uint64_t* offsets; // element offsets of the tile's transforms (in and out share them)
    const cf* tw;            // W_N^e, e < N
    float* lre;
    float* lim;
    uint32_t n;
    uint32_t count;
};

template <int R, bool INV>
__device__ __noinline__ void many_stage(const TileCtx& c, uint32_t p, bool first, bool last, int tid)
{
    constexpr int QMAX = (kTilePoints / kThreads + R - 1) / R;
    const uint32_t nr = c.n / R;
    const uint32_t total = c.count * nr;
    const uint32_t twstep = c.n / (p * R); // W_{PR} = W_N^{twstep}
    cf a[QMAX][R];
#pragma unroll
    for (int q = 0; q < QMAX; ++q)
    {
        const uint32_t g = tid + q * kThreads;
        if (g < total)
        {
            const uint32_t j = g / nr;
            const uint32_t b = g - j * nr;
            if (first)
            {
                const cf* src = c.in + c.offsets[j];
#pragma unroll
                for (int m = 0; m < R; ++m)
                {
                    a[q][m] = src[b + m * nr];
                }
            }
            else
            {
                const uint32_t base = j * c.n + b;
#pragma unroll
                for (int m = 0; m < R; ++m)
                {
                    const int o = pad16(static_cast<int>(base + m * nr));
                    a[q][m] = cf{c.lre[o], c.lim[o]};
                }
            }
        }
    }
    if (!first)
    {
        __syncthreads(); // every read of this stage precedes any write into the same LDS buffer
    }
#pragma unroll
    for (int q = 0; q < QMAX; ++q)
    {
        const uint32_t g = tid + q * kThreads;
        if (g < total)
        {
            const uint32_t j = g / nr;
            const uint32_t b = g - j * nr;
            const uint32_t k = b % p;
            if (p > 1)
            {
                const uint32_t ek = k * twstep;
#pragma unroll
                for (int m = 1; m < R; ++m)
                {
                    a[q][m] = ctw<INV>(a[q][m], c.tw[m * ek]);
                }
            }
            Dft<R, INV>::run(a[q]);
            const uint32_t ob = (b - k) * R + k;
            if (last)
            {
                cf* dst = c.out + c.offsets[j];
#pragma unroll
                for (int m = 0; m < R; ++m)
                {
                    dst[ob + m * p] = a[q][m];
                }
            }
            else
            {
                const uint32_t base = j * c.n + ob;
#pragma unroll
                for (int m = 0; m < R; ++m)
                {
                    const int o = pad16(static_cast<int>(base + m * p));
                    c.lre[o] = a[q][m].x;
                    c.lim[o] = a[q][m].y;
                }
            }
        }
    }
    if (!last)
    {
        __syncthreads();
    }
}

template <bool INV> __global__ __launch_bounds__(kThreads) void k_many(ManyArgs a)
{
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const uint32_t wg = blockIdx.x;
    const uint32_t t0 = a.wg_begin[wg];
    const uint32_t t1 = a.wg_begin[wg + 1];
    for (uint32_t t = t0; t < t1; ++t)
    {
        const ManyTile tile = a.tiles[t];
        const ManySize& sz = a.sizes[tile.size_id];
        TileCtx c{reinterpret_cast<const cf*>(a.in), reinterpret_cast<cf*>(a.out), a.offsets + tile.first,
                  reinterpret_cast<const cf*>(a.tw) + sz.tw_offset, lds, lds + kLdsFloats, sz.n, tile.count};
        uint32_t p = 1;
        const uint32_t ns = sz.nstages;
        for (uint32_t s = 0; s < ns; ++s)
        {
            const uint32_t r = sz.radix[s];
            const bool first = s == 0;
            const bool last = s + 1 == ns;
            switch (r)
            {
                case 1: many_stage<1, INV>(c, p, first, last, tid); break;
                case 2: many_stage<2, INV>(c, p, first, last, tid); break;
                case 3: many_stage<3, INV>(c, p, first, last, tid); break;
                case 4: many_stage<4, INV>(c, p, first, last, tid); break;
                case 5: many_stage<5, INV>(c, p, first, last, tid); break;
                case 7: many_stage<7, INV>(c, p, first, last, tid); break;
                case 8: many_stage<8, INV>(c, p, first, last, tid); break;
                default: many_stage<16, INV>(c, p, first, last, tid); break;
            }
            p *= r;
        }
    }
}

} // namespace

unsigned many_lds_bytes() { return 2u * kLdsFloats * 4u; }

hipError_t many_prepare()
{
    for (auto kern : {k_many<false>, k_many<true>})
    {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 static_cast<int>(many_lds_bytes()));
        if (e != hipSuccess)
        {
            return e;
        }
    }
    return hipSuccess;
}

hipError_t launch_many(const ManyArgs& a, unsigned workgroups, bool inverse, hipStream_t s)
{
    if (workgroups == 0)
    {
        return hipSuccess;
    }
    auto kern = inverse ? k_many<true> : k_many<false>;
    hipLaunchKernelGGL(kern, dim3(workgroups), dim3(kThreads), many_lds_bytes(), s, a);
    return hipGetLastError();
}

} // namespace crdfft
