// The mixed-size persistent kernel: thousands of C2C transforms of 2/3/5/7-smooth lengths (up to 16384) in one
// launch.
//
// The host (many.cpp) cuts the batch into tiles (runs of at most 16384 points of contiguous transforms of one
// length) and balances the tiles over a grid of one 1024-thread workgroup per CU (longest-processing-time first), so
// the kernel has no queue, no atomics and no inter-workgroup wait: workgroup w walks tiles [begin[w], begin[w+1]).
//
// Per tile: the tile's points, loaded into registers while the previous tile was transformed, go to LDS in their
// flat order (transform j, element i at j*n + i); the loads of the next tile are issued; then the length's radix
// plan (chosen per length by the autotuner) runs as a Stockham autosort whose radix is a runtime value dispatched to
// compile-time butterflies: stage s reads butterfly b of transform j at slots b + m*n/R, twiddles by
// W_{pR}^{m (b mod p)}, and writes slot (b - b mod p)*R + (b mod p) + m*p, into LDS or, in the last stage, to
// global memory. Stages touch no global memory but the last one's stores, so the next tile's loads (which complete
// in issue order with every other vector memory access) are not waited for until the tile is done. Twiddles come
// from a 256-entry LDS table per length, W_n^e = hi[e >> 7] * lo[e & 127] (both FP64-built), and the index
// divisions are multiply-highs by precomputed magic numbers. A butterfly looks up W_{pR}^k once and forms the
// other R - 2 powers by a product tree; the LDS holds the tile as interleaved complex slots, one pad slot per 16.
//
// Radices 3, 5, 7 and the products 6, 9, 10, 12 are new relative to the reference, which sends every length that is not a power of two to
// Bluestein (engine/numerics/hesap-fft/include/crd/hesap/fft/bluestein.hpp).
#include "buffer.hpp"
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
constexpr int kPer = kTilePoints / kThreads; // tile points per thread
constexpr int kLdsPoints = kTilePoints + (kTilePoints >> 4) + 16; // complex slots, one pad per 16

// x / d by the host's magic number (exact for x * d < 2^32).
__device__ __forceinline__ uint32_t mdiv(uint32_t x, uint32_t magic) { return magic == 0 ? x : __umulhi(x, magic); }

// The plan's read-only tables through the constant address space: their uniform reads become scalar loads, which
// complete on their own counter instead of waiting behind the prefetched tile's vector loads.
template <typename T> using cptr = const __attribute__((address_space(4))) T*;
template <typename T> __device__ __forceinline__ cptr<T> constant(const T* p)
{
    return reinterpret_cast<cptr<T>>(reinterpret_cast<uintptr_t>(p));
}

__device__ __forceinline__ ManyTile read(cptr<ManyTile> p)
{
    ManyTile t;
    t.size_id = p->size_id;
    t.count = p->count;
    t.base = p->base;
    return t;
}

__device__ __forceinline__ ManyStage read(cptr<ManyStage> p)
{
    ManyStage s;
    s.radix = p->radix;
    s.nr = p->nr;
    s.nr_magic = p->nr_magic;
    s.p = p->p;
    s.p_magic = p->p_magic;
    s.twstep = p->twstep;
    return s;
}

struct TileCtx
{
    __amdgpu_buffer_rsrc_t out; // the tile's output run
    cf* ld;        // LDS: the tile, point o at ld[pad16(o)]
    const cf* tlo; // LDS: W_n^j, j < 128
    const cf* thi; // LDS: W_n^{128 i}
    uint32_t n;
    uint32_t count;
};

// w[m] = w[1]^m for 2 <= m < R by a product tree (at most four products deep past w[1]).
template <int R> __device__ __forceinline__ void twiddle_powers(cf (&w)[R])
{
    if constexpr (R > 2)
        w[2] = cmul(w[1], w[1]);
    if constexpr (R > 3)
        w[3] = cmul(w[2], w[1]);
    if constexpr (R > 4)
        w[4] = cmul(w[2], w[2]);
    if constexpr (R > 5)
        w[5] = cmul(w[4], w[1]);
    if constexpr (R > 6)
        w[6] = cmul(w[3], w[3]);
    if constexpr (R > 7)
        w[7] = cmul(w[4], w[3]);
    if constexpr (R > 8)
    {
        w[8] = cmul(w[4], w[4]);
#pragma unroll
        for (int m = 9; m < R; ++m)
        {
            w[m] = cmul(w[8], w[m - 8]);
        }
    }
}

// The padded slot of point o (one pad slot per 16) as a byte offset, (o + o / 16) * 8, in unsigned arithmetic: a
// shift and one add-and-shift per point.
__device__ __forceinline__ uint32_t slot_bytes(uint32_t o) { return (o + (o >> 4)) << 3; }
__device__ __forceinline__ cf* slot(cf* ld, uint32_t o)
{
    return reinterpret_cast<cf*>(reinterpret_cast<char*>(ld) + slot_bytes(o));
}
__device__ __forceinline__ const cf* slot(const cf* ld, uint32_t o)
{
    return reinterpret_cast<const cf*>(reinterpret_cast<const char*>(ld) + slot_bytes(o));
}

// v[m] = point (base + m * stride) of the tile for m < R. When the stride is a multiple of 16 the padded slots are
// evenly spaced too (pad16(o + 16 s) = pad16(o) + 17 s): one address per butterfly and a scalar step.
template <int R>
__device__ __forceinline__ void lds_gather(cf (&v)[R], const cf* ld, uint32_t base, uint32_t stride)
{
    if ((stride & 15u) == 0)
    {
        const cf* p0 = ld + pad16(static_cast<int>(base));
        const uint32_t step = stride + (stride >> 4);
#pragma unroll
        for (int m = 0; m < R; ++m)
        {
            v[m] = p0[m * step];
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < R; ++m)
        {
            v[m] = *slot(ld, base + m * stride);
        }
    }
}

template <int R>
__device__ __forceinline__ void lds_scatter(const cf (&v)[R], cf* ld, uint32_t base, uint32_t stride)
{
    if ((stride & 15u) == 0)
    {
        cf* p0 = ld + pad16(static_cast<int>(base));
        const uint32_t step = stride + (stride >> 4);
#pragma unroll
        for (int m = 0; m < R; ++m)
        {
            p0[m * step] = v[m];
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < R; ++m)
        {
            *slot(ld, base + m * stride) = v[m];
        }
    }
}

template <int R, bool INV>
__device__ __forceinline__ void many_stage(const TileCtx& c, const ManyStage& st, bool last, int tid)
{
    constexpr int Q = (kPer + R - 1) / R;
    const uint32_t nr = st.nr;
    const uint32_t total = c.count * nr;
    cf a[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        const uint32_t g = tid + q * kThreads;
        if (g < total)
        {
            const uint32_t j = mdiv(g, st.nr_magic);
            lds_gather<R>(a[q], c.ld, g + j * (c.n - nr), nr); // j * n + b
        }
    }
    __syncthreads(); // every read of this stage precedes any write into the same LDS buffer
#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        const uint32_t g = tid + q * kThreads;
        if (g < total)
        {
            const uint32_t j = mdiv(g, st.nr_magic);
            const uint32_t b = g - j * nr;
            const uint32_t k = b - mdiv(b, st.p_magic) * st.p;
            if (R > 1 && st.p > 1) // a radix-1 stage (n = 1) has no twiddle
            {
                const uint32_t e = k * st.twstep; // W_{pR}^{k} = W_n^{e}
                if constexpr (R == 1)
                {
                }
                else if constexpr (R <= 8 || R == 16)
                {
                    cf w[R];
                    w[1] = cmul(c.thi[e >> 7], c.tlo[e & (kManyTwLo - 1)]);
                    twiddle_powers<R>(w);
#pragma unroll
                    for (int m = 1; m < R; ++m)
                    {
                        a[q][m] = ctw<INV>(a[q][m], w[m]);
                    }
                }
                else
                {
                    // two butterflies of 9 to 12 values per thread: one table product per value keeps the registers
                    // for the data
#pragma unroll
                    for (int m = 1; m < R; ++m)
                    {
                        const uint32_t em = e * m;
                        a[q][m] = ctw<INV>(a[q][m], cmul(c.thi[em >> 7], c.tlo[em & (kManyTwLo - 1)]));
                    }
                }
            }
            Dft<R, INV>::run(a[q]);
            const uint32_t ob = j * c.n + (b - k) * R + k;
            if (last)
            {
#pragma unroll
                for (int m = 0; m < R; ++m)
                {
                    bst<kAuxNt>(c.out, (ob + m * st.p) * 8u, 0, a[q][m]);
                }
            }
            else
            {
                lds_scatter<R>(a[q], c.ld, ob, st.p);
            }
        }
    }
    if (!last)
    {
        __syncthreads();
    }
}

// The loads of one tile into registers: point f = tid + k * threads, and (tid < 256) one twiddle entry.
__device__ __forceinline__ void fetch_tile(const ManyArgs& a, uint32_t t, int tid, cf (&pf)[kPer], cf& ptw)
{
    const ManyTile tile = read(constant(a.tiles) + t);
    const uint32_t n = constant(a.sizes)[tile.size_id].n;
    const uint64_t tw_offset = constant(a.sizes)[tile.size_id].tw_offset;
    const uint32_t points = tile.count * n;
    const __amdgpu_buffer_rsrc_t src = make_rsrc(a.in + tile.base, points * 8u);
#pragma unroll
    for (int k = 0; k < kPer; ++k)
    {
        const uint32_t f = tid + k * kThreads;
        pf[k] = f < points ? bld<kAuxNt>(src, f * 8u, 0) : cf{0.0f, 0.0f};
    }
    if (tid < kManyTwEntries)
    {
        const cfloat w = a.tw[tw_offset + tid];
        ptw = cf{w.x, w.y};
    }
}

template <bool INV> __global__ __launch_bounds__(kThreads) void k_many(ManyArgs a)
{
    // static LDS (the tile and the twiddle table, 141,440 bytes): a compile-time base, so the slot arithmetic needs no
    // extra add for it
    __shared__ cf lds_tile[kLdsPoints + kManyTwEntries];
    cf* ld = lds_tile;
    cf* ltw = ld + kLdsPoints;
    const int tid = threadIdx.x;
    const uint32_t t0 = constant(a.wg_begin)[blockIdx.x];
    const uint32_t t1 = constant(a.wg_begin)[blockIdx.x + 1];
    // LDS slot of point tid + k * threads: pad16 of it is o0 + k * (threads + threads / 16)
    const int o0 = pad16(tid);
    constexpr int kStep = kThreads + kThreads / 16;
    cf pf[kPer];
    cf ptw = cf{1.0f, 0.0f};
    if (t0 < t1)
    {
        fetch_tile(a, t0, tid, pf, ptw);
    }
#pragma unroll 1
    for (uint32_t t = t0; t < t1; ++t)
    {
        const ManyTile tile = read(constant(a.tiles) + t);
        cptr<ManySize> sz = constant(a.sizes) + tile.size_id;
        const uint32_t n = sz->n;
        const uint32_t points = tile.count * n;
        __syncthreads(); // the previous tile's last LDS reads are done
#pragma unroll
        for (int k = 0; k < kPer; ++k)
        {
            const uint32_t f = tid + k * kThreads;
            if (f < points)
            {
                ld[o0 + k * kStep] = pf[k];
            }
        }
        if (tid < kManyTwEntries)
        {
            ltw[tid] = ptw;
        }
        __syncthreads();
        if (t + 1 < t1)
        {
            fetch_tile(a, t + 1, tid, pf, ptw);
        }
        const TileCtx c{make_rsrc(a.out + tile.base, points * 8u), ld, ltw, ltw + kManyTwLo, n, tile.count};
        const uint32_t ns = sz->nstages;
#pragma unroll 1
        for (uint32_t s = 0; s < ns; ++s)
        {
            const ManyStage st = read(&sz->stage[s]);
            const bool last = s + 1 == ns;
            switch (st.radix)
            {
                case 1: many_stage<1, INV>(c, st, last, tid); break;
                case 2: many_stage<2, INV>(c, st, last, tid); break;
                case 3: many_stage<3, INV>(c, st, last, tid); break;
                case 4: many_stage<4, INV>(c, st, last, tid); break;
                case 5: many_stage<5, INV>(c, st, last, tid); break;
                case 7: many_stage<7, INV>(c, st, last, tid); break;
                case 8: many_stage<8, INV>(c, st, last, tid); break;
                case 6: many_stage<6, INV>(c, st, last, tid); break;
                case 9: many_stage<9, INV>(c, st, last, tid); break;
                case 10: many_stage<10, INV>(c, st, last, tid); break;
                case 12: many_stage<12, INV>(c, st, last, tid); break;
                default: many_stage<16, INV>(c, st, last, tid); break;
            }
        }
    }
}

} // namespace

hipError_t launch_many(const ManyArgs& a, unsigned workgroups, bool inverse, hipStream_t s)
{
    if (workgroups == 0)
    {
        return hipSuccess;
    }
    auto kern = inverse ? k_many<true> : k_many<false>;
    hipLaunchKernelGGL(kern, dim3(workgroups), dim3(kThreads), 0, s, a); // the tile is static LDS
    return hipGetLastError();
}

} // namespace crdfft
