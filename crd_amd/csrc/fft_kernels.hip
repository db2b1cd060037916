// HIP kernels of the plain FP32 batched FFT path for gfx950 (CDNA4, wave64) and their launchers.
//
//   k_rows     one or more contiguous N-point transforms per workgroup, N <= 16384: coalesced loads straight into
//              the Stockham core's slots, radix-E register stages with LDS exchanges, coalesced stores. One HBM
//              read and one write per point. (The reference's single-transform engines: Stockham, in-place-3,
//              standalone-hier, deep-split; engine/numerics/hesap-fft/kernels/fft_kernels.cpp:187-585, 1603-1670.)
//   k_cols     column transforms of a row-major n1 x ncols matrix, C adjacent columns per workgroup (C*8 bytes of
//              every row per load), optionally times the four-step twiddle W_N^{n2 k1} from two FP64-built tables
//              (the reference's four-step phase 1 and its factored twiddle, fft_kernels.cpp:2363-, fft.hpp:859-891),
//              and the element-major batch (fft.hpp:462-470, fft_kernels.cpp:593-784).
//   k_rowtr    row transforms with the four-step's transposed store out[k1 + n1 k2], R rows per workgroup so every
//              store is a run of R*8 bytes (the reference's phase 2 + scatter).
//   k_blue     the whole Bluestein transform (chirp, FFT_m, product with FFT_m(b), IFFT_m, chirp) for m <= 16384 in
//              the k_rows layout: one HBM read and one write per point of x (bluestein.hpp:82-135).
//   k_roots*   FP64 table builders (the reference builds W_n in f64 and narrows, fft.hpp:251-260).
//   k_ref_*    the radix-2 reference transform (fft.hpp:417-460), a differential check only.
//
// Every kernel guards its own range (t < count, column < ncols) and runs its barriers unconditionally, so a
// partial last workgroup is safe.
#include "buffer.hpp"
#include "fft_core.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace crdfft
{
using namespace dev;

namespace
{

// Workgroups are dispatched round-robin over the 8 XCDs (block b on XCD b % 8). Renumber them so that each XCD
// owns one contiguous range of logical blocks, in dispatch order: neighbouring tiles (which share 128-byte lines
// when a tile row is 64 bytes) then meet in the same L2.
__device__ __forceinline__ uint64_t xcd_block()
{
    const uint32_t b = blockIdx.x;
    const uint32_t nb = gridDim.x;
    const uint32_t full = nb & ~7u;
    if (b < full)
    {
        return static_cast<uint64_t>(b & 7u) * (full >> 3) + (b >> 3);
    }
    return b;
}

constexpr int round_up(int v, int m) { return (v + m - 1) / m * m; }

template <bool NT> __device__ __forceinline__ cf ldg(const cf* p)
{
    if constexpr (NT)
    {
        const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
        return cf{v.x, v.y};
    }
    else
    {
        return *p;
    }
}

template <bool NT> __device__ __forceinline__ void stg(cf* p, cf v)
{
    if constexpr (NT)
    {
        f2v w;
        w.x = v.x;
        w.y = v.y;
        __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
    }
    else
    {
        *p = v;
    }
}

// v[m] = src[m * stride] for m < E, with or without the non-temporal hint (a uniform branch).
template <int E>
__device__ __forceinline__ void load_strided(cf (&v)[E], const cf* src, uint32_t stride, bool nt)
{
    if (nt)
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = ldg<true>(src + static_cast<uint32_t>(m) * stride);
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = ldg<false>(src + static_cast<uint32_t>(m) * stride);
        }
    }
}

template <int E>
__device__ __forceinline__ void store_strided(const cf (&v)[E], cf* dst, uint32_t stride, float scale, bool nt)
{
    if (nt)
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            stg<true>(dst + static_cast<uint32_t>(m) * stride, cf{v[m].x * scale, v[m].y * scale});
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            stg<false>(dst + static_cast<uint32_t>(m) * stride, cf{v[m].x * scale, v[m].y * scale});
        }
    }
}

// v[m] = element (off + m * stride) of the resource's array, for m < E (offsets in elements).
template <int E>
__device__ __forceinline__ void buf_load_strided(cf (&v)[E], __amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t stride,
                                                 bool nt)
{
    if (nt)
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = bld<kAuxNt>(r, off * 8u, static_cast<uint32_t>(m) * stride * 8u);
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = bld<0>(r, off * 8u, static_cast<uint32_t>(m) * stride * 8u);
        }
    }
}

template <int E>
__device__ __forceinline__ void buf_store_strided(const cf (&v)[E], __amdgpu_buffer_rsrc_t r, uint32_t off,
                                                  uint32_t stride, float scale, bool nt)
{
    if (nt)
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            bst<kAuxNt>(r, off * 8u, static_cast<uint32_t>(m) * stride * 8u, cf{v[m].x * scale, v[m].y * scale});
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            bst<0>(r, off * 8u, static_cast<uint32_t>(m) * stride * 8u, cf{v[m].x * scale, v[m].y * scale});
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// k_rows
// ---------------------------------------------------------------------------------------------------------------
template <int N, int E> struct RowCfg
{
    static constexpr int T = N / E;
    static constexpr int M = T >= 256 ? 1 : 256 / T;
    static constexpr int threads = T * M;
    static constexpr int LF = lds_floats(N);
    // per-transform region: re then im, rounded to a 64-bank row and offset so that the transforms sharing a
    // wave land on different banks
    static constexpr int RS = (N == E) ? 0 : round_up(2 * LF, 64) + (T < 64 ? T : 0);
    static constexpr int lds_bytes = RS * M * 4;
};

template <int N, int E, bool INV>
__global__ __launch_bounds__((RowCfg<N, E>::threads)) void k_rows(RowArgs a)
{
    using C = RowCfg<N, E>;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int local = tid / C::T;
    const int i = tid - local * C::T;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * C::M + local;
    const bool active = t < a.count;
    cf v[E];
    if (active)
    {
        load_strided<E>(v, reinterpret_cast<const cf*>(a.in) + t * a.idist + i, static_cast<uint32_t>(C::T),
                        (a.hints & kHintNtLoad) != 0);
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = cf{0.0f, 0.0f};
        }
    }
    const LdsView<1> view{lds + local * C::RS, C::LF};
    fft_core<N, E, INV, 1>(v, i, view, reinterpret_cast<const cf*>(a.tw));
    if (active)
    {
        store_strided<E>(v, reinterpret_cast<cf*>(a.out) + t * a.odist + i, static_cast<uint32_t>(C::T), a.scale,
                         (a.hints & kHintNtStore) != 0);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// k_cols
// ---------------------------------------------------------------------------------------------------------------
template <int N1, int E, int CC> struct ColCfg
{
    static constexpr int T = N1 / E;
    static constexpr int threads = T * CC;
    static constexpr int LF = lds_floats(N1);
    static constexpr int IM = LF * CC + 32; // im array after the re array
    static constexpr int lds_bytes = (N1 == E) ? 0 : (IM + LF * CC) * 4;
};

// PRE (C2R, INV): the column values are formed from the h + 1 input bins as pre_scale * Z[j] (c2r_z) on load, which
// replaces the separate pre-rotation pass; X[h-j] of a column lies in the mirrored column, read in the same
// coalesced pattern backwards.
template <int N1, int E, int CC, bool INV, bool PRE = false>
__global__ __launch_bounds__((ColCfg<N1, E, CC>::threads)) void k_cols(ColArgs a)
{
    using C = ColCfg<N1, E, CC>;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int c = tid % CC;
    const int i = tid / CC;
    const uint64_t groups = (a.ncols + CC - 1) / CC;
    const uint64_t lb = xcd_block();
    const uint64_t t = lb / groups;
    const uint64_t col = (lb - t * groups) * CC + c;
    const bool active = t < a.count && col < a.ncols;
    cf v[E];
    const uint32_t ncols = static_cast<uint32_t>(a.ncols); // n1 * ncols points per matrix fit 32 bits
    const uint32_t off = static_cast<uint32_t>(col) + static_cast<uint32_t>(i) * ncols;
    if (active)
    {
        if constexpr (PRE)
        {
            const cf* x = reinterpret_cast<const cf*>(a.in) + t * a.idist;
            const cf* w = reinterpret_cast<const cf*>(a.wreal);
            const uint32_t h = static_cast<uint32_t>(N1) * ncols;
            // W_n^j for j = col + ncols (i + T m) as W_n^{col} W_n^{ncols i} W_n^{ncols T m}: two table reads per
            // thread and one workgroup-uniform read per value instead of a gather of the h-entry table
            const cf wb = cmul(w[col], w[static_cast<uint32_t>(i) * ncols]);
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                const uint32_t j = off + static_cast<uint32_t>(m) * (C::T * ncols);
                const cf wj = m == 0 ? wb : cmul(wb, w[static_cast<uint32_t>(m) * (C::T * ncols)]);
                v[m] = c2r_z(x[j], x[h - j], wj, a.pre_scale);
            }
        }
        else
        {
            load_strided<E>(v, reinterpret_cast<const cf*>(a.in) + t * a.idist + off, C::T * ncols,
                            (a.hints & kHintNtLoad) != 0);
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = cf{0.0f, 0.0f};
        }
    }
    const LdsView<CC> view{lds + c, C::IM};
    fft_core<N1, E, INV, CC>(v, i, view, reinterpret_cast<const cf*>(a.tw));
    if (!active)
    {
        return;
    }
    if (a.tfull != nullptr)
    {
        // W_N^{col k1} from the full table at [k1 * ncols + col]: one coalesced load per value
        const cf* tf = reinterpret_cast<const cf*>(a.tfull) + off;
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = ctw<INV>(v[m], tf[static_cast<uint32_t>(m) * (C::T * ncols)]);
        }
    }
    else if (a.thi != nullptr)
    {
        // W_N^{col (i + T m)} = W_N^{col i} * W_N^{col T m}: one factored lookup per thread, then one coalesced
        // table load per value (lanes of adjacent columns read adjacent entries)
        const cf* thi = reinterpret_cast<const cf*>(a.thi);
        const cf* tlo = reinterpret_cast<const cf*>(a.tlo);
        const cf* tstep = reinterpret_cast<const cf*>(a.tstep) + col;
        const uint32_t lmask = (1u << a.hshift) - 1;
        const uint32_t e0 = static_cast<uint32_t>(col) * static_cast<uint32_t>(i);
        const cf w0 = cmul(thi[e0 >> a.hshift], tlo[e0 & lmask]);
        v[0] = ctw<INV>(v[0], w0);
#pragma unroll
        for (int m = 1; m < E; ++m)
        {
            const cf w = cmul(w0, tstep[static_cast<uint32_t>(m) * ncols]);
            v[m] = ctw<INV>(v[m], w);
        }
    }
    // tiled intermediate: row slot m of column group g is the contiguous run [g*T*CC + (m*ncols/CC)*T*CC, +T*CC),
    // point (i, c) at tid
    const uint32_t soff = (a.hints & kHintTiled) != 0
                              ? static_cast<uint32_t>(lb - t * groups) * static_cast<uint32_t>(C::T * CC) +
                                    static_cast<uint32_t>(tid)
                              : off;
    store_strided<E>(v, reinterpret_cast<cf*>(a.out) + t * a.odist + soff, C::T * ncols, 1.0f,
                     (a.hints & kHintNtStore) != 0);
}

// ---------------------------------------------------------------------------------------------------------------
// k_cols_em: one pass of the element-major four-step (strided columns, separate input and output row strides)
// ---------------------------------------------------------------------------------------------------------------
// Element i = i1*n2 + i2 of transform t sits at (i1*n2 + i2)*B + t. Pass 1 transforms over i1 for every (i2, t):
// matrix i2 holds rows i1 at stride n2*B and the B transforms as adjacent columns, and its outputs k1 are twiddled
// by W_n^{i2 k1}. Pass 2 transforms over i2 for every (k1, t): matrix k1 is contiguous (rows i2 at stride B), and
// output k2 goes to element k1 + n1*k2, i.e. row stride n1*B from base k1*B. A workgroup owns CC adjacent columns
// of one matrix; its loads and stores are runs of CC*8 bytes.
template <int N1, int E, int CC, bool INV, bool TW>
__global__ __launch_bounds__((ColCfg<N1, E, CC>::threads)) void k_cols_em(EmArgs a)
{
    using C = ColCfg<N1, E, CC>;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int c = tid % CC;
    const int i = tid / CC;
    const uint64_t groups = (a.ncols + CC - 1) / CC;
    const uint64_t lb = blockIdx.x;
    const uint64_t t = lb / groups;
    const uint64_t col = (lb - t * groups) * CC + c;
    const bool active = t < a.count && col < a.ncols;
    cf v[E];
    if (active)
    {
        const cf* src = reinterpret_cast<const cf*>(a.in) + t * a.idist + col + static_cast<uint64_t>(i) * a.in_ld;
        const uint64_t step = static_cast<uint64_t>(C::T) * a.in_ld;
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = src[static_cast<uint64_t>(m) * step];
        }
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = cf{0.0f, 0.0f};
        }
    }
    const LdsView<CC> view{lds + c, C::IM};
    fft_core<N1, E, INV, CC>(v, i, view, reinterpret_cast<const cf*>(a.tw));
    if (!active)
    {
        return;
    }
    if constexpr (TW)
    {
        const cf* thi = reinterpret_cast<const cf*>(a.thi);
        const cf* tlo = reinterpret_cast<const cf*>(a.tlo);
        const uint32_t lmask = (1u << a.hshift) - 1;
        const uint32_t tt = static_cast<uint32_t>(t + a.t0); // (t0 + t) * k1 < n <= 2^24
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            const uint32_t e = tt * static_cast<uint32_t>(i + C::T * m);
            v[m] = ctw<INV>(v[m], cmul(thi[e >> a.hshift], tlo[e & lmask]));
        }
    }
    cf* dst = reinterpret_cast<cf*>(a.out) + t * a.odist + col + static_cast<uint64_t>(i) * a.out_ld;
    const uint64_t ostep = static_cast<uint64_t>(C::T) * a.out_ld;
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        dst[static_cast<uint64_t>(m) * ostep] = v[m];
    }
}

// ---------------------------------------------------------------------------------------------------------------
// k_cols_pf: the column pass as a persistent launch that loads the next tile while it transforms this one
// ---------------------------------------------------------------------------------------------------------------
// For long columns (n1 >= 1024) one k_cols workgroup fills a CU's LDS, so nothing overlaps its load phase with its
// compute phase and the CU alternates between waiting on HBM and running LDS stages. Here the grid is the resident
// capacity; workgroup b runs tiles (matrix t, column group g) of one contiguous per-XCD range in steps of the
// XCD's workgroup count, and issues the loads of its next tile before the FFT of the current one. The stage
// twiddles sit in LDS (global loads complete in order: a table read from memory would wait for the prefetched tile).
template <int N1, int E, int CC> struct ColPfCfg
{
    using C = ColCfg<N1, E, CC>;
    static constexpr int tw_entries = static_cast<int>(Geometry<N1, E>::twiddle_entries());
    static constexpr int lds_bytes = C::lds_bytes + tw_entries * 8;
};

// TWM: the four-step twiddle of the tile, 1 = full table (a.tfull), 2 = factored (a.thi, a.tlo, a.tstep)
template <int N1, int E, int CC, int TWM, bool INV>
__global__ __launch_bounds__((ColCfg<N1, E, CC>::threads)) void k_cols_pf(ColArgs a)
{
    using C = ColCfg<N1, E, CC>;
    using P = ColPfCfg<N1, E, CC>;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int c = tid % CC;
    const int i = tid / CC;
    cf* tw = reinterpret_cast<cf*>(lds + C::lds_bytes / 4);
    for (int j = tid; j < P::tw_entries; j += C::threads)
    {
        tw[j] = reinterpret_cast<const cf*>(a.tw)[j];
    }
    __syncthreads();
    // tiles of this XCD: [lo, hi), taken in steps of the XCD's workgroup count (the launcher makes the grid a
    // multiple of 8; workgroup b runs on XCD b % 8)
    const uint32_t groups = static_cast<uint32_t>(a.ncols / CC);
    const uint64_t tiles = a.count * groups;
    const uint32_t xcd = blockIdx.x & 7u;
    const uint32_t per_xcd = gridDim.x >> 3;
    const uint64_t lo = tiles * xcd / 8;
    const uint64_t hi = tiles * (xcd + 1) / 8;
    const uint32_t ncols = static_cast<uint32_t>(a.ncols);
    const uint32_t stride = static_cast<uint32_t>(C::T) * ncols;
    const uint32_t ioff = static_cast<uint32_t>(c) + static_cast<uint32_t>(i) * ncols; // + g * CC
    const bool ntl = (a.hints & kHintNtLoad) != 0;
    const bool nts = (a.hints & kHintNtStore) != 0;
    const cf* in = reinterpret_cast<const cf*>(a.in);
    cf* out = reinterpret_cast<cf*>(a.out);
    const LdsView<CC> view{lds + c, C::IM};

    uint64_t tile = lo + (blockIdx.x >> 3);
    cf nx[E];
    if (tile < hi)
    {
        const uint64_t t = tile / groups;
        const uint32_t g = static_cast<uint32_t>(tile - t * groups);
        buf_load_strided<E>(nx, make_rsrc(in + t * a.idist), ioff + g * CC, stride, ntl);
    }
#pragma unroll 1
    for (; tile < hi; tile += per_xcd)
    {
        const uint64_t t = tile / groups;
        const uint32_t g = static_cast<uint32_t>(tile - t * groups);
        cf v[E];
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = nx[m];
        }
        const uint64_t next = tile + per_xcd;
        if (next < hi)
        {
            const uint64_t tn = next / groups;
            const uint32_t gn = static_cast<uint32_t>(next - tn * groups);
            buf_load_strided<E>(nx, make_rsrc(in + tn * a.idist), ioff + gn * CC, stride, ntl);
        }
        fft_core<N1, E, INV, CC>(v, i, view, tw);
        const uint32_t col = g * CC + c;
        const uint32_t off = ioff + g * CC;
        if constexpr (TWM == 1)
        {
            const __amdgpu_buffer_rsrc_t tf = make_rsrc(a.tfull);
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                v[m] = ctw<INV>(v[m], bld<0>(tf, off * 8u, static_cast<uint32_t>(m) * stride * 8u));
            }
        }
        else
        {
            const cf* thi = reinterpret_cast<const cf*>(a.thi);
            const cf* tlo = reinterpret_cast<const cf*>(a.tlo);
            const __amdgpu_buffer_rsrc_t ts = make_rsrc(a.tstep);
            const uint32_t lmask = (1u << a.hshift) - 1;
            const uint32_t e0 = col * static_cast<uint32_t>(i);
            const cf w0 = cmul(thi[e0 >> a.hshift], tlo[e0 & lmask]);
            v[0] = ctw<INV>(v[0], w0);
#pragma unroll
            for (int m = 1; m < E; ++m)
            {
                const cf w = cmul(w0, bld<0>(ts, col * 8u, static_cast<uint32_t>(m) * ncols * 8u));
                v[m] = ctw<INV>(v[m], w);
            }
        }
        buf_store_strided<E>(v, make_rsrc(out + t * a.odist), off, stride, 1.0f, nts);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// k_rowtr
// ---------------------------------------------------------------------------------------------------------------
template <int N2, int E, int RR> struct RowTrCfg
{
    static constexpr int T = N2 / E;
    static constexpr int threads = T * RR;
    static constexpr int LF = lds_floats(N2);
    static constexpr int RS = (N2 == E) ? 0 : round_up(2 * LF, 64) + (T < 64 ? T : 0);
    static constexpr int fft_floats = RS * RR;
    static constexpr int TP = RR + 1;              // transpose row pitch (odd: conflict-free column writes)
    static constexpr int tr_floats = 2 * N2 * TP;  // re then im
    static constexpr int lds_bytes = (fft_floats > tr_floats ? fft_floats : tr_floats) * 4;
};

template <int N2, int E, int RR, bool INV>
__global__ __launch_bounds__((RowTrCfg<N2, E, RR>::threads)) void k_rowtr(RowTrArgs a)
{
    using C = RowTrCfg<N2, E, RR>;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int r = tid / C::T;
    const int i = tid - r * C::T;
    const uint64_t groups = a.n1 / RR;
    const uint64_t lb = xcd_block();
    const uint64_t t = lb / groups;
    const uint64_t g = lb - t * groups;
    const bool active = t < a.count;
    cf v[E];
    if (active)
    {
        // tiled intermediate: rows g*RR .. g*RR+RR-1 are one contiguous block, slot m of every row a run of RR*T
        // points (row r, column i at tid)
        const bool tiled = (a.hints & kHintTiled) != 0;
        load_strided<E>(v,
                        reinterpret_cast<const cf*>(a.in) + t * a.idist +
                            (tiled ? g * RR * N2 + static_cast<uint64_t>(tid) : (g * RR + r) * N2 + i),
                        static_cast<uint32_t>(tiled ? C::threads : C::T), (a.hints & kHintNtLoad) != 0);
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = cf{0.0f, 0.0f};
        }
    }
    const LdsView<1> view{lds + r * C::RS, C::LF};
    fft_core<N2, E, INV, 1>(v, i, view, reinterpret_cast<const cf*>(a.tw));
    // fft_core's last LDS read is behind a barrier (or it used no LDS): the buffer is free for the transpose
    if constexpr (N2 == E)
    {
        __syncthreads();
    }
    float* tre = lds;
    float* tim = lds + N2 * C::TP;
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const int k2 = i + C::T * m;
        tre[k2 * C::TP + r] = v[m].x;
        tim[k2 * C::TP + r] = v[m].y;
    }
    __syncthreads();
    if (!active)
    {
        return;
    }
    const int r2 = tid % RR;
    const int j2 = tid / RR;
    cf w[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const int k2 = j2 + C::T * m;
        w[m] = cf{tre[k2 * C::TP + r2], tim[k2 * C::TP + r2]};
    }
    const uint32_t n1 = static_cast<uint32_t>(a.n1);
    store_strided<E>(w, reinterpret_cast<cf*>(a.out) + t * a.odist + g * RR + r2 + static_cast<uint32_t>(j2) * n1,
                     C::T * n1, a.scale, (a.hints & kHintNtStore) != 0);
}

// ---------------------------------------------------------------------------------------------------------------
// k_rowtr_r2c: the row pass of a real transform's half-length four-step with the Hermitian recombine fused in
// ---------------------------------------------------------------------------------------------------------------
// The rfft of n reals is the h = n/2 point FFT Z of z[j] = x[2j] + i x[2j+1] followed by
//   X[k] = E + W_n^k O,  E = (Z[k] + conj Z[h-k]) / 2,  O = (Z[k] - conj Z[h-k]) / (2i),  X[h] = Re Z0 - Im Z0
// (real_fft.hpp:41-68). With h = n1 * n2 and Z[k1 + n1 k2] the row FFT of intermediate row k1, the partner of
// (k1, k2) is (n1 - k1, n2 - 1 - k2) for k1 > 0 and (0, -k2 mod n2) for k1 = 0; row n1/2 is its own mirror. So a
// workgroup takes Q = RR/2 rows a = pQ .. pQ+Q-1 of the lower half and their mirrors n1 - a (row n1/2 in place of
// the mirror of row 0, which is row 0 itself): every partner is in the tile, the recombine runs on the transposed
// LDS staging, and each X[k] is stored once, in runs of Q elements. This replaces the separate recombine pass (one
// read and one write of the spectrum).
template <int N2, int E, int RR>
__global__ __launch_bounds__((RowTrCfg<N2, E, RR>::threads)) void k_rowtr_r2c(RowTrArgs a)
{
    using C = RowTrCfg<N2, E, RR>;
    constexpr int Q = RR / 2;
    static_assert(RR % 2 == 0, "rows come in mirror pairs");
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int r = tid / C::T;
    const int i = tid - r * C::T;
    const uint64_t groups = a.n1 / RR;
    const uint64_t lb = xcd_block();
    const uint64_t t = lb / groups;
    const uint64_t p = lb - t * groups;
    const bool active = t < a.count;
    const uint32_t n1 = static_cast<uint32_t>(a.n1);
    // the intermediate row of tile slot q: the lower rows, then their mirrors (row n1/2 for the mirror of row 0)
    auto row_of = [&](int q) -> uint32_t {
        const uint32_t lo = static_cast<uint32_t>(p) * Q + static_cast<uint32_t>(q < Q ? q : q - Q);
        return q < Q ? lo : (lo == 0 ? n1 / 2 : n1 - lo);
    };
    cf v[E];
    if (active)
    {
        // plain loads and stores whatever the plan's hints: the intermediate is read back from the caches, and the
        // output goes out in runs of Q elements, which non-temporal stores write as partial lines (measured 1.6x
        // slower at n = 2^16)
        load_strided<E>(v, reinterpret_cast<const cf*>(a.in) + t * a.idist + static_cast<uint64_t>(row_of(r)) * N2 + i,
                        static_cast<uint32_t>(C::T), false);
    }
    else
    {
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            v[m] = cf{0.0f, 0.0f};
        }
    }
    const LdsView<1> view{lds + r * C::RS, C::LF};
    fft_core<N2, E, false, 1>(v, i, view, reinterpret_cast<const cf*>(a.tw));
    if constexpr (N2 == E)
    {
        __syncthreads();
    }
    float* tre = lds;
    float* tim = lds + N2 * C::TP;
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const int k2 = i + C::T * m;
        tre[k2 * C::TP + r] = v[m].x;
        tim[k2 * C::TP + r] = v[m].y;
    }
    __syncthreads();
    if (!active)
    {
        return;
    }
    const int r2 = tid % RR;
    const int j2 = tid / RR;
    const uint32_t k1 = row_of(r2);
    // the partner's tile slot and how its column follows from k2
    const bool self = (k1 == 0 || 2 * k1 == n1);
    const int rp = self ? r2 : (r2 < Q ? r2 + Q : r2 - Q);
    const uint64_t h = static_cast<uint64_t>(n1) * N2;
    // W_n^k for k = k1 + n1 (j2 + T m) as W_n^{k1} W_n^{n1 j2} W_n^{n1 T m}: two table reads per thread and one
    // workgroup-uniform (scalar) read per value, instead of a gather of the h-entry table per output
    const cf* w = reinterpret_cast<const cf*>(a.wreal);
    const cf wb = cmul(w[k1], w[n1 * static_cast<uint32_t>(j2)]);
    const __amdgpu_buffer_rsrc_t out = make_rsrc(reinterpret_cast<cf*>(a.out) + t * a.odist);
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const int k2 = j2 + C::T * m;
        const int k2p = k1 == 0 ? ((N2 - k2) & (N2 - 1)) : (N2 - 1 - k2);
        const cf za = cf{tre[k2 * C::TP + r2], tim[k2 * C::TP + r2]};
        const cf zb = cf{tre[k2p * C::TP + rp], tim[k2p * C::TP + rp]};
        const cf e = {(za.x + zb.x) * 0.5f, (za.y - zb.y) * 0.5f};
        const cf o = {(za.y + zb.y) * 0.5f, (zb.x - za.x) * 0.5f};
        const uint32_t k = k1 + n1 * static_cast<uint32_t>(k2);
        const cf wk = m == 0 ? wb : cmul(wb, w[n1 * static_cast<uint32_t>(C::T * m)]);
        bst<0>(out, k * 8u, 0, cadd(e, cmul(o, wk)));
        if (m == 0 && k == 0) // k >= n1 * T for m >= 1
        {
            bst<0>(out, static_cast<uint32_t>(h) * 8u, 0, cf{za.x - za.y, 0.0f});
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// k_fused: both four-step passes of one transform in one workgroup
// ---------------------------------------------------------------------------------------------------------------
// For n = N1 * N2 up to 2^17 the four-step intermediate of one transform (n * 8 bytes) is small enough to keep per
// workgroup. A persistent grid of G workgroups, each owning one n-point slot of the workspace, runs transforms
// t = block, block + G, ...:
//   phase A  N2/CC tiles of CC adjacent columns: column FFT of N1 points, twiddle W_n^{col k1} (full table), store
//            into the slot at [k1 * N2 + col];
//   phase B  after a barrier, N1/RR tiles of RR rows of the slot: row FFT of N2 points, transposed store
//            out[k1 + N1 k2] through LDS in runs of RR * 8 bytes.
// The slot is written and read back by the same CU within one transform (G slots = 128-256 MiB in all), so it is
// served by the L2 and the Infinity Cache and HBM sees one read and one write per point, where the two-kernel
// four-step reads and writes every point twice. Each column tile's loads are issued before the previous tile is
// transformed (and the next transform's first tile during the last column tile), so the HBM reads of the input
// overlap the LDS stages of the workgroup's own compute.
template <int N1, int E1, int N2, int E2, int TH> struct FusedCfg
{
    static constexpr int T1 = N1 / E1;
    static constexpr int T2 = N2 / E2;
    static constexpr int CC = TH / T1; // columns per phase-A tile
    static constexpr int RR = TH / T2; // rows per phase-B tile
    using A = ColCfg<N1, E1, CC>;
    using B = RowTrCfg<N2, E2, RR>;
    static_assert(A::threads == TH && B::threads == TH, "both phases use every thread");
    static_assert(N2 % CC == 0 && N1 % RR == 0, "whole tiles");
    static constexpr int data_bytes = A::lds_bytes > B::lds_bytes ? A::lds_bytes : B::lds_bytes;
    // the stage twiddles of both lengths follow the data in LDS: global loads complete in order, so a table read
    // from memory would make every stage wait for the prefetched tile as well
    static constexpr int tw1_entries = static_cast<int>(Geometry<N1, E1>::twiddle_entries());
    static constexpr int tw2_entries = static_cast<int>(Geometry<N2, E2>::twiddle_entries());
    static constexpr int lds_bytes = data_bytes + (tw1_entries + tw2_entries) * 8;
};

template <int N1, int E1, int N2, int E2, int TH, bool INV>
__global__ __launch_bounds__(TH) void k_fused(FusedArgs a)
{
    using F = FusedCfg<N1, E1, N2, E2, TH>;
    constexpr int CC = F::CC;
    constexpr int RR = F::RR;
    constexpr int T1 = F::T1;
    constexpr int T2 = F::T2;
    constexpr uint32_t kN = static_cast<uint32_t>(N1) * N2;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int c = tid % CC; // phase A: column within the tile, slots i1 + T1 m
    const int i1 = tid / CC;
    const int r = tid / T2; // phase B: row within the tile, slots i2 + T2 m
    const int i2 = tid - r * T2;
    const int r2 = tid % RR; // phase B store: row r2, elements j2 + T2 m
    const int j2 = tid / RR;
    const LdsView<CC> va{lds + c, F::A::IM};
    const LdsView<1> vb{lds + r * F::B::RS, F::B::LF};
    float* tre = lds;
    float* tim = lds + N2 * F::B::TP;
    const cf* in = reinterpret_cast<const cf*>(a.in);
    cf* out = reinterpret_cast<cf*>(a.out);
    const __amdgpu_buffer_rsrc_t slot = make_rsrc(reinterpret_cast<cf*>(a.ws) + static_cast<uint64_t>(blockIdx.x) * kN);
    const __amdgpu_buffer_rsrc_t tfull = make_rsrc(a.tfull);
    cf* tw1 = reinterpret_cast<cf*>(lds + F::data_bytes / 4);
    cf* tw2 = tw1 + F::tw1_entries;
    for (int j = tid; j < F::tw1_entries; j += TH)
    {
        tw1[j] = reinterpret_cast<const cf*>(a.tw1)[j];
    }
    for (int j = tid; j < F::tw2_entries; j += TH)
    {
        tw2[j] = reinterpret_cast<const cf*>(a.tw2)[j];
    }
    __syncthreads();
    const bool ntl = (a.hints & kHintNtLoad) != 0;
    const bool nts = (a.hints & kHintNtStore) != 0;
    const uint32_t a_off = static_cast<uint32_t>(c) + static_cast<uint32_t>(i1) * N2; // + tile * CC
    constexpr uint32_t a_stride = static_cast<uint32_t>(T1) * N2;
    const uint32_t b_off = static_cast<uint32_t>(r) * N2 + i2;                 // + tile * RR * N2
    const uint32_t s_off = static_cast<uint32_t>(r2) + static_cast<uint32_t>(j2) * N1; // + tile * RR
    constexpr uint32_t s_stride = static_cast<uint32_t>(T2) * N1;

    uint64_t t = blockIdx.x;
    cf nx[E1];
    if (t < a.count)
    {
        buf_load_strided<E1>(nx, make_rsrc(in + t * a.idist), a_off, a_stride, ntl);
    }
#pragma unroll 1
    for (; t < a.count; t += gridDim.x)
    {
        const __amdgpu_buffer_rsrc_t src = make_rsrc(in + t * a.idist);
        const __amdgpu_buffer_rsrc_t dst = make_rsrc(out + t * a.odist);
#pragma unroll 1
        for (int ct = 0; ct < N2 / CC; ++ct)
        {
            cf v[E1];
#pragma unroll
            for (int m = 0; m < E1; ++m)
            {
                v[m] = nx[m];
            }
            if (ct + 1 < N2 / CC)
            {
                buf_load_strided<E1>(nx, src, a_off + (ct + 1) * CC, a_stride, ntl);
            }
            else if (t + gridDim.x < a.count)
            {
                buf_load_strided<E1>(nx, make_rsrc(in + (t + gridDim.x) * a.idist), a_off, a_stride, ntl);
            }
            fft_core<N1, E1, INV, CC>(v, i1, va, tw1);
            const uint32_t off = a_off + static_cast<uint32_t>(ct) * CC;
#pragma unroll
            for (int m = 0; m < E1; ++m)
            {
                v[m] = ctw<INV>(v[m], bld<0>(tfull, off * 8u, static_cast<uint32_t>(m) * a_stride * 8u));
            }
            buf_store_strided<E1>(v, slot, off, a_stride, 1.0f, false);
        }
        __syncthreads(); // the slot is complete (and phase A's last LDS reads are done)
#pragma unroll 1
        for (int rt = 0; rt < N1 / RR; ++rt)
        {
            cf v[E2];
            buf_load_strided<E2>(v, slot, b_off + static_cast<uint32_t>(rt * RR) * N2, static_cast<uint32_t>(T2),
                                 false);
            fft_core<N2, E2, INV, 1>(v, i2, vb, tw2);
            if constexpr (N2 == E2)
            {
                __syncthreads();
            }
#pragma unroll
            for (int m = 0; m < E2; ++m)
            {
                const int k2 = i2 + T2 * m;
                tre[k2 * F::B::TP + r] = v[m].x;
                tim[k2 * F::B::TP + r] = v[m].y;
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E2; ++m)
            {
                const int k2 = j2 + T2 * m;
                v[m] = cf{tre[k2 * F::B::TP + r2], tim[k2 * F::B::TP + r2]};
            }
            buf_store_strided<E2>(v, dst, s_off + static_cast<uint32_t>(rt * RR), s_stride, a.scale, nts);
            __syncthreads(); // LDS (and, after the last tile, the slot) free for the next tile or transform
        }
    }
}

// ---------------------------------------------------------------------------------------------------------------
// k_blue: the Bluestein transform of one sequence per workgroup slot, m = N <= 16384 (the k_rows layout)
// ---------------------------------------------------------------------------------------------------------------
// The five passes of the chirp-z route (bluestein.hpp:82-135) with the m-point intermediate kept in registers and
// LDS: the only HBM traffic is x in and X out (n points each) plus the L2-resident c and FFT_m(b) tables. The
// zero padding and the k < n store guard are the buffer resources' range checks: reads at or past n return 0 and
// writes there are dropped, as are all accesses of an inactive slot (num_records 0).
template <int N, int E, bool INV>
__global__ __launch_bounds__((RowCfg<N, E>::threads)) void k_blue(BlueArgs a)
{
    using C = RowCfg<N, E>;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int local = tid / C::T;
    const int i = tid - local * C::T;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * C::M + local;
    const bool active = t < a.count;
    const uint32_t bytes = static_cast<uint32_t>(a.n * sizeof(cf));
    const __amdgpu_buffer_rsrc_t rx = make_rsrc(a.in + (active ? t * a.idist : 0), active ? bytes : 0u);
    const __amdgpu_buffer_rsrc_t rc = make_rsrc(a.chirp, bytes);
    const __amdgpu_buffer_rsrc_t rb = make_rsrc(a.bhat, static_cast<uint32_t>(N * sizeof(cf)));
    const uint32_t o0 = static_cast<uint32_t>(i) * sizeof(cf);
    constexpr uint32_t kStep = C::T * sizeof(cf);
    cf v[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        cf x = bld<0>(rx, o0, m * kStep);
        if constexpr (INV)
        {
            x.y = -x.y; // IDFT(x) = conj(DFT(conj x)) / n
        }
        v[m] = cmulc(x, bld<0>(rc, o0, m * kStep));
    }
    const LdsView<1> view{lds + local * C::RS, C::LF};
    const cf* tw = reinterpret_cast<const cf*>(a.tw);
    fft_core<N, E, false, 1>(v, i, view, tw);
    // the table offsets pass through an empty asm after each transform, so the compiler cannot issue the table
    // loads early and hold them in registers across the transform (at m = 16384 that spills)
    uint32_t o1 = o0;
    asm volatile("" : "+v"(o1));
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        v[m] = cmul(v[m], bld<0>(rb, o1, m * kStep));
    }
    // fft_core ends every LDS use behind a barrier, so the second transform may reuse the region at once; its lane
    // index is laundered too, so that it recomputes its LDS addresses instead of keeping the first one's live
    int i2 = i;
    asm volatile("" : "+v"(i2));
    fft_core<N, E, true, 1>(v, i2, view, tw);
    uint32_t o2 = o0;
    asm volatile("" : "+v"(o2));
    const __amdgpu_buffer_rsrc_t ry = make_rsrc(a.out + (active ? t * a.odist : 0), active ? bytes : 0u);
    const float cm = 1.0f / static_cast<float>(N);
    const float cn = 1.0f / static_cast<float>(a.n);
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        cf y = cmulc(v[m], bld<0>(rc, o2, m * kStep));
        y.x *= cm;
        y.y *= cm;
        if constexpr (INV)
        {
            y.x *= cn;
            y.y = -y.y * cn;
        }
        bst<0>(ry, o2, m * kStep, y);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// FP64 table builders
// ---------------------------------------------------------------------------------------------------------------
// W_d^e as floats from FP64: the angle is reduced to (-d/2, d/2] before sincospi, so it is exact in units of pi.
__device__ __forceinline__ cf root_fp64(uint64_t e, uint64_t d)
{
    e %= d;
    int64_t s = static_cast<int64_t>(e);
    if (2 * e > d)
    {
        s -= static_cast<int64_t>(d);
    }
    double sn, cs;
    sincospi(2.0 * static_cast<double>(s) / static_cast<double>(d), &sn, &cs);
    return cf{static_cast<float>(cs), static_cast<float>(-sn)};
}

__global__ void k_roots_lin(cf* out, uint64_t count, uint64_t mult, uint64_t d)
{
    const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < count)
    {
        out[j] = root_fp64((j * mult) % d, d);
    }
}

// Full four-step twiddle table: out[k1 * ncols + c] = W_n^{c k1} for k1 < rows.
__global__ void k_roots_full(cf* out, uint64_t rows, uint64_t ncols, uint64_t n)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < rows * ncols)
    {
        const uint64_t k1 = idx / ncols;
        const uint64_t c = idx - k1 * ncols;
        out[idx] = root_fp64(c * k1, n);
    }
}

// Step table of the column pass: out[m * ncols + n2] = W_n^{n2 * step * m}.
__global__ void k_roots_step(cf* out, uint64_t ncols, uint64_t e, uint64_t step, uint64_t n)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < ncols * e)
    {
        const uint64_t m = idx / ncols;
        const uint64_t n2 = idx - m * ncols;
        out[idx] = root_fp64((n2 * step % n) * m, n);
    }
}

// Stage table of stage P (entries k*R + j, k < P, j < R): W_{PR}^{j k}.
__global__ void k_roots_stage(cf* out, uint64_t p, uint64_t r)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < p * r)
    {
        const uint64_t k = idx / r;
        const uint64_t j = idx - k * r;
        out[idx] = root_fp64(j * k, p * r);
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Radix-2 reference and copies
// ---------------------------------------------------------------------------------------------------------------
__global__ void k_ref_bitrev(const cf* in, cf* out, uint64_t n, uint32_t lg, uint64_t count, uint64_t idist,
                             uint64_t odist)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n * count)
    {
        return;
    }
    const uint64_t t = idx / n;
    const uint64_t i = idx - t * n;
    const uint64_t r = lg == 0 ? 0 : (__brevll(i) >> (64 - lg));
    out[t * odist + r] = in[t * idist + i];
}

// The in-place bit reversal (fft.hpp:421-431 swaps data[i] and data[rev(i)] for i < rev(i)): the thread of the
// smaller index of each pair swaps it, so every element is read and written by one thread only.
__global__ void k_ref_bitrev_inplace(cf* data, uint64_t n, uint32_t lg, uint64_t count, uint64_t dist)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n * count)
    {
        return;
    }
    const uint64_t t = idx / n;
    const uint64_t i = idx - t * n;
    const uint64_t r = lg == 0 ? 0 : (__brevll(i) >> (64 - lg));
    if (i < r)
    {
        cf* base = data + t * dist;
        const cf a = base[i];
        const cf b = base[r];
        base[i] = b;
        base[r] = a;
    }
}

template <bool INV>
__global__ void k_ref_stage(cf* data, const cf* w, uint64_t n, uint64_t len, uint64_t count, uint64_t dist)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t half_n = n / 2;
    if (idx >= half_n * count)
    {
        return;
    }
    const uint64_t t = idx / half_n;
    const uint64_t bfly = idx - t * half_n;
    const uint64_t half = len / 2;
    const uint64_t grp = bfly / half;
    const uint64_t k = bfly - grp * half;
    cf* base = data + t * dist + grp * len;
    const cf tw = w[k * (n / len)];
    const cf b = ctw<INV>(base[k + half], tw);
    const cf av = base[k];
    base[k] = cadd(av, b);
    base[k + half] = csub(av, b);
}

// One 16-byte element per work-item over a grid that covers the buffer: every wave's loads are issued at once and
// the dispatcher keeps the CUs full (a grid-stride loop of one load per iteration leaves too few bytes in flight and
// measures ~20% below this on MI355X).
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ in, float4* __restrict__ out,
                                                     uint64_t count)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < count)
    {
        out[i] = in[i];
    }
}

__global__ void k_copy(const cf* in, cf* out, uint64_t n, uint64_t count, uint64_t idist, uint64_t odist)
{
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n * count)
    {
        return;
    }
    const uint64_t t = idx / n;
    const uint64_t i = idx - t * n;
    out[t * odist + i] = in[t * idist + i];
}

// ---------------------------------------------------------------------------------------------------------------
// instantiation tables
// ---------------------------------------------------------------------------------------------------------------
#define CRDFFT_ROW_LIST(X)                                                                                             \
    X(2, 2) X(4, 4) X(8, 8) X(16, 16) X(16, 4) X(32, 16) X(32, 32) X(32, 4) X(64, 16) X(64, 64) X(64, 4) X(64, 8)      \
        X(128, 16) X(128, 4) X(128, 8) X(256, 16) X(256, 4) X(256, 8) X(512, 16) X(512, 32) X(512, 8) X(1024, 16)      \
            X(1024, 32) X(1024, 8) X(2048, 16) X(2048, 32) X(2048, 8) X(4096, 16) X(4096, 32) X(4096, 64) X(4096, 8)   \
                X(8192, 16) X(8192, 32) X(8192, 8) X(16384, 16) X(16384, 32)

#define CRDFFT_COL_LIST(X)                                                                                             \
    X(2, 2, 64) X(4, 4, 64) X(8, 8, 64) X(16, 16, 64) X(16, 16, 16) X(32, 32, 64) X(32, 16, 16) X(64, 64, 64)          \
        X(64, 16, 16) X(128, 16, 16) X(128, 16, 8) X(128, 8, 16) X(256, 16, 16) X(256, 16, 8) X(256, 8, 16)            \
            X(512, 16, 16) X(512, 16, 8) X(512, 8, 16) X(1024, 16, 16) X(1024, 16, 8) X(1024, 32, 16) X(2048, 16, 8)   \
                X(2048, 16, 4) X(4096, 16, 4)

#define CRDFFT_ROWTR_LIST(X)                                                                                           \
    X(16, 16, 16) X(32, 16, 16) X(64, 16, 16) X(128, 16, 16) X(128, 8, 16) X(128, 8, 32) X(128, 16, 32)               \
        X(256, 16, 16)                                                                                                 \
        X(256, 16, 8) X(256, 8, 16) X(256, 16, 32) X(256, 8, 32) X(512, 16, 16) X(512, 16, 8) X(512, 8, 16)            \
            X(1024, 16, 16) X(1024, 16, 8) X(1024, 32, 16) X(2048, 16, 8) X(2048, 16, 4) X(4096, 16, 4)

// (N1, E, C) of the C2R column pass with the pre-rotation fused in (the planner's splits of half lengths 2^15..2^24)
#define CRDFFT_COL_C2R_LIST(X) X(128, 16, 16) X(256, 16, 16) X(1024, 16, 8) X(2048, 16, 8) X(4096, 16, 4)

// (N2, E, R) of the real-transform row pass (the four-step splits the planner gives half lengths 2^15 .. 2^24)
#define CRDFFT_ROWTR_R2C_LIST(X)                                                                                       \
    X(128, 8, 32) X(128, 16, 32) X(256, 16, 16) X(256, 16, 32) X(256, 8, 32) X(512, 8, 16) X(1024, 16, 8)             \
        X(2048, 16, 8) X(4096, 16, 4)

// (N1, E, C) of the persistent prefetching column pass
#define CRDFFT_COL_PF_LIST(X)                                                                                          \
    X(256, 16, 16) X(512, 16, 16) X(1024, 16, 8) X(1024, 16, 16) X(1024, 32, 16) X(2048, 16, 8)

// (N1, E1, N2, E2, threads) of the fused four-step
#define CRDFFT_FUSED_LIST(X)                                                                                           \
    X(128, 16, 256, 16, 1024) X(128, 16, 256, 16, 512) X(256, 16, 256, 16, 1024) X(256, 16, 256, 16, 512)            \
        X(256, 16, 512, 16, 1024)

// (m, E) of the one-kernel Bluestein transform: 4 / 8 values per thread up to m = 8192 (the occupancy the two
// m-point cores per sequence need), 16 at 16384
#define CRDFFT_BLUE_LIST(X)                                                                                            \
    X(4, 4) X(8, 8) X(16, 4) X(32, 4) X(64, 4) X(128, 4) X(256, 4) X(512, 8) X(1024, 8) X(2048, 8) X(4096, 8)          \
        X(8192, 8) X(16384, 16)

constexpr uint64_t kMaxItems = 0xFFFFFFFFull; // work-items per launch

template <typename K> hipError_t set_lds(K kernel, unsigned bytes)
{
    if (bytes > 64 * 1024)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(bytes));
    }
    return hipSuccess;
}

} // namespace

bool row_kernel_shape(uint64_t n, int e, LaunchShape* out)
{
#define X(N, E)                                                                                                        \
    if (n == N && e == E)                                                                                              \
    {                                                                                                                  \
        out->threads = RowCfg<N, E>::threads;                                                                          \
        out->per_group = RowCfg<N, E>::M;                                                                              \
        out->lds_bytes = RowCfg<N, E>::lds_bytes;                                                                      \
        return true;                                                                                                   \
    }
    CRDFFT_ROW_LIST(X)
#undef X
    return false;
}

bool col_kernel_shape(uint64_t n1, int e, int c, LaunchShape* out)
{
#define X(N, E, CC)                                                                                                    \
    if (n1 == N && e == E && c == CC)                                                                                  \
    {                                                                                                                  \
        out->threads = ColCfg<N, E, CC>::threads;                                                                      \
        out->per_group = CC;                                                                                           \
        out->lds_bytes = ColCfg<N, E, CC>::lds_bytes;                                                                  \
        return true;                                                                                                   \
    }
    CRDFFT_COL_LIST(X)
#undef X
    return false;
}

bool fused_kernel_shape(uint64_t n1, int e1, uint64_t n2, int e2, int threads, LaunchShape* out)
{
#define X(N1, E1, N2, E2, TH)                                                                                          \
    if (n1 == N1 && e1 == E1 && n2 == N2 && e2 == E2 && threads == TH)                                                 \
    {                                                                                                                  \
        out->threads = TH;                                                                                             \
        out->per_group = 1;                                                                                            \
        out->lds_bytes = FusedCfg<N1, E1, N2, E2, TH>::lds_bytes;                                                      \
        return true;                                                                                                   \
    }
    CRDFFT_FUSED_LIST(X)
#undef X
    return false;
}

bool rowtr_kernel_shape(uint64_t n2, int e, int r, LaunchShape* out)
{
#define X(N, E, RR)                                                                                                    \
    if (n2 == N && e == E && r == RR)                                                                                  \
    {                                                                                                                  \
        out->threads = RowTrCfg<N, E, RR>::threads;                                                                    \
        out->per_group = RR;                                                                                           \
        out->lds_bytes = RowTrCfg<N, E, RR>::lds_bytes;                                                                \
        return true;                                                                                                   \
    }
    CRDFFT_ROWTR_LIST(X)
#undef X
    return false;
}

hipError_t launch_rows(uint64_t n, int e, bool inverse, const RowArgs& a, hipStream_t s)
{
#define X(N, E)                                                                                                        \
    if (n == N && e == E)                                                                                              \
    {                                                                                                                  \
        using C = RowCfg<N, E>;                                                                                        \
        auto kern = inverse ? k_rows<N, E, true> : k_rows<N, E, false>;                                                \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t max_groups = kMaxItems / C::threads;                                                            \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            RowArgs part = a;                                                                                          \
            const uint64_t left = a.count - done;                                                                      \
            const uint64_t groups = std::min<uint64_t>((left + C::M - 1) / C::M, max_groups);                         \
            part.count = std::min<uint64_t>(left, groups * C::M);                                                      \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups)), dim3(C::threads), C::lds_bytes, s, part);   \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_ROW_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_cols(uint64_t n1, int e, int c, bool inverse, const ColArgs& a, hipStream_t s)
{
#define X(N, E, CC)                                                                                                    \
    if (n1 == N && e == E && c == CC)                                                                                  \
    {                                                                                                                  \
        using C = ColCfg<N, E, CC>;                                                                                    \
        if (static_cast<uint64_t>(N) * a.ncols >= (uint64_t{1} << 32))                                                 \
            return hipErrorInvalidValue; /* the kernel's in-matrix offsets are 32-bit */                               \
        auto kern = inverse ? k_cols<N, E, CC, true> : k_cols<N, E, CC, false>;                                        \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t groups_per = (a.ncols + CC - 1) / CC;                                                           \
        const uint64_t max_t = std::max<uint64_t>(1, (kMaxItems / C::threads) / groups_per);                          \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            ColArgs part = a;                                                                                          \
            part.count = std::min<uint64_t>(a.count - done, max_t);                                                    \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            const uint64_t groups = groups_per * part.count;                                                           \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups)), dim3(C::threads), C::lds_bytes, s, part);   \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_COL_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

// (N1, E, CC) of the element-major four-step's column kernels: every length a split of 2^13..2^24 gives
#define CRDFFT_COL_EM_LIST(X)                                                                                          \
    X(64, 16, 16) X(128, 16, 16) X(256, 16, 16) X(512, 16, 16) X(1024, 16, 8) X(2048, 16, 8) X(4096, 16, 4)

bool em_col_shape(uint64_t len, int* e, int* c)
{
#define X(N, E, CC)                                                                                                    \
    if (len == N)                                                                                                      \
    {                                                                                                                  \
        *e = E;                                                                                                        \
        *c = CC;                                                                                                       \
        return true;                                                                                                   \
    }
    CRDFFT_COL_EM_LIST(X)
#undef X
    return false;
}

hipError_t launch_cols_em(uint64_t len, bool inverse, const EmArgs& a, hipStream_t s)
{
#define X(N, E, CC)                                                                                                    \
    if (len == N)                                                                                                      \
    {                                                                                                                  \
        using C = ColCfg<N, E, CC>;                                                                                    \
        const bool tw = a.thi != nullptr;                                                                              \
        auto kern = inverse ? (tw ? k_cols_em<N, E, CC, true, true> : k_cols_em<N, E, CC, true, false>)                \
                            : (tw ? k_cols_em<N, E, CC, false, true> : k_cols_em<N, E, CC, false, false>);             \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t groups_per = (a.ncols + CC - 1) / CC;                                                           \
        const uint64_t max_t = std::max<uint64_t>(1, (kMaxItems / C::threads) / groups_per);                          \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            EmArgs part = a;                                                                                           \
            part.count = std::min<uint64_t>(a.count - done, max_t);                                                    \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            part.t0 = a.t0 + done;                                                                                     \
            const uint64_t groups = groups_per * part.count;                                                           \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups)), dim3(C::threads), C::lds_bytes, s, part);   \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_COL_EM_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

bool have_cols_pf(uint64_t n1, int e, int c)
{
#define X(N, E, CC)                                                                                                    \
    if (n1 == N && e == E && c == CC)                                                                                  \
        return true;
    CRDFFT_COL_PF_LIST(X)
#undef X
    return false;
}

hipError_t launch_cols_pf(uint64_t n1, int e, int c, bool inverse, const ColArgs& a, hipStream_t s)
{
#define X(N, E, CC)                                                                                                    \
    if (n1 == N && e == E && c == CC)                                                                                  \
    {                                                                                                                  \
        using P = ColPfCfg<N, E, CC>;                                                                                  \
        using C = ColCfg<N, E, CC>;                                                                                    \
        /* whole column groups, byte offsets of one matrix below 4 GiB */                                              \
        if (a.ncols % CC != 0 || static_cast<uint64_t>(N) * a.ncols * 8 >= (uint64_t{1} << 32) ||                      \
            a.count * (a.ncols / CC) >= (uint64_t{1} << 40))                                                           \
            return hipErrorInvalidValue;                                                                               \
        if (a.tfull == nullptr && a.thi == nullptr)                                                                    \
            return hipErrorInvalidValue; /* four-step columns only */                                                  \
        auto kern = a.tfull != nullptr ? (inverse ? k_cols_pf<N, E, CC, 1, true> : k_cols_pf<N, E, CC, 1, false>)      \
                                       : (inverse ? k_cols_pf<N, E, CC, 2, true> : k_cols_pf<N, E, CC, 2, false>);     \
        hipError_t err = set_lds(kern, P::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        int dev = 0, cus = 0, per_cu = 0;                                                                              \
        err = hipGetDevice(&dev);                                                                                      \
        if (err == hipSuccess)                                                                                         \
            err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);                             \
        if (err == hipSuccess)                                                                                         \
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, C::threads, P::lds_bytes);              \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        if (per_cu < 1)                                                                                                \
            return hipErrorInvalidValue;                                                                               \
        /* the tile walk needs a multiple of 8 workgroups (one range per XCD) */                                       \
        const uint64_t grid = std::max<uint64_t>(8, (static_cast<uint64_t>(cus) * per_cu) & ~uint64_t{7});             \
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(C::threads), P::lds_bytes, s, a);            \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_COL_PF_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

bool have_cols_c2r(uint64_t n1, int e, int c)
{
#define X(N, E, CC)                                                                                                    \
    if (n1 == N && e == E && c == CC)                                                                                  \
        return true;
    CRDFFT_COL_C2R_LIST(X)
#undef X
    return false;
}

hipError_t launch_cols_c2r(uint64_t n1, int e, int c, const ColArgs& a, hipStream_t s)
{
    if (a.wreal == nullptr || a.ncols % static_cast<uint64_t>(c) != 0)
    {
        return hipErrorInvalidValue;
    }
#define X(N, E, CC)                                                                                                    \
    if (n1 == N && e == E && c == CC)                                                                                  \
    {                                                                                                                  \
        using C = ColCfg<N, E, CC>;                                                                                    \
        auto kern = k_cols<N, E, CC, true, true>;                                                                      \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t groups_per = a.ncols / CC;                                                                      \
        const uint64_t max_t = std::max<uint64_t>(1, (kMaxItems / C::threads) / groups_per);                          \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            ColArgs part = a;                                                                                          \
            part.count = std::min<uint64_t>(a.count - done, max_t);                                                    \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups_per * part.count)), dim3(C::threads),          \
                               C::lds_bytes, s, part);                                                                 \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_COL_C2R_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

bool have_rowtr_r2c(uint64_t n2, int e, int r)
{
#define X(N, E, RR)                                                                                                    \
    if (n2 == N && e == E && r == RR)                                                                                  \
        return true;
    CRDFFT_ROWTR_R2C_LIST(X)
#undef X
    return false;
}

hipError_t launch_rowtr_r2c(uint64_t n2, int e, int r, const RowTrArgs& a, hipStream_t s)
{
    if (a.wreal == nullptr || a.n1 % static_cast<uint64_t>(r) != 0 || a.n1 < static_cast<uint64_t>(r))
    {
        return hipErrorInvalidValue;
    }
#define X(N, E, RR)                                                                                                    \
    if (n2 == N && e == E && r == RR)                                                                                  \
    {                                                                                                                  \
        using C = RowTrCfg<N, E, RR>;                                                                                  \
        auto kern = k_rowtr_r2c<N, E, RR>;                                                                             \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t groups = a.n1 / RR;                                                                             \
        const uint64_t max_count = std::max<uint64_t>(1, 0xFFFFFFFFull / (groups * C::threads));                      \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            RowTrArgs part = a;                                                                                        \
            part.count = std::min<uint64_t>(a.count - done, max_count);                                                \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups * part.count)), dim3(C::threads), C::lds_bytes, \
                               s, part);                                                                               \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_ROWTR_R2C_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_rowtr(uint64_t n2, int e, int r, bool inverse, const RowTrArgs& a, hipStream_t s)
{
#define X(N, E, RR)                                                                                                    \
    if (n2 == N && e == E && r == RR)                                                                                  \
    {                                                                                                                  \
        using C = RowTrCfg<N, E, RR>;                                                                                  \
        if (a.n1 % RR != 0)                                                                                            \
            return hipErrorInvalidValue;                                                                               \
        auto kern = inverse ? k_rowtr<N, E, RR, true> : k_rowtr<N, E, RR, false>;                                      \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t groups_per = a.n1 / RR;                                                                         \
        const uint64_t max_t = std::max<uint64_t>(1, (kMaxItems / C::threads) / groups_per);                          \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            RowTrArgs part = a;                                                                                        \
            part.count = std::min<uint64_t>(a.count - done, max_t);                                                    \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            const uint64_t groups = groups_per * part.count;                                                           \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups)), dim3(C::threads), C::lds_bytes, s, part);   \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_ROWTR_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t fused_blocks_per_cu(uint64_t n1, int e1, uint64_t n2, int e2, int threads, int* blocks)
{
#define X(N1, E1, N2, E2, TH)                                                                                          \
    if (n1 == N1 && e1 == E1 && n2 == N2 && e2 == E2 && threads == TH)                                                 \
    {                                                                                                                  \
        using F = FusedCfg<N1, E1, N2, E2, TH>;                                                                        \
        auto kern = k_fused<N1, E1, N2, E2, TH, false>;                                                                \
        hipError_t err = set_lds(kern, F::lds_bytes);                                                                  \
        if (err == hipSuccess)                                                                                         \
            err = set_lds(k_fused<N1, E1, N2, E2, TH, true>, F::lds_bytes);                                            \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, kern, TH, F::lds_bytes);                           \
    }
    CRDFFT_FUSED_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_fused(uint64_t n1, int e1, uint64_t n2, int e2, int threads, bool inverse, const FusedArgs& a,
                        uint64_t grid, hipStream_t s)
{
#define X(N1, E1, N2, E2, TH)                                                                                          \
    if (n1 == N1 && e1 == E1 && n2 == N2 && e2 == E2 && threads == TH)                                                 \
    {                                                                                                                  \
        using F = FusedCfg<N1, E1, N2, E2, TH>;                                                                        \
        if (grid == 0 || grid > a.slots)                                                                               \
            return hipErrorInvalidValue; /* one workspace slot per workgroup */                                        \
        auto kern = inverse ? k_fused<N1, E1, N2, E2, TH, true> : k_fused<N1, E1, N2, E2, TH, false>;                  \
        hipError_t err = set_lds(kern, F::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(std::min<uint64_t>(grid, a.count))), dim3(TH),          \
                           F::lds_bytes, s, a);                                                                        \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_FUSED_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

int blue_fused_e(uint64_t m)
{
#define X(N, E)                                                                                                        \
    if (m == N)                                                                                                        \
        return E;
    CRDFFT_BLUE_LIST(X)
#undef X
    return 0;
}

hipError_t launch_blue_fused(uint64_t m, int e, bool inverse, const BlueArgs& a, hipStream_t s)
{
    if (a.n == 0 || 2 * a.n - 1 > m)
    {
        return hipErrorInvalidValue;
    }
#define X(N, E)                                                                                                        \
    if (m == N && e == E)                                                                                              \
    {                                                                                                                  \
        using C = RowCfg<N, E>;                                                                                        \
        auto kern = inverse ? k_blue<N, E, true> : k_blue<N, E, false>;                                                \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t max_groups = kMaxItems / C::threads;                                                            \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            BlueArgs part = a;                                                                                         \
            const uint64_t left = a.count - done;                                                                      \
            const uint64_t groups = std::min<uint64_t>((left + C::M - 1) / C::M, max_groups);                         \
            part.count = std::min<uint64_t>(left, groups * C::M);                                                      \
            part.in = a.in + done * a.idist;                                                                           \
            part.out = a.out + done * a.odist;                                                                         \
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups)), dim3(C::threads), C::lds_bytes, s, part);   \
            done += part.count;                                                                                        \
        }                                                                                                              \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_BLUE_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

uint64_t stage_twiddle_entries(uint64_t n, int e)
{
    if (n <= 1 || e <= 1 || static_cast<uint64_t>(e) >= n)
    {
        return 0;
    }
    const int lg = ilog2c(static_cast<long>(n));
    const int le = ilog2c(e);
    const uint64_t r0 = (lg % le) != 0 ? (uint64_t{1} << (lg % le)) : static_cast<uint64_t>(e);
    uint64_t total = 0;
    for (uint64_t p = r0; p < n; p *= static_cast<uint64_t>(e))
    {
        total += p * static_cast<uint64_t>(e);
    }
    return total;
}

hipError_t build_stage_twiddles(cfloat* table, uint64_t n, int e, hipStream_t s)
{
    if (stage_twiddle_entries(n, e) == 0)
    {
        return hipSuccess;
    }
    const int lg = ilog2c(static_cast<long>(n));
    const int le = ilog2c(e);
    const uint64_t r0 = (lg % le) != 0 ? (uint64_t{1} << (lg % le)) : static_cast<uint64_t>(e);
    uint64_t off = 0;
    for (uint64_t p = r0; p < n; p *= static_cast<uint64_t>(e))
    {
        const uint64_t cnt = p * static_cast<uint64_t>(e);
        const unsigned blocks = static_cast<unsigned>((cnt + 255) / 256);
        hipLaunchKernelGGL(k_roots_stage, dim3(blocks), dim3(256), 0, s, reinterpret_cast<cf*>(table + off), p,
                           static_cast<uint64_t>(e));
        off += cnt;
    }
    return hipGetLastError();
}

hipError_t build_roots(cfloat* out, uint64_t count, uint64_t mult, uint64_t d, hipStream_t s)
{
    if (count == 0)
    {
        return hipSuccess;
    }
    const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(k_roots_lin, dim3(blocks), dim3(256), 0, s, reinterpret_cast<cf*>(out), count, mult, d);
    return hipGetLastError();
}

hipError_t build_full_twiddle(cfloat* out, uint64_t rows, uint64_t ncols, uint64_t n, hipStream_t s)
{
    const uint64_t count = rows * ncols;
    if (count == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(k_roots_full, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<cf*>(out), rows, ncols, n);
    return hipGetLastError();
}

hipError_t build_step_table(cfloat* out, uint64_t ncols, int e, uint64_t step, uint64_t n, hipStream_t s)
{
    const uint64_t count = ncols * static_cast<uint64_t>(e);
    if (count == 0)
    {
        return hipSuccess;
    }
    hipLaunchKernelGGL(k_roots_step, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<cf*>(out), ncols, static_cast<uint64_t>(e), step, n);
    return hipGetLastError();
}

hipError_t launch_reference(const cfloat* in, cfloat* out, const cfloat* w, uint64_t n, uint64_t count,
                            uint64_t idist, uint64_t odist, bool inverse, hipStream_t s)
{
    if (n == 0 || count == 0)
    {
        return hipSuccess;
    }
    if (in == out && idist != odist)
    {
        return hipErrorInvalidValue; // in place needs one layout
    }
    const uint32_t lg = static_cast<uint32_t>(ilog2c(static_cast<long>(n)));
    const uint64_t items = n * count;
    if (items > kMaxItems)
    {
        return hipErrorInvalidValue;
    }
    if (in == out)
    {
        hipLaunchKernelGGL(k_ref_bitrev_inplace, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<cf*>(out), n, lg, count, odist);
    }
    else
    {
        hipLaunchKernelGGL(k_ref_bitrev, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<const cf*>(in), reinterpret_cast<cf*>(out), n, lg, count, idist, odist);
    }
    const uint64_t bfly = (n / 2) * count;
    for (uint64_t len = 2; len <= n; len <<= 1)
    {
        auto kern = inverse ? k_ref_stage<true> : k_ref_stage<false>;
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>((bfly + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<cf*>(out), reinterpret_cast<const cf*>(w), n, len, count, odist);
    }
    return hipGetLastError();
}

hipError_t launch_stream_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s)
{
    if (bytes % 16 != 0)
    {
        return hipErrorInvalidValue;
    }
    const uint64_t total = bytes / 16;
    const uint64_t max_items = (kMaxItems / 256) * 256;
    for (uint64_t done = 0; done < total; done += max_items)
    {
        const uint64_t count = std::min(total - done, max_items);
        hipLaunchKernelGGL(k_stream_copy, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, s,
                           static_cast<const float4*>(src) + done, static_cast<float4*>(dst) + done, count);
    }
    return hipGetLastError();
}

hipError_t launch_copy(const cfloat* in, cfloat* out, uint64_t n, uint64_t count, uint64_t idist, uint64_t odist,
                       hipStream_t s)
{
    const uint64_t items = n * count;
    if (items == 0)
    {
        return hipSuccess;
    }
    if (items > kMaxItems)
    {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_copy, dim3(static_cast<unsigned>((items + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const cf*>(in), reinterpret_cast<cf*>(out), n, count, idist, odist);
    return hipGetLastError();
}

} // namespace crdfft
