// k_coop: one square transform n = n1 * n1 per group of workgroups of one XCD, the four-step intermediate shared
// through that XCD's L2 instead of a round trip through HBM.
//
// The two-kernel four-step (k_cols + k_rowtr) moves every point across the memory fabric twice (input in,
// intermediate out, intermediate in, output out) and runs at the copy ceiling for that traffic. Here the G = n1 / C
// workgroups of a group each run one C-column tile of phase A (column FFT, twiddle W_n^{col k1}, store into the
// group's slot) and, after a group barrier, one C-row tile of phase B (row FFT of the slot rows, transposed store),
// so HBM sees one read and one write per point while the slot (n * 8 bytes per group) stays in the L2 of the XCD all
// members run on. The reference's four-step is fft_kernels.cpp:2363- (phase 1 + twiddle, phase 2 + scatter).
//
// Coherence, by the memory model's agent-scope release/acquire: every thread of a workgroup runs a release fence after
// its slot stores, the workgroup barrier follows, and one thread counts the workgroup in with a release atomic; a
// waiter polls with acquire loads, and after the barrier that releases its workgroup every thread runs an acquire
// fence before it reads the slot. Members share one XCD (so the slot stays in that XCD's L2 and the fences' cache
// maintenance stays local in practice): groups are formed at launch from the XCC_ID register each workgroup reads,
// a per-XCD ticket, a grid-wide arrival count, and a mode word decided once by compare-and-swap. When every workgroup has arrived (the launch is resident
// as a whole) each XCD's workgroups form groups of G and the remainder runs alone; when a bounded wait for the grid
// gives up, or kHintCoopSolo is set, every workgroup runs alone on its own slot (the k_fused scheme). Every wait is a
// bounded poll; one that gives up counts an error in the control words and lets the workgroup finish without
// waiting again, so every wave exits.
#include "buffer.hpp"
#include "fft_core.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace crdfft
{
using namespace dev;

namespace
{

constexpr int round_up(int v, int m) { return (v + m - 1) / m * m; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int N1, int E, int TH> struct CoopCfg
{
    static constexpr int T = N1 / E;  // threads per column (phase A) and per row (phase B)
    static constexpr int CC = TH / T; // columns per phase-A tile
    static constexpr int RR = TH / T; // rows per phase-B tile
    static constexpr int G = N1 / CC; // tiles per phase = workgroups per group
    static constexpr int LF = lds_floats(N1);
    static constexpr int IM = LF * CC + 32; // phase A: CC interleaved columns, re block then im block
    static constexpr int col_floats = IM + LF * CC;
    static constexpr int RS = round_up(2 * LF, 64) + (T < 64 ? T : 0); // phase B: one region per row
    static constexpr int row_floats = RS * RR;
    static constexpr int TP = RR + 1; // transposed-store staging pitch
    static constexpr int tr_floats = 2 * N1 * TP;
    static constexpr int data_floats = cmax(col_floats, cmax(row_floats, tr_floats));
    static constexpr int tw_entries = static_cast<int>(Geometry<N1, E>::twiddle_entries());
    static constexpr int lds_bytes = data_floats * 4 + (tw_entries + 2 * N1) * 8;
    static_assert(CC * T == TH && N1 % CC == 0 && data_floats % 2 == 0, "whole tiles, aligned tables");
};

__device__ __forceinline__ uint32_t xcc_id()
{
    uint32_t x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x & 15u;
}

__device__ __forceinline__ uint32_t atom_add(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// counts this workgroup in after its stores (release: they are visible to whoever acquires the count)
__device__ __forceinline__ void atom_add_release(uint32_t* p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t load_acquire(uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// the value before the exchange
__device__ __forceinline__ uint32_t atom_cas(uint32_t* p, uint32_t expected, uint32_t desired)
{
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;
}

// Polls *p (acquire loads) until it reaches target, at most `limit` times with a short sleep between polls.
__device__ bool wait_at_least(uint32_t* p, uint32_t target, uint32_t limit)
{
    for (uint32_t it = 0; it < limit; ++it)
    {
        if (load_acquire(p) >= target)
        {
            return true;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return false;
}

// v[m] = element (off + m * stride) of the slot (plain loads: the caller has run its acquire fence)
template <int E>
__device__ __forceinline__ void load_slot(cf (&v)[E], __amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t stride)
{
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        v[m] = bld<0>(r, off * 8u, static_cast<uint32_t>(m) * stride * 8u);
    }
}

template <int E>
__device__ __forceinline__ void load_in(cf (&v)[E], __amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t stride, bool nt)
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
__device__ __forceinline__ void store_out(const cf (&v)[E], __amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t stride,
                                          float scale, bool nt)
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

template <int N1, int E, int TH, bool INV>
__global__ __launch_bounds__(TH) void k_coop(CoopArgs a)
{
    using C = CoopCfg<N1, E, TH>;
    constexpr int T = C::T;
    constexpr int CC = C::CC;
    constexpr int RR = C::RR;
    constexpr int G = C::G;
    constexpr uint32_t kN = static_cast<uint32_t>(N1) * N1;
    constexpr int L1 = ilog2c(N1);
    extern __shared__ float lds[];
    __shared__ uint32_t s_ctl[5]; // unit, rank, solo, units, gave up
    const int tid = threadIdx.x;
    const int c = tid % CC; // phase A: column within the tile, slots i1 + T m
    const int i1 = tid / CC;
    const int r = tid / T; // phase B: row within the tile, slots i2 + T m
    const int i2 = tid - r * T;
    const int r2 = tid % RR; // phase B store: row r2, elements j2 + T m
    const int j2 = tid / RR;
    cf* tw = reinterpret_cast<cf*>(lds + C::data_floats);
    cf* thi = tw + C::tw_entries;
    cf* tlo = thi + N1;
    for (int j = tid; j < C::tw_entries; j += TH)
    {
        tw[j] = reinterpret_cast<const cf*>(a.tw)[j];
    }
    for (int j = tid; j < N1; j += TH)
    {
        thi[j] = reinterpret_cast<const cf*>(a.thi)[j];
        tlo[j] = reinterpret_cast<const cf*>(a.tlo)[j];
    }
    uint32_t* ctl = a.ctl;
    if (tid == 0)
    {
        uint32_t unit = blockIdx.x, rank = 0, solo = 1, units = gridDim.x;
        if ((a.hints & kHintCoopSolo) == 0)
        {
            const uint32_t x = xcc_id();
            const uint32_t ticket = atom_add(ctl + kCoopCtlXcd + 16 * x, 1u);
            atom_add_release(ctl + kCoopCtlGrid, 1u); // the ticket is counted before the arrival
            const bool all = wait_at_least(ctl + kCoopCtlGrid, gridDim.x, a.spin_limit);
            const uint32_t old = atom_cas(ctl + kCoopCtlMode, 0u, all ? 1u : 2u);
            const uint32_t mode = old == 0u ? (all ? 1u : 2u) : old;
            if (mode == 1u)
            {
                // every ticket is counted: groups of G per XCD, in XCD order, then the remainders alone
                uint32_t full_before = 0, solo_before = 0, full_all = 0, solo_all = 0, full_x = 0;
                for (uint32_t y = 0; y < 16; ++y)
                {
                    const uint32_t cnt = atom_add(ctl + kCoopCtlXcd + 16 * y, 0u);
                    const uint32_t f = cnt / G;
                    const uint32_t s = cnt - f * G;
                    if (y < x)
                    {
                        full_before += f;
                        solo_before += s;
                    }
                    if (y == x)
                    {
                        full_x = f;
                    }
                    full_all += f;
                    solo_all += s;
                }
                units = full_all + solo_all;
                if (ticket < full_x * G)
                {
                    unit = full_before + ticket / G;
                    rank = ticket % G;
                    solo = 0;
                }
                else
                {
                    unit = full_all + solo_before + (ticket - full_x * G);
                }
            }
        }
        s_ctl[0] = unit;
        s_ctl[1] = rank;
        s_ctl[2] = solo;
        s_ctl[3] = units;
        s_ctl[4] = 0;
    }
    __syncthreads();
    const uint32_t unit = s_ctl[0];
    const int rank = static_cast<int>(s_ctl[1]);
    const bool solo = s_ctl[2] != 0;
    const uint64_t units = s_ctl[3];
    const int first = solo ? 0 : rank;
    const int step = solo ? 1 : G;
    uint32_t* a_done = ctl + kCoopCtlUnits + 64 * unit;
    uint32_t* b_done = a_done + 32;
    // all threads: a bounded wait by thread 0, then the workgroup barrier that releases the others
    auto group_wait = [&](uint32_t* p, uint32_t target) {
        if (tid == 0 && s_ctl[4] == 0 && !wait_at_least(p, target, a.spin_limit))
        {
            atom_add(ctl + kCoopCtlError, 1u);
            s_ctl[4] = 1; // no further waits: finish the loop and exit
        }
        __syncthreads();
        fence_acquire(); // every thread: the members' released stores are visible to its loads
    };
    // all threads: this workgroup's stores (or reads) of the slot are done; one thread counts it in
    auto count_in = [&](uint32_t* p) {
        fence_release();
        __syncthreads();
        if (tid == 0)
        {
            atom_add_release(p, 1u);
        }
    };
    const __amdgpu_buffer_rsrc_t slot = make_rsrc(a.slots + static_cast<uint64_t>(unit) * kN);
    const LdsView<CC> va{lds + c, C::IM};
    const LdsView<1> vb{lds + r * C::RS, C::LF};
    float* tre = lds;
    float* tim = lds + N1 * C::TP;
    const bool ntl = (a.hints & kHintNtLoad) != 0;
    const bool nts = (a.hints & kHintNtStore) != 0;
    const uint32_t a_off = static_cast<uint32_t>(c) + static_cast<uint32_t>(i1) * N1; // + tile * CC
    constexpr uint32_t a_stride = static_cast<uint32_t>(T) * N1;
    const uint32_t b_off = static_cast<uint32_t>(r) * N1 + i2;                   // + tile * RR * N1
    const uint32_t s_off = static_cast<uint32_t>(r2) + static_cast<uint32_t>(j2) * N1; // + tile * RR
    constexpr uint32_t s_stride = static_cast<uint32_t>(T) * N1;
    const cfloat* in = a.in;
    cfloat* out = a.out;

    uint64_t t = unit;
    cf nx[E];
    if (t < a.count)
    {
        load_in<E>(nx, make_rsrc(in + t * a.idist), a_off + first * CC, a_stride, ntl);
    }
    uint32_t k = 0;
#pragma unroll 1
    for (; t < a.count; t += units, ++k)
    {
        const __amdgpu_buffer_rsrc_t src = make_rsrc(in + t * a.idist);
        const __amdgpu_buffer_rsrc_t dst = make_rsrc(out + t * a.odist);
        // phase A: column FFTs of this member's tiles into the slot
#pragma unroll 1
        for (int ct = first; ct < G; ct += step)
        {
            cf v[E];
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                v[m] = nx[m];
            }
            if (ct + step < G)
            {
                load_in<E>(nx, src, a_off + (ct + step) * CC, a_stride, ntl);
            }
            fft_core<N1, E, INV, CC>(v, i1, va, tw);
            const uint32_t col = static_cast<uint32_t>(ct * CC + c);
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                const uint32_t e = col * static_cast<uint32_t>(i1 + T * m); // W_n^{col k1} = hi[e / n1] lo[e % n1]
                v[m] = ctw<INV>(v[m], cmul(thi[e >> L1], tlo[e & (N1 - 1)]));
            }
            if (!solo && k > 0)
            {
                group_wait(b_done, static_cast<uint32_t>(G) * k); // every member has read the previous transform
            }
            store_out<E>(v, slot, a_off + ct * CC, a_stride, 1.0f, false);
        }
        if (!solo)
        {
            count_in(a_done);
        }
        else
        {
            __syncthreads();
        }
        if (t + units < a.count)
        {
            load_in<E>(nx, make_rsrc(in + (t + units) * a.idist), a_off + first * CC, a_stride, ntl);
        }
        if (!solo)
        {
            group_wait(a_done, static_cast<uint32_t>(G) * (k + 1));
        }
        // phase B: row FFTs of this member's tiles of the slot, transposed store
#pragma unroll 1
        for (int rt = first; rt < G; rt += step)
        {
            cf v[E];
            load_slot<E>(v, slot, b_off + static_cast<uint32_t>(rt * RR) * N1, static_cast<uint32_t>(T));
            if (!solo)
            {
                count_in(b_done); // this member is done reading the slot (its loads are ordered before the count)
            }
            fft_core<N1, E, INV, 1>(v, i2, vb, tw);
            if constexpr (N1 == E)
            {
                __syncthreads();
            }
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                const int k2 = i2 + T * m;
                tre[k2 * C::TP + r] = v[m].x;
                tim[k2 * C::TP + r] = v[m].y;
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                const int k2 = j2 + T * m;
                v[m] = cf{tre[k2 * C::TP + r2], tim[k2 * C::TP + r2]};
            }
            store_out<E>(v, dst, s_off + static_cast<uint32_t>(rt * RR), s_stride, a.scale, nts);
            __syncthreads(); // LDS free for the next tile or transform
        }
    }
}

__global__ void k_coop_clear(uint32_t* ctl, uint32_t words, uint32_t solo)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < words)
    {
        ctl[j] = (j == kCoopCtlMode && solo != 0) ? 2u : 0u;
    }
}

// (n1, E, threads): n = n1^2
#define CRDFFT_COOP_LIST(X) X(64, 4, 256) X(256, 16, 512)

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

bool coop_kernel_shape(uint64_t n1, int e, int threads, LaunchShape* out)
{
#define X(N1, E, TH)                                                                                                   \
    if (n1 == N1 && e == E && threads == TH)                                                                           \
    {                                                                                                                  \
        out->threads = TH;                                                                                             \
        out->per_group = CoopCfg<N1, E, TH>::G;                                                                        \
        out->lds_bytes = CoopCfg<N1, E, TH>::lds_bytes;                                                                \
        return true;                                                                                                   \
    }
    CRDFFT_COOP_LIST(X)
#undef X
    return false;
}

hipError_t coop_blocks_per_cu(uint64_t n1, int e, int threads, int* blocks)
{
#define X(N1, E, TH)                                                                                                   \
    if (n1 == N1 && e == E && threads == TH)                                                                           \
    {                                                                                                                  \
        using C = CoopCfg<N1, E, TH>;                                                                                  \
        hipError_t err = set_lds(k_coop<N1, E, TH, false>, C::lds_bytes);                                             \
        if (err == hipSuccess)                                                                                         \
            err = set_lds(k_coop<N1, E, TH, true>, C::lds_bytes);                                                      \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, k_coop<N1, E, TH, false>, TH, C::lds_bytes);      \
    }
    CRDFFT_COOP_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_coop(uint64_t n1, int e, int threads, bool inverse, const CoopArgs& a, unsigned grid, hipStream_t s)
{
    if (grid == 0 || a.count == 0)
    {
        return hipSuccess;
    }
#define X(N1, E, TH)                                                                                                   \
    if (n1 == N1 && e == E && threads == TH)                                                                           \
    {                                                                                                                  \
        using C = CoopCfg<N1, E, TH>;                                                                                  \
        auto kern = inverse ? k_coop<N1, E, TH, true> : k_coop<N1, E, TH, false>;                                     \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint32_t words = static_cast<uint32_t>(coop_ctl_words(grid));                                           \
        hipLaunchKernelGGL(k_coop_clear, dim3((words + 255) / 256), dim3(256), 0, s, a.ctl, words,                    \
                           (a.hints & kHintCoopSolo) != 0 ? 1u : 0u);                                                  \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(TH), C::lds_bytes, s, a);                                           \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_COOP_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

} // namespace crdfft
