// The spectral front end over the real FFT, fused (the reference's hesap-dsp spectral.hpp:43-126, :140-172,
// :334-386: Welch PSD, spectrogram, periodogram and STFT, which there run frame by frame as detrend + window loops,
// a RealFftPlan rfft and a |X|^2 loop).
//
// k_spec<H, E, MODE>: one real frame of N = 2H samples per group of T = H/E threads (several frames per workgroup
// below 256 threads). Each thread loads the sample pairs of its complex slots j = i + T m straight from the signal
// (frame s starts at s * hop, any hop), subtracts the frame mean (a workgroup tree reduction, fixed order) when
// asked, multiplies by the window, runs the H-point Stockham core on z[j] = x[2j] + i x[2j+1], and does the
// Hermitian recombine X[k] = E[k] + W_N^k O[k], E = (Z[k] + conj Z[H-k])/2, O = (Z[k] - conj Z[H-k])/(2i), with
// Z[H-k] read back through LDS; X[H] = Re Z[0] - Im Z[0] comes from the k = 0 slot. The epilogue then writes
//   MODE 0  the complex bins (STFT; with no window, no detrend and hop = the row distance, a batch of R2C rows),
//   MODE 1  the density periodogram scale |X|^2, doubled except DC and Nyquist (spectrogram rows),
//   MODE 2  nothing per frame: lane group g sums the periodograms of frames g, g + G, g + 2G, ... in registers and
//           writes one partial row; k_spec_sum adds the G partial rows in order (Welch, bit-identical run to run).
//           Frames of 32768 samples run MODE 1 in chunks of rows instead, each chunk added in order.
// HBM traffic: the signal once per frame that covers it (twice at 50% overlap) and the output; no framed copy of
// the signal, no spectrum round trip.
#include "buffer.hpp"
#include "fft_core.hpp"
#include "kernels.hpp"
#include "spectral.hpp"

#include <algorithm>

namespace crdfft
{
using namespace dev;

namespace
{

constexpr int round_up(int v, int m) { return (v + m - 1) / m * m; }

template <int H, int E> struct SpecCfg
{
    static constexpr int T = H / E;
    static constexpr int M = T >= 256 ? 1 : 256 / T; // frames per workgroup
    static constexpr int threads = T * M;
    static constexpr int LF = lds_floats(H);
    static constexpr int RS = round_up(2 * LF, 64) + (T < 64 ? T : 0); // per-frame region (re then im)
    static constexpr int lds_floats_total = RS * M + threads;         // + the mean reduction
    static constexpr int lds_bytes = lds_floats_total * 4;
};

template <int H, int E, int MODE>
__global__ __launch_bounds__((SpecCfg<H, E>::threads)) void k_spec(SpecArgs a)
{
    using C = SpecCfg<H, E>;
    constexpr int T = C::T;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int local = tid / T;
    const int i = tid - local * T;
    const LdsView<1> view{lds + local * C::RS, C::LF};
    float* red = lds + C::RS * C::M;
    const uint64_t group = static_cast<uint64_t>(blockIdx.x) * C::M + local;
    const uint64_t groups = static_cast<uint64_t>(gridDim.x) * C::M;
    // Welch accumulators (MODE 2 exists for H <= 8192: the largest core needs every register of its 1024 threads)
    static_assert(MODE != 2 || H <= kSpecWelchMaxH, "Welch lane groups up to H = 8192");
    float acc[E + 1];
#pragma unroll
    for (int m = 0; m <= E; ++m)
    {
        acc[m] = 0.0f;
    }
    // MODE 0/1: one frame per lane group; MODE 2: frames group, group + groups, ... (uniform trip count per
    // workgroup: every lane group runs the same number of iterations, idle ones on a masked frame)
    const uint64_t iters = MODE == 2 ? (a.frames + groups - 1) / groups : 1;
    // 8-byte sample pairs when every frame starts 8-byte aligned (an even hop from an 8-byte aligned signal)
    const bool pairs = (reinterpret_cast<uintptr_t>(a.x) & 7u) == 0 && (a.hop & 1u) == 0;
#pragma unroll 1
    for (uint64_t it = 0; it < iters; ++it)
    {
        const uint64_t f = group + it * groups;
        const bool active = f < a.frames;
        // the tables are the same for every frame; an empty asm hides that, so the compiler reloads them per frame
        // (L1/L2 hits) instead of holding every butterfly's twiddles in registers across the frame loop
        const cf* tw = reinterpret_cast<const cf*>(a.tw);
        const cf* wk = reinterpret_cast<const cf*>(a.wtab);
        const float* win = a.window;
        asm volatile("" : "+s"(tw), "+s"(wk), "+s"(win));
        // the frame through a buffer resource: an inactive frame gets no records, so its loads read zeros and touch
        // no memory
        const __amdgpu_buffer_rsrc_t xr = make_rsrc(a.x + (active ? f * a.hop : 0), active ? 8u * H : 0u);
        cf v[E];
        if (pairs)
        {
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                v[m] = bld<0>(xr, 8u * static_cast<uint32_t>(i + T * m), 0);
            }
        }
        else
        {
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                const uint32_t j = static_cast<uint32_t>(i + T * m);
                v[m] = cf{bld32<0>(xr, 8u * j, 0), bld32<0>(xr, 8u * j + 4u, 0)};
            }
        }
        if (a.detrend != 0)
        {
            // the frame mean, in a fixed order: a butterfly over the frame's lanes of one wave (each step adds the
            // same two partial sums in both lanes, so every lane ends with the same bits), then the waves' sums in
            // wave order through LDS
            float part = 0.0f;
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                part += v[m].x + v[m].y;
            }
            constexpr int W = T < 64 ? T : 64;
#pragma unroll
            for (int o = W / 2; o > 0; o >>= 1)
            {
                part += __shfl_xor(part, o, 64);
            }
            if constexpr (T > 64)
            {
                if ((tid & 63) == 0)
                {
                    red[tid >> 6] = part;
                }
                __syncthreads();
                const int w0 = (local * T) >> 6;
                part = 0.0f;
#pragma unroll
                for (int w = 0; w < T / 64; ++w)
                {
                    part += red[w0 + w];
                }
                __syncthreads();
            }
            const float mean = part / static_cast<float>(2 * H);
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                v[m].x -= mean;
                v[m].y -= mean;
            }
        }
        if (a.window != nullptr)
        {
            const __amdgpu_buffer_rsrc_t wr = make_rsrc(win, 8u * H);
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                const cf wv = bld<0>(wr, 8u * static_cast<uint32_t>(i + T * m), 0);
                v[m].x *= wv.x;
                v[m].y *= wv.y;
            }
        }
        fft_core<H, E, false, 1>(v, i, view, tw);
        // Z through LDS for the recombine (fft_core's last LDS read is behind a barrier; with H == E it used none)
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            view.put(i + T * m, v[m]);
        }
        __syncthreads();
        // recombine and epilogue per slot (nothing but the accumulators outlives the frame)
        const uint64_t row = f * a.ostride;
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            const int k = i + T * m;
            const cf zk = v[m];
            const cf zh = view.get(k == 0 ? 0 : H - k);
            const cf e = {(zk.x + zh.x) * 0.5f, (zk.y - zh.y) * 0.5f};
            const cf o = {(zk.y + zh.y) * 0.5f, (zh.x - zk.x) * 0.5f};
            const cf xk = cadd(e, cmul(o, wk[k]));
            // k == 0 also yields the Nyquist bin: W_N^H = -1, X[H] = Re Z0 - Im Z0 (read only where k == 0)
            const cf xn = {zk.x - zk.y, 0.0f};
            if constexpr (MODE == 0)
            {
                if (active)
                {
                    cf* out = reinterpret_cast<cf*>(a.out) + row;
                    out[k] = xk;
                    if (k == 0)
                    {
                        out[H] = xn;
                    }
                }
            }
            else
            {
                const float p = a.scale * (xk.x * xk.x + xk.y * xk.y) * (k == 0 ? 1.0f : 2.0f);
                const float pn = a.scale * (xn.x * xn.x);
                if constexpr (MODE == 1)
                {
                    if (active)
                    {
                        float* out = reinterpret_cast<float*>(a.out) + row;
                        out[k] = p;
                        if (k == 0)
                        {
                            out[H] = pn;
                        }
                    }
                }
                else
                {
                    acc[m] += active ? p : 0.0f;
                    if (k == 0)
                    {
                        acc[E] += active ? pn : 0.0f;
                    }
                }
            }
        }
        __syncthreads(); // LDS free for the next frame
    }
    if constexpr (MODE == 2)
    {
        // one partial row per lane group
        float* prow = reinterpret_cast<float*>(a.out) + group * static_cast<uint64_t>(H + 1);
#pragma unroll
        for (int m = 0; m < E; ++m)
        {
            prow[i + T * m] = acc[m];
        }
        if (i == 0)
        {
            prow[H] = acc[E];
        }
    }
}

// acc[k] = ((first ? 0 : acc[k]) + sum over the rows of rows[r][k]) * scale, in a fixed order: a workgroup owns 64
// bins, its 16 row groups sum rows g, g + 16, ... (four loads in flight per step), and the 16 partial sums are added
// in group order, so the result does not depend on timing.
constexpr int kSumBins = 64;
constexpr int kSumGroups = 16;

__global__ __launch_bounds__(kSumBins * kSumGroups) void k_spec_sum(const float* rows, float* acc, uint64_t nfreq,
                                                                    uint64_t nrows, int first, float scale)
{
    __shared__ float part[kSumGroups][kSumBins];
    const int kk = threadIdx.x % kSumBins;
    const int g = threadIdx.x / kSumBins;
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kSumBins + kk;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    if (k < nfreq)
    {
        uint64_t r = g;
        for (; r + 3 * kSumGroups < nrows; r += 4 * kSumGroups)
        {
            s0 += rows[r * nfreq + k];
            s1 += rows[(r + kSumGroups) * nfreq + k];
            s2 += rows[(r + 2 * kSumGroups) * nfreq + k];
            s3 += rows[(r + 3 * kSumGroups) * nfreq + k];
        }
        for (; r < nrows; r += kSumGroups)
        {
            s0 += rows[r * nfreq + k];
        }
    }
    part[g][kk] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (g == 0 && k < nfreq)
    {
        float s = first != 0 ? 0.0f : acc[k];
        for (int q = 0; q < kSumGroups; ++q)
        {
            s += part[q][kk];
        }
        acc[k] = s * scale;
    }
}

// (H, E): the per-length values per thread of the real-transform cores (the same choice as the DCT kernels', whose
// load / core / recombine structure this kernel shares)
#define CRDFFT_SPEC_LIST(X)                                                                                            \
    X(2, 2) X(4, 4) X(8, 8) X(16, 4) X(32, 4) X(64, 4) X(128, 4) X(256, 4) X(512, 8) X(1024, 8) X(2048, 8) X(4096, 8)  \
        X(8192, 8) X(16384, 16)

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

bool spec_shape(uint64_t h, int* e, unsigned* frames_per_group)
{
#define X(H, E)                                                                                                        \
    if (h == H)                                                                                                        \
    {                                                                                                                  \
        *e = E;                                                                                                        \
        *frames_per_group = SpecCfg<H, E>::M;                                                                          \
        return true;                                                                                                   \
    }
    CRDFFT_SPEC_LIST(X)
#undef X
    return false;
}

hipError_t launch_spec(uint64_t h, int mode, const SpecArgs& a, uint64_t groups, hipStream_t s)
{
#define X(H, E)                                                                                                        \
    if (h == H)                                                                                                        \
    {                                                                                                                  \
        using C = SpecCfg<H, E>;                                                                                       \
        if (mode == 2 && H > kSpecWelchMaxH)                                                                           \
            return hipErrorInvalidValue;                                                                               \
        auto kern = mode == 0 ? k_spec<H, E, 0>                                                                        \
                              : (mode == 1 ? k_spec<H, E, 1> : k_spec<H, E, (H <= kSpecWelchMaxH ? 2 : 1)>);        \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        if (groups == 0 || groups > 0xFFFFFFFFull)                                                                     \
            return hipErrorInvalidValue;                                                                               \
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(groups)), dim3(C::threads), C::lds_bytes, s, a);        \
        return hipGetLastError();                                                                                      \
    }
    CRDFFT_SPEC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t spec_welch_residency(uint64_t h, int* per_cu)
{
#define X(H, E)                                                                                                        \
    if (h == H)                                                                                                        \
    {                                                                                                                  \
        using C = SpecCfg<H, E>;                                                                                       \
        if (H > kSpecWelchMaxH)                                                                                        \
            return hipErrorInvalidValue;                                                                               \
        auto kern = k_spec<H, E, (H <= kSpecWelchMaxH ? 2 : 1)>;                                                       \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, kern, C::threads, C::lds_bytes);                   \
    }
    CRDFFT_SPEC_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_spec_sum(const float* rows, float* acc, uint64_t nfreq, uint64_t nrows, bool first, float scale,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_spec_sum, dim3(static_cast<unsigned>((nfreq + kSumBins - 1) / kSumBins)),
                       dim3(kSumBins * kSumGroups), 0, s, rows, acc, nfreq, nrows, first ? 1 : 0, scale);
    return hipGetLastError();
}

} // namespace crdfft
