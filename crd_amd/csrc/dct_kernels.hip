// DCT-II/III and DST-II/III (scipy norm=None) of power-of-two lengths on gfx950, by Makhoul's reduction of each to
// one complex FFT of half the length (the reference's DctPlan, engine/numerics/hesap-fft/include/crd/hesap/fft/
// dct.hpp:34-170, which runs the same reduction through its CPU real FFT).
//
//   k_dct_fwd   DCT-II / DST-II of n = 2h <= 32768, one row per workgroup slot (the k_rows layout of fft_core):
//               16-byte loads x[4j..4j+3] give the half-length complex sequence z[j] = (w[2j], w[2j+1]) of the
//               shuffled row w (even samples up, odd samples reversed) at slots j and h-1-j, the second half crossing
//               through LDS; the h-point FFT; the real-FFT recombination of Z[k] with Z[h-k] through LDS; and the
//               half-angle rotation straight into the two outputs each bin feeds. x read once, y written once.
//   k_dct_inv   DCT-III / DST-III: x staged in LDS, the rotated Hermitian half spectrum b[k] = e^{i theta_k}
//               (x[k] - i x[n-k]) folded into the h-point C2R pre-pass, the inverse h-point FFT, and the unshuffle as
//               16-byte stores y[4j..4j+3] = (v[2j], v[n-1-2j], v[2j+1], v[n-2-2j]).
//   k_dct_n2    n = 2.
//   k_dct_*     the separate passes around the library's R2C / C2R plans for longer n.
//
// Tables (FP64-built, narrowed): cs[k] = (cos, sin)(pi k / 2n) for k <= h, rec[k] = W_n^k for k < h.
#include "buffer.hpp"
#include "dct.hpp"
#include "fft_core.hpp"

#include <algorithm>

namespace crdfft
{
using namespace dev;

namespace
{

constexpr int round_up(int v, int m) { return (v + m - 1) / m * m; }

// One h-point core per slot, M slots per workgroup; unlike k_rows every slot has an LDS region, since the
// shuffle and the recombination exchange values between threads even where the core itself needs none.
template <int H, int E> struct DctCfg
{
    static constexpr int T = H / E;
    static constexpr int M = T >= 256 ? 1 : 256 / T;
    static constexpr int threads = T * M;
    static constexpr int LF = lds_floats(H);
    static constexpr int RS = round_up(2 * LF, 64) + (T < 64 ? T : 0);
    static constexpr int lds_bytes = RS * M * 4;
};

__device__ __forceinline__ cf conjf2(cf a) { return {a.x, -a.y}; }

template <int H, int E, bool SINE>
__global__ __launch_bounds__((DctCfg<H, E>::threads)) void k_dct_fwd(DctArgs a)
{
    using C = DctCfg<H, E>;
    constexpr int N = 2 * H;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int local = tid / C::T;
    const int i = tid - local * C::T;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * C::M + local;
    const bool active = t < a.count;
    const uint32_t row_bytes = active ? static_cast<uint32_t>(N * sizeof(float)) : 0u;
    const __amdgpu_buffer_rsrc_t rx = make_rsrc(a.in + (active ? t * a.idist : 0), row_bytes);
    const LdsView<1> view{lds + local * C::RS, C::LF};
    cf v[E];
    // x[4j..4j+3] = (w[2j], w[n-1-2j], w[2j+1], w[n-2-2j]) for j < h/2: z[j] = (x0, x2) and z[h-1-j] = (x3, x1),
    // the odd samples negated for DST-II (dct.hpp:114-118)
#pragma unroll
    for (int m = 0; m < E / 2; ++m)
    {
        const int j = i + C::T * m;
        const f4v q = bld128<0>(rx, static_cast<uint32_t>(j) * 16u, 0);
        v[m] = cf{q.x, q.z};
        view.put(H - 1 - j, SINE ? cf{-q.w, -q.y} : cf{q.w, q.y});
    }
    __syncthreads();
#pragma unroll
    for (int m = E / 2; m < E; ++m)
    {
        v[m] = view.get(i + C::T * m);
    }
    __syncthreads();
    fft_core<H, E, false, 1>(v, i, view, reinterpret_cast<const cf*>(a.tw));
    // Z[k] beside Z[(h-k) mod h]: the rfft of w is W[k] = E[k] + W_n^k O[k] (dct.hpp:74-95 takes it from the rfft)
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        view.put(i + C::T * m, v[m]);
    }
    __syncthreads();
    int i2 = i;
    asm volatile("" : "+v"(i2));
    const __amdgpu_buffer_rsrc_t ry = make_rsrc(a.out + (active ? t * a.odist : 0), row_bytes);
    const __amdgpu_buffer_rsrc_t rc = make_rsrc(a.cs, static_cast<uint32_t>((H + 1) * sizeof(cf)));
    const __amdgpu_buffer_rsrc_t rr = make_rsrc(a.rec, static_cast<uint32_t>(H * sizeof(cf)));
    const cf csh = bld<0>(rc, 0, H * sizeof(cf)); // (cos, sin)(pi / 4)
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const int k = i2 + C::T * m;
        const cf z = v[m];
        const cf zp = conjf2(view.get((H - k) & (H - 1)));
        const cf A = cf{z.x + zp.x, z.y + zp.y};
        const cf B = cf{z.x - zp.x, z.y - zp.y};
        const cf rb = cmul(B, bld<0>(rr, static_cast<uint32_t>(k) * 8u, 0));
        const cf w2 = cf{A.x + rb.y, A.y - rb.x}; // 2 W[k]
        const cf cs = bld<0>(rc, static_cast<uint32_t>(k) * 8u, 0);
        const float wh = 2.0f * (z.x - z.y);      // 2 W[h] (real), used by slot 0 only
        uint32_t k1, k2;
        float y1, y2;
        if constexpr (!SINE)
        {
            // y[k] = c Re 2W + s Im 2W; y[n-k] = s Re 2W - c Im 2W; slot 0 also y[h] = cos(pi/4) 2W[h]
            k1 = static_cast<uint32_t>(k);
            y1 = cs.x * w2.x + cs.y * w2.y;
            k2 = k > 0 ? static_cast<uint32_t>(N - k) : static_cast<uint32_t>(H);
            y2 = k > 0 ? cs.y * w2.x - cs.x * w2.y : csh.x * wh;
        }
        else
        {
            // y[k-1] = s Re 2W - c Im 2W; y[n-1-k] = c Re 2W + s Im 2W; slot 0: y[n-1] = 2W[0], y[h-1] = sin(pi/4) 2W[h]
            k1 = k > 0 ? static_cast<uint32_t>(k - 1) : static_cast<uint32_t>(N - 1);
            y1 = k > 0 ? cs.y * w2.x - cs.x * w2.y : w2.x;
            k2 = k > 0 ? static_cast<uint32_t>(N - 1 - k) : static_cast<uint32_t>(H - 1);
            y2 = k > 0 ? cs.x * w2.x + cs.y * w2.y : csh.y * wh;
        }
        bst32<0>(ry, k1 * 4u, 0, y1);
        bst32<0>(ry, k2 * 4u, 0, y2);
    }
}

template <int H, int E, bool SINE>
__global__ __launch_bounds__((DctCfg<H, E>::threads)) void k_dct_inv(DctArgs a)
{
    using C = DctCfg<H, E>;
    constexpr int N = 2 * H;
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int local = tid / C::T;
    const int i = tid - local * C::T;
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * C::M + local;
    const bool active = t < a.count;
    const uint32_t row_bytes = active ? static_cast<uint32_t>(N * sizeof(float)) : 0u;
    const __amdgpu_buffer_rsrc_t rx = make_rsrc(a.in + (active ? t * a.idist : 0), row_bytes);
    const LdsView<1> view{lds + local * C::RS, C::LF};
    // the row as pairs (x[2j], x[2j+1]) in the slot-j cells of the LDS region
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const int j = i + C::T * m;
        view.put(j, bld<0>(rx, static_cast<uint32_t>(j) * 8u, 0));
    }
    __syncthreads();
    // x'[q]: the row itself for DCT-III, reversed for DST-III (dst3's b[j] takes x[n-1-j] and x[j-1], dct.hpp:160-165)
    auto xr = [&](int q) {
        const int p = SINE ? N - 1 - q : q;
        return view.re[pad16(p >> 1) + ((p & 1) ? view.im_off : 0)];
    };
    const __amdgpu_buffer_rsrc_t rc = make_rsrc(a.cs, static_cast<uint32_t>((H + 1) * sizeof(cf)));
    const __amdgpu_buffer_rsrc_t rr = make_rsrc(a.rec, static_cast<uint32_t>(H * sizeof(cf)));
    cf v[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
    {
        const int k = i + C::T * m;
        // b[k] = e^{i theta_k} (x'[k] - i x'[n-k]), b[0] = x'[0] (twice dct.hpp's G, its final x2 folded in)
        const float xa = xr(k);
        const float xb = xr(k > 0 ? N - k : 0);
        const cf c1 = bld<0>(rc, static_cast<uint32_t>(k) * 8u, 0);
        const cf bk = k > 0 ? cf{c1.x * xa + c1.y * xb, c1.y * xa - c1.x * xb} : cf{xa, 0.0f};
        // b[h-k], 1 <= h-k <= h: x'[h-k] and x'[h+k]
        const float xc = xr(H - k);
        const float xd = xr(H + k);
        const cf c2 = bld<0>(rc, static_cast<uint32_t>(H - k) * 8u, 0);
        const cf bh = cf{c2.x * xc + c2.y * xd, -(c2.y * xc - c2.x * xd)}; // conj b[h-k]
        // Z[k] = (b[k] + conj b[h-k]) + i W_n^{-k} (b[k] - conj b[h-k]): z = IDFT_h(Z) = v[2j] + i v[2j+1]
        const cf A = cf{bk.x + bh.x, bk.y + bh.y};
        const cf B = cf{bk.x - bh.x, bk.y - bh.y};
        const cf rb = cmulc(B, bld<0>(rr, static_cast<uint32_t>(k) * 8u, 0));
        v[m] = cf{A.x - rb.y, A.y + rb.x};
    }
    __syncthreads(); // every read of the staged row is done before the core writes the region
    int i2 = i;
    asm volatile("" : "+v"(i2));
    fft_core<H, E, true, 1>(v, i2, view, reinterpret_cast<const cf*>(a.tw));
    // y[2k] = v[k], y[2k+1] = v[n-1-k] (DST-III: -v[n-1-k]); slot j < h/2 stores y[4j..4j+3] with z[h-1-j]
#pragma unroll
    for (int m = E / 2; m < E; ++m)
    {
        view.put(i2 + C::T * m, v[m]);
    }
    __syncthreads();
    const __amdgpu_buffer_rsrc_t ry = make_rsrc(a.out + (active ? t * a.odist : 0), row_bytes);
    const float sg = SINE ? -1.0f : 1.0f;
#pragma unroll
    for (int m = 0; m < E / 2; ++m)
    {
        const int j = i2 + C::T * m;
        const cf zq = view.get(H - 1 - j);
        f4v o;
        o.x = v[m].x;
        o.y = sg * zq.y;
        o.z = v[m].y;
        o.w = sg * zq.x;
        bst128<0>(ry, static_cast<uint32_t>(j) * 16u, 0, o);
    }
}

// n = 2: DCT-II (2(x0+x1), 2c(x0-x1)), DST-II (2s(x0+x1), 2(x0-x1)), DCT-III (x0 + 2c x1, x0 - 2c x1),
// DST-III (x1 + 2s x0, -x1 + 2s x0), c = cos(pi/4), s = sin(pi/4) from the table.
__global__ void k_dct_n2(DctArgs a, int kind)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= a.count)
    {
        return;
    }
    const float x0 = a.in[t * a.idist];
    const float x1 = a.in[t * a.idist + 1];
    const cf cs = reinterpret_cast<const cf*>(a.cs)[1];
    float y0, y1;
    switch (kind)
    {
        case kDct2: y0 = 2.0f * (x0 + x1); y1 = 2.0f * cs.x * (x0 - x1); break;
        case kDst2: y0 = 2.0f * cs.y * (x0 + x1); y1 = 2.0f * (x0 - x1); break;
        case kDct3: y0 = x0 + 2.0f * cs.x * x1; y1 = x0 - 2.0f * cs.x * x1; break;
        default: y0 = x1 + 2.0f * cs.y * x0; y1 = -x1 + 2.0f * cs.y * x0; break;
    }
    a.out[t * a.odist] = y0;
    a.out[t * a.odist + 1] = y1;
}

// cs[k] = (cos, sin)(pi k / 2n), k <= h; rec[k] = W_n^k, k < h; angles reduced in FP64 before sincospi
__global__ void k_dct_tables(cf* cs, cf* rec, uint64_t n)
{
    const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t h = n / 2;
    if (k <= h)
    {
        double sn, c;
        sincospi(static_cast<double>(k) / static_cast<double>(2 * n), &sn, &c);
        cs[k] = cf{static_cast<float>(c), static_cast<float>(sn)};
    }
    if (k < h)
    {
        // W_n^k = e^{-2 pi i k/n}; k < n/2 keeps the angle in [0, pi)
        double sn, c;
        sincospi(2.0 * static_cast<double>(k) / static_cast<double>(n), &sn, &c);
        rec[k] = cf{static_cast<float>(c), static_cast<float>(-sn)};
    }
}

// ---- separate passes (n > kDctOneKernelMax, or on request) ----

// w[j] = x[2j], w[n-1-j] = +-x[2j+1], j < h
__global__ void k_dct_shuffle(const float* x, float* w, uint64_t n, uint64_t count, uint64_t idist, int sine)
{
    const uint64_t h = n / 2;
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= h * count)
    {
        return;
    }
    const uint64_t t = idx / h;
    const uint64_t j = idx - t * h;
    const float e = x[t * idist + 2 * j];
    const float o = x[t * idist + 2 * j + 1];
    w[t * n + j] = e;
    w[t * n + n - 1 - j] = sine ? -o : o;
}

// From the rfft spectrum W[0..h] of the shuffled row (dct.hpp:86-95, :133-148):
//   DCT-II: y[k] = 2(c_k Re W[k] + s_k Im W[k]) for k <= h, y[n-k] = 2(s_k Re W[k] - c_k Im W[k]) for 0 < k < h
//   DST-II: y[k-1] = 2(s_k Re W[k] - c_k Im W[k]) for 0 < k <= h, y[n-1-k] = 2(c_k Re W[k] + s_k Im W[k]) for
//           0 < k < h, y[n-1] = 2 Re W[0]
__global__ void k_dct_fwd_post(const cf* spec, float* y, const cf* cs, uint64_t n, uint64_t count, uint64_t odist,
                               int sine)
{
    const uint64_t h = n / 2;
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= (h + 1) * count)
    {
        return;
    }
    const uint64_t t = idx / (h + 1);
    const uint64_t k = idx - t * (h + 1);
    const cf w = spec[t * (h + 1) + k];
    const cf c = cs[k];
    float* row = y + t * odist;
    if (!sine)
    {
        row[k] = 2.0f * (c.x * w.x + c.y * w.y);
        if (k > 0 && k < h)
        {
            row[n - k] = 2.0f * (c.y * w.x - c.x * w.y);
        }
    }
    else if (k == 0)
    {
        row[n - 1] = 2.0f * w.x;
    }
    else
    {
        row[k - 1] = 2.0f * (c.y * w.x - c.x * w.y);
        if (k < h)
        {
            row[n - 1 - k] = 2.0f * (c.x * w.x + c.y * w.y);
        }
    }
}

// spec[k] = e^{i theta_k} (x'[k] - i x'[n-k]) for 0 < k <= h, spec[0] = x'[0] (x' = x, or x reversed for DST-III)
__global__ void k_dct_inv_pre(const float* x, cf* spec, const cf* cs, uint64_t n, uint64_t count, uint64_t idist,
                              int sine)
{
    const uint64_t h = n / 2;
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= (h + 1) * count)
    {
        return;
    }
    const uint64_t t = idx / (h + 1);
    const uint64_t k = idx - t * (h + 1);
    const float* row = x + t * idist;
    auto xr = [&](uint64_t q) { return row[sine ? n - 1 - q : q]; };
    cf b;
    if (k == 0)
    {
        b = cf{xr(0), 0.0f};
    }
    else
    {
        const float xa = xr(k);
        const float xb = xr(n - k);
        const cf c = cs[k];
        b = cf{c.x * xa + c.y * xb, c.y * xa - c.x * xb};
    }
    spec[t * (h + 1) + k] = b;
}

// y[2k] = n v[k], y[2k+1] = +-n v[n-1-k] (v = irfft of the spectrum, which carries 1/n)
__global__ void k_dct_inv_post(const float* v, float* y, uint64_t n, uint64_t count, uint64_t odist, int sine)
{
    const uint64_t h = n / 2;
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= h * count)
    {
        return;
    }
    const uint64_t t = idx / h;
    const uint64_t k = idx - t * h;
    const float sc = static_cast<float>(n);
    y[t * odist + 2 * k] = sc * v[t * n + k];
    y[t * odist + 2 * k + 1] = (sine ? -sc : sc) * v[t * n + n - 1 - k];
}

// (h, E): values per thread of the h-point core, measured per length on gfx950 over 2^24-sample batches (small E
// keeps the LDS and registers per thread low, and these kernels exchange through LDS three times per row)
#define CRDFFT_DCT_LIST(X)                                                                                             \
    X(2, 2) X(4, 4) X(8, 4) X(16, 4) X(32, 4) X(64, 4) X(128, 4) X(256, 4) X(512, 8) X(1024, 8) X(2048, 8) X(4096, 8) \
        X(8192, 8) X(16384, 16)

constexpr uint64_t kMaxItems = 0xFFFFFFFFull;

template <typename K> hipError_t set_lds(K kernel, unsigned bytes)
{
    if (bytes > 64 * 1024)
    {
        return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(bytes));
    }
    return hipSuccess;
}

unsigned blocks_for(uint64_t items) { return static_cast<unsigned>((items + 255) / 256); }

} // namespace

int dct_one_kernel_e(uint64_t n)
{
#define X(H, E)                                                                                                        \
    if (n == 2 * static_cast<uint64_t>(H))                                                                             \
        return E;
    CRDFFT_DCT_LIST(X)
#undef X
    return 0;
}

hipError_t launch_dct_one_kernel(uint64_t n, int e, int kind, const DctArgs& a, hipStream_t s)
{
#define X(H, E)                                                                                                        \
    if (n == 2 * static_cast<uint64_t>(H) && e == E)                                                                   \
    {                                                                                                                  \
        using C = DctCfg<H, E>;                                                                                        \
        void (*kern)(DctArgs) = nullptr;                                                                               \
        switch (kind)                                                                                                  \
        {                                                                                                              \
            case kDct2: kern = k_dct_fwd<H, E, false>; break;                                                          \
            case kDst2: kern = k_dct_fwd<H, E, true>; break;                                                           \
            case kDct3: kern = k_dct_inv<H, E, false>; break;                                                          \
            case kDst3: kern = k_dct_inv<H, E, true>; break;                                                           \
            default: return hipErrorInvalidValue;                                                                      \
        }                                                                                                              \
        hipError_t err = set_lds(kern, C::lds_bytes);                                                                  \
        if (err != hipSuccess)                                                                                         \
            return err;                                                                                                \
        const uint64_t max_groups = kMaxItems / C::threads;                                                            \
        for (uint64_t done = 0; done < a.count;)                                                                       \
        {                                                                                                              \
            DctArgs part = a;                                                                                          \
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
    CRDFFT_DCT_LIST(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_dct_n2(int kind, const DctArgs& a, hipStream_t s)
{
    for (uint64_t done = 0; done < a.count;)
    {
        DctArgs part = a;
        part.count = std::min<uint64_t>(a.count - done, kMaxItems - 255);
        part.in = a.in + done * a.idist;
        part.out = a.out + done * a.odist;
        hipLaunchKernelGGL(k_dct_n2, dim3(blocks_for(part.count)), dim3(256), 0, s, part, kind);
        done += part.count;
    }
    return hipGetLastError();
}

hipError_t build_dct_tables(cfloat* cs, cfloat* rec, uint64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_dct_tables, dim3(blocks_for(n / 2 + 1)), dim3(256), 0, s, reinterpret_cast<cf*>(cs),
                       reinterpret_cast<cf*>(rec), n);
    return hipGetLastError();
}

hipError_t launch_dct_shuffle(const float* x, float* w, uint64_t n, uint64_t count, uint64_t idist, bool sine,
                              hipStream_t s)
{
    const uint64_t items = n / 2 * count;
    if (items == 0)
        return hipSuccess;
    if (items > kMaxItems)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dct_shuffle, dim3(blocks_for(items)), dim3(256), 0, s, x, w, n, count, idist, sine ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_dct_fwd_post(const cfloat* spec, float* y, const cfloat* cs, uint64_t n, uint64_t count,
                               uint64_t odist, bool sine, hipStream_t s)
{
    const uint64_t items = (n / 2 + 1) * count;
    if (items == 0)
        return hipSuccess;
    if (items > kMaxItems)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dct_fwd_post, dim3(blocks_for(items)), dim3(256), 0, s, reinterpret_cast<const cf*>(spec), y,
                       reinterpret_cast<const cf*>(cs), n, count, odist, sine ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_dct_inv_pre(const float* x, cfloat* spec, const cfloat* cs, uint64_t n, uint64_t count,
                              uint64_t idist, bool sine, hipStream_t s)
{
    const uint64_t items = (n / 2 + 1) * count;
    if (items == 0)
        return hipSuccess;
    if (items > kMaxItems)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dct_inv_pre, dim3(blocks_for(items)), dim3(256), 0, s, x, reinterpret_cast<cf*>(spec),
                       reinterpret_cast<const cf*>(cs), n, count, idist, sine ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_dct_inv_post(const float* v, float* y, uint64_t n, uint64_t count, uint64_t odist, bool sine,
                               hipStream_t s)
{
    const uint64_t items = n / 2 * count;
    if (items == 0)
        return hipSuccess;
    if (items > kMaxItems)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_dct_inv_post, dim3(blocks_for(items)), dim3(256), 0, s, v, y, n, count, odist,
                       sine ? 1 : 0);
    return hipGetLastError();
}

} // namespace crdfft
