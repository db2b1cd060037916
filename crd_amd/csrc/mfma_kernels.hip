// Split-precision MFMA passes of the four-step FFT of n = 256 x 256 (= 16^4): every radix-16 stage is a complex
// 16x16 DFT-matrix product on gfx950's v_mfma_f32_16x16x32_{f16,bf16}, accumulated in FP32; twiddles stay in FP32.
//
// A stage contracts X[n1][n2] (16 x 16 complex, one 256-point sequence) over n1 with F[k][n1] = W_16^{k n1}. Real
// and imaginary parts are stacked along K = 32: Re Y = [Fr | -Fi] [Xr; Xi], Im Y = [Fi | Fr] [Xr; Xi]. Operand
// layouts (checked on the device): lane l holds A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16], j < 8, and the result
// D[4(l/16)+r][l%16], r < 4. Operands are split so that the products keep FP32-class accuracy:
//   FP16 mode: x = s * value (s a power of two per sequence and stage: the largest |value| maps near 2^14),
//              x ~ xh + xl, F ~ Fh + Fl, product Fh xh + Fh xl + Fl xh (3 MFMAs per real product), divided by s;
//   BF16 mode: x ~ xh + xm + xl, F ~ Fh + Fm + Fl (bf16 keeps FP32's exponent range: no scaling), six cross
//              products down to the 2^-24 level.
// The DFT-matrix splits are built in FP64 on the device (sincospi) once per plan (k_split_tables).
//
// Pass 0 (columns): tiles of 16 columns x 256 rows, then the four-step twiddle W_N^{col k}; pass 1 (rows): 16 rows
// of 256, stored transposed (out[row + 256 k]). One workgroup = 4 waves = 16 sequences; each wave runs 4.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dft.hpp"
#include "mfma.hpp"

namespace crdfft
{
using namespace dev;

namespace
{

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kSeq = 256;              // points per sequence
constexpr int kPitch = kSeq + 16;      // LDS floats per sequence (pad: spreads the 16 sequences over banks)
constexpr int kTileSeqs = 16;          // sequences per workgroup
constexpr int kThreads = 256;

// Operand tables, per lane: [term][matrix re/im][8 values], as float (converted to the mode's type in the kernel).
// term count: FP16 2 (hi, lo), BF16 3 (hi, mid, lo); matrix 0 = [Fr | -Fi], 1 = [Fi | Fr] (forward; the inverse
// negates every Fi).
__global__ void k_split_tables(float* tab, int mode, int inverse)
{
    const int l = threadIdx.x; // 64 lanes
    const int i = l % 16;
    for (int j = 0; j < 8; ++j)
    {
        const int k = 8 * (l / 16) + j;
        const int c = k % 16;
        double sn, cs;
        // W_16^{i c} = cos(2 pi i c / 16) - i sin(...); forward Fi = -sin, inverse Fi = +sin
        sincospi(2.0 * static_cast<double>((i * c) % 16) / 16.0, &sn, &cs);
        const double fr = cs;
        const double fi = inverse ? sn : -sn;
        const double a0 = k < 16 ? fr : -fi; // [Fr | -Fi]
        const double a1 = k < 16 ? fi : fr;  // [Fi | Fr]
        const double vals[2] = {a0, a1};
        for (int mtx = 0; mtx < 2; ++mtx)
        {
            double v = vals[mtx];
            if (mode == kSplitFp16)
            {
                const _Float16 h = static_cast<_Float16>(v);
                const _Float16 lo = static_cast<_Float16>(v - static_cast<double>(h));
                tab[((0 * 2 + mtx) * 64 + l) * 8 + j] = static_cast<float>(h);
                tab[((1 * 2 + mtx) * 64 + l) * 8 + j] = static_cast<float>(lo);
            }
            else
            {
                const __bf16 h = static_cast<__bf16>(static_cast<float>(v));
                const double r1 = v - static_cast<double>(static_cast<float>(h));
                const __bf16 m = static_cast<__bf16>(static_cast<float>(r1));
                const double r2 = r1 - static_cast<double>(static_cast<float>(m));
                const __bf16 lo = static_cast<__bf16>(static_cast<float>(r2));
                tab[((0 * 2 + mtx) * 64 + l) * 8 + j] = static_cast<float>(h);
                tab[((1 * 2 + mtx) * 64 + l) * 8 + j] = static_cast<float>(m);
                tab[((2 * 2 + mtx) * 64 + l) * 8 + j] = static_cast<float>(lo);
            }
        }
    }
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
    {
        v = fmaxf(v, __shfl_xor(v, o));
    }
    return v;
}

// Power of two s with max * s in [2^13, 2^14) (1 for an all-zero sequence or a non-finite max).
__device__ __forceinline__ float fp16_scale(float maxabs)
{
    if (!(maxabs > 0.0f) || !(maxabs < 3.0e38f))
    {
        return 1.0f;
    }
    int e;
    frexpf(maxabs, &e); // maxabs = f * 2^e, f in [0.5, 1)
    return ldexpf(1.0f, 14 - e);
}

// D_re, D_im of one stage: B (8 floats per lane, already scaled) against the lane's A tables.
template <int MODE>
__device__ __forceinline__ void stage_mfma(const float (&b)[8], const float* __restrict__ tab, floatx4& dre,
                                           floatx4& dim)
{
    dre = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    dim = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const int l = threadIdx.x & 63;
    if constexpr (MODE == kSplitFp16)
    {
        half8 bh, bl, a0h, a0l, a1h, a1l;
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            const _Float16 h = static_cast<_Float16>(b[j]);
            bh[j] = h;
            bl[j] = static_cast<_Float16>(b[j] - static_cast<float>(h));
            a0h[j] = static_cast<_Float16>(tab[((0 * 2 + 0) * 64 + l) * 8 + j]);
            a1h[j] = static_cast<_Float16>(tab[((0 * 2 + 1) * 64 + l) * 8 + j]);
            a0l[j] = static_cast<_Float16>(tab[((1 * 2 + 0) * 64 + l) * 8 + j]);
            a1l[j] = static_cast<_Float16>(tab[((1 * 2 + 1) * 64 + l) * 8 + j]);
        }
        // small terms first: they add into an accumulator still near zero
        dre = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0l, bh, dre, 0, 0, 0);
        dre = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0h, bl, dre, 0, 0, 0);
        dre = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0h, bh, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1l, bh, dim, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, bl, dim, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1h, bh, dim, 0, 0, 0);
    }
    else
    {
        bf16x8 bh, bm, bl, ah[2], am[2], al[2];
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            const __bf16 h = static_cast<__bf16>(b[j]);
            const float r1 = b[j] - static_cast<float>(h);
            const __bf16 m = static_cast<__bf16>(r1);
            const float r2 = r1 - static_cast<float>(m);
            bh[j] = h;
            bm[j] = m;
            bl[j] = static_cast<__bf16>(r2);
#pragma unroll
            for (int mtx = 0; mtx < 2; ++mtx)
            {
                ah[mtx][j] = static_cast<__bf16>(tab[((0 * 2 + mtx) * 64 + l) * 8 + j]);
                am[mtx][j] = static_cast<__bf16>(tab[((1 * 2 + mtx) * 64 + l) * 8 + j]);
                al[mtx][j] = static_cast<__bf16>(tab[((2 * 2 + mtx) * 64 + l) * 8 + j]);
            }
        }
        floatx4* d[2] = {&dre, &dim};
#pragma unroll
        for (int mtx = 0; mtx < 2; ++mtx)
        {
            floatx4 acc = *d[mtx];
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[mtx], bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[mtx], bm, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mtx], bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[mtx], bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mtx], bm, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[mtx], bh, acc, 0, 0, 0);
            *d[mtx] = acc;
        }
    }
}

// One 256-point sequence (in LDS at re[0..256), im at +im_off) through both radix-16 stages; on exit the LDS holds
// X[k] at index k. Called by one whole wave; `tw` holds W_256^e.
template <int MODE, bool INV>
__device__ __forceinline__ void seq256(float* re, int im_off, const float* __restrict__ tab, const cf* __restrict__ tw)
{
    const int l = threadIdx.x & 63;
    const int col = l % 16;     // n2 (stage A) / k1 (stage B)
    const int kb = 8 * ((l / 16) & 1); // first of the lane's 8 contraction indices within a half
    const bool imag_half = (l / 16) >= 2;
    // ---- stage A: contract n1; lane reads X[n1 = kb + j][n2 = col] (re or im), index 16 n1 + n2
    float b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        b[j] = re[(imag_half ? im_off : 0) + 16 * (kb + j) + col];
    }
    float s = 1.0f;
    if constexpr (MODE == kSplitFp16)
    {
        float m = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            m = fmaxf(m, fabsf(b[j]));
        }
        s = fp16_scale(wave_max(m));
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            b[j] *= s;
        }
    }
    floatx4 dre, dim;
    stage_mfma<MODE>(b, tab, dre, dim);
    const float inv_s = 1.0f / s;
    // lane holds Y[k1 = 4(l/16) + r][n2 = col]; twiddle W_256^{n2 k1}; write Y at 16 k1 + n2
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        const int k1 = 4 * (l / 16) + r;
        cf y{dre[r] * inv_s, dim[r] * inv_s};
        y = ctw<INV>(y, tw[(col * k1) & 255]);
        re[16 * k1 + col] = y.x;
        re[im_off + 16 * k1 + col] = y.y;
    }
    __builtin_amdgcn_wave_barrier();
    // ---- stage B: contract n2; lane reads Y[k1 = col][n2 = kb + j]
#pragma unroll
    for (int j = 0; j < 8; ++j)
    {
        b[j] = re[(imag_half ? im_off : 0) + 16 * col + kb + j];
    }
    s = 1.0f;
    if constexpr (MODE == kSplitFp16)
    {
        float m = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            m = fmaxf(m, fabsf(b[j]));
        }
        s = fp16_scale(wave_max(m));
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            b[j] *= s;
        }
    }
    stage_mfma<MODE>(b, tab, dre, dim);
    const float inv_t = 1.0f / s;
    __builtin_amdgcn_wave_barrier();
    // lane holds Z^T[k2 = 4(l/16) + r][k1 = col] = X[k1 + 16 k2]
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        const int k2 = 4 * (l / 16) + r;
        re[col + 16 * k2] = dre[r] * inv_t;
        re[im_off + col + 16 * k2] = dim[r] * inv_t;
    }
    __builtin_amdgcn_wave_barrier();
}

template <int PASS, int MODE, bool INV> __global__ __launch_bounds__(kThreads) void k_split(SplitArgs a)
{
    __shared__ float lds[2 * kTileSeqs * kPitch];
    float* lre = lds;
    const int im_off = kTileSeqs * kPitch;
    const int tid = threadIdx.x;
    const uint64_t groups = 16; // 256 columns (rows) / 16 per workgroup
    const uint64_t t = blockIdx.x / groups;
    const uint64_t g = blockIdx.x - t * groups;
    if (t >= a.count)
    {
        return; // uniform per workgroup
    }
    const cf* src = reinterpret_cast<const cf*>(a.in) + t * a.idist;
    // ---- load the tile into LDS: sequence q at lre[q * kPitch + n]
    if constexpr (PASS == 0)
    {
        // columns 16g + q of the 256 x 256 row-major matrix: element n of column c at n * 256 + c
        const int q = tid % 16;
        for (int m = 0; m < 16; ++m)
        {
            const int n = tid / 16 + 16 * m;
            const cf v = src[static_cast<uint64_t>(n) * 256 + g * 16 + q];
            lre[q * kPitch + n] = v.x;
            lre[im_off + q * kPitch + n] = v.y;
        }
    }
    else
    {
        for (int m = 0; m < 16; ++m)
        {
            const cf v = src[(g * 16 + m) * 256 + tid];
            lre[m * kPitch + tid] = v.x;
            lre[im_off + m * kPitch + tid] = v.y;
        }
    }
    __syncthreads();
    const int wave = tid / 64;
    const cf* tw = reinterpret_cast<const cf*>(a.tw256);
    for (int q = wave; q < kTileSeqs; q += kThreads / 64)
    {
        seq256<MODE, INV>(lre + q * kPitch, im_off, a.tab, tw);
    }
    __syncthreads();
    cf* dst = reinterpret_cast<cf*>(a.out) + t * a.odist;
    if constexpr (PASS == 0)
    {
        // X[k] of column c = 16g + q, times W_N^{c k}, stored at k * 256 + c (same layout)
        const int q = tid % 16;
        const uint64_t c = g * 16 + q;
        const cf* thi = reinterpret_cast<const cf*>(a.thi);
        const cf* tlo = reinterpret_cast<const cf*>(a.tlo);
        const uint64_t lmask = (uint64_t{1} << a.hshift) - 1;
        for (int m = 0; m < 16; ++m)
        {
            const int k = tid / 16 + 16 * m;
            cf v{lre[q * kPitch + k], lre[im_off + q * kPitch + k]};
            const uint64_t e = c * static_cast<uint64_t>(k);
            v = ctw<INV>(v, cmul(thi[e >> a.hshift], tlo[e & lmask]));
            dst[static_cast<uint64_t>(k) * 256 + c] = v;
        }
    }
    else
    {
        // row r = 16g + q: X[k] stored at r + 256 k (transposed), 16 rows side by side per k
        const int q = tid % 16;
        for (int m = 0; m < 16; ++m)
        {
            const int k = tid / 16 + 16 * m;
            dst[g * 16 + q + static_cast<uint64_t>(k) * 256] = cf{lre[q * kPitch + k], lre[im_off + q * kPitch + k]};
        }
    }
}

} // namespace

uint64_t split_table_floats() { return 3u * 2u * 64u * 8u; }

hipError_t build_split_tables(float* tab, int mode, bool inverse, hipStream_t s)
{
    hipLaunchKernelGGL(k_split_tables, dim3(1), dim3(64), 0, s, tab, mode, inverse ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_split_pass(int pass, int mode, bool inverse, const SplitArgs& a, hipStream_t s)
{
    const uint64_t blocks = a.count * 16;
    if (blocks == 0)
    {
        return hipSuccess;
    }
    if (blocks > 0x7FFFFFFFull)
    {
        return hipErrorInvalidValue;
    }
    const dim3 grid(static_cast<unsigned>(blocks)), block(kThreads);
#define L(P, M, I) hipLaunchKernelGGL((k_split<P, M, I>), grid, block, 0, s, a)
    if (mode == kSplitFp16)
    {
        if (pass == 0)
        {
            if (inverse)
                L(0, kSplitFp16, true);
            else
                L(0, kSplitFp16, false);
        }
        else
        {
            if (inverse)
                L(1, kSplitFp16, true);
            else
                L(1, kSplitFp16, false);
        }
    }
    else
    {
        if (pass == 0)
        {
            if (inverse)
                L(0, kSplitBf16, true);
            else
                L(0, kSplitBf16, false);
        }
        else
        {
            if (inverse)
                L(1, kSplitBf16, true);
            else
                L(1, kSplitBf16, false);
        }
    }
#undef L
    return hipGetLastError();
}

} // namespace crdfft
