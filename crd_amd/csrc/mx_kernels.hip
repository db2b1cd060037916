// Split-precision MFMA passes of the four-step FFT: radix-R stages as complex R x R DFT-matrix products on gfx950's
// matrix cores, accumulated in FP32, with the sequence kept in one wave's registers from load to store.
//
// A sequence of S = R*R points (R = 16: S = 256 on v_mfma_f32_16x16x32_{f16,bf16}; R = 32: S = 1024 on
// v_mfma_f32_32x32x16_{f16,bf16}) is the R x R matrix X[a][b], point n = R*a + b. Stage A forms
// Y[b][k1] = sum_a X[a][b] W_R^{a k1} for every b and multiplies it by W_S^{b k1}; stage B forms
// Z[k1][k2] = sum_b Y[b][k1] W_R^{b k2}, and X_out[k1 + R*k2] = Z[k1][k2]. Each stage is one matrix product
// D = A * B with A = the data (rows: the index kept, K: the contracted index, real parts then imaginary parts
// stacked along K) and B = [Fr; -Fi] for the real part of the result, [Fi; Fr] for the imaginary part.
//
// Layouts (the MFMA operand and result layouts of gfx950): the result D[i][j] of a product sits in lane l, register
// r at j = l % R, i = row(l, r) (R = 16: 4*(l/16) + r; R = 32: 8*(r/4) + 4*(l/32) + r%4). A sequence is loaded into
// that same layout (i = a, j = b), so the data enter stage A and leave each stage in the result layout. An A
// operand wants lane l to hold A[l % R][k] for eight consecutive k of its lane group; the move from the result
// layout to the operand layout is a fixed exchange between lane groups with the same l % R, done in registers with
// v_permlane32_swap (and, for R = 16, v_permlane16_swap): no LDS and no barrier inside a sequence.
//
// Split products (the accuracy of FP32 from 11- or 8-bit matrix inputs):
//   FP16: per sequence and stage the operand is scaled by a power of two s (largest |value| near 2^14), split
//         x = xh + xl, the table F = Fh + Fl, and D = (xh Fh + xh Fl + xl Fh) / s (three products);
//   BF16: x = xh + xm + xl and F = Fh + Fm + Fl (bf16 keeps FP32's exponent range: no scaling), six products down
//         to the 2^-24 level.
// The DFT-matrix splits are built once per plan in FP64 on the device (k_mx_tables).
//
// Passes: PASS 0 transforms C columns of the four-step's n1 x n2 matrix (tile staged through LDS so global reads
// are C*8-byte row runs) and multiplies by W_n^{col k}; PASS 1 transforms C rows (read straight into the register
// layout) and stores them transposed, out[row + n1 k], through an LDS tile; PASS 2 is a whole transform of S
// points per sequence (n = S), stored through LDS in point order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "buffer.hpp"
#include "dft.hpp"
#include "mfma.hpp"

namespace crdfft
{
using namespace dev;

namespace
{

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256; // 4 waves per workgroup

template <int R> struct Mx;

template <> struct Mx<16>
{
    static constexpr int S = 256;
    static constexpr int P = 4;      // complex values per lane
    static constexpr int CHUNKS = 1; // K = 2R = 32 per product: one 16x16x32 instruction
    static constexpr int C = 16;     // sequences per workgroup tile
    // LDS pitch of a tile sequence (complex): 2 spare points so the C-column row writes do not collide; PASS 2
    // stages point k at k + k/R, so its pitch holds S + S/R
    static constexpr int SP = S + 2;
    static constexpr int SP2 = S + S / 16;
    using acc_t = floatx4;
    __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
    // row(l, r) = row_lane(l) + row_reg(r)
    __device__ static int row_lane(int l) { return 4 * (l >> 4); }
    static constexpr int row_reg(int r) { return r; }
    // the 16-point blocks of every odd 64-point block swapped in pairs, so a wave's two 64-point halves (lane groups
    // 0 and 1 read points 64g + 16r + c) land on different banks without growing the tile
    __device__ static int pad(int o) { return o ^ (((o >> 6) & 1) << 4); }
    // K index of operand value v of chunk c in lane l
    __device__ static int kidx(int l, int /*c*/, int v) { return 8 * (l >> 4) + v; }
};

template <> struct Mx<32>
{
    static constexpr int S = 1024;
    static constexpr int P = 16;
    static constexpr int CHUNKS = 4; // K = 2R = 64 per product: four 32x32x16 instructions
    static constexpr int C = 8;
    // pitch: PASS 2 stages point k at k + k/R (up to S + S/R - 1); the +4 keeps C-column row writes apart
    static constexpr int SP = S + S / 32 + 4;
    static constexpr int SP2 = SP;
    using acc_t = floatx16;
    __device__ static int row(int l, int r) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
    __device__ static int row_lane(int l) { return 4 * (l >> 5); }
    static constexpr int row_reg(int r) { return 8 * (r >> 2) + (r & 3); }
    __device__ static int pad(int o) { return o; }
    __device__ static int kidx(int l, int c, int v) { return 16 * c + 8 * (l >> 5) + v; }
};

template <int R> constexpr int terms(int mode) { return mode == kSplitFp16 ? 2 : 3; }
static_assert(Mx<16>::SP2 >= Mx<16>::S + Mx<16>::S / 16 && Mx<32>::SP2 >= Mx<32>::S + Mx<32>::S / 32,
              "the point-order staging of PASS 2 fits a sequence's pitch");

// Operand tables, per lane: [term][matrix re/im][chunk][lane][8] 16-bit values (f16 or bf16 by mode, so a lane
// loads its eight as one 128-bit access): matrix 0 = [Fr; -Fi] (gives the real part), 1 = [Fi; Fr] (the imaginary
// part), F = W_R^{a k} (forward: Fi = -sin; the inverse negates every Fi).
template <int R> __global__ void k_mx_tables(uint16_t* tab, int mode, int inverse)
{
    using M = Mx<R>;
    const int l = threadIdx.x; // 64 lanes
    const int j = l % R;
    for (int c = 0; c < M::CHUNKS; ++c)
    {
        for (int v = 0; v < 8; ++v)
        {
            const int k = M::kidx(l, c, v);
            const int a = k % R;
            double sn, cs;
            sincospi(2.0 * static_cast<double>((a * j) % R) / static_cast<double>(R), &sn, &cs);
            const double fr = cs;
            const double fi = inverse ? sn : -sn;
            const double vals[2] = {k < R ? fr : -fi, k < R ? fi : fr};
            for (int mtx = 0; mtx < 2; ++mtx)
            {
                const double x = vals[mtx];
                uint16_t parts[3] = {0, 0, 0};
                if (mode == kSplitFp16)
                {
                    const _Float16 h = static_cast<_Float16>(x);
                    parts[0] = __builtin_bit_cast(uint16_t, h);
                    parts[1] = __builtin_bit_cast(uint16_t, static_cast<_Float16>(x - static_cast<double>(h)));
                }
                else
                {
                    const __bf16 h = static_cast<__bf16>(static_cast<float>(x));
                    const double r1 = x - static_cast<double>(static_cast<float>(h));
                    const __bf16 m = static_cast<__bf16>(static_cast<float>(r1));
                    const double r2 = r1 - static_cast<double>(static_cast<float>(m));
                    parts[0] = __builtin_bit_cast(uint16_t, h);
                    parts[1] = __builtin_bit_cast(uint16_t, m);
                    parts[2] = __builtin_bit_cast(uint16_t, static_cast<__bf16>(static_cast<float>(r2)));
                }
                for (int t = 0; t < terms<R>(mode); ++t)
                {
                    tab[((((t * 2 + mtx) * M::CHUNKS + c) * 64) + l) * 8 + v] = parts[t];
                }
            }
        }
    }
}

// The exponent x of the power of two s = 2^x with max * s in [2^13, 2^14) (0 for an all-zero sequence or a
// non-finite max; at most 126, so s stays a finite float for sequences below 2^-112, which then keep a smaller
// scaled maximum).
__device__ __forceinline__ int fp16_scale_exp(float maxabs)
{
    if (!(maxabs > 0.0f) || !(maxabs < 3.0e38f))
    {
        return 0;
    }
    int e;
    frexpf(maxabs, &e);
    return min(14 - e, 126);
}

__device__ __forceinline__ void swap32(float& x, float& y)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, x), __builtin_bit_cast(uint32_t, y),
                                                    false, false);
    x = __builtin_bit_cast(float, static_cast<uint32_t>(r[0]));
    y = __builtin_bit_cast(float, static_cast<uint32_t>(r[1]));
}

__device__ __forceinline__ void swap16(float& x, float& y)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(uint32_t, x), __builtin_bit_cast(uint32_t, y),
                                                    false, false);
    x = __builtin_bit_cast(float, static_cast<uint32_t>(r[0]));
    y = __builtin_bit_cast(float, static_cast<uint32_t>(r[1]));
}

// Largest value over the wave, in every lane, of values >= 0: their bit patterns order as the values do (and a
// NaN above every finite value), so the reduction runs on integers: DPP within each row of 16 (quad swaps, half-row
// and row mirrors), then the row pairs and halves through the lane swaps (no LDS traffic).
__device__ __forceinline__ float wave_max(float v)
{
    const auto dpp = [](uint32_t x, auto ctrl) {
        return static_cast<uint32_t>(
            __builtin_amdgcn_update_dpp(0, static_cast<int>(x), decltype(ctrl)::value, 0xF, 0xF, true));
    };
    uint32_t u = __builtin_bit_cast(uint32_t, v);
    u = max(u, dpp(u, std::integral_constant<int, 0xB1>{}));  // quad_perm [1,0,3,2]
    u = max(u, dpp(u, std::integral_constant<int, 0x4E>{}));  // quad_perm [2,3,0,1]
    u = max(u, dpp(u, std::integral_constant<int, 0x141>{})); // row_half_mirror
    u = max(u, dpp(u, std::integral_constant<int, 0x140>{})); // row_mirror
    auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false); // rows (0, 0, 2, 2) and (1, 1, 3, 3)
    u = max(static_cast<uint32_t>(r[0]), static_cast<uint32_t>(r[1]));
    r = __builtin_amdgcn_permlane32_swap(u, u, false, false); // halves (lo, lo) and (hi, hi)
    u = max(static_cast<uint32_t>(r[0]), static_cast<uint32_t>(r[1]));
    return __builtin_bit_cast(float, u);
}

// Result layout (re[P], im[P]: i = row(l, r), j = l % R) -> A operand values op[c][v] (K index kidx(l, c, v): the
// real parts of i = 0..R-1, then the imaginary parts). Consumes re and im.
template <int R> __device__ __forceinline__ void to_operand(float (&re)[Mx<R>::P], float (&im)[Mx<R>::P],
                                                            float (&op)[Mx<R>::CHUNKS][8])
{
    if constexpr (R == 16)
    {
        // group g = l/16 holds i = 4g + r; operand group g wants K = 8g..8g+7. After the 32-lane swap the rows of
        // (re, im) are (R0, R1, I0, I1) and (R2, R3, I2, I3); after the 16-lane swap (R0, R2, I0, I2) and
        // (R1, R3, I1, I3): group 0 holds Re i = 0..7, group 1 Re i = 8..15, groups 2, 3 the same of Im.
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
            float x = re[r], y = im[r];
            swap32(x, y);
            swap16(x, y);
            op[0][r] = x;
            op[0][4 + r] = y;
        }
    }
    else
    {
        // half h = l/32 holds i = 8(r/4) + 4h + r%4; chunk c (q = c % 2) wants i = 16q + 8h + v, v < 8: swapping the
        // upper half of register block 2q with the lower half of block 2q + 1 puts i = 16q + 8h + (0..3) in block 2q
        // and i = 16q + 8h + (4..7) in block 2q + 1 of half h.
#pragma unroll
        for (int q = 0; q < 2; ++q)
        {
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                float x = re[8 * q + j], y = re[8 * q + 4 + j];
                swap32(x, y);
                op[q][j] = x;
                op[q][4 + j] = y;
                float u = im[8 * q + j], w = im[8 * q + 4 + j];
                swap32(u, w);
                op[2 + q][j] = u;
                op[2 + q][4 + j] = w;
            }
        }
    }
}

template <int R> __device__ __forceinline__ floatx4 mfma(half8 a, half8 b, floatx4 acc)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
}
template <int R> __device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 acc)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}
template <int R> __device__ __forceinline__ floatx16 mfma(half8 a, half8 b, floatx16 acc)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
}
template <int R> __device__ __forceinline__ floatx16 mfma(bf16x8 a, bf16x8 b, floatx16 acc)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
}

template <int MODE> struct Elem;
template <> struct Elem<kSplitFp16>
{
    using vec = half8;
    using scalar = _Float16;
    static constexpr int T = 2;
};
template <> struct Elem<kSplitBf16>
{
    using vec = bf16x8;
    using scalar = __bf16;
    static constexpr int T = 3;
};

// The B operands of this lane: [term][matrix][chunk], converted from the float tables once per kernel.
template <int R, int MODE> struct Tables
{
    using V = typename Elem<MODE>::vec;
    V b[Elem<MODE>::T][2][Mx<R>::CHUNKS];
    __device__ __forceinline__ void load(const uint16_t* __restrict__ tab)
    {
        const int l = threadIdx.x & 63;
#pragma unroll
        for (int t = 0; t < Elem<MODE>::T; ++t)
        {
#pragma unroll
            for (int m = 0; m < 2; ++m)
            {
#pragma unroll
                for (int c = 0; c < Mx<R>::CHUNKS; ++c)
                {
                    b[t][m][c] = *reinterpret_cast<const V*>(tab + ((((t * 2 + m) * Mx<R>::CHUNKS + c) * 64) + l) * 8);
                }
            }
        }
    }
};

// One stage of NS independent sequences at once (their products interleave, so one sequence's matrix-core latency
// hides behind the others'): re/im in the result layout (i = the index kept, j = l % R = the index transformed away
// next) in, D_re/D_im of (re + i im)^T F out, in the same layout with j the new frequency index. carry[u] is the
// power-of-two exponent sequence u's values already carry (FP16: each stage scales its operands by 2^sx): with
// UNSCALE the outputs are brought back to scale 1 (ldexp by -(carry + sx), exact) and carry is reset, else the
// outputs keep the factor and carry grows by sx (the next stage's scan and scale see the scaled values, so the
// factor costs nothing until the last stage removes it).
template <int R, int MODE, int NS, bool UNSCALE>
__device__ __forceinline__ void stage(float (&re)[NS][Mx<R>::P], float (&im)[NS][Mx<R>::P], int (&carry)[NS],
                                      const Tables<R, MODE>& tb)
{
    using M = Mx<R>;
    using E = Elem<MODE>;
    float op[NS][M::CHUNKS][8];
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
        to_operand<R>(re[u], im[u], op[u]);
    }
    int sx[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
        sx[u] = 0;
        if constexpr (MODE == kSplitFp16)
        {
            float m = 0.0f;
#pragma unroll
            for (int c = 0; c < M::CHUNKS; ++c)
            {
#pragma unroll
                for (int v = 0; v < 8; ++v)
                {
                    m = fmaxf(m, fabsf(op[u][c][v]));
                }
            }
            sx[u] = fp16_scale_exp(wave_max(m));
        }
    }
    typename M::acc_t dre[NS], dim[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
        dre[u] = typename M::acc_t{};
        dim[u] = typename M::acc_t{};
    }
#pragma unroll
    for (int c = 0; c < M::CHUNKS; ++c)
    {
        typename E::vec xh[NS], xm[NS], xl[NS];
#pragma unroll
        for (int u = 0; u < NS; ++u)
        {
            if constexpr (MODE == kSplitFp16)
            {
                const float su = ldexpf(1.0f, sx[u]);
#pragma unroll
                for (int v = 0; v < 8; v += 2)
                {
                    // two values at a time: one packed conversion per part (su is a power of two: x is exact)
                    const float x0 = op[u][c][v] * su;
                    const float x1 = op[u][c][v + 1] * su;
                    const half2v h = __builtin_convertvector((float2v{x0, x1}), half2v);
                    const half2v lo = __builtin_convertvector(
                        (float2v{x0 - static_cast<float>(h[0]), x1 - static_cast<float>(h[1])}), half2v);
                    xh[u][v] = h[0];
                    xh[u][v + 1] = h[1];
                    xl[u][v] = lo[0];
                    xl[u][v + 1] = lo[1];
                }
            }
            else
            {
                // two values at a time: one packed round-to-nearest conversion per level, whose 32-bit word holds
                // both bf16 parts, and their FP32 values read back from it (a bf16 is the top half of its float)
#pragma unroll
                for (int v = 0; v < 8; v += 2)
                {
                    const float x0 = op[u][c][v];
                    const float x1 = op[u][c][v + 1];
                    const bf16x2 h = __builtin_convertvector((float2v{x0, x1}), bf16x2);
                    const uint32_t hb = __builtin_bit_cast(uint32_t, h);
                    const float r0 = x0 - __builtin_bit_cast(float, hb << 16);
                    const float r1 = x1 - __builtin_bit_cast(float, hb & 0xFFFF0000u);
                    const bf16x2 m = __builtin_convertvector((float2v{r0, r1}), bf16x2);
                    const uint32_t mb = __builtin_bit_cast(uint32_t, m);
                    const float l0 = r0 - __builtin_bit_cast(float, mb << 16);
                    const float l1 = r1 - __builtin_bit_cast(float, mb & 0xFFFF0000u);
                    const bf16x2 lo = __builtin_convertvector((float2v{l0, l1}), bf16x2);
                    xh[u][v] = h[0];
                    xh[u][v + 1] = h[1];
                    xm[u][v] = m[0];
                    xm[u][v + 1] = m[1];
                    xl[u][v] = lo[0];
                    xl[u][v + 1] = lo[1];
                }
            }
        }
        // small terms first, into accumulators still near zero; sequences innermost, so consecutive products are
        // independent
        if constexpr (MODE == kSplitFp16)
        {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
            {
#pragma unroll
                for (int u = 0; u < NS; ++u)
                {
                    typename M::acc_t& d = mt == 0 ? dre[u] : dim[u];
                    d = mfma<R>(xl[u], tb.b[0][mt][c], d);
                }
#pragma unroll
                for (int u = 0; u < NS; ++u)
                {
                    typename M::acc_t& d = mt == 0 ? dre[u] : dim[u];
                    d = mfma<R>(xh[u], tb.b[1][mt][c], d);
                }
#pragma unroll
                for (int u = 0; u < NS; ++u)
                {
                    typename M::acc_t& d = mt == 0 ? dre[u] : dim[u];
                    d = mfma<R>(xh[u], tb.b[0][mt][c], d);
                }
            }
        }
        else
        {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
            {
#pragma unroll
                for (int pr = 0; pr < 6; ++pr)
                {
#pragma unroll
                    for (int u = 0; u < NS; ++u)
                    {
                        typename M::acc_t& d = mt == 0 ? dre[u] : dim[u];
                        // (data term, table term): (h, l), (m, m), (l, h), (h, m), (m, h), (h, h)
                        const typename E::vec& x = pr == 1 || pr == 4 ? xm[u] : (pr == 2 ? xl[u] : xh[u]);
                        const int tt = pr == 0 ? 2 : (pr == 1 || pr == 3 ? 1 : 0);
                        d = mfma<R>(x, tb.b[tt][mt][c], d);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
        if constexpr (UNSCALE && MODE == kSplitFp16)
        {
            const int back = -(carry[u] + sx[u]);
#pragma unroll
            for (int r = 0; r < M::P; ++r)
            {
                re[u][r] = ldexpf(dre[u][r], back);
                im[u][r] = ldexpf(dim[u][r], back);
            }
            carry[u] = 0;
        }
        else
        {
#pragma unroll
            for (int r = 0; r < M::P; ++r)
            {
                re[u][r] = dre[u][r];
                im[u][r] = dim[u][r];
            }
            carry[u] += sx[u];
        }
    }
}

// The stage twiddles W_S^{b k1} of this lane (b = row(l, r), k1 = l % R), from the FP64-built table W_S^e:
// R = 16 keeps all four; R = 32 keeps W^{(4h + j) k1} (j < 4) and W^{8 q k1} (q < 4) and multiplies them per use.
template <int R> struct StageTw
{
    cf w[R == 16 ? 4 : 8];
    __device__ __forceinline__ void load(const cf* __restrict__ tw)
    {
        const int l = threadIdx.x & 63;
        const int k1 = l % R;
        if constexpr (R == 16)
        {
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                w[r] = tw[(Mx<16>::row(l, r) * k1) & 255];
            }
        }
        else
        {
            const int h = l >> 5;
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                w[j] = tw[((4 * h + j) * k1) & 1023];
                w[4 + j] = tw[(8 * j * k1) & 1023];
            }
        }
    }
    __device__ __forceinline__ cf at(int r) const
    {
        if constexpr (R == 16)
        {
            return w[r];
        }
        else
        {
            return cmul(w[r & 3], w[4 + (r >> 2)]);
        }
    }
};

// NS whole sequences: re/im hold X[a = row(l, r)][b = l % R] on entry, X_out[k1 + R k2] at k1 = row(l, r),
// k2 = l % R on exit.
template <int R, int MODE, bool INV, int NS>
__device__ __forceinline__ void sequences(float (&re)[NS][Mx<R>::P], float (&im)[NS][Mx<R>::P],
                                          const Tables<R, MODE>& tb, const StageTw<R>& tw)
{
    int carry[NS];
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
        carry[u] = 0;
    }
    stage<R, MODE, NS, false>(re, im, carry, tb); // now i = b, j = k1 (values times 2^carry)
#pragma unroll
    for (int u = 0; u < NS; ++u)
    {
#pragma unroll
        for (int r = 0; r < Mx<R>::P; ++r)
        {
            const cf y = ctw<INV>(cf{re[u][r], im[u][r]}, tw.at(r));
            re[u][r] = y.x;
            im[u][r] = y.y;
        }
    }
    stage<R, MODE, NS, true>(re, im, carry, tb); // now i = k1, j = k2, at scale 1
}

// R = 16: the LDS tile allows four workgroups per CU (four waves per SIMD), so the registers are held to what four
// waves per SIMD can have (128) — except in the row pass, whose prefetched next tile (32 more VGPRs) measured faster
// at three waves per SIMD than spilled at four. R = 32: two workgroups per CU fit the LDS, so every pass is held to
// two waves per SIMD (256 registers, all VGPRs; the whole-transform BF16 pass then spills 8 and still measured 19 %
// faster than at one wave, the others fit).
template <int R, int PASS, int MODE, bool INV>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(R == 16 && PASS != 1 ? 4
                                                                          : R == 32               ? 2
                                                                                                  : 1))) void k_mx(
    SplitArgs a)
{
    using M = Mx<R>;
    constexpr int S = M::S, P = M::P, C = M::C;
    constexpr int NS = R == 16 ? 2 : 1; // sequences transformed together (interleaved products)
    extern __shared__ float lds_raw[];
    cf* lds = reinterpret_cast<cf*>(lds_raw);
    const int tid = threadIdx.x;
    const int l = tid & 63;
    const int wave = tid >> 6;
    // PASS 0: a.n_seq = n2 (columns), PASS 1: a.n_seq = n1 (rows), PASS 2: a.n_seq = 1 and t counts transforms
    const uint64_t groups = PASS == 2 ? 1 : a.n_seq / C;
    const uint64_t t = PASS == 2 ? static_cast<uint64_t>(blockIdx.x) * C : blockIdx.x / groups;
    const uint64_t g = PASS == 2 ? 0 : blockIdx.x - t * groups;
    if (t >= a.count)
    {
        return; // uniform per workgroup
    }
    Tables<R, MODE> tb;
    tb.load(a.tab);
    StageTw<R> tw;
    tw.load(reinterpret_cast<const cf*>(a.tws));
    const int row_l = l % R;

    if constexpr (PASS == 0)
    {
        // the tile: C columns (col0 + q) of the S-row column, element n of column c at n * ncols + c
        const cf* src = reinterpret_cast<const cf*>(a.in) + t * a.idist;
        const uint64_t ncols = a.n_seq;
        const uint64_t col0 = g * C;
        // store round m of thread tid writes point k = k1 + R k2 of column q = tid % C with k1 = m, k2 = tid / C
        // (kThreads / C = R): the threads of a half wave take points k2 apart by one, so their LDS reads
        // (o = R k1 + k2) fall on distinct banks
        constexpr int kStores = S * C / kThreads;
        // byte offsets within one transform (n * 8 bytes, below 4 GiB): a per-lane part and a uniform step per round
        const uint32_t ncols32 = static_cast<uint32_t>(ncols);
        {
            // every load of the tile issued before any is waited for (S*C/kThreads per thread)
            constexpr int L = S * C / kThreads;
            const int q = tid % C;
            const int n0 = tid / C;
            const __amdgpu_buffer_rsrc_t rin = make_rsrc(src);
            const uint32_t lane = (static_cast<uint32_t>(n0) * ncols32 + static_cast<uint32_t>(col0) + q) * 8u;
            const uint32_t step = (kThreads / C) * ncols32 * 8u;
            cf v[L];
#pragma unroll
            for (int m = 0; m < L; ++m)
            {
                v[m] = bld<0>(rin, lane, m * step);
            }
#pragma unroll
            for (int m = 0; m < L; ++m)
            {
                lds[q * M::SP + M::pad(n0 + (kThreads / C) * m)] = v[m];
            }
        }
        __syncthreads();
        constexpr int NQ = C / (kThreads / 64);
#pragma unroll
        for (int u0 = 0; u0 < NQ; u0 += NS)
        {
            float re[NS][P], im[NS][P];
#pragma unroll
            for (int w = 0; w < NS; ++w)
            {
                const int q = wave + (u0 + w) * (kThreads / 64);
#pragma unroll
                for (int r = 0; r < P; ++r)
                {
                    const cf v = lds[q * M::SP + M::pad(R * M::row(l, r) + row_l)];
                    re[w][r] = v.x;
                    im[w][r] = v.y;
                }
            }
            sequences<R, MODE, INV, NS>(re, im, tb, tw);
            // back in place, at o = R k1 + k2 (the store loop reads point k1 + R k2 from there)
#pragma unroll
            for (int w = 0; w < NS; ++w)
            {
                const int q = wave + (u0 + w) * (kThreads / 64);
#pragma unroll
                for (int r = 0; r < P; ++r)
                {
                    lds[q * M::SP + M::pad(R * M::row(l, r) + row_l)] = cf{re[w][r], im[w][r]};
                }
            }
        }
        __syncthreads();
        // X[k] of column c = col0 + q, times W_n^{c k}, stored at k * ncols + c; k = m + R (tid / C), so the byte
        // offset is a per-lane part plus m * ncols * 8
        static_assert(kThreads / C == R, "store round m of a thread is point m + R (tid / C)");
        const int q = tid % C;
        const uint64_t c = col0 + q;
        const __amdgpu_buffer_rsrc_t rout = make_rsrc(reinterpret_cast<cf*>(a.out) + t * a.odist);
        const uint32_t lane_k = (static_cast<uint32_t>(R * (tid / C)) * ncols32 + static_cast<uint32_t>(c)) * 8u;
        const uint32_t kstep = ncols32 * 8u;
        if (a.tfull != nullptr)
        {
            // every twiddle load issued before the first is waited for
            const __amdgpu_buffer_rsrc_t rtw = make_rsrc(a.tfull);
            cf twp[kStores];
#pragma unroll
            for (int m = 0; m < kStores; ++m)
            {
                twp[m] = bld<0>(rtw, lane_k, m * kstep);
            }
#pragma unroll
            for (int m = 0; m < kStores; ++m)
            {
                const cf v = lds[q * M::SP + M::pad(R * m + tid / C)]; // point k = m + R (tid / C)
                bst<0>(rout, lane_k, m * kstep, ctw<INV>(v, twp[m]));
            }
        }
        else if (a.tcm != nullptr)
        {
            // W_n^{c k} for k = m + R (tid / C) as W_n^{c R (tid / C)} (two table gathers, once per thread) times
            // W_n^{c m} from the R-row table (one coalesced load per value from a table the L2 holds)
            const cf* thi = reinterpret_cast<const cf*>(a.thi);
            const cf* tlo = reinterpret_cast<const cf*>(a.tlo);
            const uint64_t lmask = (uint64_t{1} << a.hshift) - 1;
            const uint64_t e0 = c * static_cast<uint64_t>(R * (tid / C));
            const cf w0 = cmul(thi[e0 >> a.hshift], tlo[e0 & lmask]);
            const __amdgpu_buffer_rsrc_t rcm = make_rsrc(a.tcm);
            const uint32_t lane_c = static_cast<uint32_t>(c) * 8u;
            cf twp[kStores];
#pragma unroll
            for (int m = 0; m < kStores; ++m)
            {
                twp[m] = bld<0>(rcm, lane_c, m * kstep);
            }
#pragma unroll
            for (int m = 0; m < kStores; ++m)
            {
                const cf v = lds[q * M::SP + M::pad(R * m + tid / C)];
                bst<0>(rout, lane_k, m * kstep, ctw<INV>(v, cmul(w0, twp[m])));
            }
        }
        else
        {
            const cf* thi = reinterpret_cast<const cf*>(a.thi);
            const cf* tlo = reinterpret_cast<const cf*>(a.tlo);
            const uint64_t lmask = (uint64_t{1} << a.hshift) - 1;
#pragma unroll 4
            for (int m = 0; m < kStores; ++m)
            {
                const int k = m + R * (tid / C);
                const cf v = lds[q * M::SP + M::pad(R * m + tid / C)];
                const uint64_t e = c * static_cast<uint64_t>(k);
                bst<0>(rout, lane_k, m * kstep, ctw<INV>(v, cmul(thi[e >> a.hshift], tlo[e & lmask])));
            }
        }
    }
    else if constexpr (PASS == 1)
    {
        // rows row0 + q of the n1 x S matrix, read straight into the register layout (R consecutive points per
        // lane group: 8R-byte runs). Tile = (transform, group of C rows). At R = 16 the launch is a resident grid
        // that walks the tiles in steps of its size, each workgroup loading its next tile before it transforms the
        // current one; at R = 32 it is one tile per workgroup with no prefetch (a second 1024-point tile in
        // registers would halve the waves per SIMD).
        constexpr bool walk = R == 16;
        constexpr int NQ = C / (kThreads / 64);
        const uint64_t tiles = a.count * groups;
        const uint64_t n1 = a.n_seq;
        const auto load_tile = [&](uint64_t tile, cf (&dst)[NQ][P]) {
            const uint64_t tt = tile / groups;
            const uint32_t row0 = static_cast<uint32_t>(tile - tt * groups) * C;
            if constexpr (R == 16)
            {
                // byte offsets within one transform (n * 8 bytes, below 4 GiB): point R row(l, r) + row_l of row
                // row0 + wave + u (kThreads / 64) splits into the lane's part and a uniform rest
                const __amdgpu_buffer_rsrc_t rin = make_rsrc(reinterpret_cast<const cf*>(a.in) + tt * a.idist);
                const uint32_t lane_in = (static_cast<uint32_t>(wave) * S + R * M::row_lane(l) + row_l) * 8u;
#pragma unroll
                for (int u = 0; u < NQ; ++u)
                {
#pragma unroll
                    for (int r = 0; r < P; ++r)
                    {
                        dst[u][r] =
                            bld<0>(rin, lane_in, ((row0 + u * (kThreads / 64)) * S + R * M::row_reg(r)) * 8u);
                    }
                }
            }
            else
            {
                // R = 32: global loads (the buffer form's schedule costs the BF16 pass 21 spilled VGPRs at two waves
                // per SIMD)
                const cf* src = reinterpret_cast<const cf*>(a.in) + tt * a.idist;
#pragma unroll
                for (int u = 0; u < NQ; ++u)
                {
#pragma unroll
                    for (int r = 0; r < P; ++r)
                    {
                        dst[u][r] = src[(static_cast<uint64_t>(row0) + wave + u * (kThreads / 64)) * S +
                                        R * M::row(l, r) + row_l];
                    }
                }
            }
        };
        // transform, stage through LDS and store one tile held in registers
        const auto process = [&](uint64_t tile, const cf (&cur)[NQ][P]) {
#pragma unroll
            for (int u0 = 0; u0 < NQ; u0 += NS)
            {
                float re[NS][P], im[NS][P];
#pragma unroll
                for (int w = 0; w < NS; ++w)
                {
#pragma unroll
                    for (int r = 0; r < P; ++r)
                    {
                        re[w][r] = cur[u0 + w][r].x;
                        im[w][r] = cur[u0 + w][r].y;
                    }
                }
                sequences<R, MODE, INV, NS>(re, im, tb, tw);
#pragma unroll
                for (int w = 0; w < NS; ++w)
                {
                    const int q = wave + (u0 + w) * (kThreads / 64);
#pragma unroll
                    for (int r = 0; r < P; ++r)
                    {
                        lds[q * M::SP + M::pad(R * M::row(l, r) + row_l)] = cf{re[w][r], im[w][r]};
                    }
                }
            }
            __syncthreads();
            // point k of row row0 + q goes to out[row0 + q + n1 k]: C consecutive complex per k; round m of a thread
            // stores k = m + R (tid / C), at a per-lane byte offset plus (row0 + m n1) * 8
            const uint64_t tt = tile / groups;
            const uint32_t row0 = static_cast<uint32_t>(tile - tt * groups) * C;
            const __amdgpu_buffer_rsrc_t rout = make_rsrc(reinterpret_cast<cf*>(a.out) + tt * a.odist);
            const int q = tid % C;
            constexpr int per = kThreads / C;
            static_assert(per == R, "store round m of a thread is point m + R (tid / C)");
            const uint32_t n1_32 = static_cast<uint32_t>(n1);
            const uint32_t lane_out = (static_cast<uint32_t>(q) + n1_32 * static_cast<uint32_t>(R * (tid / C))) * 8u;
#pragma unroll 4
            for (int m = 0; m < S / per; ++m)
            {
                const cf v = lds[q * M::SP + M::pad(R * m + tid / C)]; // point k1 + R k2 with k1 = m, k2 = tid / C
                bst<0>(rout, lane_out, (row0 + m * n1_32) * 8u, cf{v.x * a.scale, v.y * a.scale});
            }
        };
        uint64_t tile = blockIdx.x;
        cf ta[NQ][P];
        load_tile(tile, ta);
        if constexpr (!walk)
        {
            process(tile, ta); // one tile per workgroup
        }
        else
        {
            // two register tiles in alternation (no copies between them): the next tile's loads are in flight
            // while the current one is transformed and stored; the workgroup-uniform exits keep the barriers matched
            const uint64_t step = gridDim.x;
            cf tb2[NQ][P];
#pragma unroll 1
            for (;;)
            {
                if (tile + step < tiles)
                {
                    load_tile(tile + step, tb2);
                }
                process(tile, ta);
                tile += step;
                if (tile >= tiles)
                {
                    break;
                }
                __syncthreads(); // the tile's LDS reads are done before the next tile's writes
                if (tile + step < tiles)
                {
                    load_tile(tile + step, ta);
                }
                process(tile, tb2);
                tile += step;
                if (tile >= tiles)
                {
                    break;
                }
                __syncthreads();
            }
        }
    }
    else
    {
        // whole transforms t0 + q (n = S): straight loads; the result goes through the wave's LDS sequence so the
        // store is in point order. Layout there: point k at k + k / R (the lanes' stride-R writes spread over the
        // banks, the point-order reads stay nearly contiguous).
        constexpr int NQ = C / (kThreads / 64);
        // R = 32 BF16: the sequence bases as scalars (the wave index read as uniform) and buffer accesses, which
        // frees the registers of 64-bit addresses (measured 6.5 % faster there, slower in the other forms)
        constexpr bool scalar_base = R == 32 && MODE == kSplitBf16;
        const int wv = scalar_base ? __builtin_amdgcn_readfirstlane(wave) : wave;
        cf pre[NQ][P];
#pragma unroll
        for (int u = 0; u < NQ; ++u)
        {
            const uint64_t tq = t + wv + u * (kThreads / 64);
#pragma unroll
            for (int r = 0; r < P; ++r)
            {
                pre[u][r] = cf{0.0f, 0.0f};
            }
            if (tq < a.count) // uniform per wave
            {
                if constexpr (scalar_base)
                {
                    const __amdgpu_buffer_rsrc_t rin = make_rsrc(reinterpret_cast<const cf*>(a.in) + tq * a.idist);
                    const uint32_t lane = static_cast<uint32_t>(R * M::row_lane(l) + row_l) * 8u;
#pragma unroll
                    for (int r = 0; r < P; ++r)
                    {
                        pre[u][r] = bld<0>(rin, lane, R * M::row_reg(r) * 8u);
                    }
                }
                else
                {
                    const cf* rp = reinterpret_cast<const cf*>(a.in) + tq * a.idist;
#pragma unroll
                    for (int r = 0; r < P; ++r)
                    {
                        pre[u][r] = rp[R * M::row(l, r) + row_l];
                    }
                }
            }
        }
#pragma unroll
        for (int u0 = 0; u0 < NQ; u0 += NS)
        {
            if (t + wv + u0 * (kThreads / 64) >= a.count)
            {
                break; // uniform per wave: no sequence of this group exists
            }
            float re[NS][P], im[NS][P];
#pragma unroll
            for (int w = 0; w < NS; ++w)
            {
#pragma unroll
                for (int r = 0; r < P; ++r)
                {
                    re[w][r] = pre[u0 + w][r].x;
                    im[w][r] = pre[u0 + w][r].y;
                }
            }
            sequences<R, MODE, INV, NS>(re, im, tb, tw);
#pragma unroll
            for (int w = 0; w < NS; ++w)
            {
                const int q = wv + (u0 + w) * (kThreads / 64);
                const uint64_t tq = t + q;
                if (tq >= a.count)
                {
                    break; // uniform per wave: the group's last sequences past the batch are not stored
                }
                cf* ws = lds + q * M::SP2;
#pragma unroll
                for (int r = 0; r < P; ++r)
                {
                    const int k = M::row(l, r) + R * row_l;
                    ws[k + k / R] = cf{re[w][r] * a.scale, im[w][r] * a.scale};
                }
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if constexpr (scalar_base)
                {
                    const __amdgpu_buffer_rsrc_t rout = make_rsrc(reinterpret_cast<cf*>(a.out) + tq * a.odist);
#pragma unroll
                    for (int m = 0; m < P; ++m)
                    {
                        const int k = l + 64 * m;
                        bst<0>(rout, static_cast<uint32_t>(l) * 8u, 512u * m, ws[k + k / R]);
                    }
                }
                else
                {
                    cf* dst = reinterpret_cast<cf*>(a.out) + tq * a.odist;
#pragma unroll
                    for (int m = 0; m < P; ++m)
                    {
                        const int k = l + 64 * m;
                        dst[k] = ws[k + k / R];
                    }
                }
            }
        }
    }
}

template <int R, int PASS, int MODE> hipError_t launch_one(bool inverse, const SplitArgs& a, hipStream_t s)
{
    using M = Mx<R>;
    const uint64_t blocks = PASS == 2 ? (a.count + M::C - 1) / M::C : a.count * (a.n_seq / M::C);
    if (blocks == 0)
    {
        return hipSuccess;
    }
    if (blocks > 0x7FFFFFFFull || (PASS != 2 && a.n_seq % M::C != 0))
    {
        return hipErrorInvalidValue;
    }
    const unsigned lds = static_cast<unsigned>(M::C * (PASS == 2 ? M::SP2 : M::SP) * sizeof(cf));
    auto kern = inverse ? k_mx<R, PASS, MODE, true> : k_mx<R, PASS, MODE, false>;
    hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (err != hipSuccess)
    {
        return err;
    }
    uint64_t grid = blocks;
    if constexpr (PASS == 1 && R == 16)
    {
        // the row pass as a resident grid that prefetches its next tile (the kernel walks the tiles in steps of the
        // grid)
        int dev = 0, cus = 0, per_cu = 0;
        err = hipGetDevice(&dev);
        if (err == hipSuccess)
            err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (err == hipSuccess)
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), kThreads,
                                                               lds);
        if (err != hipSuccess)
        {
            return err;
        }
        grid = std::min<uint64_t>(blocks, static_cast<uint64_t>(std::max(1, cus * std::max(1, per_cu))));
    }
    hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(kThreads), lds, s, a);
    return hipGetLastError();
}

template <int R, int PASS> hipError_t launch_mode(int mode, bool inverse, const SplitArgs& a, hipStream_t s)
{
    return mode == kSplitFp16 ? launch_one<R, PASS, kSplitFp16>(inverse, a, s)
                              : launch_one<R, PASS, kSplitBf16>(inverse, a, s);
}

} // namespace

bool split_radix_of(uint64_t len, int* radix)
{
    if (len == 256 || len == 1024)
    {
        *radix = len == 256 ? 16 : 32;
        return true;
    }
    return false;
}

uint64_t split_table_entries(int radix)
{
    return static_cast<uint64_t>(3 * 2 * (radix == 16 ? 1 : 4) * 64 * 8);
}

hipError_t build_split_tables(uint16_t* tab, int radix, int mode, bool inverse, hipStream_t s)
{
    if (radix == 16)
    {
        hipLaunchKernelGGL(k_mx_tables<16>, dim3(1), dim3(64), 0, s, tab, mode, inverse ? 1 : 0);
    }
    else if (radix == 32)
    {
        hipLaunchKernelGGL(k_mx_tables<32>, dim3(1), dim3(64), 0, s, tab, mode, inverse ? 1 : 0);
    }
    else
    {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_split_pass(int pass, int radix, int mode, bool inverse, const SplitArgs& a, hipStream_t s)
{
    if (radix == 16)
    {
        return pass == 0 ? launch_mode<16, 0>(mode, inverse, a, s)
                         : (pass == 1 ? launch_mode<16, 1>(mode, inverse, a, s) : launch_mode<16, 2>(mode, inverse, a, s));
    }
    if (radix == 32)
    {
        return pass == 0 ? launch_mode<32, 0>(mode, inverse, a, s)
                         : (pass == 1 ? launch_mode<32, 1>(mode, inverse, a, s) : launch_mode<32, 2>(mode, inverse, a, s));
    }
    return hipErrorInvalidValue;
}

} // namespace crdfft
