// Register-resident small DFTs for gfx950 kernels: complex helpers, compile-time roots of unity and the radix-R
// butterflies (R = 2..64, powers of two, plus 3, 5 and 7 for the 7-smooth sizes) that every FFT kernel of this
// library composes. All loops have compile-time trip counts, so after unrolling every value lives in a VGPR.
//
// Sign convention (reference engine/numerics/hesap-fft/include/crd/hesap/fft/fft.hpp:16, :251-260): forward uses
// W_n^k = exp(-2 pi i k / n); INV uses the conjugate; neither direction scales.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace crdfft::dev
{

struct alignas(8) cf
{
    float x, y;
};

__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
// a * w
__device__ __forceinline__ cf cmul(cf a, cf w) { return {__fmaf_rn(a.x, w.x, -a.y * w.y), __fmaf_rn(a.x, w.y, a.y * w.x)}; }
// a * conj(w)
__device__ __forceinline__ cf cmulc(cf a, cf w) { return {__fmaf_rn(a.x, w.x, a.y * w.y), __fmaf_rn(a.y, w.x, -a.x * w.y)}; }
template <bool INV> __device__ __forceinline__ cf ctw(cf a, cf w) { return INV ? cmulc(a, w) : cmul(a, w); }

// The C2R pre-rotation of one bin (real_fft.hpp:70-93): from X[k], X[h-k] and W_n^k, scale * Z[k] of the half-length
// sequence whose inverse FFT holds the reals as z[j] = x[2j] + i x[2j+1]:
//   Z = E + i O,  E = (X[k] + conj X[h-k]) / 2,  O = W_n^{-k} (X[k] - conj X[h-k]) / 2.
__device__ __forceinline__ cf c2r_z(cf xk, cf xh, cf wk, float scale)
{
    const cf e = {(xk.x + xh.x) * 0.5f, (xk.y - xh.y) * 0.5f};
    const cf d = {(xk.x - xh.x) * 0.5f, (xk.y + xh.y) * 0.5f};
    const cf o = cmulc(d, wk); // W^{-k} D
    return {(e.x - o.y) * scale, (e.y + o.x) * scale};
}
// a * (-i) forward, a * (+i) inverse: multiplication by W_4^1
template <bool INV> __device__ __forceinline__ cf mul_w4(cf a) { return INV ? cf{-a.y, a.x} : cf{a.y, -a.x}; }

// ---- compile-time trigonometry (FP64, octant reduction + Taylor series) ----
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x)
{
    double term = x, sum = x;
    for (int k = 1; k < 14; ++k)
    {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}

constexpr double taylor_cos(double x)
{
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k)
    {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}

// cos and sin of 2 pi e / n for integers, reduced exactly to [0, pi/4] by octant symmetry.
struct cs
{
    double c, s;
};

constexpr cs root(long e, long n)
{
    e %= n;
    if (e < 0)
    {
        e += n;
    }
    // angle = 2 pi e / n = (pi/4) * (8 e / n); octant index and remainder in units of n
    const long e8 = 8 * e;
    const long oct = e8 / n;
    const long rem = e8 - oct * n; // in [0, n)
    // within the octant: t = (pi/4) rem / n ; mirror odd octants
    double c = 0, s = 0;
    const bool odd = (oct & 1) != 0;
    const double t = kPi / 4.0 * static_cast<double>(odd ? (n - rem) : rem) / static_cast<double>(n);
    const double ct = taylor_cos(t), st = taylor_sin(t);
    switch (oct)
    {
        case 0: c = ct; s = st; break;
        case 1: c = st; s = ct; break;
        case 2: c = -st; s = ct; break;
        case 3: c = -ct; s = st; break;
        case 4: c = -ct; s = -st; break;
        case 5: c = -st; s = -ct; break;
        case 6: c = st; s = -ct; break;
        default: c = ct; s = -st; break;
    }
    return {c, s};
}

// W_n^e = exp(-2 pi i e / n) as a float pair.
template <long E, long N> struct W
{
    static constexpr cs v = root(E, N);
    static constexpr float re = static_cast<float>(v.c);
    static constexpr float im = static_cast<float>(-v.s);
};

// a * W_N^E (conjugated when INV), with the trivial angles resolved at compile time.
template <long E, long N, bool INV> __device__ __forceinline__ cf twc(cf a)
{
    constexpr long e = ((E % N) + N) % N;
    if constexpr (e == 0)
    {
        return a;
    }
    else if constexpr (4 * e == N)
    {
        return mul_w4<INV>(a);
    }
    else if constexpr (2 * e == N)
    {
        return {-a.x, -a.y};
    }
    else if constexpr (4 * e == 3 * N)
    {
        return mul_w4<!INV>(a);
    }
    else if constexpr (8 * e == N || 8 * e == 3 * N || 8 * e == 5 * N || 8 * e == 7 * N)
    {
        // (+-1 +- i)/sqrt2: two adds and two multiplies
        constexpr float h = 0.70710678118654752440f;
        constexpr float wr = W<e, N>::re;
        constexpr float wi = INV ? -W<e, N>::im : W<e, N>::im;
        constexpr float sr = wr > 0 ? 1.0f : -1.0f;
        constexpr float si = wi > 0 ? 1.0f : -1.0f;
        // (a.x + i a.y)(sr + i si) h
        return {h * (sr * a.x - si * a.y), h * (sr * a.y + si * a.x)};
    }
    else
    {
        constexpr float wr = W<e, N>::re;
        constexpr float wi = INV ? -W<e, N>::im : W<e, N>::im;
        return {__fmaf_rn(a.x, wr, -a.y * wi), __fmaf_rn(a.x, wi, a.y * wr)};
    }
}

// ---- the butterflies: DFT of R values in place, natural order in and out ----
template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<1, INV>
{
    static __device__ __forceinline__ void run(cf*) {}
};

template <bool INV> struct Dft<2, INV>
{
    static __device__ __forceinline__ void run(cf* a)
    {
        const cf t = a[1];
        a[1] = csub(a[0], t);
        a[0] = cadd(a[0], t);
    }
};

template <bool INV> struct Dft<4, INV>
{
    static __device__ __forceinline__ void run(cf* a)
    {
        const cf s0 = cadd(a[0], a[2]), d0 = csub(a[0], a[2]);
        const cf s1 = cadd(a[1], a[3]), d1 = mul_w4<INV>(csub(a[1], a[3]));
        a[0] = cadd(s0, s1);
        a[2] = csub(s0, s1);
        a[1] = cadd(d0, d1);
        a[3] = csub(d0, d1);
    }
};

// Odd radices 3, 5, 7: the direct O(R^2) form with constant roots (real and imaginary symmetric pairs).
template <int R, bool INV> struct DftOdd
{
    // a_J W^{JK} + a_{R-J} W^{-JK} added into acc, with the root folded at compile time
    template <int K, int J> static __device__ __forceinline__ void pair(cf& acc, const cf* a)
    {
        constexpr cs r = root(static_cast<long>(J) * K, R);
        constexpr float c = static_cast<float>(r.c);
        constexpr float s = INV ? static_cast<float>(r.s) : static_cast<float>(-r.s);
        const cf p = cadd(a[J], a[R - J]);
        const cf m = csub(a[J], a[R - J]);
        acc.x += p.x * c - m.y * s;
        acc.y += p.y * c + m.x * s;
    }
    template <int K, int... J> static __device__ __forceinline__ cf out_k(const cf* a, std::integer_sequence<int, J...>)
    {
        cf acc = a[0];
        (pair<K, J + 1>(acc, a), ...);
        return acc;
    }
    template <int... K> static __device__ __forceinline__ void all(cf* a, std::integer_sequence<int, K...>)
    {
        const cf out[R] = {out_k<K>(a, std::make_integer_sequence<int, R / 2>{})...};
#pragma unroll
        for (int k = 0; k < R; ++k)
        {
            a[k] = out[k];
        }
    }
    static __device__ __forceinline__ void run(cf* a) { all(a, std::make_integer_sequence<int, R>{}); }
};

template <bool INV> struct Dft<3, INV> : DftOdd<3, INV>
{
};
template <bool INV> struct Dft<5, INV> : DftOdd<5, INV>
{
};
template <bool INV> struct Dft<7, INV> : DftOdd<7, INV>
{
};

// R = R1 * R2 in registers: R2 DFTs of size R1 over x[R2 n1 + n2], twiddle W_R^{n2 k1}, R1 DFTs of size R2,
// output X[k1 + R1 k2].
template <int R1, int R2, bool INV> struct DftSplit
{
    static constexpr int R = R1 * R2;
    template <int N2, int K1> static __device__ __forceinline__ cf tw(cf a) { return twc<static_cast<long>(N2) * K1, R, INV>(a); }

    template <int N2, int... K1> static __device__ __forceinline__ void twrow(cf* b, std::integer_sequence<int, K1...>)
    {
        ((b[K1] = tw<N2, K1>(b[K1])), ...);
    }
    template <int... N2> static __device__ __forceinline__ void twall(cf (*t)[R1], std::integer_sequence<int, N2...>)
    {
        (twrow<N2>(t[N2], std::make_integer_sequence<int, R1>{}), ...);
    }

    static __device__ __forceinline__ void run(cf* a)
    {
        cf t[R2][R1];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2)
        {
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1)
            {
                t[n2][n1] = a[R2 * n1 + n2];
            }
            Dft<R1, INV>::run(t[n2]);
        }
        twall(t, std::make_integer_sequence<int, R2>{});
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1)
        {
            cf u[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2)
            {
                u[n2] = t[n2][k1];
            }
            Dft<R2, INV>::run(u);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2)
            {
                a[k1 + R1 * k2] = u[k2];
            }
        }
    }
};

template <bool INV> struct Dft<8, INV> : DftSplit<2, 4, INV>
{
};
template <bool INV> struct Dft<16, INV> : DftSplit<4, 4, INV>
{
};
template <bool INV> struct Dft<32, INV> : DftSplit<8, 4, INV>
{
};
template <bool INV> struct Dft<64, INV> : DftSplit<8, 8, INV>
{
};
template <bool INV> struct Dft<6, INV> : DftSplit<2, 3, INV>
{
};
template <bool INV> struct Dft<9, INV> : DftSplit<3, 3, INV>
{
};
template <bool INV> struct Dft<10, INV> : DftSplit<2, 5, INV>
{
};
template <bool INV> struct Dft<12, INV> : DftSplit<4, 3, INV>
{
};
template <bool INV> struct Dft<14, INV> : DftSplit<2, 7, INV>
{
};
template <bool INV> struct Dft<15, INV> : DftSplit<3, 5, INV>
{
};

} // namespace crdfft::dev
