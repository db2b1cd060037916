// The LDS-staged Stockham core every power-of-two kernel of this library runs for one sequence of N points.
//
// Thread i of the T = N/E threads that own a sequence holds E values in registers, always the slots
// i + T*m (m < E): that is the layout it loads, and after the last stage it holds X[i + T*m], so global loads and
// stores are unit-stride across lanes. Each stage (sub-DFT size P grown to P*R, radix R dividing E) runs E/R
// radix-R butterflies per thread: butterfly b = i + q*T reads slots b + j*N/R (registers v[q + j*E/R]), multiplies
// by W_{PR}^{j (b mod P)} from the stage's twiddle table, runs the register DFT and scatters to
// (b - b mod P)*R + (b mod P) + j*P, which for the last stage (P*R == N) is the thread's own slots again. Between
// stages the values cross through LDS: split re/im float arrays, one pad float every 16, so the stride-R scatter of
// the first stages spreads over all banks.
//
// This is the Stockham autosort of the reference's CPU engine (engine/numerics/hesap-fft/kernels/
// fft_kernels.cpp:266-353, radix-4..32 passes on split buffers with per-pass twiddle tables, fft.hpp:355-403) laid
// out for wave64: one radix-E register pass per stage instead of one memory pass, and the ping-pong buffers are LDS.
#pragma once

#include "dft.hpp"

namespace crdfft::dev
{

constexpr int ilog2c(long n)
{
    int l = 0;
    while ((1L << l) < n)
    {
        ++l;
    }
    return l;
}

// One pad float after every 16 (bank spreading for the stride-16 scatter).
__device__ __forceinline__ int pad16(int o) { return o + (o >> 4); }
constexpr int lds_floats(int n) { return n + (n >> 4) + 1; }

// Stage geometry of an N-point transform at E values per thread: a first stage of radix R0 (the bits of log2 N
// that do not fill a radix-E stage, else E) and radix-E stages after it.
template <int N, int E> struct Geometry
{
    static_assert((N & (N - 1)) == 0 && (E & (E - 1)) == 0 && E <= N, "pow2 N, pow2 E <= N");
    static constexpr int L = ilog2c(N);
    static constexpr int e = ilog2c(E);
    static constexpr int T = N / E;
    static constexpr int R0 = (e == 0) ? 1 : ((L % e) != 0 ? (1 << (L % e)) : E);
    // twiddle entries of every stage after the first: stage with P takes P*R entries W_{PR}^{j k} at [k*R + j]
    static constexpr long twiddle_entries()
    {
        long total = 0;
        for (long p = R0; p < N; p *= E)
        {
            total += p * E;
        }
        return total;
    }
};

// Strided LDS view: element o of the sequence lives at base + pad16(o) * STRIDE in the re array and at the same
// offset in the im array (im = re + im_off).
template <int STRIDE> struct LdsView
{
    float* re;
    int im_off;
    __device__ __forceinline__ void put(int o, cf v) const
    {
        const int a = pad16(o) * STRIDE;
        re[a] = v.x;
        re[a + im_off] = v.y;
    }
    __device__ __forceinline__ cf get(int o) const
    {
        const int a = pad16(o) * STRIDE;
        return {re[a], re[a + im_off]};
    }
};

template <int N, int E, int P, bool INV, int STRIDE> struct Stage
{
    using G = Geometry<N, E>;
    static constexpr int R = (P == 1) ? G::R0 : E;
    static constexpr int Q = E / R;
    static constexpr int T = G::T;
    static constexpr bool kLast = (static_cast<long>(P) * R == N);

    static __device__ __forceinline__ void run(cf (&v)[E], int i, const LdsView<STRIDE>& lds, const cf* __restrict__ tw)
    {
#pragma unroll
        for (int q = 0; q < Q; ++q)
        {
            const int b = i + q * T;
            const int k = b & (P - 1);
            cf a[R];
#pragma unroll
            for (int j = 0; j < R; ++j)
            {
                a[j] = v[q + j * Q];
            }
            if constexpr (P > 1)
            {
                const cf* w = tw + k * R;
#pragma unroll
                for (int j = 1; j < R; ++j)
                {
                    a[j] = ctw<INV>(a[j], w[j]);
                }
            }
            Dft<R, INV>::run(a);
            if constexpr (kLast)
            {
#pragma unroll
                for (int j = 0; j < R; ++j)
                {
                    v[q + j * Q] = a[j];
                }
            }
            else
            {
                const int base = (b - k) * R + k;
#pragma unroll
                for (int j = 0; j < R; ++j)
                {
                    lds.put(base + j * P, a[j]);
                }
            }
        }
        if constexpr (!kLast)
        {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < E; ++m)
            {
                v[m] = lds.get(i + T * m);
            }
            __syncthreads();
            Stage<N, E, P * R, INV, STRIDE>::run(v, i, lds, tw + (P > 1 ? P * R : 0));
        }
    }
};

// The whole transform: v holds x[i + T*m] on entry and X[i + T*m] on exit. Every thread of the workgroup must
// call it (it synchronises the workgroup between stages). `tw` is the table of the stages after the first.
template <int N, int E, bool INV, int STRIDE>
__device__ __forceinline__ void fft_core(cf (&v)[E], int i, const LdsView<STRIDE>& lds, const cf* __restrict__ tw)
{
    Stage<N, E, 1, INV, STRIDE>::run(v, i, lds, tw);
}

} // namespace crdfft::dev
