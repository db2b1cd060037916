// Test infrastructure only: the CPU oracle of the batched FFT path (declarations and pinning in oracle.h).
//
// Nothing here is product code: libcrdfft.so never links it. tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load liboracle.so to check the GPU path and to time the CPU baseline.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace
{

using cd = std::complex<double>;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kPi = 3.14159265358979323846264338327950288;

bool is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }

unsigned ilog2(size_t n)
{
    unsigned l = 0;
    while ((size_t{1} << l) < n)
    {
        ++l;
    }
    return l;
}

int resolve_threads(int threads)
{
    if (threads > 0)
    {
        return threads;
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return hw == 0 ? 1 : static_cast<int>(hw);
}

// Runs fn(first, last) over [0, count) in `jobs` contiguous ranges, job j taking [j*count/jobs, (j+1)*count/jobs)
// (reference detail/batch_arms.hpp:59-77); one job runs on the calling thread.
template <typename F> void run_batch(int jobs, size_t count, const F& fn)
{
    if (jobs <= 1 || count < 2)
    {
        fn(size_t{0}, count);
        return;
    }
    if (static_cast<size_t>(jobs) > count)
    {
        jobs = static_cast<int>(count);
    }
    std::vector<std::thread> pool;
    pool.reserve(static_cast<size_t>(jobs));
    for (int j = 0; j < jobs; ++j)
    {
        const size_t first = static_cast<size_t>(j) * count / static_cast<size_t>(jobs);
        const size_t last = static_cast<size_t>(j + 1) * count / static_cast<size_t>(jobs);
        pool.emplace_back([&fn, first, last] { fn(first, last); });
    }
    for (auto& t : pool)
    {
        t.join();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// FP64 truth
// ------------------------------------------------------------------------------------------------------------------

// W_n^e = exp(-2 pi i e / n) for e in [0, n), each entry from its own cos/sin call.
std::vector<cd> roots64(size_t n)
{
    std::vector<cd> w(n);
    for (size_t e = 0; e < n; ++e)
    {
        const double ang = -kTwoPi * static_cast<double>(e) / static_cast<double>(n);
        w[e] = cd(std::cos(ang), std::sin(ang));
    }
    return w;
}

// Radix-2 decimation in time with a bit-reversal first, the method of the reference's execute_reference
// (fft.hpp:417-460), in FP64 with every twiddle read from the full table.
void fft64_pow2(cd* x, size_t n, bool inverse, const std::vector<cd>& w)
{
    if (n <= 1)
    {
        return;
    }
    const unsigned lg = ilog2(n);
    for (size_t i = 0; i < n; ++i)
    {
        size_t r = 0;
        for (unsigned b = 0; b < lg; ++b)
        {
            r |= ((i >> b) & 1U) << (lg - 1U - b);
        }
        if (i < r)
        {
            std::swap(x[i], x[r]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1)
    {
        const size_t half = len >> 1;
        const size_t step = n / len;
        for (size_t base = 0; base < n; base += len)
        {
            for (size_t k = 0; k < half; ++k)
            {
                cd t = w[k * step];
                if (inverse)
                {
                    t = std::conj(t);
                }
                const cd v = t * x[base + k + half];
                const cd a = x[base + k];
                x[base + k] = a + v;
                x[base + k + half] = a - v;
            }
        }
    }
}

size_t smallest_factor(size_t n)
{
    for (size_t p = 2; p * p <= n; ++p)
    {
        if (n % p == 0)
        {
            return p;
        }
    }
    return n;
}

size_t largest_prime_factor(size_t n)
{
    size_t best = 1;
    while (n > 1)
    {
        const size_t p = smallest_factor(n);
        best = std::max(best, p);
        n /= p;
    }
    return best;
}

// Mixed-radix decimation in time: n = p*m with p the smallest prime factor; the p sub-DFTs of x[s + p*t] are
// combined with W_n^{s(k + m q)} read from the top-level table `w` (size top, stride top/n).
void mixed64(const cd* in, size_t stride, cd* out, size_t n, bool inverse, const std::vector<cd>& w, size_t top)
{
    if (n == 1)
    {
        out[0] = in[0];
        return;
    }
    const size_t p = smallest_factor(n);
    const size_t m = n / p;
    for (size_t s = 0; s < p; ++s)
    {
        mixed64(in + s * stride, stride * p, out + s * m, m, inverse, w, top);
    }
    const size_t tws = top / n;
    std::vector<cd> tmp(p);
    for (size_t k = 0; k < m; ++k)
    {
        for (size_t q = 0; q < p; ++q)
        {
            cd acc(0.0, 0.0);
            for (size_t s = 0; s < p; ++s)
            {
                const size_t e = (s * (k + m * q)) % n;
                cd t = w[e * tws];
                if (inverse)
                {
                    t = std::conj(t);
                }
                acc += t * out[s * m + k];
            }
            tmp[q] = acc;
        }
        for (size_t q = 0; q < p; ++q)
        {
            out[k + m * q] = tmp[q];
        }
    }
}

// Bluestein in FP64: X = conj(b) . ((x conj(b)) (*) b) with b[k] = exp(+i pi k^2 / n), the convolution at a pow2
// length M >= 2n-1 (the identity of the reference's bluestein.hpp:1-11). The inverse conjugates in and out.
void bluestein64(cd* x, size_t n, bool inverse)
{
    size_t m = 1;
    while (m < 2 * n - 1)
    {
        m <<= 1;
    }
    const std::vector<cd> wm = roots64(m);
    std::vector<cd> chirp(n);
    for (size_t k = 0; k < n; ++k)
    {
        const size_t kk = static_cast<size_t>((static_cast<unsigned __int128>(k) * k) % (2 * n));
        const double th = kPi * static_cast<double>(kk) / static_cast<double>(n);
        chirp[k] = cd(std::cos(th), std::sin(th));
    }
    std::vector<cd> a(m, cd(0, 0));
    std::vector<cd> b(m, cd(0, 0));
    for (size_t k = 0; k < n; ++k)
    {
        const cd xv = inverse ? std::conj(x[k]) : x[k];
        a[k] = xv * std::conj(chirp[k]);
    }
    b[0] = chirp[0];
    for (size_t k = 1; k < n; ++k)
    {
        b[k] = chirp[k];
        b[m - k] = chirp[k];
    }
    fft64_pow2(a.data(), m, false, wm);
    fft64_pow2(b.data(), m, false, wm);
    for (size_t k = 0; k < m; ++k)
    {
        a[k] *= b[k];
    }
    fft64_pow2(a.data(), m, true, wm);
    const double inv_m = 1.0 / static_cast<double>(m);
    for (size_t k = 0; k < n; ++k)
    {
        cd y = a[k] * inv_m * std::conj(chirp[k]);
        x[k] = inverse ? std::conj(y) : y;
    }
}

// One FP64 transform of any n, with a table cache for the batch driver.
struct Fft64
{
    size_t n = 0;
    std::vector<cd> w;
    std::vector<cd> scratch;
    bool pow2 = false;
    bool blue = false;

    explicit Fft64(size_t size) : n(size)
    {
        pow2 = is_pow2(n);
        blue = !pow2 && n > 1 && largest_prime_factor(n) > 64;
        if (!blue && n > 0)
        {
            w = roots64(n);
        }
        scratch.resize(n);
    }

    void run(cd* x, bool inverse)
    {
        if (n <= 1)
        {
            return;
        }
        if (pow2)
        {
            fft64_pow2(x, n, inverse, w);
        }
        else if (blue)
        {
            bluestein64(x, n, inverse);
        }
        else
        {
            std::copy(x, x + n, scratch.begin());
            mixed64(scratch.data(), 1, x, n, inverse, w, n);
        }
    }
};

// ------------------------------------------------------------------------------------------------------------------
// FP32 restatement of the reference's CPU path
// ------------------------------------------------------------------------------------------------------------------

// The small DFT of a Stockham pass, radix R in {2, 4, 8, 16}, in place on split arrays, forward or inverse.
// Radix 8 and 16 split into two radix-(R/2) halves (even and odd inputs) combined with W_R^q, the constants
// read from a table filled in FP64 and narrowed.
template <int R> struct RootsR
{
    float re[R];
    float im[R];
    RootsR()
    {
        for (int q = 0; q < R; ++q)
        {
            const double ang = -kTwoPi * q / R;
            re[q] = static_cast<float>(std::cos(ang));
            im[q] = static_cast<float>(std::sin(ang));
        }
    }
};

template <int R, bool INV> inline void small_dft(float* ar, float* ai, const float* cr, const float* ci)
{
    if constexpr (R == 1)
    {
        return;
    }
    else if constexpr (R == 2)
    {
        const float xr = ar[0] - ar[1];
        const float xi = ai[0] - ai[1];
        ar[0] += ar[1];
        ai[0] += ai[1];
        ar[1] = xr;
        ai[1] = xi;
    }
    else if constexpr (R == 4)
    {
        const float s0r = ar[0] + ar[2], s0i = ai[0] + ai[2];
        const float d0r = ar[0] - ar[2], d0i = ai[0] - ai[2];
        const float s1r = ar[1] + ar[3], s1i = ai[1] + ai[3];
        const float d1r = ar[1] - ar[3], d1i = ai[1] - ai[3];
        // forward: X1 = d0 - i d1, X3 = d0 + i d1; inverse swaps them
        ar[0] = s0r + s1r;
        ai[0] = s0i + s1i;
        ar[2] = s0r - s1r;
        ai[2] = s0i - s1i;
        if (!INV)
        {
            ar[1] = d0r + d1i;
            ai[1] = d0i - d1r;
            ar[3] = d0r - d1i;
            ai[3] = d0i + d1r;
        }
        else
        {
            ar[1] = d0r - d1i;
            ai[1] = d0i + d1r;
            ar[3] = d0r + d1i;
            ai[3] = d0i - d1r;
        }
    }
    else
    {
        constexpr int H = R / 2;
        float er[H], ei[H], orr[H], oi[H];
        for (int s = 0; s < H; ++s)
        {
            er[s] = ar[2 * s];
            ei[s] = ai[2 * s];
            orr[s] = ar[2 * s + 1];
            oi[s] = ai[2 * s + 1];
        }
        // the halves use every second root of the radix-R table
        float hr[H], hi[H];
        for (int q = 0; q < H; ++q)
        {
            hr[q] = cr[2 * q];
            hi[q] = ci[2 * q];
        }
        small_dft<H, INV>(er, ei, hr, hi);
        small_dft<H, INV>(orr, oi, hr, hi);
        for (int q = 0; q < H; ++q)
        {
            const float wr = cr[q];
            const float wi = INV ? -ci[q] : ci[q];
            const float tr = wr * orr[q] - wi * oi[q];
            const float ti = wr * oi[q] + wi * orr[q];
            ar[q] = er[q] + tr;
            ai[q] = ei[q] + ti;
            ar[q + H] = er[q] - tr;
            ai[q + H] = ei[q] - ti;
        }
    }
}

struct Plan32
{
    size_t n = 0;
    unsigned log2 = 0;
    bool planted = false;
    std::vector<float> tw_re, tw_im; // W_n^k, k < n, from FP64 narrowed (fft.hpp:251-260)
    std::vector<size_t> rev;         // bit reversal (execute_reference)
    // Stockham pass list: (log2 of sub-DFT size before the pass, radix bits)
    std::vector<std::pair<unsigned, unsigned>> passes;
    std::vector<float> ptw_re, ptw_im; // per-pass twiddles in (j, m) order (fft.hpp:355-403)
    // four-step (fft.hpp:859-891): n = n1 * n2, factored twiddles W_n^a = hi[a >> h] * lo[a & (2^h - 1)]
    bool four_step = false;
    size_t n1 = 0, n2 = 0;
    unsigned ftw_h = 0;
    std::vector<float> hi_re, hi_im, lo_re, lo_im;
    std::unique_ptr<Plan32> p1, p2;
    // scratch: one plan per thread, as the reference's (fft.hpp:17-18)
    std::vector<float> re0, im0, re1, im1;
    std::vector<float> tbuf, blk;

    Plan32(size_t size, bool plant) : n(size), planted(plant)
    {
        log2 = ilog2(n);
        tw_re.resize(n);
        tw_im.resize(n);
        for (size_t k = 0; k < n; ++k)
        {
            const double ang = -kTwoPi * static_cast<double>(k) / static_cast<double>(n);
            tw_re[k] = static_cast<float>(std::cos(ang));
            tw_im[k] = static_cast<float>(std::sin(ang)) * (planted ? -1.0f : 1.0f);
        }
        rev.resize(n);
        for (size_t i = 0; i < n; ++i)
        {
            size_t r = 0;
            for (unsigned b = 0; b < log2; ++b)
            {
                r |= ((i >> b) & 1U) << (log2 - 1U - b);
            }
            rev[i] = r;
        }
        // four-step from 2^17 (four_step_planned, fft.hpp:63-66), split per four_step_split (fft.hpp:72-85)
        if (n >= (size_t{1} << 17))
        {
            four_step = true;
            if (n == (size_t{1} << 17) || n == (size_t{1} << 18))
            {
                n1 = 1024;
            }
            else if (n == (size_t{1} << 19))
            {
                n1 = 2048;
            }
            else
            {
                n1 = size_t{1} << ((log2 + 1U) / 2U);
            }
            n2 = n / n1;
            ftw_h = log2 / 2U;
            const size_t m = size_t{1} << ftw_h;
            const size_t nh = n >> ftw_h;
            hi_re.resize(nh);
            hi_im.resize(nh);
            lo_re.resize(m);
            lo_im.resize(m);
            for (size_t j = 0; j < nh; ++j)
            {
                const double ang = -kTwoPi * static_cast<double>(j * m) / static_cast<double>(n);
                hi_re[j] = static_cast<float>(std::cos(ang));
                hi_im[j] = static_cast<float>(std::sin(ang)) * (planted ? -1.0f : 1.0f);
            }
            for (size_t j = 0; j < m; ++j)
            {
                const double ang = -kTwoPi * static_cast<double>(j) / static_cast<double>(n);
                lo_re[j] = static_cast<float>(std::cos(ang));
                lo_im[j] = static_cast<float>(std::sin(ang)) * (planted ? -1.0f : 1.0f);
            }
            p1 = std::make_unique<Plan32>(n1, plant);
            p2 = std::make_unique<Plan32>(n2, plant);
            tbuf.resize(2 * n);
        }
        build_passes();
    }

    // The pass sequence of the reference's Stockham loop (fft_kernels.cpp:266-353): a radix-2 first pass when log2
    // is odd, else radix-4; combine passes of at most 4 bits that never strand a 1-bit remainder; a radix-4 last.
    void build_passes()
    {
        passes.clear();
        if (n <= 1)
        {
            return;
        }
        if (log2 == 1U)
        {
            passes.push_back({0U, 1U});
        }
        else if (log2 == 2U)
        {
            passes.push_back({0U, 2U});
        }
        else
        {
            unsigned t = (log2 & 1U) ? 1U : 2U;
            passes.push_back({0U, t});
            const unsigned rmax = 4U;
            while (t < log2 - 2U)
            {
                const unsigned rem = (log2 - 2U) - t;
                unsigned b = std::min(rmax, rem);
                while (b > 2U && rem - b == 1U)
                {
                    --b;
                }
                passes.push_back({t, b});
                t += b;
            }
            passes.push_back({t, 2U});
        }
        // per-pass twiddles W_n^{m j r}, (j, m) order
        ptw_re.clear();
        ptw_im.clear();
        for (auto [t, b] : passes)
        {
            const size_t ls = size_t{1} << t;
            const size_t radix = size_t{1} << b;
            const size_t r = n >> (t + b);
            for (size_t j = 0; j < ls; ++j)
            {
                for (size_t m = 0; m < radix; ++m)
                {
                    ptw_re.push_back(tw_re[m * j * r]);
                    ptw_im.push_back(tw_im[m * j * r]);
                }
            }
        }
    }

    // One Stockham pass over split arrays holding `lanes` interleaved batch lanes per element: sub-DFTs of size
    // ls = 2^t grow to ls * R. Input element (j*R*r + m*r + k), twiddled by W_n^{m j r}; output element
    // (j*r + q*ls*r + k).
    template <int R, bool INV>
    void pass(const float* xr, const float* xi, float* yr, float* yi, size_t ls, size_t r, size_t lanes,
              const float* twr, const float* twi) const
    {
        static const RootsR<R> roots;
        const size_t rl = r * lanes;
        for (size_t j = 0; j < ls; ++j)
        {
            float wr[R], wi[R];
            for (int m = 0; m < R; ++m)
            {
                wr[m] = twr[j * R + m];
                wi[m] = INV ? -twi[j * R + m] : twi[j * R + m];
            }
            const float* sr = xr + j * R * rl;
            const float* si = xi + j * R * rl;
            float* dr = yr + j * rl;
            float* di = yi + j * rl;
            for (size_t k = 0; k < rl; ++k)
            {
                float ar[R], ai[R];
                for (int m = 0; m < R; ++m)
                {
                    const float vr = sr[m * rl + k];
                    const float vi = si[m * rl + k];
                    if (j == 0)
                    {
                        ar[m] = vr;
                        ai[m] = vi;
                    }
                    else
                    {
                        ar[m] = wr[m] * vr - wi[m] * vi;
                        ai[m] = wr[m] * vi + wi[m] * vr;
                    }
                }
                small_dft<R, INV>(ar, ai, roots.re, roots.im);
                for (int q = 0; q < R; ++q)
                {
                    dr[q * ls * rl + k] = ar[q];
                    di[q * ls * rl + k] = ai[q];
                }
            }
        }
    }

    template <bool INV>
    void run_pass(unsigned b, const float* xr, const float* xi, float* yr, float* yi, size_t ls, size_t r,
                  size_t lanes, const float* twr, const float* twi) const
    {
        switch (b)
        {
            case 1U:
                pass<2, INV>(xr, xi, yr, yi, ls, r, lanes, twr, twi);
                break;
            case 2U:
                pass<4, INV>(xr, xi, yr, yi, ls, r, lanes, twr, twi);
                break;
            case 3U:
                pass<8, INV>(xr, xi, yr, yi, ls, r, lanes, twr, twi);
                break;
            default:
                pass<16, INV>(xr, xi, yr, yi, ls, r, lanes, twr, twi);
                break;
        }
    }

    // Stockham over `lanes` element-major lanes: data[i*lanes + t] (lanes == 1: one contiguous transform).
    void stockham(float* data, size_t lanes, bool inverse)
    {
        const size_t total = n * lanes;
        if (re0.size() < total)
        {
            re0.resize(total);
            im0.resize(total);
            re1.resize(total);
            im1.resize(total);
        }
        float* xr = re0.data();
        float* xi = im0.data();
        float* yr = re1.data();
        float* yi = im1.data();
        for (size_t i = 0; i < total; ++i)
        {
            xr[i] = data[2 * i];
            xi[i] = data[2 * i + 1];
        }
        const float* twr = ptw_re.data();
        const float* twi = ptw_im.data();
        for (auto [t, b] : passes)
        {
            const size_t ls = size_t{1} << t;
            const size_t r = n >> (t + b);
            if (inverse)
            {
                run_pass<true>(b, xr, xi, yr, yi, ls, r, lanes, twr, twi);
            }
            else
            {
                run_pass<false>(b, xr, xi, yr, yi, ls, r, lanes, twr, twi);
            }
            twr += ls << b;
            twi += ls << b;
            std::swap(xr, yr);
            std::swap(xi, yi);
        }
        for (size_t i = 0; i < total; ++i)
        {
            data[2 * i] = xr[i];
            data[2 * i + 1] = xi[i];
        }
    }

    // Bailey's four-step (fft_kernels.cpp:2363-): column sub-FFTs over blocks of columns (element-major batched),
    // the factored twiddle, a store into the transposed intermediate, then row sub-FFTs and a scatter.
    void four_step_run(float* data, bool inverse)
    {
        const size_t block = std::max<size_t>(1, std::min(n2, (size_t{1} << 20) / (8 * n1))); // block_width.hpp
        blk.resize(2 * n1 * block);
        const size_t lmask = (size_t{1} << ftw_h) - 1;
        const float sgn = inverse ? -1.0f : 1.0f;
        for (size_t c0 = 0; c0 < n2; c0 += block)
        {
            const size_t bw = std::min(block, n2 - c0);
            for (size_t i1 = 0; i1 < n1; ++i1)
            {
                std::memcpy(&blk[2 * i1 * bw], &data[2 * (i1 * n2 + c0)], 2 * bw * sizeof(float));
            }
            p1->execute_batched(blk.data(), bw, inverse);
            for (size_t k1 = 0; k1 < n1; ++k1)
            {
                for (size_t c = 0; c < bw; ++c)
                {
                    const size_t a = k1 * (c0 + c);
                    const size_t hi = a >> ftw_h, lo = a & lmask;
                    const float wr = hi_re[hi] * lo_re[lo] - hi_im[hi] * lo_im[lo];
                    const float wi = sgn * (hi_re[hi] * lo_im[lo] + hi_im[hi] * lo_re[lo]);
                    const float vr = blk[2 * (k1 * bw + c)];
                    const float vi = blk[2 * (k1 * bw + c) + 1];
                    tbuf[2 * (k1 * n2 + c0 + c)] = wr * vr - wi * vi;
                    tbuf[2 * (k1 * n2 + c0 + c) + 1] = wr * vi + wi * vr;
                }
            }
        }
        for (size_t k1 = 0; k1 < n1; ++k1)
        {
            float* row = &tbuf[2 * k1 * n2];
            p2->execute(row, inverse);
            for (size_t k2 = 0; k2 < n2; ++k2)
            {
                data[2 * (k1 + n1 * k2)] = row[2 * k2];
                data[2 * (k1 + n1 * k2) + 1] = row[2 * k2 + 1];
            }
        }
    }

    void execute(float* data, bool inverse)
    {
        if (n <= 1)
        {
            return;
        }
        if (four_step)
        {
            four_step_run(data, inverse);
        }
        else
        {
            stockham(data, 1, inverse);
        }
    }

    void execute_batched(float* data, size_t batch, bool inverse)
    {
        if (n <= 1 || batch == 0)
        {
            return;
        }
        stockham(data, batch, inverse);
    }

    void execute_reference(float* data, bool inverse) const
    {
        if (n <= 1)
        {
            return;
        }
        for (size_t i = 0; i < n; ++i)
        {
            const size_t j = rev[i];
            if (i < j)
            {
                std::swap(data[2 * i], data[2 * j]);
                std::swap(data[2 * i + 1], data[2 * j + 1]);
            }
        }
        const float isign = inverse ? -1.0f : 1.0f;
        for (size_t len = 2; len <= n; len <<= 1)
        {
            const size_t half = len >> 1;
            const size_t step = n / len;
            for (size_t base = 0; base < n; base += len)
            {
                size_t tw = 0;
                for (size_t k = 0; k < half; ++k)
                {
                    const float wr = tw_re[tw];
                    const float wi = isign * tw_im[tw];
                    float* a = &data[2 * (base + k)];
                    float* b = &data[2 * (base + k + half)];
                    const float vr = wr * b[0] - wi * b[1];
                    const float vi = wr * b[1] + wi * b[0];
                    const float ar = a[0], ai = a[1];
                    a[0] = ar + vr;
                    a[1] = ai + vi;
                    b[0] = ar - vr;
                    b[1] = ai - vi;
                    tw += step;
                }
            }
        }
    }
};

// Bluestein over a pow2 FP32 plan (bluestein.hpp:32-136), chirp from FP64 narrowed.
struct Bluestein32
{
    size_t n, m;
    std::unique_ptr<Plan32> plan;
    std::vector<float> chirp, bfft, work;

    explicit Bluestein32(size_t size) : n(size), m(1)
    {
        while (m < 2 * n - 1)
        {
            m <<= 1;
        }
        plan = std::make_unique<Plan32>(m, false);
        chirp.resize(2 * n);
        for (size_t k = 0; k < n; ++k)
        {
            const size_t kk = static_cast<size_t>((static_cast<unsigned __int128>(k) * k) % (2 * n));
            const double th = kPi * static_cast<double>(kk) / static_cast<double>(n);
            chirp[2 * k] = static_cast<float>(std::cos(th));
            chirp[2 * k + 1] = static_cast<float>(std::sin(th));
        }
        bfft.assign(2 * m, 0.0f);
        bfft[0] = chirp[0];
        bfft[1] = chirp[1];
        for (size_t k = 1; k < n; ++k)
        {
            bfft[2 * k] = chirp[2 * k];
            bfft[2 * k + 1] = chirp[2 * k + 1];
            bfft[2 * (m - k)] = chirp[2 * k];
            bfft[2 * (m - k) + 1] = chirp[2 * k + 1];
        }
        plan->execute(bfft.data(), false);
        work.resize(2 * m);
    }

    void execute(float* x, bool inverse)
    {
        float* w = work.data();
        for (size_t k = 0; k < n; ++k)
        {
            const float xr = x[2 * k];
            const float xi = inverse ? -x[2 * k + 1] : x[2 * k + 1];
            const float cr = chirp[2 * k], ci = chirp[2 * k + 1];
            w[2 * k] = xr * cr + xi * ci;
            w[2 * k + 1] = xi * cr - xr * ci;
        }
        std::fill(w + 2 * n, w + 2 * m, 0.0f);
        plan->execute(w, false);
        for (size_t k = 0; k < m; ++k)
        {
            const float ar = w[2 * k], ai = w[2 * k + 1];
            const float br = bfft[2 * k], bi = bfft[2 * k + 1];
            w[2 * k] = ar * br - ai * bi;
            w[2 * k + 1] = ar * bi + ai * br;
        }
        plan->execute(w, true);
        const float cm = 1.0f / static_cast<float>(m);
        for (size_t k = 0; k < n; ++k)
        {
            const float cr = chirp[2 * k], ci = chirp[2 * k + 1];
            float yr = (w[2 * k] * cr + w[2 * k + 1] * ci) * cm;
            float yi = (w[2 * k + 1] * cr - w[2 * k] * ci) * cm;
            if (inverse)
            {
                const float s = 1.0f / static_cast<float>(n);
                yr *= s;
                yi = -yi * s;
            }
            x[2 * k] = yr;
            x[2 * k + 1] = yi;
        }
    }
};

// RealFftPlan restatement (real_fft.hpp:18-101): pack, half-size complex plan, Hermitian recombine.
struct Real32
{
    size_t n, h;
    Plan32 half;
    std::vector<float> z, wr, wi;

    explicit Real32(size_t size) : n(size), h(size / 2), half(size / 2, false)
    {
        z.resize(2 * h);
        wr.resize(h + 1);
        wi.resize(h + 1);
        for (size_t k = 0; k <= h; ++k)
        {
            const double ang = -kTwoPi * static_cast<double>(k) / static_cast<double>(n);
            wr[k] = static_cast<float>(std::cos(ang));
            wi[k] = static_cast<float>(std::sin(ang));
        }
    }

    void rfft(const float* x, float* out)
    {
        std::memcpy(z.data(), x, 2 * h * sizeof(float));
        half.execute(z.data(), false);
        for (size_t k = 0; k <= h; ++k)
        {
            const size_t a = k % h, b = (h - k) % h;
            const float zr = z[2 * a], zi = z[2 * a + 1];
            const float nr = z[2 * b], ni = z[2 * b + 1];
            const float er = (zr + nr) * 0.5f;
            const float ei = (zi - ni) * 0.5f;
            const float orr = (zi + ni) * 0.5f;
            const float oii = (nr - zr) * 0.5f;
            out[2 * k] = er + (wr[k] * orr - wi[k] * oii);
            out[2 * k + 1] = ei + (wr[k] * oii + wi[k] * orr);
        }
    }

    void irfft(const float* in, float* x)
    {
        for (size_t k = 0; k < h; ++k)
        {
            const float xr = in[2 * k], xi = in[2 * k + 1];
            const float hr = in[2 * (h - k)], hi = in[2 * (h - k) + 1];
            const float er = (xr + hr) * 0.5f;
            const float ei = (xi - hi) * 0.5f;
            const float dr = (xr - hr) * 0.5f;
            const float di = (xi + hi) * 0.5f;
            const float orr = wr[k] * dr + wi[k] * di;
            const float oii = wr[k] * di - wi[k] * dr;
            z[2 * k] = er - oii;
            z[2 * k + 1] = ei + orr;
        }
        half.execute(z.data(), true);
        const float inv = 1.0f / static_cast<float>(h);
        for (size_t j = 0; j < 2 * h; ++j)
        {
            x[j] = z[j] * inv;
        }
    }
};

uint64_t lcg_next(uint64_t& s)
{
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return s;
}

double lcg_unit(uint64_t& s)
{
    return (static_cast<double>(lcg_next(s) >> 11) / static_cast<double>(1ULL << 53)) * 2.0 - 1.0;
}

} // namespace

struct oracle_plan32
{
    Plan32 plan;
    oracle_plan32(size_t n, bool planted) : plan(n, planted) {}
};

extern "C" {

void oracle_lcg_fill64(double* out, size_t count, uint64_t seed)
{
    uint64_t s = seed;
    for (size_t i = 0; i < count; ++i)
    {
        out[i] = lcg_unit(s);
    }
}

void oracle_lcg_fill32(float* out, size_t count, uint64_t seed)
{
    uint64_t s = seed;
    for (size_t i = 0; i < count; ++i)
    {
        out[i] = static_cast<float>(lcg_unit(s));
    }
}

void oracle_normal_fill32(float* out, size_t count, uint64_t seed)
{
    uint64_t s = seed;
    for (size_t i = 0; i < count; i += 2)
    {
        const double u1 = static_cast<double>((lcg_next(s) >> 11) + 1) / static_cast<double>(1ULL << 53);
        const double u2 = static_cast<double>(lcg_next(s) >> 11) / static_cast<double>(1ULL << 53);
        const double rad = std::sqrt(-2.0 * std::log(u1));
        out[i] = static_cast<float>(rad * std::cos(kTwoPi * u2));
        if (i + 1 < count)
        {
            out[i + 1] = static_cast<float>(rad * std::sin(kTwoPi * u2));
        }
    }
}

void oracle_dft64(const double* in, double* out, size_t n, int inverse)
{
    const double sgn = inverse ? 1.0 : -1.0;
    for (size_t k = 0; k < n; ++k)
    {
        double re = 0.0, im = 0.0;
        for (size_t j = 0; j < n; ++j)
        {
            const double ang = sgn * kTwoPi * static_cast<double>((j * k) % n) / static_cast<double>(n);
            const double c = std::cos(ang), s = std::sin(ang);
            re += in[2 * j] * c - in[2 * j + 1] * s;
            im += in[2 * j] * s + in[2 * j + 1] * c;
        }
        out[2 * k] = re;
        out[2 * k + 1] = im;
    }
}

void oracle_fft64(double* data, size_t n, int inverse)
{
    Fft64 f(n);
    f.run(reinterpret_cast<cd*>(data), inverse != 0);
}

void oracle_fft64_batch(double* data, size_t n, size_t batch, size_t dist, int inverse, int threads)
{
    run_batch(resolve_threads(threads), batch,
              [&](size_t first, size_t last)
              {
                  Fft64 f(n);
                  for (size_t b = first; b < last; ++b)
                  {
                      f.run(reinterpret_cast<cd*>(data + 2 * b * dist), inverse != 0);
                  }
              });
}

void oracle_c2c_f32_to_f64(const float* in, double* out, size_t n, size_t batch, int inverse, int threads)
{
    run_batch(resolve_threads(threads), batch,
              [&](size_t first, size_t last)
              {
                  Fft64 f(n);
                  for (size_t b = first; b < last; ++b)
                  {
                      double* o = out + 2 * b * n;
                      const float* x = in + 2 * b * n;
                      for (size_t i = 0; i < 2 * n; ++i)
                      {
                          o[i] = static_cast<double>(x[i]);
                      }
                      f.run(reinterpret_cast<cd*>(o), inverse != 0);
                  }
              });
}

void oracle_rfft64(const double* x, double* out, size_t n)
{
    std::vector<cd> z(n);
    for (size_t j = 0; j < n; ++j)
    {
        z[j] = cd(x[j], 0.0);
    }
    Fft64 f(n);
    f.run(z.data(), false);
    for (size_t k = 0; k <= n / 2; ++k)
    {
        out[2 * k] = z[k].real();
        out[2 * k + 1] = z[k].imag();
    }
}

void oracle_irfft64(const double* in, double* x, size_t n)
{
    std::vector<cd> z(n);
    for (size_t k = 0; k <= n / 2; ++k)
    {
        z[k] = cd(in[2 * k], in[2 * k + 1]);
    }
    for (size_t k = n / 2 + 1; k < n; ++k)
    {
        z[k] = std::conj(z[n - k]);
    }
    // the bins a real signal forces to be real
    z[0] = cd(z[0].real(), 0.0);
    if (n % 2 == 0)
    {
        z[n / 2] = cd(z[n / 2].real(), 0.0);
    }
    Fft64 f(n);
    f.run(z.data(), true);
    for (size_t j = 0; j < n; ++j)
    {
        x[j] = z[j].real() / static_cast<double>(n);
    }
}

oracle_plan32* oracle_plan32_create(size_t n)
{
    if (!is_pow2(n))
    {
        return nullptr;
    }
    return new oracle_plan32(n, false);
}

oracle_plan32* oracle_plan32_create_planted(size_t n)
{
    if (!is_pow2(n))
    {
        return nullptr;
    }
    return new oracle_plan32(n, true);
}

void oracle_plan32_destroy(oracle_plan32* plan) { delete plan; }

void oracle_plan32_execute(oracle_plan32* plan, float* data, int inverse) { plan->plan.execute(data, inverse != 0); }

void oracle_plan32_execute_batched(oracle_plan32* plan, float* data, size_t batch, int inverse)
{
    plan->plan.execute_batched(data, batch, inverse != 0);
}

void oracle_plan32_execute_reference(oracle_plan32* plan, float* data, int inverse)
{
    plan->plan.execute_reference(data, inverse != 0);
}

void oracle_c2c32_batch(float* data, size_t n, size_t batch, size_t dist, int inverse, int threads)
{
    const bool pow2 = is_pow2(n);
    run_batch(resolve_threads(threads), batch,
              [&](size_t first, size_t last)
              {
                  if (pow2)
                  {
                      Plan32 plan(n, false);
                      for (size_t b = first; b < last; ++b)
                      {
                          plan.execute(data + 2 * b * dist, inverse != 0);
                      }
                  }
                  else
                  {
                      Bluestein32 plan(n);
                      for (size_t b = first; b < last; ++b)
                      {
                          plan.execute(data + 2 * b * dist, inverse != 0);
                      }
                  }
              });
}

void oracle_rfft32_batch(const float* in, float* out, size_t n, size_t batch, int threads)
{
    run_batch(resolve_threads(threads), batch,
              [&](size_t first, size_t last)
              {
                  Real32 plan(n);
                  for (size_t b = first; b < last; ++b)
                  {
                      plan.rfft(in + b * n, out + 2 * b * (n / 2 + 1));
                  }
              });
}

void oracle_irfft32_batch(const float* in, float* out, size_t n, size_t batch, int threads)
{
    run_batch(resolve_threads(threads), batch,
              [&](size_t first, size_t last)
              {
                  Real32 plan(n);
                  for (size_t b = first; b < last; ++b)
                  {
                      plan.irfft(in + 2 * b * (n / 2 + 1), out + b * n);
                  }
              });
}

void oracle_bluestein32(float* data, size_t n, int inverse)
{
    Bluestein32 plan(n);
    plan.execute(data, inverse != 0);
}

int oracle_hardware_threads(void) { return resolve_threads(0); }

} // extern "C"
