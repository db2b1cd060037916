// The planner (planner.hpp): host-only choice of kernels, tiles and chunking.
#include "planner.hpp"

#include <algorithm>

#include "kernels.hpp"

namespace crdfft
{

bool is_pow2(uint64_t n) { return n != 0 && (n & (n - 1)) == 0; }

uint32_t log2_exact(uint64_t n)
{
    uint32_t l = 0;
    while ((uint64_t{1} << l) < n)
    {
        ++l;
    }
    return l;
}

bool have_row_kernel(uint64_t n, int e)
{
    LaunchShape s;
    return row_kernel_shape(n, e, &s);
}

bool have_col_kernel(uint64_t n1, int e, int c)
{
    LaunchShape s;
    return col_kernel_shape(n1, e, c, &s);
}

bool have_rowtr_kernel(uint64_t n2, int e, int r)
{
    LaunchShape s;
    return rowtr_kernel_shape(n2, e, r, &s);
}

bool have_fused_kernel(uint64_t n1, int e1, uint64_t n2, int e2, int threads, unsigned* lds_bytes)
{
    LaunchShape s;
    if (!fused_kernel_shape(n1, e1, n2, e2, threads, &s))
    {
        return false;
    }
    if (lds_bytes != nullptr)
    {
        *lds_bytes = s.lds_bytes;
    }
    return true;
}

bool coop_shape_of(uint64_t n, uint64_t* n1, int* e, int* threads, unsigned* lds_bytes)
{
    static const struct
    {
        uint64_t n1;
        int e, threads;
    } kCoop[] = {{64, 4, 256}, {256, 16, 512}};
    for (const auto& c : kCoop)
    {
        LaunchShape s;
        if (c.n1 * c.n1 == n && (*e == 0 || *e == c.e) && coop_kernel_shape(c.n1, c.e, c.threads, &s))
        {
            *n1 = c.n1;
            *e = c.e;
            *threads = c.threads;
            *lds_bytes = s.lds_bytes;
            return true;
        }
    }
    return false;
}

namespace
{

// A tuned hints field: CRDFFT_HINTS_NONE asks for none, other bits are kept where the algorithm has them.
uint32_t tuning_hints(uint32_t h, uint32_t mask) { return (h & CRDFFT_HINTS_NONE) != 0 ? 0u : (h & mask); }

// Values per thread of the single kernel. C2C: the fastest per length in an autotune sweep over 2^22-point batches
// on gfx950 (fewer values per thread keep more waves resident where a transform is short, and for 1024-4096 points
// the extra LDS round trip of radix-8 stages costs less than the occupancy radix-16 gives up). R2C / C2R (n = the
// half length here): the values the one-kernel real forms are instantiated at, so their stage table is the plan's.
int default_e(uint64_t n, int32_t type)
{
    if (type == CRDFFT_C2C)
    {
        switch (n)
        {
            case 16: case 32: case 64: return 4;
            case 128: case 1024: case 2048: case 4096: return 8;
            case 8192: return 32;
            default: break;
        }
    }
    else if (n >= 16 && n <= 8192)
    {
        return n <= 256 ? 4 : 8;
    }
    return n <= 8 ? static_cast<int>(n) : 16;
}

// Four-step split of n = 2^L: n1 = 2^floor(L/2) up to 2^17, from 2^18 one bit more (measured on MI355X at 2^20:
// 2048 x 512 ran 1.17x faster than 1024 x 1024, the row pass keeping its transposed store to 512-point rows), at most
// 4096; tiles sized so one workgroup's LDS holds at most 16K points.
void default_four_step(uint64_t n, PlanShape* s)
{
    const uint32_t lg = log2_exact(n);
    const uint32_t l1 = lg >= 18 ? std::min<uint32_t>(lg / 2 + 1, 12) : lg / 2;
    s->n1 = uint64_t{1} << l1;
    s->n2 = n / s->n1;
    s->e1 = 16;
    // 512-point rows at 8 values per thread (1024 threads, twice the waves per CU in the same LDS): measured 8 %
    // faster for n = 2^20 (2048 x 512); the other passes measured no better at 8
    s->e2 = s->n2 == 512 ? 8 : 16;
    s->tile1 = s->n1 <= 512 ? 16 : (s->n1 <= 2048 ? 8 : 4);
    s->tile2 = s->n2 <= 512 ? 16 : (s->n2 <= 2048 ? 8 : 4);
}

} // namespace

bool element_major_shape(uint64_t n, uint64_t batch, int* e, int* c)
{
    if (n <= 16)
    {
        *e = static_cast<int>(n);
        *c = 64;
    }
    else if (n <= 64)
    {
        *e = 16;
        *c = 16;
    }
    else if (n <= 1024)
    {
        *e = 16;
        *c = batch >= 16 ? 16 : 8;
        if (!have_col_kernel(n, *e, *c))
        {
            *c = 16;
        }
    }
    else if (n <= 4096)
    {
        *e = 16;
        *c = n == 4096 ? 4 : 8;
    }
    else
    {
        return false;
    }
    return have_col_kernel(n, *e, *c);
}

crdfft_status choose_shape(uint64_t n, uint64_t batch, int32_t type, int32_t precision, const crdfft_tuning* tuning,
                           PlanShape* out)
{
    if (out == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    if (type != CRDFFT_C2C && type != CRDFFT_R2C && type != CRDFFT_C2R)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    if (precision != CRDFFT_FP32 && precision != CRDFFT_SPLIT_BF16 && precision != CRDFFT_SPLIT_FP16)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    if (!is_pow2(n))
    {
        return CRDFFT_INVALID_SIZE;
    }
    if (type != CRDFFT_C2C && n < 4)
    {
        return CRDFFT_INVALID_SIZE; // RealFftPlan: a power of two >= 4 (real_fft.hpp:24)
    }
    PlanShape s;
    s.n = n;
    s.batch = batch;
    s.type = type;
    s.precision = precision;
    s.complex_n = type == CRDFFT_C2C ? n : n / 2;
    const uint64_t cn = s.complex_n;
    if (cn > kFourStepMax)
    {
        return CRDFFT_INVALID_SIZE;
    }
    if (precision != CRDFFT_FP32)
    {
        // the split-precision MFMA passes run sequences of 256 (radix-16 stages) and 1024 points (radix-32): one
        // sequence per transform at n = 256, 1024, else the four-step of n1 x n2 with n1, n2 in {256, 1024}
        // (n = 2^16, 2^18, 2^20); tile1/tile2 hold the radices of the two passes
        if (type != CRDFFT_C2C)
        {
            return CRDFFT_UNSUPPORTED;
        }
        if (cn == 256 || cn == 1024)
        {
            s.algorithm = CRDFFT_ALGO_SINGLE;
            s.e = cn == 256 ? 16 : 32;
            s.chunk = std::max<uint64_t>(batch, 1);
            s.workspace_bytes = 0;
            *out = s;
            return CRDFFT_OK;
        }
        if (cn != (uint64_t{1} << 16) && cn != (uint64_t{1} << 18) && cn != (uint64_t{1} << 20))
        {
            return CRDFFT_UNSUPPORTED;
        }
        s.algorithm = CRDFFT_ALGO_FOUR_STEP;
        s.n1 = cn == (uint64_t{1} << 20) ? 1024 : 256;
        s.n2 = cn / s.n1;
        s.e1 = s.e2 = 0;
        s.tile1 = s.n1 == 256 ? 16 : 32;
        s.tile2 = s.n2 == 256 ? 16 : 32;
        // half-size chunks (128 MiB of intermediate) over radix-16 column passes: measured 2.6-4.5 % faster at
        // n = 2^16 and 5.5-8 % at 2^18; full chunks at 2^20 (half: 6 % slower)
        s.chunk = std::max<uint64_t>(1, (s.n1 == 256 ? kChunkBytes / 2 : kChunkBytes) / (cn * 8));
        if (tuning != nullptr && tuning->chunk != 0)
        {
            s.chunk = tuning->chunk;
        }
        s.chunk = std::min<uint64_t>(s.chunk, std::max<uint64_t>(batch, 1));
        s.workspace_bytes = s.chunk * cn * 8;
        *out = s;
        return CRDFFT_OK;
    }
    if (cn == 1)
    {
        s.algorithm = CRDFFT_ALGO_COPY;
    }
    else if (cn <= kSingleMax && (tuning == nullptr || tuning->algorithm != CRDFFT_ALGO_COOP))
    {
        s.algorithm = CRDFFT_ALGO_SINGLE;
        s.e = default_e(cn, type);
        // batches of 256 points and up stream through once: non-temporal loads and stores measured 8-18% faster
        // (2^22-point batches, n = 256..16384); below that the default policy was faster
        s.hints = cn >= 256 ? kDefaultFourStepHints : 0u;
        if (tuning != nullptr && tuning->e != 0)
        {
            s.e = tuning->e;
        }
        if (!have_row_kernel(cn, s.e) ||
            (tuning != nullptr && tuning->algorithm != 0 && tuning->algorithm != CRDFFT_ALGO_SINGLE))
        {
            return CRDFFT_UNSUPPORTED;
        }
        if (tuning != nullptr && tuning->hints != 0)
        {
            s.hints = tuning_hints(tuning->hints, 3u);
        }
    }
    else
    {
        s.algorithm = CRDFFT_ALGO_FOUR_STEP;
        default_four_step(cn, &s);
        if (type == CRDFFT_C2R && cn == (uint64_t{1} << 15))
        {
            // the C2R column pass reads each bin twice (X[j] and X[h - j]); 256-point columns over 128-point rows
            // measured 2-3 % faster than 128 x 256 for n = 2^16, and 8 values per thread in those rows 1.6 % more
            s.n1 = 256;
            s.n2 = 128;
            s.e2 = 8;
        }
        if (type == CRDFFT_R2C && cn == (uint64_t{1} << 15))
        {
            // the R2C row pass stores runs of tile2 / 2 bins: 32 rows per workgroup (runs of 16, 128 bytes) at 8
            // values per thread measured 3.5 % faster than 16 rows for n = 2^16; 256-point columns over 128-point
            // rows (the C2C column pass of 2^16) 5 % faster than 128 x 256
            s.n1 = 256;
            s.n2 = 128;
            s.e2 = 8;
            s.tile2 = 32;
        }
        if (tuning != nullptr)
        {
            if (tuning->n1 != 0)
            {
                if (!is_pow2(tuning->n1) || tuning->n1 >= cn)
                {
                    return CRDFFT_INVALID_ARGUMENT;
                }
                s.n1 = tuning->n1;
                s.n2 = cn / s.n1;
                s.e2 = s.n2 == 512 ? 8 : 16;
                s.tile1 = s.n1 <= 512 ? 16 : (s.n1 <= 2048 ? 8 : 4);
                s.tile2 = s.n2 <= 512 ? 16 : (s.n2 <= 2048 ? 8 : 4);
            }
            if (tuning->e1 != 0)
                s.e1 = tuning->e1;
            if (tuning->tile1 != 0)
                s.tile1 = tuning->tile1;
            if (tuning->e2 != 0)
                s.e2 = tuning->e2;
            if (tuning->tile2 != 0)
                s.tile2 = tuning->tile2;
        }
        const uint64_t per = cn * 8;
        if (tuning != nullptr && tuning->algorithm == CRDFFT_ALGO_COOP)
        {
            uint64_t n1 = 0;
            int e = tuning->e1, threads = 0;
            unsigned lds = 0;
            if (!coop_shape_of(cn, &n1, &e, &threads, &lds))
            {
                return CRDFFT_UNSUPPORTED;
            }
            s.algorithm = CRDFFT_ALGO_COOP;
            s.n1 = s.n2 = n1;
            s.e1 = s.e2 = e;
            s.threads = threads;
            s.tile1 = s.tile2 = threads / (static_cast<int>(n1) / e);
            // one slot per resident workgroup (the plan recounts them on its device)
            const uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(2048 / threads, (160u * 1024u) / lds));
            s.chunk = kDescribeCUs * per_cu;
            s.workspace_bytes = s.chunk * per + coop_ctl_words(s.chunk) * 4;
            s.hints = kDefaultFourStepHints;
            if (tuning->hints != 0)
            {
                s.hints = tuning_hints(tuning->hints, 3u | kHintCoopSolo);
            }
            *out = s;
            return CRDFFT_OK;
        }
        if (tuning != nullptr && tuning->algorithm == CRDFFT_ALGO_FUSED)
        {
            // both passes in one workgroup: tile1 columns per phase-A tile set the workgroup size, the rows per
            // phase-B tile follow from it
            if (cn > kFusedMax)
            {
                return CRDFFT_UNSUPPORTED;
            }
            // the fused kernel is instantiated at 16 values per thread in both passes unless tuned otherwise
            s.e1 = tuning->e1 != 0 ? tuning->e1 : 16;
            s.e2 = tuning->e2 != 0 ? tuning->e2 : 16;
            const int t1 = static_cast<int>(s.n1) / s.e1;
            const int t2 = static_cast<int>(s.n2) / s.e2;
            const int threads = (tuning->tile1 != 0) ? t1 * tuning->tile1 : kFusedThreads;
            unsigned lds = 0;
            if (threads % t2 != 0 || !have_fused_kernel(s.n1, s.e1, s.n2, s.e2, threads, &lds))
            {
                return CRDFFT_UNSUPPORTED;
            }
            s.algorithm = CRDFFT_ALGO_FUSED;
            s.threads = threads;
            s.tile1 = threads / t1;
            s.tile2 = threads / t2;
            // one slot per resident workgroup (the plan recounts them on its device)
            const uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(2048 / threads, (160u * 1024u) / lds));
            s.chunk = std::min<uint64_t>(kDescribeCUs * per_cu, std::max<uint64_t>(batch, 1));
            if (tuning->chunk != 0)
            {
                s.chunk = std::min(s.chunk, tuning->chunk); // fewer slots than resident workgroups
            }
            s.workspace_bytes = s.chunk * per;
            s.hints = kDefaultFourStepHints;
            if (tuning->hints != 0)
            {
                s.hints = tuning_hints(tuning->hints, 3u);
            }
            *out = s;
            return CRDFFT_OK;
        }
        if (tuning != nullptr && tuning->algorithm != 0 && tuning->algorithm != CRDFFT_ALGO_FOUR_STEP)
        {
            return CRDFFT_INVALID_ARGUMENT;
        }
        if (!have_col_kernel(s.n1, s.e1, s.tile1) || !have_rowtr_kernel(s.n2, s.e2, s.tile2) ||
            (s.n1 % static_cast<uint64_t>(s.tile2)) != 0)
        {
            return CRDFFT_UNSUPPORTED;
        }
        // R2C: the plain cache policy for the column pass's loads too (its row pass keeps plain accesses whatever
        // the hints): measured 1 % faster at n = 2^16 and 6 % at n = 2^20
        s.hints = type == CRDFFT_R2C ? 0u : kDefaultFourStepHints;
        if (type == CRDFFT_C2C && s.n1 >= kTwoPipesMinN1 && batch > std::max<uint64_t>(1, kChunkBytes / per / 2))
        {
            s.hints |= kHintTwoPipes;
        }
        if (tuning != nullptr && tuning->hints != 0)
        {
            s.hints = tuning_hints(tuning->hints, 11u | kHintTwoPipes);
        }
        // two pipelines keep two chunks in flight: each gets half the workspace budget. C2R, whose column pass reads
        // every bin twice (X[j] and X[h - j]), also runs half-size chunks: measured 3-11 % faster at n = 2^16..2^22;
        // so does R2C from 1024-point columns (n = 2^20: 5 % faster; 2^22: equal) and at n = 2^16 over its 256 x 128
        // split (1.7 % faster)
        const bool half = (s.hints & kHintTwoPipes) != 0 || type == CRDFFT_C2R ||
                          (type == CRDFFT_R2C && (s.n1 >= 1024 || cn == (uint64_t{1} << 15)));
        const uint64_t budget = half ? kChunkBytes / 2 : kChunkBytes;
        s.chunk = std::max<uint64_t>(1, budget / per);
        if (tuning != nullptr && tuning->chunk != 0)
        {
            s.chunk = tuning->chunk;
        }
        s.chunk = std::min<uint64_t>(s.chunk, std::max<uint64_t>(batch, 1));
        s.workspace_bytes = s.chunk * per;
        if ((s.hints & kHintPersistentColumns) != 0 &&
            (!have_cols_pf(s.n1, s.e1, s.tile1) || (s.n2 % static_cast<uint64_t>(s.tile1)) != 0))
        {
            return CRDFFT_UNSUPPORTED;
        }
    }
    *out = s;
    return CRDFFT_OK;
}

void shape_to_info(const PlanShape& s, crdfft_plan_info* info)
{
    *info = crdfft_plan_info{};
    info->n = s.n;
    info->batch = s.batch;
    info->type = s.type;
    info->precision = s.precision;
    info->algorithm = s.algorithm;
    info->complex_n = s.complex_n;
    info->n1 = s.n1;
    info->n2 = s.n2;
    info->e = s.e;
    info->e1 = s.e1;
    info->tile1 = s.tile1;
    info->e2 = s.e2;
    info->tile2 = s.tile2;
    info->chunk = s.chunk;
    info->workspace_bytes = s.workspace_bytes;
    info->hints = s.hints;
}

} // namespace crdfft
