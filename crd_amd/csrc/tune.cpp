// The autotuner of the power-of-two plans: candidate enumeration, on-device timing, timing records and the plan
// feature vector the learned cost model reads.
//
// The reference has no runtime tuning: its plan is a pure function of the size, the CPU level and the determinism mode,
// and the per-level choices come from a board measured offline and checked in (engine/numerics/hesap-fft/include/crd/
// hesap/fft/fft.hpp:1-8, detail/level_plans_gen.hpp). Here the same split (a measured table behind a deterministic
// planner) is kept, but the measuring runs on the device the plan will run on, and a learned model can rank the
// candidates so only the most promising are timed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "planner.hpp"

namespace
{

bool same_shape(const crdfft::PlanShape& a, const crdfft::PlanShape& b)
{
    return a.algorithm == b.algorithm && a.n1 == b.n1 && a.e == b.e && a.e1 == b.e1 && a.tile1 == b.tile1 &&
           a.e2 == b.e2 && a.tile2 == b.tile2 && a.chunk == b.chunk && a.hints == b.hints && a.threads == b.threads;
}

crdfft_tuning tuning_of(const crdfft::PlanShape& s)
{
    crdfft_tuning t{};
    t.n1 = s.n1;
    t.e = s.e;
    t.e1 = s.e1;
    t.tile1 = s.tile1;
    t.e2 = s.e2;
    t.tile2 = s.tile2;
    t.chunk = s.chunk;
    t.hints = s.hints == 0 ? CRDFFT_HINTS_NONE : s.hints; // a tuning's 0 would mean the planner's default
    t.algorithm = s.algorithm;
    return t;
}

std::vector<crdfft_tuning> enumerate(uint64_t n, uint64_t batch, int32_t type)
{
    std::vector<crdfft_tuning> out;
    std::vector<crdfft::PlanShape> shapes;
    auto add = [&](const crdfft_tuning* t) {
        crdfft::PlanShape s;
        if (crdfft::choose_shape(n, batch, type, CRDFFT_FP32, t, &s) != CRDFFT_OK)
        {
            return;
        }
        for (const crdfft::PlanShape& q : shapes)
        {
            if (same_shape(q, s))
            {
                return;
            }
        }
        shapes.push_back(s);
        out.push_back(tuning_of(s));
    };
    add(nullptr); // candidate 0: the planner's default
    crdfft::PlanShape def;
    if (crdfft::choose_shape(n, batch, type, CRDFFT_FP32, nullptr, &def) != CRDFFT_OK)
    {
        return out;
    }
    const uint64_t cn = def.complex_n;
    if (def.algorithm == CRDFFT_ALGO_SINGLE)
    {
        for (int e : {4, 8, 16, 32, 64})
        {
            for (uint32_t hints : {CRDFFT_HINTS_NONE, 3u})
            {
                crdfft_tuning t{};
                t.e = e;
                t.hints = hints;
                add(&t);
            }
        }
    }
    else if (def.algorithm == CRDFFT_ALGO_FOUR_STEP || def.algorithm == CRDFFT_ALGO_FUSED ||
             def.algorithm == CRDFFT_ALGO_COOP)
    {
        const uint32_t lg = crdfft::log2_exact(cn);
        // the cooperative single-pass form where it is instantiated
        uint64_t cn1 = 0;
        int ce = 0, cth = 0; // the first instantiated shape
        unsigned clds = 0;
        if (crdfft::coop_shape_of(cn, &cn1, &ce, &cth, &clds))
        {
            for (uint32_t hints : {1u, 3u})
            {
                crdfft_tuning t{};
                t.algorithm = CRDFFT_ALGO_COOP;
                t.hints = hints;
                add(&t);
            }
        }
        // the fused four-step at its instantiated workgroup sizes (tile1 = columns per phase-A tile)
        for (int tile1 : {0, 32, 64, 128})
        {
            for (uint32_t hints : {1u, 3u})
            {
                crdfft_tuning t{};
                t.algorithm = CRDFFT_ALGO_FUSED;
                t.tile1 = tile1;
                t.hints = hints;
                add(&t);
            }
        }
        for (int d = -1; d <= 2; ++d)
        {
            const int l1 = static_cast<int>(lg / 2) + d;
            if (l1 < 1 || l1 > 12 || static_cast<uint32_t>(l1) >= lg)
            {
                continue;
            }
            // from 2048-point columns also the two-pipeline schedule
            const std::vector<uint32_t> hint_set =
                l1 >= 11 ? std::vector<uint32_t>{1u, 3u, 3u | crdfft::kHintTwoPipes} : std::vector<uint32_t>{1u, 3u};
            for (uint64_t mib : {64u, 256u, 1024u})
            {
                for (uint32_t hints : hint_set)
                {
                    for (int variant = 0; variant < 6; ++variant)
                    {
                        crdfft_tuning t{};
                        t.algorithm = CRDFFT_ALGO_FOUR_STEP;
                        t.n1 = uint64_t{1} << l1;
                        t.chunk = std::max<uint64_t>(1, (mib << 20) / (cn * 8));
                        t.hints = hints;
                        if (variant == 1)
                        {
                            t.e1 = 32; // 32 values per thread, 16 columns per workgroup (where instantiated)
                            t.tile1 = 16;
                        }
                        else if (variant == 2)
                        {
                            t.e2 = 32;
                            t.tile2 = 16;
                        }
                        else if (variant >= 3)
                        {
                            // 8 values per thread (more waves per CU in the same LDS) in one or both passes
                            t.e1 = variant != 4 ? 8 : 0;
                            t.e2 = variant != 3 ? 8 : 0;
                            t.tile1 = t.e1 != 0 ? 16 : 0;
                            t.tile2 = t.e2 != 0 ? 16 : 0;
                        }
                        add(&t);
                    }
                }
            }
        }
    }
    return out;
}

struct Buf
{
    void* p = nullptr;
    ~Buf()
    {
        if (p != nullptr)
        {
            (void)hipFree(p);
        }
    }
};

float log2f_safe(double v) { return static_cast<float>(std::log2(std::max(v, 1.0))); }

} // namespace

extern "C" {

crdfft_status crdfft_plan_candidates(uint64_t n, uint64_t batch, int32_t type, crdfft_tuning* out, uint32_t capacity,
                                     uint32_t* count)
{
    if (count == nullptr || (out == nullptr && capacity != 0))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const std::vector<crdfft_tuning> c = enumerate(n, batch, type);
    *count = static_cast<uint32_t>(c.size());
    if (c.empty())
    {
        return CRDFFT_INVALID_SIZE;
    }
    for (uint32_t i = 0; i < capacity && i < c.size(); ++i)
    {
        out[i] = c[i];
    }
    return CRDFFT_OK;
}

crdfft_status crdfft_time_plan(uint64_t n, uint64_t batch, int32_t type, const crdfft_tuning* tuning, uint32_t reps,
                               crdfft_plan_record* record)
{
    if (record == nullptr || batch == 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    crdfft_plan plan = nullptr;
    crdfft_status st = crdfft_plan_create_tuned(&plan, n, batch, 0, 0, type, CRDFFT_FP32, CRDFFT_FLAG_NO_SELF_CHECK,
                                                tuning);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    const uint64_t in_bytes = type == CRDFFT_C2R ? (n / 2 + 1) * 8 * batch : (type == CRDFFT_R2C ? n * 4 * batch : n * 8 * batch);
    const uint64_t out_bytes = type == CRDFFT_R2C ? (n / 2 + 1) * 8 * batch : (type == CRDFFT_C2R ? n * 4 * batch : n * 8 * batch);
    Buf in, out;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<float> ms(std::max<uint32_t>(reps, 1));
    const int dir = type == CRDFFT_C2R ? CRDFFT_INVERSE : CRDFFT_FORWARD;
    if (hipMalloc(&in.p, in_bytes) != hipSuccess || hipMalloc(&out.p, out_bytes) != hipSuccess)
    {
        crdfft_plan_destroy(plan);
        return CRDFFT_OUT_OF_MEMORY;
    }
    if (hipMemset(in.p, 0, in_bytes) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    {
        crdfft_plan_destroy(plan);
        return CRDFFT_HIP_ERROR;
    }
    st = crdfft_execute(plan, in.p, out.p, dir, s); // warm
    for (size_t r = 0; r < ms.size() && st == CRDFFT_OK; ++r)
    {
        if (hipEventRecord(e0, s) != hipSuccess)
        {
            st = CRDFFT_HIP_ERROR;
            break;
        }
        st = crdfft_execute(plan, in.p, out.p, dir, s);
        if (st == CRDFFT_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms[r], e0, e1) != hipSuccess))
        {
            st = CRDFFT_HIP_ERROR;
        }
        if (st == CRDFFT_OK && tuning != nullptr && tuning->algorithm == CRDFFT_ALGO_COOP)
        {
            // a cooperative launch whose bounded group wait gave up skipped work: its time is not a plan's time,
            // and the candidate is dropped (crdfft_autotune_plan skips SELF_CHECK_FAILED candidates)
            uint32_t given_up = 0, mode = 0;
            st = crdfft_plan_coop_status(plan, s, &given_up, &mode);
            if (st == CRDFFT_OK && given_up != 0)
            {
                st = CRDFFT_SELF_CHECK_FAILED;
            }
        }
    }
    crdfft_plan_info info{};
    if (st == CRDFFT_OK)
    {
        st = crdfft_plan_get_info(plan, &info);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    crdfft_plan_destroy(plan);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    std::nth_element(ms.begin(), ms.begin() + ms.size() / 2, ms.end());
    *record = crdfft_plan_record{};
    record->n = n;
    record->batch = batch;
    record->type = type;
    record->precision = CRDFFT_FP32;
    record->ndim = 1;
    record->dims[0] = n;
    record->tuning.n1 = info.n1;
    record->tuning.e = info.e;
    record->tuning.e1 = info.e1;
    record->tuning.tile1 = info.tile1;
    record->tuning.e2 = info.e2;
    record->tuning.tile2 = info.tile2;
    record->tuning.chunk = info.chunk;
    record->tuning.hints = info.hints == 0 ? CRDFFT_HINTS_NONE : info.hints; // replayable: 0 means the default
    record->tuning.algorithm = info.algorithm;
    record->algorithm = info.algorithm;
    record->ms = ms[ms.size() / 2];
    return CRDFFT_OK;
}

crdfft_status crdfft_autotune_plan(uint64_t n, uint64_t batch, int32_t type, uint32_t reps, crdfft_tuning* best,
                                   crdfft_plan_record* records, uint32_t record_capacity, uint32_t* nrecords)
{
    if (best == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const std::vector<crdfft_tuning> cands = enumerate(n, batch, type);
    if (cands.empty())
    {
        return CRDFFT_INVALID_SIZE;
    }
    double best_ms = 1e300;
    uint32_t nrec = 0;
    for (const crdfft_tuning& t : cands)
    {
        crdfft_plan_record rec{};
        const crdfft_status st = crdfft_time_plan(n, batch, type, &t, reps, &rec);
        if (st == CRDFFT_OUT_OF_MEMORY || st == CRDFFT_SELF_CHECK_FAILED)
        {
            continue; // a candidate whose workspace does not fit, or whose cooperative waits gave up, is not one
        }
        if (st != CRDFFT_OK)
        {
            return st;
        }
        if (records != nullptr && nrec < record_capacity)
        {
            records[nrec] = rec;
        }
        ++nrec;
        if (rec.ms < best_ms)
        {
            best_ms = rec.ms;
            *best = rec.tuning;
        }
    }
    if (nrecords != nullptr)
    {
        *nrecords = std::min(nrec, record_capacity);
    }
    return best_ms < 1e300 ? CRDFFT_OK : CRDFFT_OUT_OF_MEMORY;
}

crdfft_status crdfft_plan_features(const crdfft_plan_record* records, uint64_t count, float* out)
{
    if ((records == nullptr || out == nullptr) && count != 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    for (uint64_t i = 0; i < count; ++i)
    {
        const crdfft_plan_record& r = records[i];
        float* f = out + i * CRDFFT_FEATURES;
        std::memset(f, 0, sizeof(float) * CRDFFT_FEATURES);
        uint64_t m = r.n;
        uint32_t ex[4] = {0, 0, 0, 0};
        const uint64_t primes[4] = {2, 3, 5, 7};
        for (int k = 0; k < 4; ++k)
        {
            while (m > 1 && m % primes[k] == 0)
            {
                m /= primes[k];
                ++ex[k];
            }
        }
        f[0] = static_cast<float>(ex[0]) / 24.0f; // factorization of n: exponents of 2, 3, 5, 7
        f[1] = static_cast<float>(ex[1]) / 8.0f;
        f[2] = static_cast<float>(ex[2]) / 6.0f;
        f[3] = static_cast<float>(ex[3]) / 5.0f;
        f[4] = log2f_safe(static_cast<double>(r.n)) / 24.0f;
        f[5] = log2f_safe(static_cast<double>(r.batch)) / 16.0f;
        f[6] = log2f_safe(static_cast<double>(r.n) * static_cast<double>(r.batch)) / 32.0f;
        f[7] = r.type == CRDFFT_C2C ? 1.0f : 0.0f;
        f[8] = r.type == CRDFFT_R2C ? 1.0f : 0.0f;
        f[9] = r.type == CRDFFT_C2R ? 1.0f : 0.0f;
        f[10] = r.precision == CRDFFT_FP32 ? 1.0f : 0.0f;
        f[11] = r.precision == CRDFFT_SPLIT_BF16 ? 1.0f : 0.0f;
        f[12] = r.precision == CRDFFT_SPLIT_FP16 ? 1.0f : 0.0f;
        f[13] = r.stride > 1 ? 1.0f : 0.0f;
        f[14] = r.algorithm == CRDFFT_ALGO_SINGLE ? 1.0f : 0.0f;
        f[15] = static_cast<float>(r.tuning.e) / 64.0f;
        f[16] = log2f_safe(static_cast<double>(r.tuning.n1)) / 12.0f;
        f[17] = r.tuning.n1 != 0 ? log2f_safe(static_cast<double>(r.n / std::max<uint64_t>(r.tuning.n1, 1))) / 12.0f : 0.0f;
        f[18] = static_cast<float>(r.tuning.tile1) / 16.0f;
        f[19] = static_cast<float>(r.tuning.tile2) / 16.0f;
        f[20] = log2f_safe(static_cast<double>(r.tuning.chunk) + 1.0) / 12.0f;
        f[21] = (r.tuning.hints & 1u) ? 1.0f : 0.0f;
        f[22] = (r.tuning.hints & 2u) ? 1.0f : 0.0f;
        // the schedule bit: N-D every axis through transposed lines, four-step two chunk pipelines
        f[23] = (r.tuning.hints & (4u | crdfft::kHintTwoPipes)) ? 1.0f : 0.0f;
        f[24] = static_cast<float>(r.tuning.e1) / 32.0f;
        f[25] = static_cast<float>(r.tuning.e2) / 32.0f;
        f[26] = static_cast<float>(r.ndim) / 3.0f;
        f[27] = r.ndim >= 2 ? log2f_safe(static_cast<double>(r.dims[r.ndim - 1])) / 12.0f : 0.0f;
        f[28] = r.ndim >= 2 ? log2f_safe(static_cast<double>(r.dims[0])) / 12.0f : 0.0f;
        f[29] = r.ndim >= 3 ? log2f_safe(static_cast<double>(r.dims[1])) / 12.0f : 0.0f;
        f[30] = r.algorithm == CRDFFT_ALGO_FOUR_STEP ? 1.0f : 0.0f;
        f[31] = r.algorithm == CRDFFT_ALGO_FUSED ? 1.0f : (r.algorithm == CRDFFT_ALGO_COOP ? 0.5f : 0.0f);
    }
    return CRDFFT_OK;
}

} // extern "C"
