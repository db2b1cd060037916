// Host side of the mixed-size persistent launch (many_kernels.hip): radix-plan candidates, tiling, the
// longest-processing-time balance over one workgroup per CU, hipGraph capture, and the autotuner that picks a radix
// plan per length by timing every candidate on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <numeric>
#include <queue>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "many.hpp"
#include "scratch.hpp"

namespace
{

crdfft_status hip_st(hipError_t e)
{
    if (e == hipSuccess)
    {
        return CRDFFT_OK;
    }
    return e == hipErrorOutOfMemory ? CRDFFT_OUT_OF_MEMORY : CRDFFT_HIP_ERROR;
}

#define MANY_TRY(expr)                                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        const hipError_t many_e_ = (expr);                                                                             \
        if (many_e_ != hipSuccess)                                                                                     \
            return hip_st(many_e_);                                                                                    \
    } while (0)

struct Exponents
{
    uint32_t a = 0, b = 0, c = 0, d = 0; // 2^a 3^b 5^c 7^d
};

bool smooth(uint64_t n, Exponents* e)
{
    if (n == 0)
    {
        return false;
    }
    Exponents x;
    while (n % 2 == 0)
    {
        n /= 2;
        ++x.a;
    }
    while (n % 3 == 0)
    {
        n /= 3;
        ++x.b;
    }
    while (n % 5 == 0)
    {
        n /= 5;
        ++x.c;
    }
    while (n % 7 == 0)
    {
        n /= 7;
        ++x.d;
    }
    *e = x;
    return n == 1;
}

// Transforms of length n in one full tile.
uint32_t max_group(uint64_t n) { return static_cast<uint32_t>(crdfft::kManyTilePoints / n); }

// The radices of one candidate: the power-of-two part in radices of at most 2^capbits (one smaller radix first for
// the remaining bits), the odd part as 7s, 5s and 3s; merged, pairs of 3s become 9s and a leftover 3 or 5 takes a
// 2 or 4 of the power-of-two part (6, 12, 10), so the length runs in fewer stages.
void candidate_radices(const Exponents& e, uint32_t capbits, bool merge, std::vector<uint32_t>* p2_out,
                       std::vector<uint32_t>* odd_out)
{
    std::vector<uint32_t> p2;
    uint32_t a = e.a;
    while (a >= capbits)
    {
        p2.push_back(1u << capbits);
        a -= capbits;
    }
    if (a > 0)
    {
        p2.insert(p2.begin(), 1u << a);
    }
    std::vector<uint32_t> odd;
    uint32_t b = e.b, c = e.c, d = e.d;
    if (merge)
    {
        for (; b >= 2; b -= 2)
        {
            odd.push_back(9u);
        }
    }
    std::vector<uint32_t> singles;
    singles.insert(singles.end(), d, 7u);
    singles.insert(singles.end(), c, 5u);
    singles.insert(singles.end(), b, 3u);
    for (uint32_t r : singles)
    {
        bool done = false;
        if (merge)
        {
            for (uint32_t& q : p2)
            {
                if ((q == 2 && r != 7) || (q == 4 && r == 3))
                {
                    q *= r; // 6, 10, 12
                    done = true;
                    break;
                }
            }
        }
        if (!done)
        {
            odd.push_back(r);
        }
    }
    *p2_out = p2;
    *odd_out = odd;
}

// Candidates: merged and unmerged radices, capbits 4, 3, 2, the power-of-two part first or the odd part first, full
// tiles or half tiles (candidate 0, the planner's default: merged, radix 16, powers of two first, full tiles).
std::vector<crdfft_radix_plan> candidates(uint64_t n)
{
    std::vector<crdfft_radix_plan> out;
    Exponents e;
    if (!smooth(n, &e) || n > crdfft::kManyTilePoints)
    {
        return out;
    }
    const uint32_t full = max_group(n);
    for (int merge = 1; merge >= 0; --merge)
    {
        for (uint32_t capbits : {4u, 3u, 2u})
        {
            std::vector<uint32_t> p2, odd;
            candidate_radices(e, capbits, merge != 0, &p2, &odd);
            for (int order = 0; order < 2; ++order)
            {
                std::vector<uint32_t> r = order == 0 ? p2 : odd;
                const std::vector<uint32_t>& tail = order == 0 ? odd : p2;
                r.insert(r.end(), tail.begin(), tail.end());
                if (r.empty())
                {
                    r.push_back(1u); // n == 1
                }
                if (r.size() > static_cast<size_t>(crdfft::kManyMaxStages))
                {
                    continue;
                }
                for (uint32_t group : {full, full / 2})
                {
                    if (group == 0 || (group != full && group == full / 2 && full / 2 == full))
                    {
                        continue;
                    }
                    crdfft_radix_plan p{};
                    p.n = n;
                    p.nstages = static_cast<uint32_t>(r.size());
                    std::copy(r.begin(), r.end(), p.radix);
                    p.group = group;
                    const bool dup = std::any_of(out.begin(), out.end(), [&](const crdfft_radix_plan& q) {
                        return q.nstages == p.nstages && q.group == p.group &&
                               std::equal(q.radix, q.radix + q.nstages, p.radix);
                    });
                    if (!dup)
                    {
                        out.push_back(p);
                    }
                }
            }
        }
    }
    return out;
}

struct DevBuf
{
    void* p = nullptr;
    hipError_t alloc(size_t bytes) { return crdfft::device_malloc(&p, std::max<size_t>(bytes, 16)); }
    ~DevBuf()
    {
        if (p != nullptr)
        {
            (void)hipFree(p);
        }
    }
};

} // namespace

struct crdfft_many_plan_s
{
    crdfft_many_info info{};
    DevBuf sizes, tiles, tw, wg_begin;
    unsigned workgroups = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~crdfft_many_plan_s()
    {
        if (exec != nullptr)
        {
            (void)hipGraphExecDestroy(exec);
        }
        if (graph != nullptr)
        {
            (void)hipGraphDestroy(graph);
        }
    }
};

namespace
{

crdfft_status build_many(crdfft_many_plan_s* p, const uint64_t* lengths, const uint64_t* offsets, uint64_t count,
                         const crdfft_radix_plan* choices, uint64_t nchoices)
{
    // distinct lengths in order of first appearance, with their radix plans
    std::map<uint64_t, uint32_t> id_of;
    std::vector<crdfft::ManySize> sizes;
    std::vector<uint32_t> group_of;
    std::vector<uint32_t> size_of(count);
    std::vector<uint64_t> offs(count);
    uint64_t run = 0;
    for (uint64_t t = 0; t < count; ++t)
    {
        offs[t] = offsets != nullptr ? offsets[t] : run;
        run += lengths[t];
    }
    uint64_t tw_total = 0;
    for (uint64_t t = 0; t < count; ++t)
    {
        const uint64_t n = lengths[t];
        auto it = id_of.find(n);
        if (it == id_of.end())
        {
            const std::vector<crdfft_radix_plan> cands = candidates(n);
            if (cands.empty())
            {
                return CRDFFT_INVALID_SIZE;
            }
            crdfft_radix_plan chosen = cands[0];
            for (uint64_t c = 0; c < nchoices; ++c)
            {
                if (choices[c].n == n)
                {
                    chosen = choices[c];
                }
            }
            // a supplied plan must multiply out to n, use only known radices and fit a tile
            uint64_t prod = 1;
            for (uint32_t s = 0; s < chosen.nstages && s < static_cast<uint32_t>(crdfft::kManyMaxStages); ++s)
            {
                const uint32_t r = chosen.radix[s];
                if (!crdfft::many_radix_supported(r))
                {
                    return CRDFFT_INVALID_ARGUMENT;
                }
                prod *= r;
            }
            const uint32_t full = max_group(n);
            if (chosen.nstages == 0 || chosen.nstages > static_cast<uint32_t>(crdfft::kManyMaxStages) || prod != n ||
                chosen.group > full)
            {
                return CRDFFT_INVALID_ARGUMENT;
            }
            crdfft::ManySize sz{};
            sz.n = static_cast<uint32_t>(n);
            sz.n_magic = crdfft::many_magic(sz.n);
            sz.nstages = chosen.nstages;
            uint32_t pp = 1;
            for (uint32_t s = 0; s < chosen.nstages; ++s)
            {
                crdfft::ManyStage& st = sz.stage[s];
                st.radix = chosen.radix[s];
                st.nr = sz.n / st.radix;
                st.nr_magic = crdfft::many_magic(st.nr);
                st.p = pp;
                st.p_magic = crdfft::many_magic(pp);
                st.twstep = sz.n / (pp * st.radix);
                pp *= st.radix;
            }
            sz.tw_offset = tw_total;
            tw_total += crdfft::kManyTwEntries;
            it = id_of.emplace(n, static_cast<uint32_t>(sizes.size())).first;
            sizes.push_back(sz);
            group_of.push_back(chosen.group == 0 ? full : chosen.group);
        }
        size_of[t] = it->second;
    }
    // tiles: runs of at most `group` transforms of one length that are contiguous in memory, in the given order
    struct HostTile
    {
        crdfft::ManyTile tile;
        double cost;
    };
    std::vector<HostTile> tiles;
    uint64_t points = 0;
    for (uint64_t t = 0; t < count;)
    {
        const uint32_t id = size_of[t];
        const uint64_t n = sizes[id].n;
        uint64_t cnt = 1;
        while (t + cnt < count && cnt < group_of[id] && size_of[t + cnt] == id && offs[t + cnt] == offs[t] + cnt * n)
        {
            ++cnt;
        }
        HostTile h;
        h.tile.size_id = id;
        h.tile.count = static_cast<uint32_t>(cnt);
        h.tile.base = offs[t];
        // cost: memory and stages per point plus a fixed tile latency, in point units
        h.cost = static_cast<double>(cnt * n) * (1.0 + 0.05 * sizes[id].nstages) + 2048.0;
        points += cnt * n;
        tiles.push_back(h);
        t += cnt;
    }
    int dev = 0, cus = 0;
    MANY_TRY(hipGetDevice(&dev));
    MANY_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned wgs = static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(cus, tiles.size())));
    // longest processing time first onto the least loaded workgroup
    std::vector<size_t> order(tiles.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return tiles[x].cost > tiles[y].cost; });
    using Load = std::pair<double, unsigned>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (unsigned w = 0; w < wgs; ++w)
    {
        heap.push({0.0, w});
    }
    std::vector<std::vector<size_t>> per_wg(wgs);
    std::vector<double> load(wgs, 0.0);
    for (size_t idx : order)
    {
        Load l = heap.top();
        heap.pop();
        per_wg[l.second].push_back(idx);
        l.first += tiles[idx].cost;
        load[l.second] = l.first;
        heap.push(l);
    }
    std::vector<crdfft::ManyTile> flat_tiles;
    std::vector<uint32_t> begin(wgs + 1, 0);
    for (unsigned w = 0; w < wgs; ++w)
    {
        begin[w] = static_cast<uint32_t>(flat_tiles.size());
        for (size_t idx : per_wg[w])
        {
            flat_tiles.push_back(tiles[idx].tile);
        }
    }
    begin[wgs] = static_cast<uint32_t>(flat_tiles.size());
    MANY_TRY(p->sizes.alloc(sizes.size() * sizeof(crdfft::ManySize)));
    MANY_TRY(p->tiles.alloc(flat_tiles.size() * sizeof(crdfft::ManyTile)));
    MANY_TRY(p->tw.alloc(tw_total * sizeof(crdfft::cfloat)));
    MANY_TRY(p->wg_begin.alloc(begin.size() * sizeof(uint32_t)));
    MANY_TRY(hipMemcpy(p->sizes.p, sizes.data(), sizes.size() * sizeof(crdfft::ManySize), hipMemcpyHostToDevice));
    MANY_TRY(hipMemcpy(p->tiles.p, flat_tiles.data(), flat_tiles.size() * sizeof(crdfft::ManyTile),
                       hipMemcpyHostToDevice));
    MANY_TRY(hipMemcpy(p->wg_begin.p, begin.data(), begin.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    for (const crdfft::ManySize& sz : sizes)
    {
        // lo[j] = W_n^j, hi[i] = W_n^{128 i}, both from FP64
        crdfft::cfloat* t = static_cast<crdfft::cfloat*>(p->tw.p) + sz.tw_offset;
        MANY_TRY(crdfft::build_roots(t, crdfft::kManyTwLo, 1, sz.n, nullptr));
        MANY_TRY(crdfft::build_roots(t + crdfft::kManyTwLo, crdfft::kManyTwLo, crdfft::kManyTwLo, sz.n, nullptr));
    }
    MANY_TRY(hipStreamSynchronize(nullptr));
    p->workgroups = wgs;
    const double mean = std::accumulate(load.begin(), load.end(), 0.0) / wgs;
    p->info.transforms = count;
    p->info.lengths = sizes.size();
    p->info.tiles = flat_tiles.size();
    p->info.points = points;
    p->info.workgroups = wgs;
    p->info.imbalance = mean > 0 ? *std::max_element(load.begin(), load.end()) / mean : 1.0;
    return CRDFFT_OK;
}

crdfft::ManyArgs args_of(const crdfft_many_plan_s* p, const void* in, void* out)
{
    return crdfft::ManyArgs{static_cast<const crdfft::cfloat*>(in),
                            static_cast<crdfft::cfloat*>(out),
                            static_cast<const crdfft::ManySize*>(p->sizes.p),
                            static_cast<const crdfft::ManyTile*>(p->tiles.p),
                            static_cast<const crdfft::cfloat*>(p->tw.p),
                            static_cast<const uint32_t*>(p->wg_begin.p)};
}

// Median time of `reps` launches (after a warm one) of a mixed plan on its own stream, hipEvents.
struct Timer
{
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t init()
    {
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess)
        {
            e = hipEventCreate(&e0);
        }
        if (e == hipSuccess)
        {
            e = hipEventCreate(&e1);
        }
        return e;
    }
    ~Timer()
    {
        if (e0 != nullptr)
        {
            (void)hipEventDestroy(e0);
        }
        if (e1 != nullptr)
        {
            (void)hipEventDestroy(e1);
        }
        if (s != nullptr)
        {
            (void)hipStreamDestroy(s);
        }
    }
    crdfft_status time(const uint64_t* lengths, uint64_t count, const crdfft_radix_plan* choices, uint64_t nchoices,
                       void* in, void* out, uint32_t reps, double* median_ms)
    {
        crdfft_many_plan mp = nullptr;
        crdfft_status st = crdfft_many_plan_create(&mp, lengths, nullptr, count, choices, nchoices, 0);
        if (st != CRDFFT_OK)
        {
            return st;
        }
        std::vector<float> ms(reps);
        st = crdfft_many_execute(mp, in, out, CRDFFT_FORWARD, s); // warm
        for (uint32_t r = 0; r < reps && st == CRDFFT_OK; ++r)
        {
            if (hipEventRecord(e0, s) != hipSuccess)
            {
                st = CRDFFT_HIP_ERROR;
                break;
            }
            st = crdfft_many_execute(mp, in, out, CRDFFT_FORWARD, s);
            if (st == CRDFFT_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                    hipEventElapsedTime(&ms[r], e0, e1) != hipSuccess))
            {
                st = CRDFFT_HIP_ERROR;
            }
        }
        crdfft_many_plan_destroy(mp);
        if (st == CRDFFT_OK)
        {
            std::nth_element(ms.begin(), ms.begin() + reps / 2, ms.end());
            *median_ms = ms[reps / 2];
        }
        return st;
    }
};

} // namespace

extern "C" {

crdfft_status crdfft_radix_candidates(uint64_t n, crdfft_radix_plan* out, uint32_t capacity, uint32_t* count)
{
    if (count == nullptr || (out == nullptr && capacity != 0))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const std::vector<crdfft_radix_plan> c = candidates(n);
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

crdfft_status crdfft_many_plan_create(crdfft_many_plan* plan, const uint64_t* lengths, const uint64_t* offsets,
                                      uint64_t count, const crdfft_radix_plan* choices, uint64_t nchoices,
                                      uint32_t flags)
{
    (void)flags;
    if (plan == nullptr || (lengths == nullptr && count != 0) || (choices == nullptr && nchoices != 0))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *plan = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    {
        return CRDFFT_NO_DEVICE;
    }
    auto* p = new (std::nothrow) crdfft_many_plan_s;
    if (p == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    const crdfft_status st = build_many(p, lengths, offsets, count, choices, nchoices);
    if (st != CRDFFT_OK)
    {
        delete p;
        return st;
    }
    *plan = p;
    return CRDFFT_OK;
}

crdfft_status crdfft_many_plan_destroy(crdfft_many_plan plan)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    delete plan;
    return CRDFFT_OK;
}

crdfft_status crdfft_many_plan_get_info(crdfft_many_plan plan, crdfft_many_info* info)
{
    if (plan == nullptr || info == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *info = plan->info;
    return CRDFFT_OK;
}

crdfft_status crdfft_many_execute(crdfft_many_plan plan, const void* in, void* out, int32_t direction, void* stream)
{
    if (plan == nullptr || in == nullptr || out == nullptr ||
        (direction != CRDFFT_FORWARD && direction != CRDFFT_INVERSE))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    MANY_TRY(crdfft::launch_many(args_of(plan, in, out), plan->workgroups, direction == CRDFFT_INVERSE,
                                 static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_many_graph_capture(crdfft_many_plan plan, const void* in, void* out, int32_t direction,
                                        void* stream)
{
    if (plan == nullptr || in == nullptr || out == nullptr || stream == nullptr ||
        (direction != CRDFFT_FORWARD && direction != CRDFFT_INVERSE))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (plan->exec != nullptr)
    {
        MANY_TRY(hipGraphExecDestroy(plan->exec));
        plan->exec = nullptr;
    }
    if (plan->graph != nullptr)
    {
        MANY_TRY(hipGraphDestroy(plan->graph));
        plan->graph = nullptr;
    }
    MANY_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const hipError_t le =
        crdfft::launch_many(args_of(plan, in, out), plan->workgroups, direction == CRDFFT_INVERSE, s);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    if (le != hipSuccess || ce != hipSuccess)
    {
        if (g != nullptr)
        {
            (void)hipGraphDestroy(g);
        }
        return hip_st(le != hipSuccess ? le : ce);
    }
    plan->graph = g;
    MANY_TRY(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0));
    return CRDFFT_OK;
}

crdfft_status crdfft_many_graph_launch(crdfft_many_plan plan, void* stream)
{
    if (plan == nullptr || plan->exec == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    MANY_TRY(hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_many_autotune(const uint64_t* lengths, uint64_t count, uint32_t reps, crdfft_radix_plan* choices,
                                   uint64_t capacity, uint64_t* nchoices, crdfft_tune_record* records,
                                   uint64_t record_capacity, uint64_t* nrecords)
{
    if ((lengths == nullptr && count != 0) || choices == nullptr || nchoices == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    std::vector<uint64_t> distinct;
    uint64_t total = 0;
    for (uint64_t t = 0; t < count; ++t)
    {
        total += lengths[t];
        if (std::find(distinct.begin(), distinct.end(), lengths[t]) == distinct.end())
        {
            distinct.push_back(lengths[t]);
        }
    }
    if (distinct.size() > capacity)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    reps = std::max<uint32_t>(reps, 1);
    // each candidate runs on 2^24 points of its length: a thousand full tiles, several per CU
    constexpr uint64_t kPoints = uint64_t{1} << 24;
    DevBuf buf;
    MANY_TRY(buf.alloc(kPoints * 2 * sizeof(crdfft::cfloat)));
    MANY_TRY(hipMemset(buf.p, 0, kPoints * 2 * sizeof(crdfft::cfloat)));
    Timer tm;
    MANY_TRY(tm.init());
    uint64_t nrec = 0;
    crdfft_status result = CRDFFT_OK;
    for (size_t i = 0; i < distinct.size() && result == CRDFFT_OK; ++i)
    {
        const uint64_t n = distinct[i];
        const std::vector<crdfft_radix_plan> cands = candidates(n);
        if (cands.empty())
        {
            result = CRDFFT_INVALID_SIZE;
            break;
        }
        const uint64_t tcount = std::max<uint64_t>(1, kPoints / n);
        std::vector<uint64_t> lens(tcount, n);
        double best = 1e300;
        crdfft_radix_plan best_plan = cands[0];
        for (size_t c = 0; c < cands.size() && result == CRDFFT_OK; ++c)
        {
            double ms = 0.0;
            result = tm.time(lens.data(), tcount, &cands[c], 1, buf.p, static_cast<crdfft::cfloat*>(buf.p) + kPoints,
                             reps, &ms);
            if (result != CRDFFT_OK)
            {
                break;
            }
            const double us = 1000.0 * ms / static_cast<double>(tcount);
            if (records != nullptr && nrec < record_capacity)
            {
                crdfft_tune_record& rec = records[nrec];
                rec.plan = cands[c];
                rec.candidate = static_cast<uint32_t>(c);
                rec.transforms = static_cast<uint32_t>(tcount);
                rec.us_per_transform = us;
            }
            ++nrec;
            if (us < best)
            {
                best = us;
                best_plan = cands[c];
            }
        }
        choices[i] = best_plan;
    }
    // the per-length winners are measured alone; the batch as given decides: if the planner's defaults run it
    // faster, they are returned instead (skipped when the batch does not fit beside the tuning buffers)
    if (result == CRDFFT_OK && count > 0)
    {
        DevBuf win, wout;
        if (win.alloc(total * sizeof(crdfft::cfloat)) == hipSuccess &&
            wout.alloc(total * sizeof(crdfft::cfloat)) == hipSuccess &&
            hipMemset(win.p, 0, total * sizeof(crdfft::cfloat)) == hipSuccess)
        {
            double tuned = 0.0, plain = 0.0;
            result = tm.time(lengths, count, choices, distinct.size(), win.p, wout.p, reps, &tuned);
            if (result == CRDFFT_OK)
            {
                result = tm.time(lengths, count, nullptr, 0, win.p, wout.p, reps, &plain);
            }
            if (result == CRDFFT_OK && plain < tuned)
            {
                for (size_t i = 0; i < distinct.size(); ++i)
                {
                    choices[i] = candidates(distinct[i])[0];
                }
            }
        }
        else
        {
            (void)hipGetLastError();
        }
    }
    *nchoices = distinct.size();
    if (nrecords != nullptr)
    {
        *nrecords = std::min(nrec, record_capacity);
    }
    return result;
}

} // extern "C"
