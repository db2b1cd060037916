// The planner, the plan object and the C-ABI (include/crdfft.h, include/crdfft_runtime.h) of the batched FFT path.
//
// Host code only (C++20, compiled by hipcc for its HIP runtime calls); every transform runs in the HIP kernels of
// fft_kernels.hip and real_kernels.hip. There is no CPU fallback: a plan that cannot be built on the device is an
// error returned to the caller.
#include <bit>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "scratch.hpp"
#include "../../include/crdfft_runtime.h"
#include "kernels.hpp"
#include "mfma.hpp"
#include "planner.hpp"
#include "real.hpp"
#include "spectral.hpp"

using crdfft::DeviceBuffer;
using crdfft::PerStream;

namespace
{

thread_local int32_t g_last_hip_error = 0;

crdfft_status hip_status(hipError_t e)
{
    if (e == hipSuccess)
    {
        return CRDFFT_OK;
    }
    g_last_hip_error = static_cast<int32_t>(e);
    if (e == hipErrorOutOfMemory)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    {
        return CRDFFT_NO_DEVICE;
    }
    return CRDFFT_HIP_ERROR;
}

#define CRDFFT_TRY(expr)                                                                                               \
    do                                                                                                                 \
    {                                                                                                                  \
        const hipError_t crdfft_e_ = (expr);                                                                           \
        if (crdfft_e_ != hipSuccess)                                                                                   \
            return hip_status(crdfft_e_);                                                                              \
    } while (0)

using crdfft::cfloat;

// The second pipeline of a two-pipeline four-step (kHintTwoPipes), one per caller stream: an internal stream at the
// default priority, its own workspace, and the events that fork it from the caller's stream and join it back. It is
// created on the first execute on that stream outside a graph capture (it allocates; a captured execute without one
// runs both halves on the captured stream); the fork and join are ordinary event records and waits, so later
// executes on the stream can be captured with it.
struct AuxPipe
{
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    hipEvent_t join = nullptr;
    DeviceBuffer ws;
    ~AuxPipe()
    {
        if (fork != nullptr)
        {
            (void)hipEventDestroy(fork);
        }
        if (join != nullptr)
        {
            (void)hipEventDestroy(join);
        }
        if (stream != nullptr)
        {
            (void)hipStreamDestroy(stream);
        }
    }
};

struct AuxPipes
{
    std::mutex mutex;
    std::vector<std::pair<hipStream_t, std::unique_ptr<AuxPipe>>> all;
    static constexpr size_t kMaxStreams = 64;

    // The pipeline of caller stream s with a workspace of `bytes`; nullptr when it cannot be created, or when it does
    // not exist yet and `create` is false.
    AuxPipe* get(hipStream_t s, uint64_t bytes, bool create)
    {
        std::lock_guard<std::mutex> lock(mutex);
        for (auto& e : all)
        {
            if (e.first == s)
            {
                return e.second.get();
            }
        }
        if (!create || all.size() >= kMaxStreams)
        {
            return nullptr;
        }
        auto a = std::make_unique<AuxPipe>();
        if (hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&a->fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&a->join, hipEventDisableTiming) != hipSuccess || a->ws.alloc(bytes) != hipSuccess)
        {
            (void)hipGetLastError();
            return nullptr;
        }
        all.emplace_back(s, std::move(a));
        return all.back().second.get();
    }
};
} // namespace

struct crdfft_plan_s
{
    crdfft::PlanShape shape;
    uint64_t idist = 0;
    uint64_t odist = 0;
    uint32_t flags = 0;
    int device = 0;
    // twiddle tables
    DeviceBuffer tw_single; // stage table of the single kernel (complex_n)
    DeviceBuffer tw1, tw2;  // stage tables of the column (n1) and row (n2) kernels
    DeviceBuffer thi, tlo;  // factored four-step twiddle
    DeviceBuffer tstep;     // column pass step table (e1 * n2 entries)
    DeviceBuffer tfull;     // column pass full twiddle table (n entries, small n only)
    uint32_t hshift = 0;
    DeviceBuffer wreal;     // W_n^k, k <= n/2, for R2C/C2R
    DeviceBuffer wref;      // W_cn^k, k < cn/2, for the radix-2 reference (built on first use)
    DeviceBuffer workspace; // four-step intermediate, chunk * complex_n
    DeviceBuffer tw_em;     // stage table of the element-major kernel (built on first use)
    // element-major batches of n > 4096 (built on first use): stage tables of both column lengths, the factored
    // W_n, and a grow-only workspace per stream
    DeviceBuffer em_tw1, em_tw2, em_thi, em_tlo;
    uint32_t em_hshift = 0;
    std::vector<std::pair<hipStream_t, std::unique_ptr<DeviceBuffer>>> em_work;
    // split-precision MFMA passes: operand splits of each pass's radix per direction ([0] column or single pass,
    // [1] row pass) and the stage twiddles W_S of each pass's sequence length
    DeviceBuffer split_fwd[2], split_inv[2], split_tw[2];
    DeviceBuffer split_tcm; // radix-32 column pass: W_n^{c m} at [m * n2 + c], m < 32
    // Bluestein: the power-of-two plan of m, chirp c[k], FFT_m(b), and the chunk workspace
    crdfft_plan inner = nullptr;
    crdfft_many_plan many = nullptr; // Bluestein plans of a 2/3/5/7-smooth n: the mixed-radix launch instead
    DeviceBuffer chirp, bhat, bwork;
    uint64_t blue_chunk = 0;
    uint64_t coop_ctl_offset = 0; // COOP: byte offset of the control words in the workspace
    int blue_e = 0; // > 0: the one-kernel form at this many values per thread (its stage table in tw_single)
    bool floor = false;           // demoted: every transform runs the radix-2 reference kernels (crdfft_check_status)
    int32_t check = CRDFFT_CHECK_NOT_RUN;
    PerStream workspace_of, bwork_of; // per-stream copies of workspace / bwork
    AuxPipes pipes;                   // per caller stream: the second pipeline of a two-pipeline four-step
    std::mutex ref_mutex;
    ~crdfft_plan_s()
    {
        if (inner != nullptr)
        {
            crdfft_plan_destroy(inner);
        }
        if (many != nullptr)
        {
            crdfft_many_plan_destroy(many);
        }
    }
};

namespace
{

// Event pairs around launches for crdfft_execute_profiled; a null profiler records nothing.
struct Profiler
{
    crdfft_phase_times* acc = nullptr;
    hipStream_t stream = nullptr;
    struct Mark
    {
        hipEvent_t a, b;
        int phase;
        uint64_t points;
    };
    std::vector<Mark> marks;
    hipEvent_t open_event = nullptr;

    hipError_t before()
    {
        if (acc == nullptr)
        {
            return hipSuccess;
        }
        hipError_t e = hipEventCreate(&open_event);
        if (e != hipSuccess)
        {
            return e;
        }
        return hipEventRecord(open_event, stream);
    }
    hipError_t after(int phase, uint64_t points)
    {
        if (acc == nullptr)
        {
            return hipSuccess;
        }
        hipEvent_t b = nullptr;
        hipError_t e = hipEventCreate(&b);
        if (e != hipSuccess)
        {
            return e;
        }
        e = hipEventRecord(b, stream);
        marks.push_back({open_event, b, phase, points});
        open_event = nullptr;
        return e;
    }
    hipError_t finish()
    {
        if (acc == nullptr)
        {
            return hipSuccess;
        }
        hipError_t err = hipStreamSynchronize(stream);
        for (const Mark& m : marks)
        {
            float ms = 0.0f;
            if (err == hipSuccess)
            {
                err = hipEventElapsedTime(&ms, m.a, m.b);
            }
            acc->ms[m.phase] += ms;
            acc->launches[m.phase] += 1;
            acc->points[m.phase] += m.points;
            (void)hipEventDestroy(m.a);
            (void)hipEventDestroy(m.b);
        }
        marks.clear();
        return err;
    }
};

crdfft_status build_tables(crdfft_plan_s* p, hipStream_t s)
{
    const crdfft::PlanShape& sh = p->shape;
    const uint64_t cn = sh.complex_n;
    if (sh.precision != CRDFFT_FP32)
    {
        const bool single = sh.algorithm == CRDFFT_ALGO_SINGLE;
        const uint64_t lens[2] = {single ? cn : sh.n1, single ? 0 : sh.n2};
        for (int ps = 0; ps < 2; ++ps)
        {
            int radix = 0;
            if (lens[ps] == 0)
            {
                continue;
            }
            if (!crdfft::split_radix_of(lens[ps], &radix))
            {
                return CRDFFT_UNSUPPORTED;
            }
            const uint64_t fl = crdfft::split_table_entries(radix);
            CRDFFT_TRY(p->split_fwd[ps].alloc(fl * sizeof(uint16_t)));
            CRDFFT_TRY(p->split_inv[ps].alloc(fl * sizeof(uint16_t)));
            CRDFFT_TRY(crdfft::build_split_tables(p->split_fwd[ps].as<uint16_t>(), radix, sh.precision, false, s));
            CRDFFT_TRY(crdfft::build_split_tables(p->split_inv[ps].as<uint16_t>(), radix, sh.precision, true, s));
            CRDFFT_TRY(p->split_tw[ps].alloc(lens[ps] * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_roots(p->split_tw[ps].as<cfloat>(), lens[ps], 1, lens[ps], s));
        }
        if (!single)
        {
            // the four-step twiddle W_n^e = hi[e >> hshift] * lo[e & (2^hshift - 1)]
            const uint32_t lg = crdfft::log2_exact(cn);
            p->hshift = lg / 2;
            const uint64_t lo = uint64_t{1} << p->hshift, hi = cn >> p->hshift;
            CRDFFT_TRY(p->thi.alloc(hi * sizeof(cfloat)));
            CRDFFT_TRY(p->tlo.alloc(lo * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_roots(p->thi.as<cfloat>(), hi, lo, cn, s));
            CRDFFT_TRY(crdfft::build_roots(p->tlo.as<cfloat>(), lo, 1, cn, s));
            if (sh.n1 == 1024)
            {
                // radix-32 column pass (n = 2^20): W_n^{c k} = W_n^{c 32 (tid / 8)} W_n^{c m}, the second factor
                // from a 32-row table (256 KiB) that stays in the L2
                CRDFFT_TRY(p->split_tcm.alloc(32 * sh.n2 * sizeof(cfloat)));
                CRDFFT_TRY(crdfft::build_full_twiddle(p->split_tcm.as<cfloat>(), 32, sh.n2, cn, s));
            }
            else
            {
                // the full table W_n^{c k} at [k * n2 + c]: one coalesced load per stored value
                CRDFFT_TRY(p->tfull.alloc(cn * sizeof(cfloat)));
                CRDFFT_TRY(crdfft::build_full_twiddle(p->tfull.as<cfloat>(), sh.n1, sh.n2, cn, s));
            }
            CRDFFT_TRY(p->workspace.alloc(sh.workspace_bytes));
        }
        return CRDFFT_OK;
    }
    if (sh.algorithm == CRDFFT_ALGO_SINGLE)
    {
        const uint64_t entries = crdfft::stage_twiddle_entries(cn, sh.e);
        CRDFFT_TRY(p->tw_single.alloc(entries * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_stage_twiddles(p->tw_single.as<cfloat>(), cn, sh.e, s));
    }
    else if (sh.algorithm == CRDFFT_ALGO_FUSED)
    {
        const uint64_t e1n = crdfft::stage_twiddle_entries(sh.n1, sh.e1);
        const uint64_t e2n = crdfft::stage_twiddle_entries(sh.n2, sh.e2);
        CRDFFT_TRY(p->tw1.alloc(e1n * sizeof(cfloat)));
        CRDFFT_TRY(p->tw2.alloc(e2n * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_stage_twiddles(p->tw1.as<cfloat>(), sh.n1, sh.e1, s));
        CRDFFT_TRY(crdfft::build_stage_twiddles(p->tw2.as<cfloat>(), sh.n2, sh.e2, s));
        CRDFFT_TRY(p->tfull.alloc(cn * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_full_twiddle(p->tfull.as<cfloat>(), sh.n1, sh.n2, cn, s));
        // one workspace slot per workgroup that can be resident on this device
        int per_cu = 0;
        hipDeviceProp_t d{};
        CRDFFT_TRY(crdfft::fused_blocks_per_cu(sh.n1, sh.e1, sh.n2, sh.e2, sh.threads, &per_cu));
        CRDFFT_TRY(hipGetDeviceProperties(&d, p->device));
        if (per_cu <= 0)
        {
            return CRDFFT_UNSUPPORTED;
        }
        // (the planner's count assumed 256 CUs and may also carry a tuned cap)
        const uint64_t planned = sh.chunk;
        p->shape.chunk = std::min<uint64_t>(static_cast<uint64_t>(d.multiProcessorCount) * per_cu,
                                            std::max<uint64_t>(sh.batch, 1));
        if (planned != 0 && planned < p->shape.chunk)
        {
            p->shape.chunk = planned;
        }
        p->shape.workspace_bytes = p->shape.chunk * cn * sizeof(cfloat);
        CRDFFT_TRY(p->workspace.alloc(p->shape.workspace_bytes));
    }
    else if (sh.algorithm == CRDFFT_ALGO_COOP)
    {
        const uint64_t e1n = crdfft::stage_twiddle_entries(sh.n1, sh.e1);
        CRDFFT_TRY(p->tw1.alloc(e1n * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_stage_twiddles(p->tw1.as<cfloat>(), sh.n1, sh.e1, s));
        CRDFFT_TRY(p->thi.alloc(sh.n1 * sizeof(cfloat)));
        CRDFFT_TRY(p->tlo.alloc(sh.n1 * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_roots(p->thi.as<cfloat>(), sh.n1, sh.n1, cn, s));
        CRDFFT_TRY(crdfft::build_roots(p->tlo.as<cfloat>(), sh.n1, 1, cn, s));
        // one workgroup (and slot) per resident workgroup of this device, so the launch can be resident as a whole
        int per_cu = 0;
        hipDeviceProp_t d{};
        CRDFFT_TRY(crdfft::coop_blocks_per_cu(sh.n1, sh.e1, sh.threads, &per_cu));
        CRDFFT_TRY(hipGetDeviceProperties(&d, p->device));
        if (per_cu <= 0)
        {
            return CRDFFT_UNSUPPORTED;
        }
        p->shape.chunk = static_cast<uint64_t>(d.multiProcessorCount) * static_cast<uint64_t>(per_cu);
        p->coop_ctl_offset = (p->shape.chunk * cn * sizeof(cfloat) + 255) / 256 * 256;
        p->shape.workspace_bytes = p->coop_ctl_offset + crdfft::coop_ctl_words(p->shape.chunk) * sizeof(uint32_t);
        CRDFFT_TRY(p->workspace.alloc(p->shape.workspace_bytes));
    }
    else if (sh.algorithm == CRDFFT_ALGO_FOUR_STEP)
    {
        const uint64_t e1n = crdfft::stage_twiddle_entries(sh.n1, sh.e1);
        const uint64_t e2n = crdfft::stage_twiddle_entries(sh.n2, sh.e2);
        CRDFFT_TRY(p->tw1.alloc(e1n * sizeof(cfloat)));
        CRDFFT_TRY(p->tw2.alloc(e2n * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_stage_twiddles(p->tw1.as<cfloat>(), sh.n1, sh.e1, s));
        CRDFFT_TRY(crdfft::build_stage_twiddles(p->tw2.as<cfloat>(), sh.n2, sh.e2, s));
        const uint32_t lg = crdfft::log2_exact(cn);
        p->hshift = lg / 2;
        const uint64_t lo = uint64_t{1} << p->hshift;
        const uint64_t hi = cn >> p->hshift;
        CRDFFT_TRY(p->thi.alloc(hi * sizeof(cfloat)));
        CRDFFT_TRY(p->tlo.alloc(lo * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_roots(p->thi.as<cfloat>(), hi, lo, cn, s));
        CRDFFT_TRY(crdfft::build_roots(p->tlo.as<cfloat>(), lo, 1, cn, s));
        CRDFFT_TRY(p->tstep.alloc(static_cast<uint64_t>(sh.e1) * sh.n2 * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_step_table(p->tstep.as<cfloat>(), sh.n2, sh.e1, sh.n1 / sh.e1, cn, s));
        if (cn <= crdfft::kFullTwiddleMax)
        {
            CRDFFT_TRY(p->tfull.alloc(cn * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_full_twiddle(p->tfull.as<cfloat>(), sh.n1, sh.n2, cn, s));
        }
        CRDFFT_TRY(p->workspace.alloc(sh.workspace_bytes));
    }
    if (sh.type != CRDFFT_C2C)
    {
        const uint64_t h = sh.n / 2;
        CRDFFT_TRY(p->wreal.alloc((h + 1) * sizeof(cfloat)));
        CRDFFT_TRY(crdfft::build_roots(p->wreal.as<cfloat>(), h + 1, 1, sh.n, s));
    }
    return CRDFFT_OK;
}

// The complex transform of the plan over `count` transforms (complex distances), on `s`.
// r2c_w (R2C plans on the four-step): the row pass also does the Hermitian recombine with this W_n^k table and
// writes h + 1 bins per transform at odist. c2r_w (C2R plans on the four-step, inverse): the input rows are h + 1
// bins at idist and the column pass forms (1/h) Z[j] from them on load.
crdfft_status run_complex(crdfft_plan_s* p, const cfloat* in, cfloat* out, uint64_t count, uint64_t idist,
                          uint64_t odist, bool inverse, hipStream_t s, Profiler& prof, float scale = 1.0f,
                          const cfloat* r2c_w = nullptr, const cfloat* c2r_w = nullptr)
{
    const crdfft::PlanShape& sh = p->shape;
    const uint64_t cn = sh.complex_n;
    if (p->floor && cn >= 2)
    {
        // the demotion floor: the radix-2 reference kernels over device memory (in place when in == out), then the
        // plan's scale on the rows only (the padding of a padded layout is the caller's)
        CRDFFT_TRY(prof.before());
        CRDFFT_TRY(crdfft::launch_reference(in, out, p->wref.as<cfloat>(), cn, count, idist, odist, inverse, s));
        if (scale != 1.0f)
        {
            CRDFFT_TRY(crdfft::launch_scale_rows(out, cn, count, odist, scale, s));
        }
        CRDFFT_TRY(prof.after(0, cn * count));
        return CRDFFT_OK;
    }
    switch (sh.algorithm)
    {
        case CRDFFT_ALGO_COPY:
            if (in != out)
            {
                CRDFFT_TRY(prof.before());
                CRDFFT_TRY(crdfft::launch_copy(in, out, cn, count, idist, odist, s));
                CRDFFT_TRY(prof.after(3, cn * count));
            }
            return CRDFFT_OK;
        case CRDFFT_ALGO_SINGLE:
        {
            crdfft::RowArgs a{in, out, p->tw_single.as<cfloat>(), count, idist, odist, scale, sh.hints};
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_rows(cn, sh.e, inverse, a, s));
            CRDFFT_TRY(prof.after(0, cn * count));
            return CRDFFT_OK;
        }
        case CRDFFT_ALGO_FOUR_STEP:
        {
            const crdfft::ScratchOverride* ov = crdfft::scratch_override();
            uint64_t chunk = sh.chunk;
            cfloat* ws = nullptr;
            if (ov != nullptr)
            {
                // the caller-free form: the library's static scratch, in chunks that fit it, one pipeline
                chunk = std::min<uint64_t>(sh.chunk, ov->bytes / (cn * sizeof(cfloat)));
                if (chunk == 0)
                {
                    return CRDFFT_INVALID_SIZE;
                }
                ws = static_cast<cfloat*>(ov->ptr);
            }
            else
            {
                ws = static_cast<cfloat*>(p->workspace_of.get(p->workspace, s));
            }
            if (ws == nullptr)
            {
                return CRDFFT_OUT_OF_MEMORY;
            }
            // two pipelines: odd chunks run on the caller stream's second pipeline with its own workspace, forked
            // from and joined back to the caller's stream (profiled executes keep one stream: their events time
            // each launch on it)
            AuxPipe* aux = nullptr;
            if ((sh.hints & crdfft::kHintTwoPipes) != 0 && count > chunk && prof.acc == nullptr && ov == nullptr)
            {
                // a stream being captured cannot allocate its pipeline: without one it runs both halves in order
                hipStreamCaptureStatus capturing = hipStreamCaptureStatusNone;
                CRDFFT_TRY(hipStreamIsCapturing(s, &capturing));
                const bool create = capturing == hipStreamCaptureStatusNone;
                aux = p->pipes.get(s, p->workspace.bytes, create);
                if (aux == nullptr && create)
                {
                    return CRDFFT_OUT_OF_MEMORY;
                }
                if (aux != nullptr)
                {
                    CRDFFT_TRY(hipEventRecord(aux->fork, s));
                    CRDFFT_TRY(hipStreamWaitEvent(aux->stream, aux->fork, 0));
                }
            }
            const hipStream_t caller = s;
            // tiled intermediate (kHintTiled) where the column pass's threads per column equal the row pass's rows
            // per workgroup and its columns per workgroup the row pass's threads per row (n = 2^16: 16 and 16)
            const bool plain_pair = c2r_w == nullptr && r2c_w == nullptr && (sh.hints & crdfft::kHintPersistent) == 0;
            const uint32_t tiled =
                plain_pair && sh.e1 > 0 && sh.e2 > 0 && static_cast<int>(sh.n1 / sh.e1) == sh.tile2 &&
                        static_cast<int>(sh.n2 / sh.e2) == sh.tile1 && sh.n2 % static_cast<uint64_t>(sh.tile1) == 0
                    ? crdfft::kHintTiled
                    : 0u;
            const crdfft_status st = [&]() -> crdfft_status {
                for (uint64_t c0 = 0; c0 < count; c0 += chunk)
                {
                    const uint64_t cnt = std::min(chunk, count - c0);
                    const bool second = aux != nullptr && ((c0 / chunk) & 1) != 0;
                    const hipStream_t cs = second ? aux->stream : caller;
                    cfloat* const ws_c = second ? aux->ws.as<cfloat>() : ws;
                    crdfft::ColArgs ca{in + c0 * idist,
                                       ws_c,
                                       p->tw1.as<cfloat>(),
                                       p->thi.as<cfloat>(),
                                       p->tlo.as<cfloat>(),
                                       p->tstep.as<cfloat>(),
                                       p->tfull.as<cfloat>(),
                                       p->hshift,
                                       sh.n2,
                                       cnt,
                                       idist,
                                       cn,
                                       (sh.hints & crdfft::kHintNtLoad) | tiled,
                                       c2r_w,
                                       1.0f / static_cast<float>(cn)};
                    CRDFFT_TRY(prof.before());
                    if (c2r_w != nullptr)
                    {
                        CRDFFT_TRY(crdfft::launch_cols_c2r(sh.n1, sh.e1, sh.tile1, ca, cs));
                    }
                    else if ((sh.hints & crdfft::kHintPersistent) != 0)
                    {
                        CRDFFT_TRY(crdfft::launch_cols_pf(sh.n1, sh.e1, sh.tile1, inverse, ca, cs));
                    }
                    else
                    {
                        CRDFFT_TRY(crdfft::launch_cols(sh.n1, sh.e1, sh.tile1, inverse, ca, cs));
                    }
                    CRDFFT_TRY(prof.after(0, cn * cnt));
                    crdfft::RowTrArgs ra{ws_c, out + c0 * odist, p->tw2.as<cfloat>(), sh.n1, cnt,
                                         cn, odist, scale, (sh.hints & crdfft::kHintNtStore) | tiled, r2c_w};
                    CRDFFT_TRY(prof.before());
                    if (r2c_w != nullptr)
                    {
                        CRDFFT_TRY(crdfft::launch_rowtr_r2c(sh.n2, sh.e2, sh.tile2, ra, cs));
                    }
                    else
                    {
                        CRDFFT_TRY(crdfft::launch_rowtr(sh.n2, sh.e2, sh.tile2, inverse, ra, cs));
                    }
                    CRDFFT_TRY(prof.after(1, cn * cnt));
                }
                return CRDFFT_OK;
            }();
            if (aux != nullptr)
            {
                // joined back even after a failed launch, so nothing on the second stream outlives the caller's order
                const hipError_t e1 = hipEventRecord(aux->join, aux->stream);
                const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(caller, aux->join, 0) : e1;
                if (st == CRDFFT_OK && e2 != hipSuccess)
                {
                    return hip_status(e2);
                }
            }
            return st;
        }
        case CRDFFT_ALGO_COOP:
        {
            char* ws = static_cast<char*>(p->workspace_of.get(p->workspace, s));
            if (ws == nullptr)
            {
                return CRDFFT_OUT_OF_MEMORY;
            }
            const crdfft::CoopArgs ca{in,
                                      out,
                                      reinterpret_cast<cfloat*>(ws),
                                      reinterpret_cast<uint32_t*>(ws + p->coop_ctl_offset),
                                      p->tw1.as<cfloat>(),
                                      p->thi.as<cfloat>(),
                                      p->tlo.as<cfloat>(),
                                      count,
                                      idist,
                                      odist,
                                      scale,
                                      sh.hints,
                                      crdfft::kCoopSpinLimit};
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_coop(sh.n1, sh.e1, sh.threads, inverse, ca, static_cast<unsigned>(sh.chunk), s));
            CRDFFT_TRY(prof.after(0, cn * count));
            return CRDFFT_OK;
        }
        case CRDFFT_ALGO_FUSED:
        {
            cfloat* ws = static_cast<cfloat*>(p->workspace_of.get(p->workspace, s));
            if (ws == nullptr)
            {
                return CRDFFT_OUT_OF_MEMORY;
            }
            crdfft::FusedArgs fa{in,
                                 out,
                                 ws,
                                 p->tw1.as<cfloat>(),
                                 p->tw2.as<cfloat>(),
                                 p->tfull.as<cfloat>(),
                                 sh.chunk,
                                 count,
                                 idist,
                                 odist,
                                 scale,
                                 sh.hints};
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_fused(sh.n1, sh.e1, sh.n2, sh.e2, sh.threads, inverse, fa, sh.chunk, s));
            CRDFFT_TRY(prof.after(0, cn * count));
            return CRDFFT_OK;
        }
        default:
            return CRDFFT_UNSUPPORTED;
    }
}

crdfft_status execute_impl(crdfft_plan p, const void* in, void* out, int32_t direction, void* stream, Profiler& prof)
{
    if (p == nullptr || in == nullptr || out == nullptr ||
        (direction != CRDFFT_FORWARD && direction != CRDFFT_INVERSE))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const crdfft::PlanShape& sh = p->shape;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool inverse = direction == CRDFFT_INVERSE;
    if (sh.batch == 0)
    {
        return CRDFFT_OK;
    }
    if (sh.precision != CRDFFT_FP32 && !p->floor)
    {
        // split-precision MFMA: one sequence per transform (n = 256, 1024), else the four-step's column pass into
        // the workspace and its row pass out, per chunk
        const cfloat* x = static_cast<const cfloat*>(in);
        cfloat* y = static_cast<cfloat*>(out);
        const float scale =
            (inverse && (p->flags & CRDFFT_FLAG_NORMALIZE_INVERSE) != 0) ? 1.0f / static_cast<float>(sh.n) : 1.0f;
        const uint16_t* tab0 = inverse ? p->split_inv[0].as<uint16_t>() : p->split_fwd[0].as<uint16_t>();
        if (sh.algorithm == CRDFFT_ALGO_SINGLE)
        {
            if (x == y && p->idist != p->odist)
            {
                return CRDFFT_INVALID_ARGUMENT;
            }
            const crdfft::SplitArgs a{x,       y,        tab0,     p->split_tw[0].as<cfloat>(), nullptr, nullptr, nullptr,
                                      0,       sh.batch, p->idist, p->odist,                    1,       scale};
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_split_pass(2, sh.e, sh.precision, inverse, a, s));
            CRDFFT_TRY(prof.after(0, sh.complex_n * sh.batch));
            return CRDFFT_OK;
        }
        const uint16_t* tab1 = inverse ? p->split_inv[1].as<uint16_t>() : p->split_fwd[1].as<uint16_t>();
        cfloat* ws = static_cast<cfloat*>(p->workspace_of.get(p->workspace, s));
        if (ws == nullptr)
        {
            return CRDFFT_OUT_OF_MEMORY;
        }
        for (uint64_t c0 = 0; c0 < sh.batch; c0 += sh.chunk)
        {
            const uint64_t cnt = std::min(sh.chunk, sh.batch - c0);
            const crdfft::SplitArgs a{x + c0 * p->idist, ws,        tab0, p->split_tw[0].as<cfloat>(), p->thi.as<cfloat>(),
                                      p->tlo.as<cfloat>(), p->tfull.as<cfloat>(), p->hshift, cnt,  p->idist,
                                      sh.complex_n,        sh.n2,     1.0f, p->split_tcm.as<cfloat>()};
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_split_pass(0, sh.tile1, sh.precision, inverse, a, s));
            CRDFFT_TRY(prof.after(0, sh.complex_n * cnt));
            const crdfft::SplitArgs b{ws,     y + c0 * p->odist, tab1, p->split_tw[1].as<cfloat>(), nullptr, nullptr,
                                      nullptr, 0, cnt, sh.complex_n, p->odist, sh.n1, scale};
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_split_pass(1, sh.tile2, sh.precision, inverse, b, s));
            CRDFFT_TRY(prof.after(1, sh.complex_n * cnt));
        }
        return CRDFFT_OK;
    }
    if (sh.algorithm == CRDFFT_ALGO_BLUESTEIN)
    {
        const uint64_t n = sh.n, m = sh.n1;
        const cfloat* x = static_cast<const cfloat*>(in);
        cfloat* y = static_cast<cfloat*>(out);
        if (p->many != nullptr)
        {
            CRDFFT_TRY(prof.before());
            const crdfft_status st = crdfft_many_execute(p->many, in, out, direction, stream);
            if (st != CRDFFT_OK)
            {
                return st;
            }
            if (inverse)
            {
                CRDFFT_TRY(crdfft::launch_scale(y, n * sh.batch, 1.0f / static_cast<float>(n), s)); // bluestein.hpp:86
            }
            CRDFFT_TRY(prof.after(0, n * sh.batch));
            return CRDFFT_OK;
        }
        if (p->blue_e > 0)
        {
            const crdfft::BlueArgs a{x,      y,        p->chirp.as<cfloat>(), p->bhat.as<cfloat>(), p->tw_single.as<cfloat>(),
                                     n,      sh.batch, p->idist,              p->odist};
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_blue_fused(m, p->blue_e, inverse, a, s));
            CRDFFT_TRY(prof.after(0, n * sh.batch));
            return CRDFFT_OK;
        }
        cfloat* w = static_cast<cfloat*>(p->bwork_of.get(p->bwork, s));
        if (w == nullptr)
        {
            return CRDFFT_OUT_OF_MEMORY;
        }
        for (uint64_t c0 = 0; c0 < sh.batch; c0 += p->blue_chunk)
        {
            const uint64_t cnt = std::min(p->blue_chunk, sh.batch - c0);
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_blue_pre(x + c0 * p->idist, w, p->chirp.as<cfloat>(), n, m, cnt, p->idist,
                                               inverse, s));
            CRDFFT_TRY(prof.after(2, n * cnt));
            crdfft_status st = run_complex(p->inner, w, w, cnt, m, m, false, s, prof);
            if (st != CRDFFT_OK)
            {
                return st;
            }
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_blue_mul(w, p->bhat.as<cfloat>(), m, cnt, s));
            CRDFFT_TRY(prof.after(2, m * cnt));
            st = run_complex(p->inner, w, w, cnt, m, m, true, s, prof);
            if (st != CRDFFT_OK)
            {
                return st;
            }
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_blue_post(w, y + c0 * p->odist, p->chirp.as<cfloat>(), n, m, cnt, p->odist,
                                                inverse, s));
            CRDFFT_TRY(prof.after(2, n * cnt));
        }
        return CRDFFT_OK;
    }
    switch (sh.type)
    {
        case CRDFFT_C2C:
        {
            // CRDFFT_FLAG_NORMALIZE_INVERSE: ifft_normalized's 1/n (fft.hpp:1015-1029), fused into the last store
            const float scale =
                (inverse && (p->flags & CRDFFT_FLAG_NORMALIZE_INVERSE) != 0) ? 1.0f / static_cast<float>(sh.n) : 1.0f;
            return run_complex(p, static_cast<const cfloat*>(in), static_cast<cfloat*>(out), sh.batch, p->idist,
                               p->odist, inverse, s, prof, scale);
        }
        case CRDFFT_R2C:
        {
            // rfft is a forward transform: the direction argument must say so
            if (inverse || in == out)
            {
                return CRDFFT_INVALID_ARGUMENT;
            }
            const uint64_t h = sh.n / 2;
            cfloat* o = static_cast<cfloat*>(out);
            int se = 0;
            unsigned per_group = 0;
            if (sh.algorithm == CRDFFT_ALGO_SINGLE && (p->flags & CRDFFT_FLAG_SEPARATE_PASSES) == 0 && !p->floor &&
                h >= 256 &&
                crdfft::spec_shape(h, &se, &per_group) && se == sh.e)
            {
                // 512 <= n <= 32768: load, half-length core, recombine and store in one launch (the spectral
                // kernel's plain mode: no window, no detrend, frames = rows at idist); measured 1.3-1.7x the
                // separate passes there, and slower below (a few threads per row load poorly)
                const crdfft::SpecArgs a{static_cast<const float*>(in), nullptr, p->wreal.as<cfloat>(),
                                         p->tw_single.as<cfloat>(), out, sh.batch, p->idist, 1.0f, 0, p->odist};
                CRDFFT_TRY(prof.before());
                CRDFFT_TRY(crdfft::launch_spec(h, 0, a, (sh.batch + per_group - 1) / per_group, s));
                CRDFFT_TRY(prof.after(0, h * sh.batch));
                return CRDFFT_OK;
            }
            if (sh.algorithm == CRDFFT_ALGO_FOUR_STEP && (p->flags & CRDFFT_FLAG_SEPARATE_PASSES) == 0 && !p->floor &&
                (sh.hints & crdfft::kHintPersistent) == 0 && crdfft::have_rowtr_r2c(sh.n2, sh.e2, sh.tile2) &&
                sh.n1 >= static_cast<uint64_t>(sh.tile2))
            {
                // n > 32768: the column pass, then the row pass with the recombine fused in (two passes, not three)
                return run_complex(p, static_cast<const cfloat*>(in), o, sh.batch, p->idist / 2, p->odist, false, s,
                                   prof, 1.0f, p->wreal.as<cfloat>());
            }
            const crdfft_status st =
                run_complex(p, static_cast<const cfloat*>(in), o, sh.batch, p->idist / 2, p->odist, false, s, prof);
            if (st != CRDFFT_OK)
            {
                return st;
            }
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_r2c_post(o, p->wreal.as<cfloat>(), h, sh.batch, p->odist, s));
            CRDFFT_TRY(prof.after(2, (h + 1) * sh.batch));
            return CRDFFT_OK;
        }
        case CRDFFT_C2R:
        {
            if (!inverse || in == out)
            {
                return CRDFFT_INVALID_ARGUMENT;
            }
            const uint64_t h = sh.n / 2;
            cfloat* o = static_cast<cfloat*>(out);
            if (sh.algorithm == CRDFFT_ALGO_SINGLE && (p->flags & CRDFFT_FLAG_SEPARATE_PASSES) == 0 && !p->floor &&
                h >= 512 &&
                crdfft::c2r_one_kernel_e(h) == sh.e)
            {
                // 1024 <= n <= 32768: recombine, inverse half-length core and store in one launch (1.1-1.4x the
                // separate passes there, slower below)
                const crdfft::C2rArgs a{static_cast<const cfloat*>(in), o, p->wreal.as<cfloat>(),
                                        p->tw_single.as<cfloat>(), sh.batch, p->idist, p->odist / 2,
                                        1.0f / static_cast<float>(h)};
                CRDFFT_TRY(prof.before());
                CRDFFT_TRY(crdfft::launch_c2r_one_kernel(h, sh.e, a, s));
                CRDFFT_TRY(prof.after(0, h * sh.batch));
                return CRDFFT_OK;
            }
            if (sh.algorithm == CRDFFT_ALGO_FOUR_STEP && (p->flags & CRDFFT_FLAG_SEPARATE_PASSES) == 0 && !p->floor &&
                (sh.hints & crdfft::kHintPersistent) == 0 && crdfft::have_cols_c2r(sh.n1, sh.e1, sh.tile1) &&
                sh.n2 % static_cast<uint64_t>(sh.tile1) == 0)
            {
                // n > 32768: the pre-rotation in the column pass's loads, then the inverse row pass (two passes)
                return run_complex(p, static_cast<const cfloat*>(in), o, sh.batch, p->idist, p->odist / 2, true, s,
                                   prof, 1.0f, nullptr, p->wreal.as<cfloat>());
            }
            CRDFFT_TRY(prof.before());
            CRDFFT_TRY(crdfft::launch_c2r_pre(static_cast<const cfloat*>(in), o, p->wreal.as<cfloat>(), h, sh.batch,
                                              p->idist, p->odist / 2, 1.0f / static_cast<float>(h), s));
            CRDFFT_TRY(prof.after(2, (h + 1) * sh.batch));
            return run_complex(p, o, o, sh.batch, p->odist / 2, p->odist / 2, true, s, prof);
        }
        default:
            return CRDFFT_INVALID_ARGUMENT;
    }
}

crdfft_status create_plan(crdfft_plan* plan, uint64_t n, uint64_t batch, uint64_t idist, uint64_t odist,
                          int32_t type, int32_t precision, uint32_t flags, const crdfft_tuning* tuning)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *plan = nullptr;
    crdfft::PlanShape shape;
    const crdfft_status st = crdfft::choose_shape(n, batch, type, precision, tuning, &shape);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    const uint64_t in_min = type == CRDFFT_C2R ? n / 2 + 1 : n;
    const uint64_t out_min = type == CRDFFT_R2C ? n / 2 + 1 : n;
    if (idist == 0)
    {
        idist = in_min;
    }
    if (odist == 0)
    {
        odist = out_min;
    }
    if (batch > 1 && (idist < in_min || odist < out_min))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    // the real side is read/written as complex pairs: its distance must be even
    if ((type == CRDFFT_R2C && (idist % 2) != 0) || (type == CRDFFT_C2R && (odist % 2) != 0))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    {
        return CRDFFT_NO_DEVICE;
    }
    auto* p = new (std::nothrow) crdfft_plan_s;
    if (p == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    p->shape = shape;
    p->idist = idist;
    p->odist = odist;
    p->flags = flags;
    (void)hipGetDevice(&p->device);
    crdfft_status rs = build_tables(p, nullptr);
    if (rs == CRDFFT_OK)
    {
        rs = hip_status(hipStreamSynchronize(nullptr));
    }
    if (rs == CRDFFT_OK && (flags & CRDFFT_FLAG_NO_SELF_CHECK) == 0 && shape.precision == CRDFFT_FP32)
    {
        // select_checked_kernel_level (kernel_dispatch.hpp:105-145): the selected kernels, else the floor
        rs = crdfft_self_check(p);
        if (rs == CRDFFT_OK)
        {
            p->check = CRDFFT_CHECK_PASSED;
        }
        else if (rs == CRDFFT_SELF_CHECK_FAILED && shape.complex_n >= 2)
        {
            rs = crdfft::plan_use_floor(p);
            if (rs == CRDFFT_OK)
            {
                rs = hip_status(hipStreamSynchronize(nullptr));
            }
            if (rs == CRDFFT_OK)
            {
                rs = crdfft_self_check(p);
            }
            if (rs == CRDFFT_OK)
            {
                p->check = CRDFFT_CHECK_DEMOTED;
            }
        }
    }
    if (rs != CRDFFT_OK)
    {
        delete p;
        return rs;
    }
    *plan = p;
    return CRDFFT_OK;
}

} // namespace

extern "C" {

crdfft_status crdfft_plan_create(crdfft_plan* plan, uint64_t n, uint64_t batch, uint64_t idist, uint64_t odist,
                                 int32_t type, int32_t precision, uint32_t flags)
{
    return create_plan(plan, n, batch, idist, odist, type, precision, flags, nullptr);
}

crdfft_status crdfft_plan_create_tuned(crdfft_plan* plan, uint64_t n, uint64_t batch, uint64_t idist,
                                       uint64_t odist, int32_t type, int32_t precision, uint32_t flags,
                                       const crdfft_tuning* tuning)
{
    return create_plan(plan, n, batch, idist, odist, type, precision, flags, tuning);
}

crdfft_status crdfft_bluestein_plan_create(crdfft_plan* plan, uint64_t n, uint64_t batch, uint64_t idist,
                                           uint64_t odist, uint32_t flags)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *plan = nullptr;
    if (n == 0)
    {
        return CRDFFT_INVALID_SIZE;
    }
    uint64_t m = 1;
    while (m < 2 * n - 1)
    {
        m <<= 1;
    }
    if (m > crdfft::kFourStepMax)
    {
        return CRDFFT_INVALID_SIZE;
    }
    idist = idist == 0 ? n : idist;
    odist = odist == 0 ? n : odist;
    if (batch > 1 && (idist < n || odist < n))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    // A 2/3/5/7-smooth n up to 16384 in packed rows is the same DFT as one mixed-radix launch of the persistent
    // kernel (crdfft_many_*), about twice the chirp-z form's speed; CRDFFT_FLAG_CHIRP_Z keeps the reference's method.
    if ((flags & (CRDFFT_FLAG_SEPARATE_PASSES | CRDFFT_FLAG_CHIRP_Z)) == 0 && n >= 2 && n <= 16384 && batch >= 1 &&
        batch <= (uint64_t{1} << 24) && idist == n && odist == n)
    {
        crdfft_radix_plan cand{};
        uint32_t ncand = 0;
        if (crdfft_radix_candidates(n, &cand, 1, &ncand) == CRDFFT_OK && ncand > 0)
        {
            std::vector<uint64_t> lengths(batch, n);
            crdfft_many_plan mp = nullptr;
            if (crdfft_many_plan_create(&mp, lengths.data(), nullptr, batch, nullptr, 0, 0) == CRDFFT_OK)
            {
                auto* p = new (std::nothrow) crdfft_plan_s;
                if (p == nullptr)
                {
                    crdfft_many_plan_destroy(mp);
                    return CRDFFT_OUT_OF_MEMORY;
                }
                p->many = mp;
                p->shape.n = n;
                p->shape.batch = batch;
                p->shape.type = CRDFFT_C2C;
                p->shape.precision = CRDFFT_FP32;
                p->shape.algorithm = CRDFFT_ALGO_BLUESTEIN;
                p->shape.complex_n = n;
                p->shape.n1 = 0; // no convolution
                p->shape.chunk = batch;
                p->idist = idist;
                p->odist = odist;
                p->flags = flags;
                (void)hipGetDevice(&p->device);
                *plan = p;
                return CRDFFT_OK;
            }
        }
    }
    // workspace chunk: at most 256 MiB of m-point rows, at least one
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(std::max<uint64_t>(batch, 1),
                                                                    crdfft::kChunkBytes / (m * 8)));
    crdfft_plan inner = nullptr;
    const int fused_e = (flags & CRDFFT_FLAG_SEPARATE_PASSES) != 0 ? 0 : crdfft::blue_fused_e(m);
    crdfft_status st = crdfft_plan_create(&inner, m, fused_e > 0 ? 1 : chunk, m, m, CRDFFT_C2C, CRDFFT_FP32,
                                          (flags & ~CRDFFT_FLAG_SEPARATE_PASSES) | CRDFFT_FLAG_NO_SELF_CHECK);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    auto* p = new (std::nothrow) crdfft_plan_s;
    if (p == nullptr)
    {
        crdfft_plan_destroy(inner);
        return CRDFFT_OUT_OF_MEMORY;
    }
    p->inner = inner;
    p->shape.n = n;
    p->shape.batch = batch;
    p->shape.type = CRDFFT_C2C;
    p->shape.precision = CRDFFT_FP32;
    p->shape.algorithm = CRDFFT_ALGO_BLUESTEIN;
    p->shape.complex_n = n;
    p->shape.n1 = m;
    p->shape.chunk = fused_e > 0 ? std::max<uint64_t>(batch, 1) : chunk;
    p->shape.workspace_bytes = fused_e > 0 ? 0 : chunk * m * 8;
    p->shape.e = fused_e;
    p->idist = idist;
    p->odist = odist;
    p->flags = flags;
    p->blue_chunk = chunk;
    p->blue_e = fused_e;
    (void)hipGetDevice(&p->device);
    auto fail = [&](crdfft_status e) {
        delete p;
        return e;
    };
    if (hipError_t e = p->chirp.alloc(n * 8); e != hipSuccess)
        return fail(hip_status(e));
    if (hipError_t e = p->bhat.alloc(m * 8); e != hipSuccess)
        return fail(hip_status(e));
    if (fused_e > 0)
    {
        const uint64_t entries = crdfft::stage_twiddle_entries(m, fused_e);
        if (entries > 0)
        {
            if (hipError_t e = p->tw_single.alloc(entries * sizeof(cfloat)); e != hipSuccess)
                return fail(hip_status(e));
            if (hipError_t e = crdfft::build_stage_twiddles(p->tw_single.as<cfloat>(), m, fused_e, nullptr);
                e != hipSuccess)
                return fail(hip_status(e));
        }
    }
    else if (hipError_t e = p->bwork.alloc(chunk * m * 8); e != hipSuccess)
        return fail(hip_status(e));
    if (hipError_t e = crdfft::build_blue_tables(p->chirp.as<cfloat>(), p->bhat.as<cfloat>(), n, m, nullptr);
        e != hipSuccess)
        return fail(hip_status(e));
    // FFT_m(b) through the inner plan's kernels (the reference transforms b with its own plan, bluestein.hpp:67)
    Profiler none;
    st = run_complex(inner, p->bhat.as<cfloat>(), p->bhat.as<cfloat>(), 1, m, m, false, nullptr, none);
    if (st != CRDFFT_OK)
        return fail(st);
    if (hipError_t e = hipStreamSynchronize(nullptr); e != hipSuccess)
        return fail(hip_status(e));
    *plan = p;
    return CRDFFT_OK;
}

crdfft_status crdfft_plan_coop_status(crdfft_plan p, void* stream, uint32_t* waits_given_up, uint32_t* mode)
{
    if (p == nullptr || waits_given_up == nullptr || mode == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *waits_given_up = 0;
    *mode = 0;
    if (p->shape.algorithm != CRDFFT_ALGO_COOP)
    {
        return CRDFFT_OK;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(p->workspace_of.get(p->workspace, s));
    if (ws == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    CRDFFT_TRY(hipStreamSynchronize(s));
    uint32_t words[96];
    CRDFFT_TRY(hipMemcpy(words, ws + p->coop_ctl_offset, sizeof(words), hipMemcpyDeviceToHost));
    *waits_given_up = words[crdfft::kCoopCtlError];
    *mode = words[crdfft::kCoopCtlMode];
    return CRDFFT_OK;
}

crdfft_status crdfft_plan_destroy(crdfft_plan plan)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    delete plan;
    return CRDFFT_OK;
}

crdfft_status crdfft_execute(crdfft_plan p, const void* in, void* out, int32_t direction, void* stream)
{
    Profiler none;
    return execute_impl(p, in, out, direction, stream, none);
}

crdfft_status crdfft_execute_profiled(crdfft_plan p, const void* in, void* out, int32_t direction, void* stream,
                                      crdfft_phase_times* acc)
{
    if (acc == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    Profiler prof;
    prof.acc = acc;
    prof.stream = static_cast<hipStream_t>(stream);
    const crdfft_status st = execute_impl(p, in, out, direction, stream, prof);
    const hipError_t e = prof.finish();
    if (st != CRDFFT_OK)
    {
        return st;
    }
    return hip_status(e);
}

} // extern "C"

namespace crdfft
{
namespace
{
thread_local const ScratchOverride* t_override = nullptr;
}
const ScratchOverride* scratch_override() { return t_override; }
void set_scratch_override(const ScratchOverride* o) { t_override = o; }

crdfft_status plan_use_floor(crdfft_plan p)
{
    if (p == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const uint64_t cn = p->shape.complex_n;
    {
        std::lock_guard<std::mutex> lock(p->ref_mutex);
        if (p->wref.ptr == nullptr && cn >= 2)
        {
            CRDFFT_TRY(p->wref.alloc((cn / 2) * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_roots(p->wref.as<cfloat>(), cn / 2, 1, cn, nullptr));
        }
    }
    p->floor = true;
    return CRDFFT_OK;
}

bool plan_on_floor(crdfft_plan p) { return p != nullptr && p->floor; }
} // namespace crdfft

namespace
{
// Element-major batches of n = n1 * n2 > 4096 (the reference's execute_batched takes every power of two,
// fft_kernels.cpp:593-784): pass 1 transforms the n1-point strided columns of every (i2, t) into the stream's
// workspace with the twiddle W_n^{i2 k1}, pass 2 the n2-point columns of every (k1, t) back into data at
// element k1 + n1*k2. The workspace holds the whole batch (n * batch points) and grows with it.
crdfft_status execute_batched_four_step(crdfft_plan p, cfloat* data, uint64_t batch, bool inverse, hipStream_t s)
{
    const uint64_t n = p->shape.complex_n;
    const uint32_t lg = static_cast<uint32_t>(std::countr_zero(n));
    const uint64_t n1 = uint64_t{1} << ((lg + 1) / 2);
    const uint64_t n2 = n >> ((lg + 1) / 2);
    int e1 = 0, c1 = 0, e2 = 0, c2 = 0;
    if (!crdfft::em_col_shape(n1, &e1, &c1) || !crdfft::em_col_shape(n2, &e2, &c2))
    {
        return CRDFFT_UNSUPPORTED;
    }
    if (batch > (uint64_t{1} << 40) / n)
    {
        return CRDFFT_INVALID_ARGUMENT; // more than 8 TiB of data
    }
    const uint64_t bytes = n * batch * sizeof(cfloat);
    cfloat* ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(p->ref_mutex);
        if (p->em_thi.ptr == nullptr)
        {
            const uint64_t t1 = crdfft::stage_twiddle_entries(n1, e1), t2 = crdfft::stage_twiddle_entries(n2, e2);
            CRDFFT_TRY(p->em_tw1.alloc(std::max<uint64_t>(t1, 1) * sizeof(cfloat)));
            CRDFFT_TRY(p->em_tw2.alloc(std::max<uint64_t>(t2, 1) * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_stage_twiddles(p->em_tw1.as<cfloat>(), n1, e1, s));
            CRDFFT_TRY(crdfft::build_stage_twiddles(p->em_tw2.as<cfloat>(), n2, e2, s));
            const uint32_t hshift = lg / 2;
            const uint64_t lo = uint64_t{1} << hshift, hi = n >> hshift;
            CRDFFT_TRY(p->em_tlo.alloc(lo * sizeof(cfloat)));
            CRDFFT_TRY(p->em_thi.alloc(hi * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_roots(p->em_thi.as<cfloat>(), hi, lo, n, s));
            CRDFFT_TRY(crdfft::build_roots(p->em_tlo.as<cfloat>(), lo, 1, n, s));
            p->em_hshift = hshift;
        }
        std::unique_ptr<DeviceBuffer>* slot = nullptr;
        for (auto& w : p->em_work)
        {
            if (w.first == s)
            {
                slot = &w.second;
            }
        }
        if (slot == nullptr)
        {
            if (p->em_work.size() >= crdfft::PerStream::kMaxStreams)
            {
                return CRDFFT_OUT_OF_MEMORY;
            }
            p->em_work.emplace_back(s, std::make_unique<DeviceBuffer>());
            slot = &p->em_work.back().second;
        }
        if ((*slot)->bytes < bytes)
        {
            auto grown = std::make_unique<DeviceBuffer>();
            CRDFFT_TRY(grown->alloc(bytes)); // the old buffer's free (after this) waits for the device
            *slot = std::move(grown);
        }
        ws = (*slot)->as<cfloat>();
    }
    crdfft::EmArgs a1{data, ws, p->em_tw1.as<cfloat>(), p->em_thi.as<cfloat>(), p->em_tlo.as<cfloat>(), p->em_hshift,
                      batch, n2, batch, batch, n2 * batch, n2 * batch};
    CRDFFT_TRY(crdfft::launch_cols_em(n1, inverse, a1, s));
    crdfft::EmArgs a2{ws, data, p->em_tw2.as<cfloat>(), nullptr, nullptr, 0, batch, n1, n2 * batch, batch, batch,
                      n1 * batch};
    CRDFFT_TRY(crdfft::launch_cols_em(n2, inverse, a2, s));
    return CRDFFT_OK;
}
} // namespace

extern "C" {

crdfft_status crdfft_execute_batched(crdfft_plan p, void* data, uint64_t batch, int32_t direction, void* stream)
{
    if (p == nullptr || data == nullptr || (direction != CRDFFT_FORWARD && direction != CRDFFT_INVERSE))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const crdfft::PlanShape& sh = p->shape;
    if (sh.type != CRDFFT_C2C || sh.precision != CRDFFT_FP32)
    {
        return CRDFFT_UNSUPPORTED;
    }
    if (batch == 0)
    {
        return CRDFFT_OK;
    }
    if (sh.complex_n == 1)
    {
        return CRDFFT_OK;
    }
    int e = 0, c = 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool inverse = direction == CRDFFT_INVERSE;
    if (!crdfft::element_major_shape(sh.complex_n, batch, &e, &c))
    {
        return execute_batched_four_step(p, static_cast<cfloat*>(data), batch, inverse, s);
    }
    // the element-major kernel's stage table: the plan's own when the shapes agree, else built once and kept
    const cfloat* tw = nullptr;
    if (sh.algorithm == CRDFFT_ALGO_SINGLE && sh.e == e)
    {
        tw = p->tw_single.as<cfloat>();
    }
    else
    {
        std::lock_guard<std::mutex> lock(p->ref_mutex);
        if (p->tw_em.ptr == nullptr)
        {
            const uint64_t entries = crdfft::stage_twiddle_entries(sh.complex_n, e);
            CRDFFT_TRY(p->tw_em.alloc(std::max<uint64_t>(entries, 1) * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_stage_twiddles(p->tw_em.as<cfloat>(), sh.complex_n, e, s));
        }
        tw = p->tw_em.as<cfloat>();
    }
    crdfft::ColArgs a{static_cast<const cfloat*>(data), static_cast<cfloat*>(data), tw, nullptr, nullptr, nullptr, nullptr,
                      0, batch, 1, 0, 0, 0};
    CRDFFT_TRY(crdfft::launch_cols(sh.complex_n, e, c, inverse, a, s));
    return CRDFFT_OK;
}

crdfft_status crdfft_execute_reference(crdfft_plan p, const void* in, void* out, int32_t direction, void* stream)
{
    if (p == nullptr || in == nullptr || out == nullptr || (in == out && p->idist != p->odist) ||
        (direction != CRDFFT_FORWARD && direction != CRDFFT_INVERSE))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const crdfft::PlanShape& sh = p->shape;
    if (sh.type != CRDFFT_C2C)
    {
        return CRDFFT_UNSUPPORTED;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t cn = sh.complex_n;
    {
        std::lock_guard<std::mutex> lock(p->ref_mutex);
        if (p->wref.ptr == nullptr && cn >= 2)
        {
            CRDFFT_TRY(p->wref.alloc((cn / 2) * sizeof(cfloat)));
            CRDFFT_TRY(crdfft::build_roots(p->wref.as<cfloat>(), cn / 2, 1, cn, s));
        }
    }
    CRDFFT_TRY(crdfft::launch_reference(static_cast<const cfloat*>(in), static_cast<cfloat*>(out),
                                        p->wref.as<cfloat>(), cn, sh.batch, p->idist, p->odist,
                                        direction == CRDFFT_INVERSE, s));
    return CRDFFT_OK;
}

crdfft_status crdfft_plan_get_info(crdfft_plan p, crdfft_plan_info* info)
{
    if (p == nullptr || info == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    crdfft::shape_to_info(p->shape, info);
    info->check = p->check;
    info->table_bytes = p->tw_single.bytes + p->tw1.bytes + p->tw2.bytes + p->thi.bytes + p->tlo.bytes +
                        p->tstep.bytes + p->tfull.bytes + p->wreal.bytes + p->split_fwd[0].bytes + p->split_inv[0].bytes + p->split_tw[0].bytes +
                        p->split_fwd[1].bytes + p->split_inv[1].bytes + p->split_tw[1].bytes + p->split_tcm.bytes;
    return CRDFFT_OK;
}

crdfft_status crdfft_plan_describe(uint64_t n, uint64_t batch, int32_t type, int32_t precision,
                                   crdfft_plan_info* info)
{
    return crdfft_plan_describe_tuned(n, batch, type, precision, nullptr, info);
}

crdfft_status crdfft_plan_describe_tuned(uint64_t n, uint64_t batch, int32_t type, int32_t precision,
                                         const crdfft_tuning* tuning, crdfft_plan_info* info)
{
    if (info == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    crdfft::PlanShape shape;
    const crdfft_status st = crdfft::choose_shape(n, batch, type, precision, tuning, &shape);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    crdfft::shape_to_info(shape, info);
    return CRDFFT_OK;
}

const char* crdfft_status_string(crdfft_status status)
{
    switch (status)
    {
        case CRDFFT_OK:
            return "ok";
        case CRDFFT_INVALID_SIZE:
            return "invalid size";
        case CRDFFT_INVALID_ARGUMENT:
            return "invalid argument";
        case CRDFFT_OUT_OF_MEMORY:
            return "out of device memory";
        case CRDFFT_HIP_ERROR:
            return "HIP runtime error";
        case CRDFFT_SELF_CHECK_FAILED:
            return "self-check failed";
        case CRDFFT_UNSUPPORTED:
            return "unsupported layout or precision";
        case CRDFFT_NO_DEVICE:
            return "no HIP device";
    }
    return "unknown status";
}

int32_t crdfft_last_hip_error(void) { return g_last_hip_error; }

const char* crdfft_version(void) { return "crdfft 0.1 gfx950"; }

// ---------------------------------------------------------------------------------------------------------------
// runtime plumbing
// ---------------------------------------------------------------------------------------------------------------
crdfft_status crdfft_device_count(int32_t* count)
{
    if (count == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    int c = 0;
    const hipError_t e = hipGetDeviceCount(&c);
    *count = e == hipSuccess ? c : 0;
    return e == hipSuccess ? CRDFFT_OK : CRDFFT_NO_DEVICE;
}

crdfft_status crdfft_device_props_get(int32_t device, crdfft_device_props* props)
{
    if (props == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    hipDeviceProp_t d{};
    CRDFFT_TRY(hipGetDeviceProperties(&d, device));
    std::memset(props, 0, sizeof(*props));
    std::strncpy(props->name, d.name, sizeof(props->name) - 1);
    std::strncpy(props->arch, d.gcnArchName, sizeof(props->arch) - 1);
    props->compute_units = d.multiProcessorCount;
    props->clock_khz = d.clockRate;
    props->memory_clock_khz = d.memoryClockRate;
    props->memory_bus_width = d.memoryBusWidth;
    props->global_mem_bytes = d.totalGlobalMem;
    props->l2_bytes = static_cast<uint64_t>(d.l2CacheSize);
    props->max_lds_per_block = static_cast<int32_t>(d.sharedMemPerBlock);
    return CRDFFT_OK;
}

crdfft_status crdfft_malloc(void** ptr, uint64_t bytes)
{
    if (ptr == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *ptr = nullptr;
    CRDFFT_TRY(hipMalloc(ptr, bytes));
    return CRDFFT_OK;
}

crdfft_status crdfft_pointer_is_device(const void* ptr, int32_t* is_device)
{
    if (is_device == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *is_device = 0;
    if (ptr == nullptr)
    {
        return CRDFFT_OK;
    }
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess)
    {
        (void)hipGetLastError(); // pageable host memory the runtime has never seen: not device memory
        return CRDFFT_OK;
    }
    *is_device = (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) ? 1 : 0;
    return CRDFFT_OK;
}

crdfft_status crdfft_free(void* ptr)
{
    CRDFFT_TRY(hipFree(ptr));
    return CRDFFT_OK;
}

crdfft_status crdfft_memcpy_h2d(void* dst, const void* src, uint64_t bytes)
{
    CRDFFT_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CRDFFT_OK;
}

crdfft_status crdfft_memcpy_d2h(void* dst, const void* src, uint64_t bytes)
{
    CRDFFT_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CRDFFT_OK;
}

crdfft_status crdfft_memcpy_d2d(void* dst, const void* src, uint64_t bytes, void* stream)
{
    CRDFFT_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_stream_copy(void* dst, const void* src, uint64_t bytes, void* stream)
{
    if ((dst == nullptr || src == nullptr) && bytes != 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    CRDFFT_TRY(crdfft::launch_stream_copy(dst, src, bytes, static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_fill_bytes(void* dst, int32_t value, uint64_t bytes, void* stream)
{
    if (dst == nullptr && bytes != 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    CRDFFT_TRY(crdfft::launch_fill_bytes(dst, value, bytes, static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_stream_create(void** stream)
{
    if (stream == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    hipStream_t s = nullptr;
    CRDFFT_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return CRDFFT_OK;
}

crdfft_status crdfft_stream_destroy(void* stream)
{
    CRDFFT_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_stream_synchronize(void* stream)
{
    CRDFFT_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_device_synchronize(void)
{
    CRDFFT_TRY(hipDeviceSynchronize());
    return CRDFFT_OK;
}

crdfft_status crdfft_event_create(void** event)
{
    if (event == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    hipEvent_t e = nullptr;
    CRDFFT_TRY(hipEventCreate(&e));
    *event = e;
    return CRDFFT_OK;
}

crdfft_status crdfft_event_destroy(void* event)
{
    CRDFFT_TRY(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return CRDFFT_OK;
}

crdfft_status crdfft_event_record(void* event, void* stream)
{
    CRDFFT_TRY(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_event_synchronize(void* event)
{
    CRDFFT_TRY(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return CRDFFT_OK;
}

crdfft_status crdfft_event_elapsed_ms(float* ms, void* start, void* stop)
{
    if (ms == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    CRDFFT_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return CRDFFT_OK;
}

crdfft_status crdfft_graph_begin(void* stream)
{
    if (stream == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT; // the null stream cannot be captured
    }
    CRDFFT_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
    return CRDFFT_OK;
}

crdfft_status crdfft_graph_end(void* stream, void** graph)
{
    if (stream == nullptr || graph == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *graph = nullptr;
    hipGraph_t g = nullptr;
    CRDFFT_TRY(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CRDFFT_TRY(e);
    *graph = exec;
    return CRDFFT_OK;
}

crdfft_status crdfft_graph_launch(void* graph, void* stream)
{
    if (graph == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    CRDFFT_TRY(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), static_cast<hipStream_t>(stream)));
    return CRDFFT_OK;
}

crdfft_status crdfft_graph_destroy(void* graph)
{
    if (graph == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    CRDFFT_TRY(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)));
    return CRDFFT_OK;
}

} // extern "C"
