// N-dimensional plan (crdfft_nd_*): the reference's NdFftPlan (engine/numerics/hesap-fft/include/crd/hesap/fft/
// nd_fft.hpp:28-196) as a sequence of per-axis launches over device memory.
//
// The volume is row-major: axis i has d_i points at stride inner_i = prod_{j>i} d_j, repeated over outer_i blocks.
//   - the last axis (inner = 1): contiguous lines, one 1D plan over outer * batch of them (power of two: FftPlan's
//     kernels; otherwise Bluestein, as the reference does, nd_fft.hpp:66-78);
//   - a power-of-two axis up to 4096 with inner > 1: the strided column kernel in place (no gather: the reference
//     copies each strided line to a buffer first, nd_fft.hpp:156-182);
//   - any other axis: a tiled transpose to contiguous lines, the 1D plan, and the transpose back.
// The inverse runs every axis inverse and scales by 1/prod(d) (nd_fft.hpp:104); Bluestein axes carry their own 1/d.
#include <algorithm>
#include <new>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "kernels.hpp"
#include "planner.hpp"
#include "scratch.hpp"

namespace
{

crdfft_status st_of(hipError_t e)
{
    if (e == hipSuccess)
    {
        return CRDFFT_OK;
    }
    return e == hipErrorOutOfMemory ? CRDFFT_OUT_OF_MEMORY : CRDFFT_HIP_ERROR;
}

#define ND_TRY(expr)                                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        const hipError_t nd_e_ = (expr);                                                                               \
        if (nd_e_ != hipSuccess)                                                                                       \
            return st_of(nd_e_);                                                                                       \
    } while (0)

enum AxisKind
{
    kSkip = 0,
    kRows = 1,
    kColumns = 2,
    kTransposed = 3
};

struct Axis
{
    uint64_t d = 1, outer = 1, inner = 1;
    AxisKind kind = kSkip;
    bool bluestein = false;
    crdfft_plan plan = nullptr; // kRows / kTransposed
    void* tw = nullptr;         // kColumns stage table
    int e = 0, c = 0;
};

} // namespace

struct crdfft_nd_plan_s
{
    std::vector<Axis> axes;
    uint64_t total = 1;
    uint64_t batch = 1;
    crdfft::DeviceBuffer work; // transpose buffer of the axes that go through transposed lines
    crdfft::PerStream work_of;
    ~crdfft_nd_plan_s()
    {
        for (Axis& a : axes)
        {
            if (a.plan != nullptr)
            {
                crdfft_plan_destroy(a.plan);
            }
            if (a.tw != nullptr)
            {
                (void)hipFree(a.tw);
            }
        }
    }
};

namespace
{

crdfft_status line_plan(uint64_t d, uint64_t lines, crdfft_plan* out, bool* blue)
{
    *blue = !crdfft::is_pow2(d);
    if (*blue)
    {
        return crdfft_bluestein_plan_create(out, d, lines, d, d, 0);
    }
    return crdfft_plan_create(out, d, lines, d, d, CRDFFT_C2C, CRDFFT_FP32, CRDFFT_FLAG_NO_SELF_CHECK);
}

crdfft_status build(crdfft_nd_plan_s* p, const uint64_t* dims, uint32_t ndim, uint64_t batch, uint32_t flags)
{
    p->batch = batch;
    p->axes.resize(ndim);
    for (uint32_t i = 0; i < ndim; ++i)
    {
        p->axes[i].d = dims[i];
        p->total *= dims[i];
    }
    uint64_t inner = 1;
    for (uint32_t i = ndim; i-- > 0;)
    {
        Axis& a = p->axes[i];
        a.inner = inner;
        a.outer = p->total / (a.d * inner);
        inner *= a.d;
    }
    bool need_work = false;
    for (Axis& a : p->axes)
    {
        if (a.d == 1)
        {
            a.kind = kSkip;
            continue;
        }
        const uint64_t blocks = a.outer * batch;
        if (a.inner == 1)
        {
            a.kind = kRows;
            crdfft_status st = line_plan(a.d, blocks, &a.plan, &a.bluestein);
            if (st != CRDFFT_OK)
            {
                return st;
            }
        }
        else if ((flags & CRDFFT_ND_TRANSPOSE_ALL) == 0 && crdfft::is_pow2(a.d) &&
                 crdfft::element_major_shape(a.d, a.inner, &a.e, &a.c))
        {
            a.kind = kColumns;
            const uint64_t entries = std::max<uint64_t>(1, crdfft::stage_twiddle_entries(a.d, a.e));
            ND_TRY(crdfft::device_malloc(reinterpret_cast<void**>(&a.tw), entries * sizeof(crdfft::cfloat)));
            ND_TRY(crdfft::build_stage_twiddles(static_cast<crdfft::cfloat*>(a.tw), a.d, a.e, nullptr));
        }
        else
        {
            a.kind = kTransposed;
            need_work = true;
            crdfft_status st = line_plan(a.d, blocks * a.inner, &a.plan, &a.bluestein);
            if (st != CRDFFT_OK)
            {
                return st;
            }
        }
    }
    if (need_work)
    {
        ND_TRY(p->work.alloc(p->total * batch * sizeof(crdfft::cfloat)));
    }
    ND_TRY(hipStreamSynchronize(nullptr));
    return CRDFFT_OK;
}

} // namespace

extern "C" {

crdfft_status crdfft_nd_plan_create(crdfft_nd_plan* plan, const uint64_t* dims, uint32_t ndim, uint64_t batch,
                                    uint32_t flags)
{
    if (plan == nullptr || dims == nullptr || ndim == 0 || batch == 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *plan = nullptr;
    for (uint32_t i = 0; i < ndim; ++i)
    {
        if (dims[i] == 0)
        {
            return CRDFFT_INVALID_SIZE;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    {
        return CRDFFT_NO_DEVICE;
    }
    auto* p = new (std::nothrow) crdfft_nd_plan_s;
    if (p == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    const crdfft_status st = build(p, dims, ndim, batch, flags);
    if (st != CRDFFT_OK)
    {
        delete p;
        return st;
    }
    *plan = p;
    return CRDFFT_OK;
}

crdfft_status crdfft_time_nd(const uint64_t* dims, uint32_t ndim, uint64_t batch, uint32_t flags, uint32_t reps,
                            crdfft_plan_record* record)
{
    if (record == nullptr || dims == nullptr || ndim == 0 || ndim > 3)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    crdfft_nd_plan plan = nullptr;
    crdfft_status st = crdfft_nd_plan_create(&plan, dims, ndim, batch, flags);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    void* data = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const uint64_t bytes = plan->total * batch * sizeof(crdfft::cfloat);
    std::vector<float> ms(std::max<uint32_t>(reps, 1));
    if (hipMalloc(&data, bytes) != hipSuccess)
    {
        crdfft_nd_plan_destroy(plan);
        return CRDFFT_OUT_OF_MEMORY;
    }
    if (hipMemset(data, 0, bytes) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    {
        st = CRDFFT_HIP_ERROR;
    }
    if (st == CRDFFT_OK)
    {
        st = crdfft_nd_execute(plan, data, CRDFFT_FORWARD, s);
    }
    for (size_t r = 0; r < ms.size() && st == CRDFFT_OK; ++r)
    {
        if (hipEventRecord(e0, s) != hipSuccess)
        {
            st = CRDFFT_HIP_ERROR;
            break;
        }
        st = crdfft_nd_execute(plan, data, CRDFFT_FORWARD, s);
        if (st == CRDFFT_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms[r], e0, e1) != hipSuccess))
        {
            st = CRDFFT_HIP_ERROR;
        }
    }
    if (e0 != nullptr)
        (void)hipEventDestroy(e0);
    if (e1 != nullptr)
        (void)hipEventDestroy(e1);
    if (s != nullptr)
        (void)hipStreamDestroy(s);
    (void)hipFree(data);
    const uint64_t total = plan->total;
    crdfft_nd_plan_destroy(plan);
    if (st != CRDFFT_OK)
    {
        return st;
    }
    std::nth_element(ms.begin(), ms.begin() + ms.size() / 2, ms.end());
    *record = crdfft_plan_record{};
    record->n = total;
    record->batch = batch;
    record->type = CRDFFT_C2C;
    record->precision = CRDFFT_FP32;
    record->algorithm = CRDFFT_ALGO_FOUR_STEP;
    record->ndim = ndim;
    for (uint32_t i = 0; i < ndim; ++i)
    {
        record->dims[i] = dims[i];
    }
    record->stride = ndim > 1 ? dims[ndim - 1] : 1;
    record->tuning.hints = (flags & CRDFFT_ND_TRANSPOSE_ALL) ? 4u : 0u;
    record->ms = ms[ms.size() / 2];
    return CRDFFT_OK;
}

crdfft_status crdfft_nd_plan_destroy(crdfft_nd_plan plan)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    delete plan;
    return CRDFFT_OK;
}

crdfft_status crdfft_nd_execute(crdfft_nd_plan p, void* data, int32_t direction, void* stream)
{
    if (p == nullptr || data == nullptr || (direction != CRDFFT_FORWARD && direction != CRDFFT_INVERSE))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const bool inverse = direction == CRDFFT_INVERSE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto* x = static_cast<crdfft::cfloat*>(data);
    auto* w = static_cast<crdfft::cfloat*>(p->work_of.get(p->work, s));
    if (w == nullptr && p->work.bytes != 0)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    double scale = 1.0;
    for (const Axis& a : p->axes)
    {
        const uint64_t blocks = a.outer * p->batch;
        switch (a.kind)
        {
            case kSkip:
                break;
            case kRows:
            {
                const crdfft_status st = crdfft_execute(a.plan, x, x, direction, stream);
                if (st != CRDFFT_OK)
                {
                    return st;
                }
                break;
            }
            case kColumns:
            {
                crdfft::ColArgs ca{x, x, static_cast<const crdfft::cfloat*>(a.tw), nullptr, nullptr, nullptr, nullptr, 0,
                                   a.inner, blocks, a.d * a.inner, a.d * a.inner, 0};
                ND_TRY(crdfft::launch_cols(a.d, a.e, a.c, inverse, ca, s));
                break;
            }
            case kTransposed:
            {
                ND_TRY(crdfft::launch_transpose(x, w, a.d, a.inner, blocks, a.d * a.inner, a.d * a.inner, s));
                const crdfft_status st = crdfft_execute(a.plan, w, w, direction, stream);
                if (st != CRDFFT_OK)
                {
                    return st;
                }
                ND_TRY(crdfft::launch_transpose(w, x, a.inner, a.d, blocks, a.d * a.inner, a.d * a.inner, s));
                break;
            }
        }
        if (inverse && a.kind != kSkip && !a.bluestein)
        {
            scale /= static_cast<double>(a.d);
        }
    }
    if (inverse && scale != 1.0)
    {
        ND_TRY(crdfft::launch_scale(x, p->total * p->batch, static_cast<float>(scale), s));
    }
    return CRDFFT_OK;
}

} // extern "C"
