// The real-to-real transforms' C-ABI (crdfft_dct_*): the GPU form of the reference's DctPlan<f32>
// (engine/numerics/hesap-fft/include/crd/hesap/fft/dct.hpp:34-170) for batches of rows. Lengths up to 32768 run one
// kernel per call (dct_kernels.hip); longer ones, or any length with CRDFFT_FLAG_SEPARATE_PASSES, run the shuffle /
// Makhoul rotation passes around the library's own R2C and C2R plans.
#include <algorithm>
#include <new>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "dct.hpp"
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

#define DCT_TRY(expr)                                                                                                  \
    do                                                                                                                 \
    {                                                                                                                  \
        const hipError_t dct_e_ = (expr);                                                                              \
        if (dct_e_ != hipSuccess)                                                                                      \
            return st_of(dct_e_);                                                                                      \
    } while (0)

#define DCT_ST(expr)                                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        const crdfft_status dct_s_ = (expr);                                                                           \
        if (dct_s_ != CRDFFT_OK)                                                                                       \
            return dct_s_;                                                                                             \
    } while (0)

} // namespace

struct crdfft_dct_plan_s
{
    uint64_t n = 0, batch = 0, idist = 0, odist = 0;
    uint32_t flags = 0;
    int e = 0; // > 0: the one-kernel form at e values per thread
    uint64_t chunk = 0;
    uint64_t spec_offset = 0; // byte offset of the spectrum inside a workspace
    crdfft::DeviceBuffer cs, rec, tw, work;
    crdfft::PerStream work_of;
    crdfft_plan r2c = nullptr, c2r = nullptr;
    ~crdfft_dct_plan_s()
    {
        if (r2c != nullptr)
        {
            crdfft_plan_destroy(r2c);
        }
        if (c2r != nullptr)
        {
            crdfft_plan_destroy(c2r);
        }
    }
};

extern "C" {

crdfft_status crdfft_dct_plan_create(crdfft_dct_plan* plan, uint64_t n, uint64_t batch, uint64_t idist, uint64_t odist,
                                     uint32_t flags)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *plan = nullptr;
    // dct.hpp:43: a power of two >= 2; the separate passes run the library's real FFT, complex length up to 2^24
    if (n < 2 || (n & (n - 1)) != 0 || n > 2 * crdfft::kFourStepMax)
    {
        return CRDFFT_INVALID_SIZE;
    }
    if ((flags & ~(CRDFFT_FLAG_NO_SELF_CHECK | CRDFFT_FLAG_SEPARATE_PASSES)) != 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    idist = idist == 0 ? n : idist;
    odist = odist == 0 ? n : odist;
    if (batch > 1 && (idist < n || odist < n))
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    auto* p = new (std::nothrow) crdfft_dct_plan_s;
    if (p == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    p->n = n;
    p->batch = batch;
    p->idist = idist;
    p->odist = odist;
    p->flags = flags;
    auto fail = [&](crdfft_status st) {
        delete p;
        return st;
    };
    const uint64_t h = n / 2;
    if (hipError_t e = p->cs.alloc((h + 1) * sizeof(crdfft::cfloat)); e != hipSuccess)
        return fail(st_of(e));
    if (hipError_t e = p->rec.alloc(h * sizeof(crdfft::cfloat)); e != hipSuccess)
        return fail(st_of(e));
    if (hipError_t e = crdfft::build_dct_tables(p->cs.as<crdfft::cfloat>(), p->rec.as<crdfft::cfloat>(), n, nullptr);
        e != hipSuccess)
        return fail(st_of(e));
    const bool passes = n >= 4 && ((flags & CRDFFT_FLAG_SEPARATE_PASSES) != 0 || crdfft::dct_one_kernel_e(n) == 0);
    if (!passes && n >= 4)
    {
        p->e = crdfft::dct_one_kernel_e(n);
        const uint64_t entries = crdfft::stage_twiddle_entries(h, p->e);
        if (entries > 0)
        {
            if (hipError_t e = p->tw.alloc(entries * sizeof(crdfft::cfloat)); e != hipSuccess)
                return fail(st_of(e));
            if (hipError_t e = crdfft::build_stage_twiddles(p->tw.as<crdfft::cfloat>(), h, p->e, nullptr);
                e != hipSuccess)
                return fail(st_of(e));
        }
    }
    else if (passes)
    {
        // workspace per chunk of rows: the shuffled row / irfft output (n floats) and the half spectrum (h+1 bins)
        const uint64_t row_bytes = n * sizeof(float) + (h + 1) * sizeof(crdfft::cfloat);
        p->chunk = std::max<uint64_t>(1, std::min<uint64_t>(std::max<uint64_t>(batch, 1), crdfft::kChunkBytes / row_bytes));
        p->spec_offset = (p->chunk * n * sizeof(float) + 255) / 256 * 256;
        if (hipError_t e = p->work.alloc(p->spec_offset + p->chunk * (h + 1) * sizeof(crdfft::cfloat)); e != hipSuccess)
            return fail(st_of(e));
        const uint32_t inner = CRDFFT_FLAG_NO_SELF_CHECK;
        crdfft_status st = crdfft_plan_create(&p->r2c, n, p->chunk, n, h + 1, CRDFFT_R2C, CRDFFT_FP32, inner);
        if (st != CRDFFT_OK)
            return fail(st);
        st = crdfft_plan_create(&p->c2r, n, p->chunk, h + 1, n, CRDFFT_C2R, CRDFFT_FP32, inner);
        if (st != CRDFFT_OK)
            return fail(st);
    }
    if (hipError_t e = hipStreamSynchronize(nullptr); e != hipSuccess)
        return fail(st_of(e));
    *plan = p;
    return CRDFFT_OK;
}

crdfft_status crdfft_dct_plan_destroy(crdfft_dct_plan plan)
{
    if (plan == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    delete plan;
    return CRDFFT_OK;
}

crdfft_status crdfft_dct_plan_get_info(crdfft_dct_plan p, crdfft_plan_info* info)
{
    if (p == nullptr || info == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *info = crdfft_plan_info{};
    info->n = p->n;
    info->batch = p->batch;
    info->type = CRDFFT_R2C;
    info->precision = CRDFFT_FP32;
    info->algorithm = CRDFFT_ALGO_SINGLE;
    if (p->r2c != nullptr)
    {
        crdfft_plan_info inner{};
        if (crdfft_plan_get_info(p->r2c, &inner) == CRDFFT_OK)
        {
            info->algorithm = inner.algorithm; // the separate passes: the algorithm of the real FFT they run
        }
    }
    info->complex_n = p->n / 2;
    info->e = p->e;
    info->chunk = p->e > 0 || p->n == 2 ? p->batch : p->chunk;
    info->workspace_bytes = p->work.bytes;
    info->table_bytes = p->cs.bytes + p->rec.bytes + p->tw.bytes;
    return CRDFFT_OK;
}

crdfft_status crdfft_dct_execute(crdfft_dct_plan p, int32_t kind, const float* in, float* out, void* stream)
{
    if (p == nullptr || in == nullptr || out == nullptr || kind < CRDFFT_DCT2 || kind > CRDFFT_DST3)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    if (in == out && p->idist != p->odist)
    {
        return CRDFFT_INVALID_ARGUMENT; // in place needs one row layout
    }
    if (p->batch == 0)
    {
        return CRDFFT_OK;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const crdfft::DctArgs a{in,       out,     p->cs.as<crdfft::cfloat>(), p->rec.as<crdfft::cfloat>(),
                            p->tw.as<crdfft::cfloat>(), p->batch, p->idist, p->odist};
    if (p->n == 2)
    {
        DCT_TRY(crdfft::launch_dct_n2(kind, a, s));
        return CRDFFT_OK;
    }
    if (p->e > 0)
    {
        DCT_TRY(crdfft::launch_dct_one_kernel(p->n, p->e, kind, a, s));
        return CRDFFT_OK;
    }
    char* work = static_cast<char*>(p->work_of.get(p->work, s));
    if (work == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    float* w = reinterpret_cast<float*>(work);
    auto* spec = reinterpret_cast<crdfft::cfloat*>(work + p->spec_offset);
    const bool sine = kind == CRDFFT_DST2 || kind == CRDFFT_DST3;
    const bool forward = kind == CRDFFT_DCT2 || kind == CRDFFT_DST2;
    // the inner plans always run a whole chunk; rows past the last chunk's count are workspace only
    for (uint64_t r0 = 0; r0 < p->batch; r0 += p->chunk)
    {
        const uint64_t cnt = std::min(p->chunk, p->batch - r0);
        const float* x = in + r0 * p->idist;
        float* y = out + r0 * p->odist;
        if (forward)
        {
            DCT_TRY(crdfft::launch_dct_shuffle(x, w, p->n, cnt, p->idist, sine, s));
            DCT_ST(crdfft_execute(p->r2c, w, spec, CRDFFT_FORWARD, s));
            DCT_TRY(crdfft::launch_dct_fwd_post(spec, y, p->cs.as<crdfft::cfloat>(), p->n, cnt, p->odist, sine, s));
        }
        else
        {
            DCT_TRY(crdfft::launch_dct_inv_pre(x, spec, p->cs.as<crdfft::cfloat>(), p->n, cnt, p->idist, sine, s));
            DCT_ST(crdfft_execute(p->c2r, spec, w, CRDFFT_INVERSE, s));
            DCT_TRY(crdfft::launch_dct_inv_post(w, y, p->n, cnt, p->odist, sine, s));
        }
    }
    return CRDFFT_OK;
}

} // extern "C"
