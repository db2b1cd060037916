// Host side of the plan cost model (crdfft_mlp_*): parameters live on the device; prediction, the loss gradient and
// Adam training run in mlp_kernels.hip. Initialisation is He-normal from a seeded splitmix64 stream, so a seed gives
// the same network on every host.
#include <algorithm>
#include <cmath>
#include <new>
#include <numeric>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/crdfft.h"
#include "mlp.hpp"

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

#define MLP_TRY(expr)                                                                                                  \
    do                                                                                                                 \
    {                                                                                                                  \
        const hipError_t mlp_e_ = (expr);                                                                              \
        if (mlp_e_ != hipSuccess)                                                                                      \
            return st_of(mlp_e_);                                                                                      \
    } while (0)

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double unit(uint64_t& s) { return (static_cast<double>(splitmix(s) >> 11) + 0.5) / static_cast<double>(1ULL << 53); }

double normal(uint64_t& s)
{
    const double u1 = unit(s), u2 = unit(s);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Dev
{
    float* p = nullptr;
    hipError_t alloc(uint64_t count) { return hipMalloc(reinterpret_cast<void**>(&p), std::max<uint64_t>(count, 1) * 4); }
    ~Dev()
    {
        if (p != nullptr)
        {
            (void)hipFree(p);
        }
    }
};

// Row buffers for up to `rows` rows.
struct Scratch
{
    Dev x, t, h1, h2, y, g, dh2, dh1;
    hipError_t alloc(const crdfft::MlpDims& d, uint64_t rows)
    {
        hipError_t e;
        if ((e = x.alloc(rows * d.inputs)) != hipSuccess || (e = t.alloc(rows)) != hipSuccess ||
            (e = h1.alloc(rows * d.hidden)) != hipSuccess || (e = h2.alloc(rows * d.hidden)) != hipSuccess ||
            (e = y.alloc(rows)) != hipSuccess || (e = g.alloc(rows)) != hipSuccess ||
            (e = dh2.alloc(rows * d.hidden)) != hipSuccess || (e = dh1.alloc(rows * d.hidden)) != hipSuccess)
        {
            return e;
        }
        return hipSuccess;
    }
};

} // namespace

struct crdfft_mlp_s
{
    crdfft::MlpDims dims{};
    crdfft::MlpOffsets off{};
    Dev params, grad, m1, m2;
    uint32_t step = 0;
};

extern "C" {

crdfft_status crdfft_mlp_create(crdfft_mlp* mlp, uint32_t inputs, uint32_t hidden, uint64_t seed)
{
    if (mlp == nullptr || inputs == 0 || hidden == 0)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *mlp = nullptr;
    auto* m = new (std::nothrow) crdfft_mlp_s;
    if (m == nullptr)
    {
        return CRDFFT_OUT_OF_MEMORY;
    }
    m->dims = {inputs, hidden};
    m->off = crdfft::mlp_offsets(m->dims);
    const uint64_t n = m->off.total;
    std::vector<float> host(n, 0.0f);
    uint64_t s = seed;
    const double s1 = std::sqrt(2.0 / inputs), s2 = std::sqrt(2.0 / hidden);
    for (uint64_t i = m->off.w1; i < m->off.b1; ++i)
    {
        host[i] = static_cast<float>(normal(s) * s1);
    }
    for (uint64_t i = m->off.w2; i < m->off.b2; ++i)
    {
        host[i] = static_cast<float>(normal(s) * s2);
    }
    for (uint64_t i = m->off.w3; i < m->off.b3; ++i)
    {
        host[i] = static_cast<float>(normal(s) * s2);
    }
    hipError_t e;
    if ((e = m->params.alloc(n)) != hipSuccess || (e = m->grad.alloc(n)) != hipSuccess ||
        (e = m->m1.alloc(n)) != hipSuccess || (e = m->m2.alloc(n)) != hipSuccess ||
        (e = hipMemcpy(m->params.p, host.data(), n * 4, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(m->m1.p, 0, n * 4)) != hipSuccess || (e = hipMemset(m->m2.p, 0, n * 4)) != hipSuccess)
    {
        delete m;
        return st_of(e);
    }
    *mlp = m;
    return CRDFFT_OK;
}

crdfft_status crdfft_mlp_destroy(crdfft_mlp mlp)
{
    if (mlp == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    delete mlp;
    return CRDFFT_OK;
}

crdfft_status crdfft_mlp_param_count(crdfft_mlp mlp, uint64_t* count)
{
    if (mlp == nullptr || count == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    *count = mlp->off.total;
    return CRDFFT_OK;
}

crdfft_status crdfft_mlp_get_params(crdfft_mlp mlp, float* params)
{
    if (mlp == nullptr || params == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    MLP_TRY(hipMemcpy(params, mlp->params.p, mlp->off.total * 4, hipMemcpyDeviceToHost));
    return CRDFFT_OK;
}

crdfft_status crdfft_mlp_set_params(crdfft_mlp mlp, const float* params)
{
    if (mlp == nullptr || params == nullptr)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    MLP_TRY(hipMemcpy(mlp->params.p, params, mlp->off.total * 4, hipMemcpyHostToDevice));
    return CRDFFT_OK;
}

crdfft_status crdfft_mlp_predict(crdfft_mlp mlp, const float* x, uint64_t n, float* y)
{
    if (mlp == nullptr || ((x == nullptr || y == nullptr) && n != 0) || n > 0xFFFFFFFFull)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    if (n == 0)
    {
        return CRDFFT_OK;
    }
    Scratch sc;
    MLP_TRY(sc.alloc(mlp->dims, n));
    MLP_TRY(hipMemcpy(sc.x.p, x, n * mlp->dims.inputs * 4, hipMemcpyHostToDevice));
    MLP_TRY(crdfft::mlp_forward(mlp->dims, mlp->params.p, sc.x.p, static_cast<uint32_t>(n), sc.h1.p, sc.h2.p, sc.y.p,
                                nullptr));
    MLP_TRY(hipMemcpy(y, sc.y.p, n * 4, hipMemcpyDeviceToHost));
    return CRDFFT_OK;
}

crdfft_status crdfft_mlp_loss_grad(crdfft_mlp mlp, const float* x, const float* t, uint64_t n, float* loss,
                                   float* grad)
{
    if (mlp == nullptr || x == nullptr || t == nullptr || n == 0 || n > 0xFFFFFFFFull)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    Scratch sc;
    MLP_TRY(sc.alloc(mlp->dims, n));
    MLP_TRY(hipMemcpy(sc.x.p, x, n * mlp->dims.inputs * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(sc.t.p, t, n * 4, hipMemcpyHostToDevice));
    const uint32_t rows = static_cast<uint32_t>(n);
    MLP_TRY(crdfft::mlp_forward(mlp->dims, mlp->params.p, sc.x.p, rows, sc.h1.p, sc.h2.p, sc.y.p, nullptr));
    MLP_TRY(crdfft::mlp_backward(mlp->dims, mlp->params.p, sc.x.p, sc.t.p, rows, sc.h1.p, sc.h2.p, sc.y.p, sc.g.p,
                                 sc.dh2.p, sc.dh1.p, mlp->grad.p, nullptr));
    std::vector<float> y(n);
    MLP_TRY(hipMemcpy(y.data(), sc.y.p, n * 4, hipMemcpyDeviceToHost));
    if (loss != nullptr)
    {
        double acc = 0.0;
        for (uint64_t i = 0; i < n; ++i)
        {
            const double d = static_cast<double>(y[i]) - t[i];
            acc += d * d;
        }
        *loss = static_cast<float>(acc / static_cast<double>(n));
    }
    if (grad != nullptr)
    {
        MLP_TRY(hipMemcpy(grad, mlp->grad.p, mlp->off.total * 4, hipMemcpyDeviceToHost));
    }
    return CRDFFT_OK;
}

crdfft_status crdfft_mlp_train(crdfft_mlp mlp, const float* x, const float* t, uint64_t n, uint32_t epochs,
                               uint32_t batch, float lr, uint64_t seed, float* final_loss)
{
    if (mlp == nullptr || x == nullptr || t == nullptr || n == 0 || batch == 0 || n > 0xFFFFFFFFull)
    {
        return CRDFFT_INVALID_ARGUMENT;
    }
    const crdfft::MlpDims& d = mlp->dims;
    Dev dx, dt, didx;
    MLP_TRY(dx.alloc(n * d.inputs));
    MLP_TRY(dt.alloc(n));
    MLP_TRY(didx.alloc(n));
    MLP_TRY(hipMemcpy(dx.p, x, n * d.inputs * 4, hipMemcpyHostToDevice));
    MLP_TRY(hipMemcpy(dt.p, t, n * 4, hipMemcpyHostToDevice));
    const uint32_t bsz = static_cast<uint32_t>(std::min<uint64_t>(batch, n));
    Scratch sc;
    MLP_TRY(sc.alloc(d, bsz));
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    uint64_t s = seed;
    hipStream_t st = nullptr;
    MLP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipError_t err = hipSuccess;
    for (uint32_t ep = 0; ep < epochs && err == hipSuccess; ++ep)
    {
        for (uint64_t i = n - 1; i > 0; --i) // Fisher-Yates with the seeded stream
        {
            const uint64_t j = splitmix(s) % (i + 1);
            std::swap(perm[i], perm[j]);
        }
        err = hipMemcpyAsync(didx.p, perm.data(), n * 4, hipMemcpyHostToDevice, st);
        for (uint64_t b0 = 0; b0 + bsz <= n && err == hipSuccess; b0 += bsz)
        {
            const uint32_t* idx = reinterpret_cast<const uint32_t*>(didx.p) + b0;
            if ((err = crdfft::mlp_gather(dx.p, idx, sc.x.p, bsz, d.inputs, st)) != hipSuccess ||
                (err = crdfft::mlp_gather(dt.p, idx, sc.t.p, bsz, 1, st)) != hipSuccess ||
                (err = crdfft::mlp_forward(d, mlp->params.p, sc.x.p, bsz, sc.h1.p, sc.h2.p, sc.y.p, st)) !=
                    hipSuccess ||
                (err = crdfft::mlp_backward(d, mlp->params.p, sc.x.p, sc.t.p, bsz, sc.h1.p, sc.h2.p, sc.y.p, sc.g.p,
                                            sc.dh2.p, sc.dh1.p, mlp->grad.p, st)) != hipSuccess)
            {
                break;
            }
            ++mlp->step;
            err = crdfft::mlp_adam(mlp->params.p, mlp->grad.p, mlp->m1.p, mlp->m2.p, mlp->off.total, lr, mlp->step, st);
        }
        // the permutation buffer is rewritten next epoch: finish this epoch's reads of it first
        if (err == hipSuccess)
        {
            err = hipStreamSynchronize(st);
        }
    }
    (void)hipStreamDestroy(st);
    if (err != hipSuccess)
    {
        return st_of(err);
    }
    if (final_loss != nullptr)
    {
        return crdfft_mlp_loss_grad(mlp, x, t, n, final_loss, nullptr);
    }
    return CRDFFT_OK;
}

} // extern "C"
