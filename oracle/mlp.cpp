// Test infrastructure only: FP64 restatement of the plan cost model (the 3-layer ReLU perceptron of
// crd_amd/csrc/mlp_kernels.hip) and of its mean-squared-error gradient, the checker of the GPU kernels' gradients.
// Pinned by central finite differences in tests/test_oracle.py. Parameter layout: W1[h][in], b1[h], W2[h][h], b2[h],
// w3[h], b3.
#include <cstddef>
#include <cstdint>
#include <vector>

namespace
{

struct Off
{
    size_t w1, b1, w2, b2, w3, b3, total;
};

Off offsets(uint32_t in, uint32_t h)
{
    Off o{};
    o.w1 = 0;
    o.b1 = o.w1 + static_cast<size_t>(h) * in;
    o.w2 = o.b1 + h;
    o.b2 = o.w2 + static_cast<size_t>(h) * h;
    o.w3 = o.b2 + h;
    o.b3 = o.w3 + h;
    o.total = o.b3 + 1;
    return o;
}

} // namespace

extern "C" {

void oracle_mlp_forward(const double* p, uint32_t in, uint32_t h, const double* x, uint64_t n, double* y)
{
    const Off o = offsets(in, h);
    std::vector<double> h1(h), h2(h);
    for (uint64_t r = 0; r < n; ++r)
    {
        const double* xr = x + r * in;
        for (uint32_t m = 0; m < h; ++m)
        {
            double a = p[o.b1 + m];
            for (uint32_t k = 0; k < in; ++k)
            {
                a += p[o.w1 + m * in + k] * xr[k];
            }
            h1[m] = a > 0 ? a : 0.0;
        }
        for (uint32_t m = 0; m < h; ++m)
        {
            double a = p[o.b2 + m];
            for (uint32_t k = 0; k < h; ++k)
            {
                a += p[o.w2 + m * h + k] * h1[k];
            }
            h2[m] = a > 0 ? a : 0.0;
        }
        double a = p[o.b3];
        for (uint32_t k = 0; k < h; ++k)
        {
            a += p[o.w3 + k] * h2[k];
        }
        y[r] = a;
    }
}

// mean((y - t)^2) and its gradient (backpropagation through the two ReLU layers)
double oracle_mlp_loss_grad(const double* p, uint32_t in, uint32_t h, const double* x, const double* t, uint64_t n,
                            double* grad)
{
    const Off o = offsets(in, h);
    for (size_t i = 0; i < o.total; ++i)
    {
        grad[i] = 0.0;
    }
    std::vector<double> h1(h), h2(h), d2(h), d1(h);
    double loss = 0.0;
    for (uint64_t r = 0; r < n; ++r)
    {
        const double* xr = x + r * in;
        for (uint32_t m = 0; m < h; ++m)
        {
            double a = p[o.b1 + m];
            for (uint32_t k = 0; k < in; ++k)
            {
                a += p[o.w1 + m * in + k] * xr[k];
            }
            h1[m] = a > 0 ? a : 0.0;
        }
        for (uint32_t m = 0; m < h; ++m)
        {
            double a = p[o.b2 + m];
            for (uint32_t k = 0; k < h; ++k)
            {
                a += p[o.w2 + m * h + k] * h1[k];
            }
            h2[m] = a > 0 ? a : 0.0;
        }
        double y = p[o.b3];
        for (uint32_t k = 0; k < h; ++k)
        {
            y += p[o.w3 + k] * h2[k];
        }
        const double e = y - t[r];
        loss += e * e;
        const double g = 2.0 * e / static_cast<double>(n);
        grad[o.b3] += g;
        for (uint32_t k = 0; k < h; ++k)
        {
            grad[o.w3 + k] += g * h2[k];
            d2[k] = h2[k] > 0 ? g * p[o.w3 + k] : 0.0;
        }
        for (uint32_t m = 0; m < h; ++m)
        {
            grad[o.b2 + m] += d2[m];
            for (uint32_t k = 0; k < h; ++k)
            {
                grad[o.w2 + m * h + k] += d2[m] * h1[k];
            }
        }
        for (uint32_t k = 0; k < h; ++k)
        {
            double a = 0.0;
            for (uint32_t m = 0; m < h; ++m)
            {
                a += d2[m] * p[o.w2 + m * h + k];
            }
            d1[k] = h1[k] > 0 ? a : 0.0;
        }
        for (uint32_t m = 0; m < h; ++m)
        {
            grad[o.b1 + m] += d1[m];
            for (uint32_t k = 0; k < in; ++k)
            {
                grad[o.w1 + m * in + k] += d1[m] * xr[k];
            }
        }
    }
    return loss / static_cast<double>(n);
}

} // extern "C"
