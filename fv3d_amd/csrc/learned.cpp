// Host side of the learned block preconditioner: parameters, the training
// set gathered from the solver's own states and block-Jacobi blocks,
// GPU training (BF16-MFMA forward/backward, Adam) and the fvh_nn_* API.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "fvh_internal.h"
#include "krylov.h"
#include "learned.h"
#include "runtime.h"

namespace fvh {

namespace {

template <class X> X *dalloc(size_t n)
{
    void *p = nullptr;
    FVH_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(X)));
    return static_cast<X *>(p);
}

struct NetState
{
    bool init = false;
    float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    float *partials = nullptr, *loss = nullptr;
    int loss_cap = 0;
    int adam_t = 0;
    // training set
    float *feat = nullptr, *target = nullptr;
    size_t n = 0, cap = 0;
    // per-mesh scratch
    int N = 0;
    float *cfeat = nullptr, *cout = nullptr;
    double *dtau = nullptr, *minv = nullptr, *red = nullptr;
};
NetState g_nn;

void ensure_net()
{
    if (g_nn.theta)
        return;
    g_nn.theta = dalloc<float>(nn::kParams);
    g_nn.m = dalloc<float>(nn::kParams);
    g_nn.v = dalloc<float>(nn::kParams);
    g_nn.grad = dalloc<float>(nn::kParams);
    g_nn.partials = dalloc<float>((size_t)nn::kTrainBlocks * (nn::kParams + 1));
}

void ensure_mesh(int N)
{
    if (g_nn.N == N)
        return;
    for (void *p : {(void *)g_nn.cfeat, (void *)g_nn.cout, (void *)g_nn.dtau,
                    (void *)g_nn.minv, (void *)g_nn.red})
        if (p)
            (void)hipFree(p);
    g_nn.N = N;
    g_nn.cfeat = dalloc<float>((size_t)N * nn::kIn);
    g_nn.cout = dalloc<float>((size_t)N * nn::kOut);
    g_nn.dtau = dalloc<double>(N);
    g_nn.minv = dalloc<double>((size_t)N * 25);
    g_nn.red = dalloc<double>(2048);
}

void grow_data(size_t need)
{
    if (need <= g_nn.cap)
        return;
    size_t cap = std::max<size_t>(need, std::max<size_t>(2 * g_nn.cap, 1 << 16));
    float *f = dalloc<float>(cap * nn::kIn), *t = dalloc<float>(cap * nn::kOut);
    if (g_nn.n)
    {
        FVH_HIP_CHECK(hipMemcpy(f, g_nn.feat, g_nn.n * nn::kIn * 4, hipMemcpyDeviceToDevice));
        FVH_HIP_CHECK(hipMemcpy(t, g_nn.target, g_nn.n * nn::kOut * 4,
                                hipMemcpyDeviceToDevice));
    }
    if (g_nn.feat)
        (void)hipFree(g_nn.feat);
    if (g_nn.target)
        (void)hipFree(g_nn.target);
    g_nn.feat = f;
    g_nn.target = t;
    g_nn.cap = cap;
}

} // namespace

nn::Refs learned_refs(const RuntimeView &v)
{
    nn::Refs r;
    r.rho = v.rho_ref;
    r.p = v.p_ref;
    r.a = std::sqrt(v.P.kappa * v.p_ref / v.rho_ref);
    return r;
}

bool learned_ready() { return g_nn.init; }

// Learned inverse blocks of the current state at the pseudo steps dtau.
void learned_blocks(const RuntimeView &v, double cfl, const double *dtau, double *Minv)
{
    if (!g_nn.init)
        throw std::runtime_error("learned preconditioner: network not initialised");
    ensure_mesh(v.m.N);
    nn::features(v.stream, v.m, v.P, v.U, cfl, learned_refs(v), g_nn.cfeat);
    nn::forward(v.stream, v.m.N, g_nn.theta, g_nn.cfeat, g_nn.cout);
    nn::apply_blocks(v.stream, v.m.N, g_nn.cout, dtau, cfl, learned_refs(v), Minv);
}

// max and mean per-cell relative Frobenius error of A against B
void learned_block_error(const RuntimeView &v, const double *A, const double *B, double *err)
{
    ensure_mesh(v.m.N);
    nn::block_error(v.stream, v.m.N, A, B, g_nn.red, g_nn.red + 2000);
    FVH_HIP_CHECK(hipMemcpyAsync(err, g_nn.red + 2000, 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, v.stream));
    FVH_HIP_CHECK(hipStreamSynchronize(v.stream));
}

// The network and its training set outlive a mesh; the per-mesh scratch
// does not.
void learned_release_mesh()
{
    for (void *p : {(void *)g_nn.cfeat, (void *)g_nn.cout, (void *)g_nn.dtau,
                    (void *)g_nn.minv, (void *)g_nn.red})
        if (p)
            (void)hipFree(p);
    g_nn.cfeat = g_nn.cout = nullptr;
    g_nn.dtau = g_nn.minv = g_nn.red = nullptr;
    g_nn.N = 0;
}

} // namespace fvh

using namespace fvh;

#define NN_TRY(...)                                                            \
    try                                                                        \
    {                                                                          \
        __VA_ARGS__;                                                           \
    }                                                                          \
    catch (const std::exception &e)                                            \
    {                                                                          \
        set_error(e.what());                                                   \
        return FVH_ERR_STATE;                                                  \
    }                                                                          \
    return FVH_OK;

int fvh_nn_param_count(void) { return nn::kParams; }

int fvh_nn_init(unsigned long long seed)
{
    NN_TRY({
        ensure_net();
        std::mt19937_64 rng(seed);
        std::vector<float> th(nn::kParams, 0.f);
        auto fill = [&](int off, int fan_in, int rows, int cols, int used_rows, int used_cols) {
            std::normal_distribution<float> d(0.f, std::sqrt(1.f / (float)fan_in));
            for (int i = 0; i < rows; ++i)
                for (int j = 0; j < cols; ++j)
                    th[off + i * cols + j] = (i < used_rows && j < used_cols) ? d(rng) : 0.f;
        };
        fill(nn::oW1, nn::kFeat, nn::kIn, nn::kHid, nn::kFeat, nn::kHid);
        fill(nn::oW2, nn::kHid, nn::kHid, nn::kHid, nn::kHid, nn::kHid);
        fill(nn::oW3, nn::kHid, nn::kHid, nn::kOut, nn::kHid, nn::kBlk);
        for (int j = 0; j < nn::kBlk; ++j) // start near zero output
            for (int i = 0; i < nn::kHid; ++i)
                th[nn::oW3 + i * nn::kOut + j] *= 0.1f;
        FVH_HIP_CHECK(hipMemcpy(g_nn.theta, th.data(), th.size() * 4, hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemset(g_nn.m, 0, nn::kParams * 4));
        FVH_HIP_CHECK(hipMemset(g_nn.v, 0, nn::kParams * 4));
        g_nn.adam_t = 0;
        g_nn.init = true;
    });
}

int fvh_nn_get_params(float *theta)
{
    NN_TRY({
        if (!g_nn.init)
            throw std::runtime_error("network not initialised");
        FVH_HIP_CHECK(hipMemcpy(theta, g_nn.theta, nn::kParams * 4, hipMemcpyDeviceToHost));
    });
}

int fvh_nn_set_params(const float *theta)
{
    NN_TRY({
        ensure_net();
        FVH_HIP_CHECK(hipMemcpy(g_nn.theta, theta, nn::kParams * 4, hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemset(g_nn.m, 0, nn::kParams * 4));
        FVH_HIP_CHECK(hipMemset(g_nn.v, 0, nn::kParams * 4));
        g_nn.adam_t = 0;
        g_nn.init = true;
    });
}

int fvh_nn_features(double cfl, float *feat)
{
    NN_TRY({
        RuntimeView v = runtime_view();
        ensure_mesh(v.m.N);
        nn::features(v.stream, v.m, v.P, v.U, cfl, learned_refs(v), g_nn.cfeat);
        FVH_HIP_CHECK(hipStreamSynchronize(v.stream));
        FVH_HIP_CHECK(hipMemcpy(feat, g_nn.cfeat, (size_t)v.m.N * nn::kIn * 4,
                                hipMemcpyDeviceToHost));
    });
}

int fvh_nn_collect(double cfl, long long *n_total)
{
    NN_TRY({
        RuntimeView v = runtime_view();
        const int N = v.m.N;
        ensure_mesh(N);
        grow_data(g_nn.n + N);
        kry::local_dt(v.stream, v.m, v.P, v.U, cfl, g_nn.dtau);
        kry::precond_blocks(v.stream, v.m, v.P, v.flux, v.U, g_nn.dtau, 1, g_nn.minv);
        nn::features(v.stream, v.m, v.P, v.U, cfl, learned_refs(v),
                     g_nn.feat + g_nn.n * nn::kIn);
        nn::targets(v.stream, N, g_nn.minv, g_nn.dtau, cfl, learned_refs(v),
                    g_nn.target + g_nn.n * nn::kOut);
        FVH_HIP_CHECK(hipStreamSynchronize(v.stream));
        g_nn.n += N;
        if (n_total)
            *n_total = (long long)g_nn.n;
    });
}

int fvh_nn_clear(void)
{
    g_nn.n = 0;
    return FVH_OK;
}

int fvh_nn_train(int epochs, int batch, double lr, unsigned long long seed, double *loss)
{
    NN_TRY({
        if (!g_nn.init)
            throw std::runtime_error("network not initialised");
        if (g_nn.n == 0 || batch < 16 || epochs < 1)
            throw std::runtime_error("fvh_nn_train: no data or bad arguments");
        hipStream_t s = runtime_view().stream;
        const int nb = (int)((g_nn.n + batch - 1) / batch);
        if (g_nn.loss_cap < nb)
        {
            if (g_nn.loss)
                (void)hipFree(g_nn.loss);
            g_nn.loss = dalloc<float>(nb);
            g_nn.loss_cap = nb;
        }
        std::vector<int> order(nb);
        for (int b = 0; b < nb; ++b)
            order[b] = b;
        std::mt19937_64 rng(seed);
        std::vector<float> bl(nb);
        for (int e = 0; e < epochs; ++e)
        {
            std::shuffle(order.begin(), order.end(), rng);
            for (int i = 0; i < nb; ++i)
            {
                const size_t off = (size_t)order[i] * batch;
                const int n = (int)std::min<size_t>(batch, g_nn.n - off);
                nn::loss_grad(s, n, g_nn.theta, g_nn.feat + off * nn::kIn,
                              g_nn.target + off * nn::kOut, g_nn.partials, g_nn.grad,
                              g_nn.loss + i);
                nn::adam(s, g_nn.theta, g_nn.grad, g_nn.m, g_nn.v, ++g_nn.adam_t, (float)lr,
                         0.9f, 0.999f, 1e-8f);
            }
            FVH_HIP_CHECK(hipMemcpyAsync(bl.data(), g_nn.loss, nb * 4, hipMemcpyDeviceToHost,
                                         s));
            FVH_HIP_CHECK(hipStreamSynchronize(s));
            double acc = 0.0;
            for (int i = 0; i < nb; ++i)
                acc += bl[i];
            if (loss)
                loss[e] = acc / nb;
        }
    });
}

int fvh_nn_forward(const float *feat, int n, float *out)
{
    NN_TRY({
        if (!g_nn.init)
            throw std::runtime_error("network not initialised");
        hipStream_t s = runtime_view().stream;
        float *f = dalloc<float>((size_t)n * nn::kIn), *o = dalloc<float>((size_t)n * nn::kOut);
        FVH_HIP_CHECK(hipMemcpy(f, feat, (size_t)n * nn::kIn * 4, hipMemcpyHostToDevice));
        nn::forward(s, n, g_nn.theta, f, o);
        FVH_HIP_CHECK(hipStreamSynchronize(s));
        FVH_HIP_CHECK(hipMemcpy(out, o, (size_t)n * nn::kOut * 4, hipMemcpyDeviceToHost));
        (void)hipFree(f);
        (void)hipFree(o);
    });
}

int fvh_nn_loss_grad(const float *feat, const float *target, int n, float *grad, double *loss)
{
    NN_TRY({
        if (!g_nn.init)
            throw std::runtime_error("network not initialised");
        hipStream_t s = runtime_view().stream;
        float *f = dalloc<float>((size_t)n * nn::kIn), *t = dalloc<float>((size_t)n * nn::kOut);
        float *l = dalloc<float>(1);
        FVH_HIP_CHECK(hipMemcpy(f, feat, (size_t)n * nn::kIn * 4, hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(t, target, (size_t)n * nn::kOut * 4, hipMemcpyHostToDevice));
        nn::loss_grad(s, n, g_nn.theta, f, t, g_nn.partials, g_nn.grad, l);
        FVH_HIP_CHECK(hipStreamSynchronize(s));
        FVH_HIP_CHECK(hipMemcpy(grad, g_nn.grad, nn::kParams * 4, hipMemcpyDeviceToHost));
        float lh = 0.f;
        FVH_HIP_CHECK(hipMemcpy(&lh, l, 4, hipMemcpyDeviceToHost));
        if (loss)
            *loss = lh;
        (void)hipFree(f);
        (void)hipFree(t);
        (void)hipFree(l);
    });
}

int fvh_nn_validate(double cfl, double *err)
{
    NN_TRY({
        RuntimeView v = runtime_view();
        const int N = v.m.N;
        ensure_mesh(N);
        double *exact = dalloc<double>((size_t)N * 25);
        kry::local_dt(v.stream, v.m, v.P, v.U, cfl, g_nn.dtau);
        kry::precond_blocks(v.stream, v.m, v.P, v.flux, v.U, g_nn.dtau, 1, exact);
        learned_blocks(v, cfl, g_nn.dtau, g_nn.minv);
        learned_block_error(v, g_nn.minv, exact, err);
        (void)hipFree(exact);
    });
}
