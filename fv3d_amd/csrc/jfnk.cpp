// Steady Newton-Krylov (JFNK) driver: pseudo-transient Newton steps on the
// device state, each solved by right-preconditioned restarted GMRES with a
// matrix-free finite-difference operator.  Vectors live on the device; the
// small Hessenberg least-squares problem (Givens rotations) runs on the host
// from the dot products copied back once per Arnoldi step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "fvh_internal.h"
#include "krylov.h"
#include "runtime.h"

namespace fvh {

namespace {

// Device work buffers, kept between solves of the same size.
struct Workspace
{
    int N = 0, m = 0, prec = -1;
    size_t ld = 0;
    std::vector<void *> allocs;
    double *R0 = nullptr, *R0j = nullptr, *Rt = nullptr, *Ut = nullptr, *dtau = nullptr,
           *Minv = nullptr, *xacc = nullptr, *dU = nullptr, *partials = nullptr,
           *dots = nullptr, *ydev = nullptr, *Mex = nullptr;
    void *V = nullptr, *z = nullptr, *t = nullptr; // T storage
    float *U32 = nullptr, *Ut32 = nullptr;         // FP32 state copies (FP32 mode)
    double *host = nullptr;                        // pinned, kHost doubles
    int *again = nullptr;                          // device flag of the Arnoldi step
    hipEvent_t ev[2] = {nullptr, nullptr};         // Arnoldi slot copies done
};

// offsets in dots (and in each host slot)
constexpr int kH1 = 0, kH2 = kry::kMaxDots, kWW = 2 * kry::kMaxDots,
              kZZ = kWW + 1, kSums = kWW + 2, kRatio = kWW + 8, kEps = kWW + 9,
              kUU = kWW + 10, kAr = kWW + 12, kWW2 = kWW + 17;
// pinned host memory: two slots of kCopy doubles for the Arnoldi steps in
// flight (slot 0 also serves the synchronous fetches), then the GMRES
// solution y
constexpr int kCopy = kWW2 + 1, kSlot = kCopy + 15, kY = 2 * kSlot;
constexpr int kHost = kY + kry::kMaxDots;

Workspace g_ws;
std::vector<double> g_hist;

// Multicolour block-ILU structures of the current mesh (colouring and
// neighbour tables built on the host from the device connectivity) and the
// per-Newton-step blocks; the buffers belong to g_ws.
struct DiluState
{
    const int *mesh_cf = nullptr; // the mesh they were built for
    int N = 0, maxf = 0, ncol = 0;
    std::vector<int> coff;
    int *ccell = nullptr, *nbr = nullptr, *back = nullptr, *flags = nullptr;
    double *off = nullptr, *dinv = nullptr;     // build layout
    double *off64 = nullptr, *dinv64 = nullptr; // sweep layout, FP64
    kry::bf16_t *off16 = nullptr;               // sweep layout, bfloat16 blocks
    float *dinv32 = nullptr;                    //   with FP32 diagonal inverses
    kry::DiluMesh view() const { return {N, maxf, ncol, ccell, nbr, back, flags}; }
};
DiluState g_dilu;


template <class X> X *walloc(size_t n)
{
    void *p = nullptr;
    FVH_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(X)));
    g_ws.allocs.push_back(p);
    return static_cast<X *>(p);
}

void release()
{
    for (void *p : g_ws.allocs)
        (void)hipFree(p);
    if (g_ws.host)
        (void)hipHostFree(g_ws.host);
    for (hipEvent_t e : g_ws.ev)
        if (e)
            (void)hipEventDestroy(e);
    g_ws = Workspace();
    g_dilu = DiluState();
}

void prepare(int N, int m, int prec)
{
    if (g_ws.N == N && g_ws.m == m && g_ws.prec == prec)
        return;
    release();
    g_ws.N = N;
    g_ws.m = m;
    g_ws.prec = prec;
    const size_t n = (size_t)5 * N;
    g_ws.ld = (n + 63) / 64 * 64;
    const size_t es = prec == FVH_JFNK_FP32 ? sizeof(float) : sizeof(double);
    g_ws.R0 = walloc<double>(n);
    g_ws.R0j = prec == FVH_JFNK_FP32 ? walloc<double>(n) : g_ws.R0;
    if (prec == FVH_JFNK_FP32)
    {
        g_ws.U32 = walloc<float>(n);
        g_ws.Ut32 = walloc<float>(n);
    }
    g_ws.Rt = walloc<double>(n);
    g_ws.Ut = walloc<double>(n);
    g_ws.dtau = walloc<double>(N);
    g_ws.Minv = walloc<double>((size_t)25 * N);
    g_ws.xacc = walloc<double>(n);
    g_ws.dU = walloc<double>(n);
    g_ws.partials = walloc<double>((size_t)kry::kDotBlocks * kry::kMaxDots);
    g_ws.dots = walloc<double>(kHost);
    g_ws.ydev = walloc<double>(kry::kMaxDots);
    g_ws.V = walloc<char>((size_t)(m + 1) * g_ws.ld * es);
    g_ws.z = walloc<char>(g_ws.ld * es);
    g_ws.t = walloc<char>(g_ws.ld * es);
    FVH_HIP_CHECK(hipHostMalloc((void **)&g_ws.host, kHost * sizeof(double),
                                hipHostMallocDefault));
    g_ws.again = walloc<int>(1);
    for (hipEvent_t &e : g_ws.ev)
        FVH_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void fetch(hipStream_t s, int off, int count)
{
    FVH_HIP_CHECK(hipMemcpyAsync(g_ws.host + off, g_ws.dots + off, count * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    FVH_HIP_CHECK(hipStreamSynchronize(s));
}

// Greedy distance-1 colouring of the cell graph in cell order (at most
// maxf + 1 colours), colour-sorted positions, the neighbour / transposed-slot
// tables and the block buffers (FP32 copies when fp32).
void dilu_prepare(const RuntimeView &v, bool fp32)
{
    const dev::DevMesh &m = v.m;
    if (g_dilu.mesh_cf != m.cf || g_dilu.N != m.N)
    {
        const int N = m.N, F = m.F, maxf = m.maxf;
        if (maxf > 6)
            throw std::runtime_error("DILU: cells with more than 6 faces are not supported");
        std::vector<int> cf((size_t)N * maxf), ncf(N), fcell((size_t)2 * F);
        FVH_HIP_CHECK(hipMemcpy(cf.data(), m.cf, cf.size() * sizeof(int),
                                hipMemcpyDeviceToHost));
        FVH_HIP_CHECK(hipMemcpy(ncf.data(), m.ncf, ncf.size() * sizeof(int),
                                hipMemcpyDeviceToHost));
        FVH_HIP_CHECK(hipMemcpy(fcell.data(), m.fcell, fcell.size() * sizeof(int),
                                hipMemcpyDeviceToHost));
        std::vector<int> nb((size_t)N * maxf, -1);
        for (int c = 0; c < N; ++c)
            for (int t = 0; t < ncf[c]; ++t)
            {
                const int f = cf[(size_t)c * maxf + t];
                const int c0 = fcell[2 * (size_t)f], c1 = fcell[2 * (size_t)f + 1];
                const int o = c0 == c ? c1 : c0;
                if (o >= N)
                    continue; // boundary face
                if (o == c)
                    throw std::runtime_error("DILU: a cell is its own face neighbour");
                for (int u = 0; u < t; ++u)
                    if (nb[(size_t)c * maxf + u] == o)
                        throw std::runtime_error("DILU: two cells share more than one face");
                nb[(size_t)c * maxf + t] = o;
            }
        std::vector<int> col(N, -1);
        int ncol = 0;
        for (int c = 0; c < N; ++c)
        {
            unsigned used = 0;
            for (int t = 0; t < maxf; ++t)
            {
                const int o = nb[(size_t)c * maxf + t];
                if (o >= 0 && col[o] >= 0)
                    used |= 1u << col[o];
            }
            int q = 0;
            while ((used >> q) & 1u)
                ++q;
            col[c] = q;
            ncol = std::max(ncol, q + 1);
        }
        // positions: by colour, and inside a colour by the number of lower
        // neighbours (the order inside a colour does not change the factors),
        // so that the 64 positions of a tile mostly use the same sweep slots
        std::vector<int> nlow(N, 0);
        for (int c = 0; c < N; ++c)
            for (int t = 0; t < maxf; ++t)
            {
                const int o = nb[(size_t)c * maxf + t];
                if (o >= 0 && col[o] < col[c])
                    ++nlow[c];
            }
        std::vector<int> coff(ncol + 1, 0), ccell(N), pos(N);
        for (int c = 0; c < N; ++c)
            ++coff[col[c] + 1];
        for (int q = 0; q < ncol; ++q)
            coff[q + 1] += coff[q];
        {
            const int nk = maxf + 1;
            std::vector<int> boff((size_t)ncol * nk + 1, 0);
            for (int c = 0; c < N; ++c)
                ++boff[(size_t)col[c] * nk + nlow[c] + 1];
            for (size_t q = 0; q + 1 < boff.size(); ++q)
                boff[q + 1] += boff[q];
            for (int c = 0; c < N; ++c)
            {
                pos[c] = boff[(size_t)col[c] * nk + nlow[c]]++;
                ccell[pos[c]] = c;
            }
        }
        // tiled tables: whole tiles of 64 positions
        const size_t NT = ((size_t)N + 63) / 64 * 64;
        auto tiled = [](size_t p, size_t e, size_t E) { return ((p >> 6) * E + e) * 64 + (p & 63); };
        std::vector<int> nbr(NT * maxf, -1), back((size_t)N * maxf, -1), flags(N, 0);
        for (int p = 0; p < N; ++p)
        {
            const int c = ccell[p];
            for (int t = 0; t < maxf; ++t)
            {
                const int o = nb[(size_t)c * maxf + t];
                if (o < 0)
                    continue;
                const int f = cf[(size_t)c * maxf + t];
                int to = -1;
                for (int u = 0; u < ncf[o]; ++u)
                    if (cf[(size_t)o * maxf + u] == f)
                        to = u;
                if (to < 0)
                    throw std::runtime_error("DILU: inconsistent face lists");
                back[(size_t)p * maxf + t] = pos[o] * maxf + to;
                flags[p] |= col[o] < col[c] ? (1 << t) : (1 << (8 + t));
            }
            // sweep slots: the lower neighbours in face order, then the
            // higher ones
            int t2 = 0;
            for (int pass = 0; pass < 2; ++pass)
                for (int t = 0; t < maxf; ++t)
                    if ((flags[p] >> (8 * pass + t)) & 1)
                        nbr[tiled(p, t2++, maxf)] = nb[(size_t)c * maxf + t];
        }
        g_dilu = DiluState();
        g_dilu.mesh_cf = m.cf;
        g_dilu.N = N;
        g_dilu.maxf = maxf;
        g_dilu.ncol = ncol;
        g_dilu.coff = coff;
        g_dilu.ccell = walloc<int>(N);
        g_dilu.nbr = walloc<int>(NT * maxf);
        g_dilu.back = walloc<int>((size_t)N * maxf);
        g_dilu.flags = walloc<int>(N);
        FVH_HIP_CHECK(hipMemcpy(g_dilu.ccell, ccell.data(), (size_t)N * sizeof(int),
                                hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_dilu.nbr, nbr.data(), nbr.size() * sizeof(int),
                                hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_dilu.back, back.data(), back.size() * sizeof(int),
                                hipMemcpyHostToDevice));
        FVH_HIP_CHECK(hipMemcpy(g_dilu.flags, flags.data(), (size_t)N * sizeof(int),
                                hipMemcpyHostToDevice));
        g_dilu.off = walloc<double>((size_t)N * maxf * 25);
        g_dilu.dinv = walloc<double>((size_t)N * 25);
        // slots of boundary faces and unused slots stay zero (the sweeps
        // read every slot)
        FVH_HIP_CHECK(hipMemset(g_dilu.off, 0, (size_t)N * maxf * 25 * sizeof(double)));
    }
    // whole tiles, and a last tile of zero blocks for the slots the sweeps
    // skip
    const size_t NT = ((size_t)g_dilu.N + 63) / 64 * 64;
    const size_t E = (size_t)g_dilu.maxf * 25;
    if (fp32 && !g_dilu.off16)
    {
        g_dilu.off16 = walloc<kry::bf16_t>((NT + 64) * E);
        FVH_HIP_CHECK(hipMemset(g_dilu.off16 + NT * E, 0, 64 * E * sizeof(kry::bf16_t)));
        g_dilu.dinv32 = walloc<float>(NT * 25);
    }
    if (!fp32 && !g_dilu.off64)
    {
        g_dilu.off64 = walloc<double>((NT + 64) * E);
        FVH_HIP_CHECK(hipMemset(g_dilu.off64 + NT * E, 0, 64 * E * sizeof(double)));
        g_dilu.dinv64 = walloc<double>(NT * 25);
    }
}

// The coupled preconditioner at U from the inverse diagonal blocks in Minv;
// the sweep copies of the factors with bfloat16 off-diagonal blocks and FP32
// diagonal inverses (fp32), or all in FP64.
void dilu_build(const RuntimeView &v, bool fp32)
{
    dilu_prepare(v, fp32);
    const kry::DiluMesh d = g_dilu.view();
    kry::dilu_offdiag(v.stream, v.m, v.P, v.flux, d, v.U, g_dilu.off);
    for (int q = 0; q < g_dilu.ncol; ++q)
    {
        if (fp32)
            kry::dilu_factor<float, kry::bf16_t>(v.stream, d, g_dilu.coff[q],
                                                 g_dilu.coff[q + 1], g_ws.Minv, g_dilu.off,
                                                 g_dilu.dinv, g_dilu.off16, g_dilu.dinv32);
        else
            kry::dilu_factor<double, double>(v.stream, d, g_dilu.coff[q], g_dilu.coff[q + 1],
                                             g_ws.Minv, g_dilu.off, g_dilu.dinv, g_dilu.off64,
                                             g_dilu.dinv64);
    }
}

// stop norm max_j (1/V_glob) sum_i |R_ij| V_i and |R|_2 of g_ws.R0
void residual_norms(const RuntimeView &v, double &norm, double &l2)
{
    const int n = 5 * v.m.N;
    kry::residual_sums(v.stream, v.m.N, g_ws.R0, v.m.vol, g_ws.partials, g_ws.dots + kSums);
    kry::multidot<double>(v.stream, n, 1, g_ws.R0, g_ws.ld, g_ws.R0, g_ws.partials,
                          g_ws.dots + kWW);
    FVH_HIP_CHECK(hipMemcpyAsync(g_ws.host + kWW, g_ws.dots + kWW, 7 * sizeof(double),
                                 hipMemcpyDeviceToHost, v.stream));
    FVH_HIP_CHECK(hipStreamSynchronize(v.stream));
    norm = 0.0;
    for (int j = 0; j < 5; ++j)
        norm = std::max(norm, g_ws.host[kSums + j] / v.vglob);
    l2 = std::sqrt(g_ws.host[kWW]);
    if (!std::isfinite(norm) || !std::isfinite(l2))
        norm = l2 = NAN;
}

struct GmresResult
{
    int its = 0, evals = 0, reorth = 0;
    double rel = 1.0;
};

// Solve (I/dtau - J) M^-1 y = R0 approximately; xacc = y (dU = M^-1 xacc).
template <class T>
GmresResult gmres(const RuntimeView &v, const fvh_jfnk_options_t &o, double fd_num)
{
    hipStream_t s = v.stream;
    const int N = v.m.N, n = 5 * N, m = o.krylov_dims;
    const size_t ld = g_ws.ld;
    T *V = static_cast<T *>(g_ws.V), *z = static_cast<T *>(g_ws.z),
      *t = static_cast<T *>(g_ws.t);
    double *dots = g_ws.dots, *host = g_ws.host;
    GmresResult res;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    FVH_HIP_CHECK(hipMemsetAsync(g_ws.xacc, 0, (size_t)n * sizeof(double), s));
    double beta0 = 0.0;

    // w = A z into out, with |z|^2 already in dots[kZZ]; in FP32 mode the
    // difference quotient takes two FP32-arithmetic residuals of FP32 states
    // (R0j is the FP32 residual of U32, the FP32 rounding of U; the
    // perturbed state is rounded once, after the FP64 sum)
    const bool fp32 = o.precision == FVH_JFNK_FP32;
    // z = M^-1 in with |z|^2 into dots[kZZ] (the coupled preconditioner's
    // factors are stored in reduced precision in both precisions)
    const bool coupled = o.coupling == FVH_COUPLING_DILU;
    auto apply_M = [&](const T *in, T *out) {
        if (coupled)
        {
            kry::dilu_apply<T, float, kry::bf16_t>(s, g_dilu.view(), g_dilu.coff.data(),
                                                   g_dilu.off16, g_dilu.dinv32, in, out);
            kry::multidot<T>(s, n, 1, out, ld, out, g_ws.partials, dots + kZZ);
        }
        else
            kry::precond_apply_norm<T>(s, N, g_ws.Minv, in, out, g_ws.partials, dots + kZZ);
    };
    auto apply_op = [&](const T *zin, T *out) {
        if (fp32)
        {
            kry::perturb<T, float>(s, n, v.U, dots + kZZ, fd_num, zin, g_ws.Ut32, dots + kEps);
            eval_residual(g_ws.Ut, g_ws.Rt, true, g_ws.Ut32);
        }
        else
        {
            kry::perturb<T, double>(s, n, v.U, dots + kZZ, fd_num, zin, g_ws.Ut, dots + kEps);
            eval_residual(g_ws.Ut, g_ws.Rt, false);
        }
        kry::op_finish<T>(s, N, zin, g_ws.dtau, g_ws.Rt, g_ws.R0j, dots + kEps, out);
        ++res.evals;
    };

    for (int cycle = 0; cycle <= o.max_restarts; ++cycle)
    {
        // r = b - A M^-1 x into V_0
        if (cycle == 0)
            kry::convert<T>(s, n, g_ws.R0, V);
        else
        {
            kry::convert<T>(s, n, g_ws.xacc, t);
            apply_M(t, z);
            apply_op(z, V);
            kry::rhs_minus<T>(s, n, g_ws.R0, V);
        }
        kry::multidot<T>(s, n, 1, V, ld, V, g_ws.partials, dots + kWW);
        fetch(s, kWW, 1);
        const double beta = std::sqrt(host[kWW]);
        if (cycle == 0)
            beta0 = beta;
        if (!(beta > 0.0) || !std::isfinite(beta))
        {
            res.rel = beta0 > 0.0 ? beta / beta0 : 0.0;
            break;
        }
        kry::scale<T>(s, n, V, 1.0 / beta, V);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        bool done = false;
        const double tiny = 1e-300 * beta0;
        // One Arnoldi step, enqueued whole: w = A M^-1 V_j; classical
        // Gram-Schmidt in one pass with the norm by Pythagoras (w is V_{j+1},
        // so the same multi-dot gives <w, w>); a second pass, run only when
        // the device finds that the new vector kept at most the fraction
        // o.reorth of its square norm (0.5: the 1/sqrt(2) criterion of Daniel,
        // Gragg, Kaufman and Stewart); w scaled by 1/hn on the device; the dots and
        // hn copied to host slot j % 2.  Step j + 1 is enqueued before the
        // host reads step j, so the device never waits for the host; when
        // step j ends the cycle, step j + 1 is wasted work that nothing reads.
        auto launch_step = [&](int j) {
            T *w = V + (size_t)(j + 1) * ld;
            apply_M(V + (size_t)j * ld, z);
            apply_op(z, w);
            kry::multidot<T>(s, n, j + 2, V, ld, w, g_ws.partials, dots + kH1);
            kry::multiaxpy<T>(s, n, j + 1, V, ld, dots + kH1, w);
            kry::arnoldi_norm(s, j + 1, dots + kH1, dots + kH1 + j + 1, dots + kAr, g_ws.again,
                              o.reorth);
            if (o.reorth > 0.0)
            {
                kry::multidot<T>(s, n, j + 2, V, ld, w, g_ws.partials, dots + kH2, g_ws.again);
                kry::multiaxpy<T>(s, n, j + 1, V, ld, dots + kH2, w, g_ws.again);
                kry::multidot<T>(s, n, 1, w, ld, w, g_ws.partials, dots + kWW2, g_ws.again);
            }
            kry::arnoldi_finish<T>(s, n, w, g_ws.again, dots + kWW2, dots + kAr, tiny);
            FVH_HIP_CHECK(hipMemcpyAsync(host + (j & 1) * kSlot, dots, kCopy * sizeof(double),
                                         hipMemcpyDeviceToHost, s));
            FVH_HIP_CHECK(hipEventRecord(g_ws.ev[j & 1], s));
        };
        launch_step(0);
        for (int j = 0; j < m; ++j)
        {
            if (j + 1 < m)
                launch_step(j + 1);
            FVH_HIP_CHECK(hipEventSynchronize(g_ws.ev[j & 1]));
            const double *hb = host + (j & 1) * kSlot;
            ++res.its;
            const bool again = hb[kAr + 1] != 0.0;
            if (again)
                ++res.reorth;
            double *Hj = &H[(size_t)j * (m + 1)]; // column j
            for (int i = 0; i <= j; ++i)
                Hj[i] = hb[kH1 + i] + (again ? hb[kH2 + i] : 0.0);
            const double hn = hb[kAr + 2];
            Hj[j + 1] = hn;
            for (int i = 0; i < j; ++i)
            {
                const double a = Hj[i], b = Hj[i + 1];
                Hj[i] = cs[i] * a + sn[i] * b;
                Hj[i + 1] = -sn[i] * a + cs[i] * b;
            }
            const double a = Hj[j], b = Hj[j + 1];
            const double r = std::hypot(a, b);
            cs[j] = r > 0.0 ? a / r : 1.0;
            sn[j] = r > 0.0 ? b / r : 0.0;
            Hj[j] = r;
            Hj[j + 1] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            k = j + 1;
            res.rel = std::fabs(g[j + 1]) / beta0;
            if (!std::isfinite(res.rel))
                throw std::runtime_error("GMRES: non-finite Arnoldi coefficients");
            if (res.rel <= o.rtol || !(hn > tiny))
            {
                done = true;
                break;
            }
        }
        // y = H^-1 g (upper triangular, k x k)
        for (int i = k - 1; i >= 0; --i)
        {
            double a = g[i];
            for (int q = i + 1; q < k; ++q)
                a -= H[(size_t)q * (m + 1) + i] * y[q];
            y[i] = a / H[(size_t)i * (m + 1) + i];
        }
        for (int i = 0; i < k; ++i)
            host[kY + i] = y[i];
        FVH_HIP_CHECK(hipMemcpyAsync(g_ws.ydev, host + kY, k * sizeof(double),
                                     hipMemcpyHostToDevice, s));
        kry::combine<T>(s, n, k, V, ld, g_ws.ydev, t);
        kry::accumulate<T>(s, n, t, g_ws.xacc);
        FVH_HIP_CHECK(hipStreamSynchronize(s)); // host y reused next cycle
        if (done)
            break;
    }
    return res;
}

int solve(const fvh_jfnk_options_t &o, fvh_jfnk_stats_t &st)
{
    if (o.krylov_dims < 1 || o.krylov_dims + 1 > kry::kMaxDots)
        throw std::runtime_error("krylov_dims must be in [1, 63]");
    if (o.precision != FVH_JFNK_FP64 && o.precision != FVH_JFNK_FP32)
        throw std::runtime_error("unknown precision");
    if (o.precond < FVH_PRECOND_SCALAR || o.precond > FVH_PRECOND_LEARNED)
        throw std::runtime_error("unknown preconditioner");
    if (o.precond == FVH_PRECOND_LEARNED && !learned_ready())
        throw std::runtime_error("learned preconditioner: no network (fvh_nn_init / "
                                 "fvh_nn_set_params)");
    if (o.coupling != FVH_COUPLING_NONE && o.coupling != FVH_COUPLING_DILU)
        throw std::runtime_error("unknown coupling");
    if (o.coupling == FVH_COUPLING_DILU && o.precond == FVH_PRECOND_SCALAR)
        throw std::runtime_error("DILU coupling needs block or learned diagonal blocks");
    RuntimeView v = runtime_view();
    const int N = v.m.N, n = 5 * N;
    prepare(N, o.krylov_dims, o.precision);
    hipStream_t s = v.stream;
    FVH_HIP_CHECK(hipStreamSynchronize(s));
    g_hist.clear();
    st = fvh_jfnk_stats_t{};

    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    eval_residual(v.U, g_ws.R0);
    st.residual_evals = 1;
    double norm, l2;
    residual_norms(v, norm, l2);
    st.norm0 = norm;
    const double l20 = l2;
    double cfl = o.cfl0;
    bool learned = o.precond == FVH_PRECOND_LEARNED;
    if (o.verbose)
        std::printf(" newton  stop-norm      cfl        gmres  rel        alpha\n");
    int it = 0;
    for (; it < o.max_newton; ++it)
    {
        if (!std::isfinite(norm))
            throw std::runtime_error("non-finite residual in the Newton loop");
        if (norm <= o.tol)
            break;
        const auto tp = std::chrono::steady_clock::now();
        kry::local_dt(s, v.m, v.P, v.U, cfl, g_ws.dtau);
        if (learned)
        {
            learned_blocks(v, cfl, g_ws.dtau, g_ws.Minv);
            if (it == 0 || (o.learned_check > 0 && it % o.learned_check == 0))
            {
                if (!g_ws.Mex)
                    g_ws.Mex = walloc<double>((size_t)25 * N);
                kry::precond_blocks(s, v.m, v.P, v.flux, v.U, g_ws.dtau, 1, g_ws.Mex);
                double err[2];
                learned_block_error(v, g_ws.Minv, g_ws.Mex, err);
                st.learned_error = err[1];
                if (!(err[1] <= o.learned_tol))
                {
                    // fall back to block Jacobi for the rest of the solve
                    learned = false;
                    ++st.precond_fallbacks;
                    FVH_HIP_CHECK(hipMemcpyAsync(g_ws.Minv, g_ws.Mex,
                                                 (size_t)25 * N * sizeof(double),
                                                 hipMemcpyDeviceToDevice, s));
                }
            }
        }
        else
            kry::precond_blocks(s, v.m, v.P, v.flux, v.U, g_ws.dtau,
                                o.precond == FVH_PRECOND_SCALAR ? 0 : 1, g_ws.Minv);
        if (o.coupling == FVH_COUPLING_DILU)
            dilu_build(v, true);
        FVH_HIP_CHECK(hipStreamSynchronize(s));
        st.precond_seconds +=
            std::chrono::duration<double>(std::chrono::steady_clock::now() - tp).count();
        kry::multidot<double>(s, n, 1, v.U, g_ws.ld, v.U, g_ws.partials, g_ws.dots + kUU);
        fetch(s, kUU, 1);
        const bool fp32 = o.precision == FVH_JFNK_FP32;
        const double fd_num = (fp32 ? o.fd_eps32 : o.fd_eps) * (1.0 + std::sqrt(g_ws.host[kUU]));
        if (fp32)
        {
            kry::convert<float>(s, n, v.U, g_ws.U32);
            eval_residual(v.U, g_ws.R0j, true, g_ws.U32);
            ++st.residual_evals;
        }
        GmresResult gr = o.precision == FVH_JFNK_FP32 ? gmres<float>(v, o, fd_num)
                                                      : gmres<double>(v, o, fd_num);
        st.gmres_iterations += gr.its;
        st.residual_evals += gr.evals;
        if (o.coupling == FVH_COUPLING_DILU)
            kry::dilu_apply<double, float, kry::bf16_t>(s, g_dilu.view(), g_dilu.coff.data(),
                                                        g_dilu.off16, g_dilu.dinv32, g_ws.xacc,
                                                        g_ws.dU);
        else
            kry::precond_apply<double>(s, N, g_ws.Minv, g_ws.xacc, g_ws.dU);
        kry::update_ratio(s, N, v.U, g_ws.dU, g_ws.partials, g_ws.dots + kRatio);
        fetch(s, kRatio, 1);
        const double ratio = g_ws.host[kRatio];
        const double alpha = ratio > o.max_change ? o.max_change / ratio : 1.0;
        if (!std::isfinite(alpha))
            throw std::runtime_error("non-finite Newton update");
        kry::update(s, n, v.U, g_ws.dU, alpha);
        const double prev = norm;
        eval_residual(v.U, g_ws.R0);
        ++st.residual_evals;
        residual_norms(v, norm, l2);
        g_hist.insert(g_hist.end(), {prev, cfl, (double)gr.its, gr.rel, alpha, elapsed()});
        if (o.verbose)
            std::printf("%7d  %.6e  %.3e  %5d  %.3e  %.3f\n", it + 1, norm, cfl, gr.its,
                        gr.rel, alpha);
        if (!(norm < 3.0 * prev))
            cfl = std::max(cfl * o.cfl_cut, 1.0);
        else if (gr.rel > o.cfl_stall)
            cfl = std::max(cfl * o.cfl_stall_cut, 1.0);
        else
            cfl = std::min(cfl * o.cfl_growth, o.cfl_max);
    }
    FVH_HIP_CHECK(hipStreamSynchronize(s));
    st.seconds = elapsed();
    st.newton_steps = it;
    st.norm = norm;
    st.l2_drop = l20 > 0.0 ? l2 / l20 : 0.0;
    st.status = norm <= o.tol ? 0 : 1;
    state_changed();
    if (o.verbose)
        std::fflush(stdout);
    return FVH_OK;
}

} // namespace

void jfnk_release() { release(); }

} // namespace fvh

using namespace fvh;

int fvh_jfnk_default_options(fvh_jfnk_options_t *o)
{
    if (!o)
        return FVH_ERR_ARG;
    o->max_newton = 300;
    o->tol = 1e-8;
    o->cfl0 = 10.0;
    o->cfl_growth = 1.5;
    o->cfl_cut = 0.3;
    o->cfl_stall = 0.5;
    o->cfl_stall_cut = 0.7;
    o->cfl_max = 1e6;
    o->krylov_dims = 30;
    o->max_restarts = 3;
    o->rtol = 1e-2;
    o->fd_eps = 1e-7;
    o->fd_eps32 = 3e-4;
    o->max_change = 0.2;
    o->precision = FVH_JFNK_FP64;
    o->precond = FVH_PRECOND_BLOCK;
    o->verbose = 0;
    o->learned_check = 10;
    o->learned_tol = 0.2;
    o->coupling = FVH_COUPLING_NONE;
    o->reorth = 0.1;
    return FVH_OK;
}

int fvh_jfnk_solve(const fvh_jfnk_options_t *opt, fvh_jfnk_stats_t *stats)
{
    if (!opt || !stats)
    {
        set_error("fvh_jfnk_solve: null argument");
        return FVH_ERR_ARG;
    }
    try
    {
        return solve(*opt, *stats);
    }
    catch (const std::exception &e)
    {
        set_error(e.what());
        return FVH_ERR_STATE;
    }
}

int fvh_jfnk_blocks(double cfl, int precond, double *dtau, double *minv)
{
    try
    {
        if (precond != FVH_PRECOND_SCALAR && precond != FVH_PRECOND_BLOCK)
            throw std::runtime_error("fvh_jfnk_blocks: unknown preconditioner");
        RuntimeView v = runtime_view();
        const int N = v.m.N;
        prepare(N, g_ws.m > 0 ? g_ws.m : 30, g_ws.prec >= 0 ? g_ws.prec : FVH_JFNK_FP64);
        kry::local_dt(v.stream, v.m, v.P, v.U, cfl, g_ws.dtau);
        kry::precond_blocks(v.stream, v.m, v.P, v.flux, v.U, g_ws.dtau,
                            precond == FVH_PRECOND_BLOCK ? 1 : 0, g_ws.Minv);
        FVH_HIP_CHECK(hipStreamSynchronize(v.stream));
        if (dtau)
            FVH_HIP_CHECK(hipMemcpy(dtau, g_ws.dtau, (size_t)N * sizeof(double),
                                    hipMemcpyDeviceToHost));
        if (minv)
            FVH_HIP_CHECK(hipMemcpy(minv, g_ws.Minv, (size_t)25 * N * sizeof(double),
                                    hipMemcpyDeviceToHost));
    }
    catch (const std::exception &e)
    {
        set_error(e.what());
        return FVH_ERR_STATE;
    }
    return FVH_OK;
}

int fvh_jfnk_precond_apply(double cfl, int precond, int coupling, const double *v, double *z)
{
    if (!v || !z)
    {
        set_error("fvh_jfnk_precond_apply: null argument");
        return FVH_ERR_ARG;
    }
    try
    {
        if (precond < FVH_PRECOND_SCALAR || precond > FVH_PRECOND_LEARNED)
            throw std::runtime_error("fvh_jfnk_precond_apply: unknown preconditioner");
        if (coupling != FVH_COUPLING_NONE && coupling != FVH_COUPLING_DILU)
            throw std::runtime_error("fvh_jfnk_precond_apply: unknown coupling");
        if (coupling == FVH_COUPLING_DILU && precond == FVH_PRECOND_SCALAR)
            throw std::runtime_error("DILU coupling needs block or learned diagonal blocks");
        if (precond == FVH_PRECOND_LEARNED && !learned_ready())
            throw std::runtime_error("learned preconditioner: no network");
        RuntimeView rv = runtime_view();
        const int N = rv.m.N, n = 5 * N;
        prepare(N, g_ws.m > 0 ? g_ws.m : 30, g_ws.prec >= 0 ? g_ws.prec : FVH_JFNK_FP64);
        hipStream_t s = rv.stream;
        kry::local_dt(s, rv.m, rv.P, rv.U, cfl, g_ws.dtau);
        if (precond == FVH_PRECOND_LEARNED)
            learned_blocks(rv, cfl, g_ws.dtau, g_ws.Minv);
        else
            kry::precond_blocks(s, rv.m, rv.P, rv.flux, rv.U, g_ws.dtau,
                                precond == FVH_PRECOND_BLOCK ? 1 : 0, g_ws.Minv);
        FVH_HIP_CHECK(hipMemcpyAsync(g_ws.xacc, v, (size_t)n * sizeof(double),
                                     hipMemcpyHostToDevice, s));
        if (coupling == FVH_COUPLING_DILU)
        {
            dilu_build(rv, false);
            kry::dilu_apply<double, double, double>(s, g_dilu.view(), g_dilu.coff.data(),
                                                    g_dilu.off64, g_dilu.dinv64, g_ws.xacc,
                                                    g_ws.dU);
        }
        else
            kry::precond_apply<double>(s, N, g_ws.Minv, g_ws.xacc, g_ws.dU);
        FVH_HIP_CHECK(hipMemcpyAsync(z, g_ws.dU, (size_t)n * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
        FVH_HIP_CHECK(hipStreamSynchronize(s));
    }
    catch (const std::exception &e)
    {
        set_error(e.what());
        return FVH_ERR_STATE;
    }
    return FVH_OK;
}

int fvh_jfnk_history(double *rows, int max_rows)
{
    const int n = (int)(g_hist.size() / 6);
    if (rows)
        std::copy(g_hist.begin(), g_hist.begin() + (size_t)std::min(n, max_rows) * 6, rows);
    return n;
}
