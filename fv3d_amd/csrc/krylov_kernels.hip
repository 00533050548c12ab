// Kernels of the Newton-Krylov (JFNK) path: pseudo time steps, block-Jacobi
// preconditioner blocks, the finite-difference operator and the
// deterministic vector reductions of GMRES.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <stdexcept>

#include "fv_physics.h"
#include "krylov.h"

namespace fvh {
namespace kry {

using dev::DevMesh;
using dev::PhysDev;
using dev::kBlock;

namespace {

inline int blocks_for(long n) { return (int)((n + kBlock - 1) / kBlock); }

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

__global__ __launch_bounds__(kBlock) void k_local_dt(DevMesh m, PhysDev P,
                                                     const double *__restrict__ U,
                                                     double cfl,
                                                     double *__restrict__ dtau)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.N)
        return;
    double pc[10];
    dev::load_cons(U, c, pc);
    dev::con_to_prim(P, pc);
    const double *dx = m.cdx + (size_t)c * 3;
    const double s_rho = 1.0 / pc[0];
    const double cs = sqrt(P.kappa * pc[8] * s_rho);
    const double ds1 = (dev::rabs(pc[5]) + cs) * dx[0] + (dev::rabs(pc[6]) + cs) * dx[1] +
                       (dev::rabs(pc[7]) + cs) * dx[2];
    dtau[c] = cfl * (m.vol[c] / ds1);
}

// First-order inviscid residual of cell c for the own state u (neighbours
// from U, ghosts from u): r = -(sum_f +-F_f A_f) / V.
template <int FLUX, int MAXF>
__device__ void first_order_residual(const DevMesh &m, const PhysDev &P,
                                     const double *__restrict__ U, int c,
                                     const double *u10, double *r)
{
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int nf = m.ncf[c];
    const dev::ViscGrad none{};
#pragma unroll 1
    for (int s = 0; s < MAXF; ++s)
    {
        if (s >= nf)
            break;
        const int f = m.cf[(size_t)c * m.maxf + s];
        const int c0 = m.fcell[2 * f], c1 = m.fcell[2 * f + 1];
        const bool own0 = c0 == c;
        const int nb = own0 ? c1 : c0;
        dev::Frame fr;
        dev::load_frame(m, f, fr);
        double pn[10];
        if (nb < m.N)
        {
            dev::load_cons(U, nb, pn);
            dev::con_to_prim(P, pn);
        }
        else
        {
            dev::ghost_state(m, P, nb - m.N, fr, u10, pn);
        }
        double L[10], R[10];
#pragma unroll
        for (int j = 0; j < 10; ++j)
        {
            L[j] = own0 ? u10[j] : pn[j];
            R[j] = own0 ? pn[j] : u10[j];
        }
        double fl[5];
        dev::face_flux<FLUX, false>(P, fr, L, R, none, none, fl);
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const double v = fl[j] * fr.area;
            acc[j] = acc[j] + (own0 ? v : -v);
        }
    }
    const double vol = m.vol[c];
#pragma unroll
    for (int j = 0; j < 5; ++j)
        r[j] = -acc[j] / vol;
}

template <int FLUX, int MAXF>
__global__ __launch_bounds__(kBlock) void k_precond_blocks(
    DevMesh m, PhysDev P, const double *__restrict__ U,
    const double *__restrict__ dtau, int blocks, double *__restrict__ Minv)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.N)
        return;
    double *out = Minv + (size_t)c * 25;
    const double idt = 1.0 / dtau[c];
    if (!blocks)
    {
#pragma unroll
        for (int q = 0; q < 25; ++q)
            out[q] = (q % 6 == 0) ? dtau[c] : 0.0;
        return;
    }
    double uc[10];
    dev::load_cons(U, c, uc);
    double base[5];
    {
        double u[10];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            u[j] = uc[j];
        dev::con_to_prim(P, u);
        first_order_residual<FLUX, MAXF>(m, P, U, c, u, base);
    }
    double A[25];
#pragma unroll
    for (int k = 0; k < 5; ++k)
    {
        double u[10];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            u[j] = uc[j];
        const double h = 1e-7 * dev::rmax(fabs(uc[k]), 1.0);
        u[k] = uc[k] + h;
        const double hk = u[k] - uc[k];
        dev::con_to_prim(P, u);
        double r[5];
        first_order_residual<FLUX, MAXF>(m, P, U, c, u, r);
#pragma unroll
        for (int j = 0; j < 5; ++j)
            A[j * 5 + k] = (j == k ? idt : 0.0) - (r[j] - base[j]) / hk;
    }
    dev::invert5(A, out);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_precond_apply(int N,
                                                          const double *__restrict__ Minv,
                                                          const T *__restrict__ v,
                                                          T *__restrict__ z)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N)
        return;
    double x[5];
#pragma unroll
    for (int j = 0; j < 5; ++j)
        x[j] = (double)v[(size_t)c * 5 + j];
    const double *Mi = Minv + (size_t)c * 25;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 5; ++q)
            a += Mi[r * 5 + q] * x[q];
        z[(size_t)c * 5 + r] = (T)a;
    }
}

// z = Minv v and the per-block partial sums of |z|^2 (grid-stride over
// cells on the fixed dot grid; z^2 of the stored T values)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_precond_apply_norm(int N,
                                                               const double *__restrict__ Minv,
                                                               const T *__restrict__ v,
                                                               T *__restrict__ z,
                                                               double *__restrict__ partials)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x)
    {
        double x[5];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            x[j] = (double)v[(size_t)c * 5 + j];
        const double *Mi = Minv + (size_t)c * 25;
#pragma unroll
        for (int r = 0; r < 5; ++r)
        {
            double a = 0.0;
#pragma unroll
            for (int q = 0; q < 5; ++q)
                a += Mi[r * 5 + q] * x[q];
            const T zt = (T)a;
            z[(size_t)c * 5 + r] = zt;
            acc += (double)zt * (double)zt;
        }
    }
    dev::block_sum(sh, acc);
    if (threadIdx.x == 0)
        partials[(size_t)blockIdx.x * kMaxDots] = acc;
}

template <typename T, typename O>
__global__ __launch_bounds__(kBlock) void k_perturb(int n, const double *__restrict__ U,
                                                    const double *__restrict__ zz, double num,
                                                    const T *__restrict__ z,
                                                    O *__restrict__ Ut,
                                                    double *__restrict__ eps_out)
{
    // eps = num / |z| (num / 1e-300 for z = 0)
    const double nz = sqrt(zz[0]);
    const double eps = num / (nz > 1e-300 ? nz : 1e-300);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        eps_out[0] = eps;
    if (i < n)
        Ut[i] = (O)(U[i] + eps * (double)z[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_op_finish(int N, const T *__restrict__ z,
                                                      const double *__restrict__ dtau,
                                                      const double *__restrict__ Rt,
                                                      const double *__restrict__ R0,
                                                      const double *__restrict__ eps,
                                                      T *__restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * N)
        return;
    const int c = i / 5;
    w[i] = (T)((double)z[i] / dtau[c] - (Rt[i] - R0[i]) / eps[0]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_rhs_minus(int n, const double *__restrict__ b,
                                                      T *__restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        w[i] = (T)(b[i] - (double)w[i]);
}

// Arnoldi step bookkeeping on the device (one thread), so that the next
// step can be enqueued before the host has read this one:
//   arnoldi_norm: ar[0] = |w|^2 - sum_i h_i^2 (classical Gram-Schmidt with
//     the norm by Pythagoras; the sum in index order), ar[1] = 1 and
//     *again = 1 when the new vector kept at most the fraction eta of its
//     square norm (a second pass is needed), else 0;
//   arnoldi_finish: after the (skipped or run) second pass, with hn2 = |w|^2
//     of that pass when it ran, else ar[0]: ar[2] = hn = sqrt(max(hn2, 0)),
//     ar[3] the scale of w (1/hn, or 1 when hn <= tiny: breakdown, the host
//     stops), ar[4] = hn2; and w scaled in place.
__global__ void k_arnoldi_norm(int k, const double *__restrict__ h,
                               const double *__restrict__ ww, double *__restrict__ ar,
                               int *__restrict__ again, double eta)
{
    if (threadIdx.x != 0)
        return;
    double hh = 0.0;
    for (int i = 0; i < k; ++i)
        hh += h[i] * h[i];
    const double hn2 = ww[0] - hh;
    const int a = eta > 0.0 && !(hn2 > eta * ww[0]);
    ar[0] = hn2;
    ar[1] = a ? 1.0 : 0.0;
    *again = a;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_arnoldi_finish(int n, T *__restrict__ w,
                                                           const int *__restrict__ again,
                                                           const double *__restrict__ ww2,
                                                           double *__restrict__ ar, double tiny)
{
    const double hn2 = *again ? ww2[0] : ar[0];
    const double hn = sqrt(hn2 > 0.0 ? hn2 : 0.0);
    const double sc = hn > tiny ? 1.0 / hn : 1.0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
    {
        ar[2] = hn;
        ar[3] = sc;
        ar[4] = hn2;
    }
    if (i < n)
        w[i] = (T)((double)w[i] * sc);
}

// 16-byte vector of T for the streaming kernels (4 floats or 2 doubles).
template <typename T> struct Vec16;
template <> struct Vec16<float>
{
    using type = float4;
    static constexpr int n = 4;
    __device__ static void get(const type &v, double *o) { o[0] = v.x, o[1] = v.y, o[2] = v.z, o[3] = v.w; }
    __device__ static type make(const double *o) { return make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]); }
};
template <> struct Vec16<double>
{
    using type = double2;
    static constexpr int n = 2;
    __device__ static void get(const type &v, double *o) { o[0] = v.x, o[1] = v.y; }
    __device__ static type make(const double *o) { return make_double2(o[0], o[1]); }
};

// Deterministic multi-dot: a fixed grid of kDotBlocks blocks, grid-stride
// loop over 16-byte vectors (VEC: every column and w 16-byte aligned; the
// elements beyond the last whole vector follow one per thread), per-thread
// FP64 sums, then a per-block column reduction in thread order; the fold
// adds the block rows in block order.  K (1..16) columns, a compile-time
// count so that the loads of all columns are issued together.
template <typename T, int K, bool VEC>
__global__ __launch_bounds__(kBlock) void k_multidot(int n, const T *__restrict__ V,
                                                     size_t ld, const T *__restrict__ w,
                                                     double *__restrict__ partials,
                                                     const int *__restrict__ on)
{
    if (on && !*on)
        return;
    constexpr int KMAX = 16, k = K;
    __shared__ double sh[kBlock * 17];
    double acc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        acc[j] = 0.0;
    const int stride = gridDim.x * blockDim.x;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    int i0 = 0;
    if (VEC)
    {
        using Vt = typename Vec16<T>::type;
        constexpr int W = Vec16<T>::n;
        const int nv = n / W;
        for (int i = tid; i < nv; i += stride)
        {
            double wv[W], vv[W];
            Vec16<T>::get(reinterpret_cast<const Vt *>(w)[i], wv);
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (j < k)
                {
                    Vec16<T>::get(reinterpret_cast<const Vt *>(V + (size_t)j * ld)[i], vv);
#pragma unroll
                    for (int e = 0; e < W; ++e)
                        acc[j] += vv[e] * wv[e];
                }
        }
        i0 = nv * W;
    }
    for (int i = i0 + tid; i < n; i += stride)
    {
        const double wi = (double)w[i];
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
            if (j < k)
                acc[j] += (double)V[(size_t)j * ld + i] * wi;
    }
    // column sums over the block's threads, 16 columns at a time: 16 threads
    // per column each add 16 rows in order, then one thread adds the 16
    // sub-sums in order
    __shared__ double sub[kBlock];
#pragma unroll
    for (int j0 = 0; j0 < KMAX; j0 += 16)
    {
        if (j0 >= k)
            break;
#pragma unroll
        for (int q = 0; q < 16; ++q)
            sh[threadIdx.x * 17 + q] = acc[j0 + q];
        __syncthreads();
        {
            const int col = threadIdx.x & 15, part = threadIdx.x >> 4;
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t)
                s += sh[(part * 16 + t) * 17 + col];
            sub[part * 16 + col] = s;
        }
        __syncthreads();
        if (threadIdx.x < 16 && j0 + (int)threadIdx.x < k)
        {
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t)
                s += sub[t * 16 + threadIdx.x];
            partials[(size_t)blockIdx.x * kMaxDots + j0 + threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// One block per dot: thread t adds the rows t, t + 256, ... in order, then
// a fixed tree over the threads (deterministic, and short: the rows are
// read in parallel).
__global__ __launch_bounds__(kBlock) void k_fold_dots(int nblk, const double *__restrict__ partials,
                                                      double *__restrict__ dots,
                                                      const int *__restrict__ on)
{
    if (on && !*on)
        return;
    __shared__ double sh[kBlock];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock)
        s += partials[(size_t)b * kMaxDots + j];
    dev::block_sum(sh, s);
    if (threadIdx.x == 0)
        dots[j] = s;
}

// w -= sum_j h[j] V_j, one 16-byte vector per thread (VEC, aligned columns)
// or one element
template <typename T, bool VEC>
__global__ __launch_bounds__(kBlock) void k_multiaxpy(int n, int k, const T *__restrict__ V,
                                                      size_t ld, const double *__restrict__ h,
                                                      T *__restrict__ w, const int *__restrict__ on)
{
    if (on && !*on)
        return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (VEC)
    {
        using Vt = typename Vec16<T>::type;
        constexpr int W = Vec16<T>::n;
        const int nv = n / W;
        if (t < nv)
        {
            double a[W], vv[W];
            Vec16<T>::get(reinterpret_cast<const Vt *>(w)[t], a);
            for (int j = 0; j < k; ++j)
            {
                const double hj = h[j];
                Vec16<T>::get(reinterpret_cast<const Vt *>(V + (size_t)j * ld)[t], vv);
#pragma unroll
                for (int e = 0; e < W; ++e)
                    a[e] -= hj * vv[e];
            }
            reinterpret_cast<Vt *>(w)[t] = Vec16<T>::make(a);
            return;
        }
        const int i = nv * W + (t - nv);
        if (i >= n)
            return;
        double a = (double)w[i];
        for (int j = 0; j < k; ++j)
            a -= h[j] * (double)V[(size_t)j * ld + i];
        w[i] = (T)a;
        return;
    }
    const int i = t;
    if (i >= n)
        return;
    double a = (double)w[i];
    for (int j = 0; j < k; ++j)
        a -= h[j] * (double)V[(size_t)j * ld + i];
    w[i] = (T)a;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(int n, const T *__restrict__ w, double sc,
                                                  T *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = (T)((double)w[i] * sc);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(int n, int k, const T *__restrict__ V,
                                                    size_t ld, const double *__restrict__ y,
                                                    T *__restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double a = 0.0;
    for (int j = 0; j < k; ++j)
        a += y[j] * (double)V[(size_t)j * ld + i];
    x[i] = (T)a;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_accumulate(int n, const T *__restrict__ z,
                                                       double *__restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        x[i] += (double)z[i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_convert(int n, const double *__restrict__ src,
                                                    T *__restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = (T)src[i];
}

__global__ __launch_bounds__(kBlock) void k_residual_sums(int N, const double *__restrict__ R,
                                                          const double *__restrict__ vol,
                                                          double *__restrict__ partials)
{
    __shared__ double sh[kBlock];
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int stride = gridDim.x * blockDim.x;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += stride)
    {
        const double v = vol[c];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            a[j] += dev::rabs(R[(size_t)c * 5 + j]) * v;
    }
#pragma unroll
    for (int j = 0; j < 5; ++j)
        dev::block_sum(sh, a[j]);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < 5; ++j)
            partials[(size_t)blockIdx.x * kMaxDots + j] = a[j];
}

__global__ __launch_bounds__(kBlock) void k_update_ratio(int N, const double *__restrict__ U,
                                                         const double *__restrict__ dU,
                                                         double *__restrict__ partials)
{
    __shared__ double sh[kBlock];
    double a = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += stride)
    {
        const double r0 = fabs(dU[(size_t)c * 5] / U[(size_t)c * 5]);
        const double r4 = fabs(dU[(size_t)c * 5 + 4] / U[(size_t)c * 5 + 4]);
        a = fmax(a, fmax(r0, r4));
    }
    // max through the min reduction of the negated value
    double na = -a;
    dev::block_min(sh, na);
    if (threadIdx.x == 0)
        partials[(size_t)blockIdx.x * kMaxDots] = -na;
}

__global__ __launch_bounds__(kBlock) void k_fold_max(int nblk, const double *__restrict__ partials,
                                                     double *__restrict__ out)
{
    __shared__ double sh[kBlock];
    double a = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock)
        a = fmax(a, partials[(size_t)b * kMaxDots]);
    double na = -a;
    dev::block_min(sh, na);
    if (threadIdx.x == 0)
        out[0] = -na;
}

__global__ __launch_bounds__(kBlock) void k_update(int n, double *__restrict__ U,
                                                   const double *__restrict__ dU,
                                                   double alpha)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        U[i] += alpha * dU[i];
}

// ---- multicolour block ILU(0) -----------------------------------------------

// Tiled component-major index of entry e (of E per position) of position p:
// tiles of 64 positions, within a tile entry-major, so that a wavefront's
// loads of one entry are contiguous and its whole working set is one
// contiguous run of 64*E values.
__device__ __forceinline__ size_t tiled(int p, int e, int E)
{
    return ((size_t)(p >> 6) * E + e) * 64 + (p & 63);
}

// storage conversions of the sweep copies (bfloat16: the upper half of the
// FP32 pattern, rounded to nearest even)
__device__ __forceinline__ double to_d(float x) { return x; }
__device__ __forceinline__ double to_d(double x) { return x; }
__device__ __forceinline__ double to_d(bf16_t x) { return __uint_as_float((unsigned)x.v << 16); }
template <class O> __device__ __forceinline__ O from_d(double x);
template <> __device__ __forceinline__ float from_d<float>(double x) { return (float)x; }
template <> __device__ __forceinline__ double from_d<double>(double x) { return x; }
template <> __device__ __forceinline__ bf16_t from_d<bf16_t>(double x)
{
    unsigned u = __float_as_uint((float)x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return bf16_t{(unsigned short)(u >> 16)};
}


// Off-diagonal blocks A1_{nb,c} = -dR1_nb/dU_c = (s_cf A_f / V_nb) dF_f/dU_c
// (s_cf = +1 when c is the face's first cell), one thread per position; the
// column block of each interior face is written into the neighbour's row
// slot (every slot is written by exactly one thread).
template <int FLUX, int MAXF>
__global__ __launch_bounds__(kBlock) void k_dilu_offdiag(DevMesh m, PhysDev P,
                                                         DiluMesh d,
                                                         const double *__restrict__ U,
                                                         double *__restrict__ off)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.N)
        return;
    const int c = d.ccell[p];
    double uc[10];
    dev::load_cons(U, c, uc);
    const int nf = m.ncf[c];
    const dev::ViscGrad none{};
#pragma unroll 1
    for (int s = 0; s < MAXF; ++s)
    {
        if (s >= nf)
            break;
        const int dst = d.back[p * d.maxf + s];
        if (dst < 0)
            continue; // boundary face
        const int f = m.cf[(size_t)c * m.maxf + s];
        const int c0 = m.fcell[2 * f], c1 = m.fcell[2 * f + 1];
        const bool own0 = c0 == c;
        const int nb = own0 ? c1 : c0;
        dev::Frame fr;
        dev::load_frame(m, f, fr);
        double pn[10];
        dev::load_cons(U, nb, pn);
        dev::con_to_prim(P, pn);
        const double w = (own0 ? -fr.area : fr.area) / m.vol[nb];
        double f0[5];
        {
            double u[10];
#pragma unroll
            for (int j = 0; j < 5; ++j)
                u[j] = uc[j];
            dev::con_to_prim(P, u);
            double L[10], R[10];
#pragma unroll
            for (int j = 0; j < 10; ++j)
            {
                L[j] = own0 ? u[j] : pn[j];
                R[j] = own0 ? pn[j] : u[j];
            }
            dev::face_flux<FLUX, false>(P, fr, L, R, none, none, f0);
        }
        double *o = off + (size_t)dst * 25;
#pragma unroll 1
        for (int k = 0; k < 5; ++k)
        {
            double u[10];
#pragma unroll
            for (int j = 0; j < 5; ++j)
                u[j] = uc[j];
            const double h = 1e-7 * dev::rmax(fabs(uc[k]), 1.0);
            u[k] = uc[k] + h;
            const double hk = u[k] - uc[k];
            dev::con_to_prim(P, u);
            double L[10], R[10];
#pragma unroll
            for (int j = 0; j < 10; ++j)
            {
                L[j] = own0 ? u[j] : pn[j];
                R[j] = own0 ? pn[j] : u[j];
            }
            double fk[5];
            dev::face_flux<FLUX, false>(P, fr, L, R, none, none, fk);
#pragma unroll
            for (int j = 0; j < 5; ++j)
                o[j * 5 + k] = w * ((fk[j] - f0[j]) / hk);
        }
    }
}

__device__ __forceinline__ void mm5(const double *A, const double *B, double *C)
{
#pragma unroll
    for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int k = 0; k < 5; ++k)
        {
            double a = 0.0;
#pragma unroll
            for (int q = 0; q < 5; ++q)
                a += A[r * 5 + q] * B[q * 5 + k];
            C[r * 5 + k] = a;
        }
}

template <typename B, typename O>
__global__ __launch_bounds__(kBlock) void k_dilu_factor(DiluMesh d, int p0, int p1,
                                                        const double *__restrict__ Minv,
                                                        const double *__restrict__ off,
                                                        double *__restrict__ dinv,
                                                        O *__restrict__ offB,
                                                        B *__restrict__ dinvB)
{
    const int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= p1)
        return;
    const int i = d.ccell[p];
    double D[25], X[25], Y[25];
#pragma unroll
    for (int q = 0; q < 25; ++q)
        X[q] = Minv[(size_t)i * 25 + q];
    dev::invert5(X, D);
    const int lo = d.flags[p] & 0xff;
#pragma unroll 1
    for (int t = 0; t < d.maxf; ++t)
    {
        if (!((lo >> t) & 1))
            continue;
        const int kt = d.back[p * d.maxf + t];
        const int pk = kt / d.maxf;
        const double *Aik = off + ((size_t)p * d.maxf + t) * 25;
        const double *Aki = off + (size_t)kt * 25;
        const double *Dk = dinv + (size_t)pk * 25;
        double Ak[25];
#pragma unroll
        for (int q = 0; q < 25; ++q)
            Ak[q] = Aik[q];
        mm5(Ak, Dk, X); // A_ik D_k^-1
#pragma unroll
        for (int q = 0; q < 25; ++q)
            Ak[q] = Aki[q];
        mm5(X, Ak, Y); // A_ik D_k^-1 A_ki
#pragma unroll
        for (int q = 0; q < 25; ++q)
            D[q] -= Y[q];
    }
    dev::invert5(D, X);
#pragma unroll
    for (int q = 0; q < 25; ++q)
    {
        dinv[(size_t)p * 25 + q] = X[q];
        dinvB[tiled(p, q, 25)] = (B)X[q];
    }
    // sweep slots: the lower blocks in face order, then the higher ones,
    // then zeros
    const int E = d.maxf * 25;
    const int hi = (d.flags[p] >> 8) & 0xff;
    int t2 = 0;
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass)
#pragma unroll 1
        for (int t = 0; t < d.maxf; ++t)
            if (((pass ? hi : lo) >> t) & 1)
            {
                for (int q = 0; q < 25; ++q)
                    offB[tiled(p, t2 * 25 + q, E)] =
                        from_d<O>(off[((size_t)p * d.maxf + t) * 25 + q]);
                ++t2;
            }
#pragma unroll 1
    for (; t2 < d.maxf; ++t2)
        for (int q = 0; q < 25; ++q)
            offB[tiled(p, t2 * 25 + q, E)] = from_d<O>(0.0);
}

// One colour of the sweeps (blocks and neighbour table in the tiled layout,
// each position's lower blocks first, then its higher ones).  Every slot is
// loaded unconditionally (straight-line code, so that the loads of all slots
// are in flight together); a slot outside the sweep's set reads the zero
// blocks of the tables' last tile and the values of cell 0, replaced by
// zero, so it contributes exactly zero and, when no position of a wavefront
// uses it, its blocks cost no HBM traffic: positions are ordered by their
// number of lower neighbours inside a colour.  The neighbour table is read
// for every slot, so that its load does not wait for the flags.
//   FORWARD: z_i = D_i^-1 (v_i - sum_{k lower} A_ik z_k)
//   backward: z_i -= D_i^-1 sum_{k higher} A_ik z_k
// NOFF: the forward sweep of the lowest colour, whose cells have no lower
// neighbours (z_i = D_i^-1 v_i, no block or neighbour loads).
template <typename T, typename B, typename O, int MAXF, bool FORWARD, bool NOFF = false>
__global__ __launch_bounds__(kBlock) void k_dilu_sweep(DiluMesh d, int p0, int p1,
                                                       const O *__restrict__ off,
                                                       const B *__restrict__ dinv,
                                                       const T *__restrict__ v,
                                                       T *__restrict__ z)
{
    const int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= p1)
        return;
    const int fl = d.flags[p];
    const int nlo = __popc(fl & 0xff), nhi = __popc((fl >> 8) & 0xff);
    if (!FORWARD && !nhi)
        return;
    const int t0 = FORWARD ? 0 : nlo, t1 = FORWARD ? nlo : nlo + nhi;
    const int i = d.ccell[p];
    const int E = d.maxf * 25;
    const int ptail = (d.N + 63) & ~63; // first position of the last tile
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < (NOFF ? 0 : MAXF); ++t)
    {
        const bool on = t >= t0 && t < t1; // t1 <= d.maxf
        const int kn = d.nbr[tiled(p, t < d.maxf ? t : 0, d.maxf)];
        const int k = on && kn >= 0 ? kn : 0;
        double y[5];
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const double zk = (double)z[(size_t)k * 5 + j];
            y[j] = on ? zk : 0.0;
        }
        const O *blk = off + tiled(on ? p : ptail, (t < d.maxf ? t : 0) * 25, E);
#pragma unroll
        for (int a = 0; a < 5; ++a)
        {
            double acc = 0.0;
#pragma unroll
            for (int b = 0; b < 5; ++b)
                acc += to_d(blk[(a * 5 + b) * 64]) * y[b];
            r[a] += acc;
        }
    }
    if (FORWARD)
    {
#pragma unroll
        for (int j = 0; j < 5; ++j)
            r[j] = (double)v[(size_t)i * 5 + j] - r[j];
    }
    double out[5];
#pragma unroll
    for (int a = 0; a < 5; ++a)
    {
        double acc = 0.0;
#pragma unroll
        for (int b = 0; b < 5; ++b)
            acc += (double)dinv[tiled(p, a * 5 + b, 25)] * r[b];
        out[a] = acc;
    }
#pragma unroll
    for (int a = 0; a < 5; ++a)
        z[(size_t)i * 5 + a] =
            FORWARD ? (T)out[a] : (T)((double)z[(size_t)i * 5 + a] - out[a]);
}

// k_multidot<T, K, VEC> for a run-time K in 1..16
template <typename T, bool VEC, int K = 1>
void multidot_launch(hipStream_t s, int k, int n, const T *V, size_t ld, const T *w,
                     double *partials, const int *on)
{
    if (k == K)
        hipLaunchKernelGGL((k_multidot<T, K, VEC>), dim3(kDotBlocks), dim3(kBlock), 0, s, n, V,
                           ld, w, partials, on);
    else if constexpr (K < 16)
        multidot_launch<T, VEC, K + 1>(s, k, n, V, ld, w, partials, on);
}

} // namespace

void local_dt(hipStream_t s, const DevMesh &m, const PhysDev &P, const double *U,
              double cfl, double *dtau)
{
    hipLaunchKernelGGL(k_local_dt, dim3(blocks_for(m.N)), dim3(kBlock), 0, s, m, P, U,
                       cfl, dtau);
}

void precond_blocks(hipStream_t s, const DevMesh &m, const PhysDev &P, int flux,
                    const double *U, const double *dtau, int blocks, double *Minv)
{
    if (m.maxf > 6)
        throw std::runtime_error("cells with more than 6 faces are not supported");
    const dim3 g(blocks_for(m.N)), b(kBlock);
    if (flux == dev::FluxRusanov)
    {
        if (m.maxf <= 4)
            hipLaunchKernelGGL((k_precond_blocks<dev::FluxRusanov, 4>), g, b, 0, s, m, P, U,
                               dtau, blocks, Minv);
        else
            hipLaunchKernelGGL((k_precond_blocks<dev::FluxRusanov, 6>), g, b, 0, s, m, P, U,
                               dtau, blocks, Minv);
    }
    else
    {
        if (m.maxf <= 4)
            hipLaunchKernelGGL((k_precond_blocks<dev::FluxAUSM, 4>), g, b, 0, s, m, P, U,
                               dtau, blocks, Minv);
        else
            hipLaunchKernelGGL((k_precond_blocks<dev::FluxAUSM, 6>), g, b, 0, s, m, P, U,
                               dtau, blocks, Minv);
    }
}

template <typename T>
void precond_apply(hipStream_t s, int N, const double *Minv, const T *v, T *z)
{
    hipLaunchKernelGGL(k_precond_apply<T>, dim3(blocks_for(N)), dim3(kBlock), 0, s, N,
                       Minv, v, z);
}

template <typename T>
void precond_apply_norm(hipStream_t s, int N, const double *Minv, const T *v, T *z,
                        double *partials, double *zz)
{
    hipLaunchKernelGGL(k_precond_apply_norm<T>, dim3(kDotBlocks), dim3(kBlock), 0, s, N, Minv,
                       v, z, partials);
    hipLaunchKernelGGL(k_fold_dots, dim3(1), dim3(kBlock), 0, s, kDotBlocks, partials, zz,
                       (const int *)nullptr);
}

template <typename T, typename O>
void perturb(hipStream_t s, int n, const double *U, const double *zz, double num,
             const T *z, O *Ut, double *eps)
{
    hipLaunchKernelGGL((k_perturb<T, O>), dim3(blocks_for(n)), dim3(kBlock), 0, s, n, U, zz,
                       num, z, Ut, eps);
}

template void perturb<double, double>(hipStream_t, int, const double *, const double *, double,
                                      const double *, double *, double *);
template void perturb<float, double>(hipStream_t, int, const double *, const double *, double,
                                     const float *, double *, double *);
template void perturb<float, float>(hipStream_t, int, const double *, const double *, double,
                                    const float *, float *, double *);
template void perturb<double, float>(hipStream_t, int, const double *, const double *, double,
                                     const double *, float *, double *);

template <typename T>
void op_finish(hipStream_t s, int N, const T *z, const double *dtau, const double *Rt,
               const double *R0, const double *eps, T *w)
{
    hipLaunchKernelGGL(k_op_finish<T>, dim3(blocks_for(5L * N)), dim3(kBlock), 0, s, N, z,
                       dtau, Rt, R0, eps, w);
}

template <typename T>
void rhs_minus(hipStream_t s, int n, const double *b, T *w)
{
    hipLaunchKernelGGL(k_rhs_minus<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, b, w);
}

template <typename T>
void multidot(hipStream_t s, int n, int k, const T *V, size_t ld, const T *w,
              double *partials, double *dots, const int *on)
{
    if (k < 1 || k > kMaxDots)
        throw std::runtime_error("multidot: bad vector count");
    // at most 16 columns per launch (w is read once per launch): 16
    // accumulators keep the kernel at a high occupancy, which measured
    // faster than one launch with 32 or 64 accumulators; each column's sum
    // is the same in either form
    const bool vec = aligned16(V) && aligned16(w) && (ld * sizeof(T)) % 16 == 0;
    for (int j0 = 0; j0 < k; j0 += 16)
    {
        const int kk = std::min(16, k - j0);
        if (vec)
            multidot_launch<T, true>(s, kk, n, V + (size_t)j0 * ld, ld, w, partials + j0, on);
        else
            multidot_launch<T, false>(s, kk, n, V + (size_t)j0 * ld, ld, w, partials + j0, on);
    }
    hipLaunchKernelGGL(k_fold_dots, dim3(k), dim3(kBlock), 0, s, kDotBlocks, partials, dots, on);
}

template <typename T>
void multiaxpy(hipStream_t s, int n, int k, const T *V, size_t ld, const double *h, T *w,
               const int *on)
{
    if (aligned16(V) && aligned16(w) && (ld * sizeof(T)) % 16 == 0)
    {
        constexpr int W = 16 / sizeof(T);
        const int threads = n / W + n % W;
        hipLaunchKernelGGL((k_multiaxpy<T, true>), dim3(blocks_for(threads)), dim3(kBlock), 0, s,
                           n, k, V, ld, h, w, on);
    }
    else
        hipLaunchKernelGGL((k_multiaxpy<T, false>), dim3(blocks_for(n)), dim3(kBlock), 0, s, n,
                           k, V, ld, h, w, on);
}

void arnoldi_norm(hipStream_t s, int k, const double *h, const double *ww, double *ar,
                  int *again, double eta)
{
    hipLaunchKernelGGL(k_arnoldi_norm, dim3(1), dim3(64), 0, s, k, h, ww, ar, again, eta);
}

template <typename T>
void arnoldi_finish(hipStream_t s, int n, T *w, const int *again, const double *ww2, double *ar,
                    double tiny)
{
    hipLaunchKernelGGL(k_arnoldi_finish<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, w,
                       again, ww2, ar, tiny);
}

template <typename T>
void scale(hipStream_t s, int n, const T *w, double sc, T *out)
{
    hipLaunchKernelGGL(k_scale<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, w, sc, out);
}

template <typename T>
void combine(hipStream_t s, int n, int k, const T *V, size_t ld, const double *y, T *x)
{
    hipLaunchKernelGGL(k_combine<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, k, V, ld,
                       y, x);
}

template <typename T>
void accumulate(hipStream_t s, int n, const T *z, double *x)
{
    hipLaunchKernelGGL(k_accumulate<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, z, x);
}

template <typename T>
void convert(hipStream_t s, int n, const double *src, T *dst)
{
    hipLaunchKernelGGL(k_convert<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, src, dst);
}

void residual_sums(hipStream_t s, int N, const double *R, const double *vol,
                   double *partials, double *sums)
{
    hipLaunchKernelGGL(k_residual_sums, dim3(kDotBlocks), dim3(kBlock), 0, s, N, R, vol,
                       partials);
    hipLaunchKernelGGL(k_fold_dots, dim3(5), dim3(kBlock), 0, s, kDotBlocks, partials, sums,
                       (const int *)nullptr);
}

void update_ratio(hipStream_t s, int N, const double *U, const double *dU, double *partials,
                  double *out)
{
    hipLaunchKernelGGL(k_update_ratio, dim3(kDotBlocks), dim3(kBlock), 0, s, N, U, dU,
                       partials);
    hipLaunchKernelGGL(k_fold_max, dim3(1), dim3(kBlock), 0, s, kDotBlocks, partials, out);
}

void update(hipStream_t s, int n, double *U, const double *dU, double alpha)
{
    hipLaunchKernelGGL(k_update, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, U, dU, alpha);
}

void dilu_offdiag(hipStream_t s, const DevMesh &m, const PhysDev &P, int flux,
                  const DiluMesh &d, const double *U, double *off)
{
    if (m.maxf > 6)
        throw std::runtime_error("cells with more than 6 faces are not supported");
    const dim3 g(blocks_for(d.N)), b(kBlock);
    if (flux == dev::FluxRusanov)
    {
        if (m.maxf <= 4)
            hipLaunchKernelGGL((k_dilu_offdiag<dev::FluxRusanov, 4>), g, b, 0, s, m, P, d, U,
                               off);
        else
            hipLaunchKernelGGL((k_dilu_offdiag<dev::FluxRusanov, 6>), g, b, 0, s, m, P, d, U,
                               off);
    }
    else
    {
        if (m.maxf <= 4)
            hipLaunchKernelGGL((k_dilu_offdiag<dev::FluxAUSM, 4>), g, b, 0, s, m, P, d, U, off);
        else
            hipLaunchKernelGGL((k_dilu_offdiag<dev::FluxAUSM, 6>), g, b, 0, s, m, P, d, U, off);
    }
}

template <typename B, typename O>
void dilu_factor(hipStream_t s, const DiluMesh &d, int p0, int p1, const double *Minv,
                 const double *off, double *dinv, O *offB, B *dinvB)
{
    if (p1 <= p0)
        return;
    hipLaunchKernelGGL((k_dilu_factor<B, O>), dim3(blocks_for(p1 - p0)), dim3(kBlock), 0, s, d,
                       p0, p1, Minv, off, dinv, offB, dinvB);
}

template <typename T, typename B, typename O, int MAXF>
void dilu_sweeps(hipStream_t s, const DiluMesh &d, const int *coff, const O *off,
                 const B *dinv, const T *v, T *z)
{
    // colour 0 has no lower neighbours (a greedy colouring gives every cell
    // of colour q > 0 a neighbour of each lower colour, and none to colour 0)
    if (d.ncol > 0 && coff[1] > coff[0])
        hipLaunchKernelGGL((k_dilu_sweep<T, B, O, MAXF, true, true>),
                           dim3(blocks_for(coff[1] - coff[0])), dim3(kBlock), 0, s, d, coff[0],
                           coff[1], off, dinv, v, z);
    for (int q = 1; q < d.ncol; ++q)
        if (coff[q + 1] > coff[q])
            hipLaunchKernelGGL((k_dilu_sweep<T, B, O, MAXF, true>),
                               dim3(blocks_for(coff[q + 1] - coff[q])), dim3(kBlock), 0, s, d,
                               coff[q], coff[q + 1], off, dinv, v, z);
    // the highest colour has no higher neighbours
    for (int q = d.ncol - 2; q >= 0; --q)
        if (coff[q + 1] > coff[q])
            hipLaunchKernelGGL((k_dilu_sweep<T, B, O, MAXF, false>),
                               dim3(blocks_for(coff[q + 1] - coff[q])), dim3(kBlock), 0, s, d,
                               coff[q], coff[q + 1], off, dinv, v, z);
}

template <typename T, typename B, typename O>
void dilu_apply(hipStream_t s, const DiluMesh &d, const int *coff, const O *off,
                const B *dinv, const T *v, T *z)
{
    if (d.maxf <= 4)
        dilu_sweeps<T, B, O, 4>(s, d, coff, off, dinv, v, z);
    else if (d.maxf <= 6)
        dilu_sweeps<T, B, O, 6>(s, d, coff, off, dinv, v, z);
    else
        throw std::runtime_error("DILU: cells with more than 6 faces are not supported");
}

template void dilu_factor<double, double>(hipStream_t, const DiluMesh &, int, int,
                                          const double *, const double *, double *, double *,
                                          double *);
template void dilu_factor<float, float>(hipStream_t, const DiluMesh &, int, int, const double *,
                                        const double *, double *, float *, float *);
template void dilu_factor<float, bf16_t>(hipStream_t, const DiluMesh &, int, int,
                                         const double *, const double *, double *, bf16_t *,
                                         float *);
#define FVH_DILU_APPLY(T, B, O)                                                            \
    template void dilu_apply<T, B, O>(hipStream_t, const DiluMesh &, const int *, const O *, \
                                      const B *, const T *, T *);
FVH_DILU_APPLY(double, double, double)
FVH_DILU_APPLY(float, float, float)
FVH_DILU_APPLY(double, float, float)
FVH_DILU_APPLY(float, float, bf16_t)
FVH_DILU_APPLY(double, float, bf16_t)

#define FVH_KRY_INST(T)                                                                    \
    template void precond_apply<T>(hipStream_t, int, const double *, const T *, T *);      \
    template void precond_apply_norm<T>(hipStream_t, int, const double *, const T *, T *,  \
                                        double *, double *);                               \
    template void op_finish<T>(hipStream_t, int, const T *, const double *,                \
                               const double *, const double *, const double *, T *);       \
    template void rhs_minus<T>(hipStream_t, int, const double *, T *);                     \
    template void multidot<T>(hipStream_t, int, int, const T *, size_t, const T *,         \
                              double *, double *, const int *);                            \
    template void multiaxpy<T>(hipStream_t, int, int, const T *, size_t, const double *,   \
                               T *, const int *);                                          \
    template void arnoldi_finish<T>(hipStream_t, int, T *, const int *, const double *,    \
                                    double *, double);                                     \
    template void scale<T>(hipStream_t, int, const T *, double, T *);                      \
    template void combine<T>(hipStream_t, int, int, const T *, size_t, const double *,     \
                             T *);                                                         \
    template void accumulate<T>(hipStream_t, int, const T *, double *);                    \
    template void convert<T>(hipStream_t, int, const double *, T *);

FVH_KRY_INST(double)
FVH_KRY_INST(float)

} // namespace kry
} // namespace fvh
