// Kernels of the reference-semantics implicit time step (timestep_newton).
// Every value is computed with the reference's operation order
// (src/Solver/src/timedisc/implicit.c), compiled with -ffp-contract=off.
#include <cfloat>
#include <cmath>

#include "fv_physics.h"
#include "implicit.h"

namespace fvh {
namespace imp {

using dev::DevMesh;
using dev::PhysDev;
using dev::kBlock;

namespace {

constexpr int kSumBlocks = 1024;
constexpr int kSumStride = 64;

inline int blocks_for(long n) { return (int)((n + kBlock - 1) / kBlock); }

__global__ __launch_bounds__(kBlock) void k_bdf_rhs(int n, const double *__restrict__ Y,
                                                    const double *__restrict__ old0,
                                                    const double *__restrict__ old1,
                                                    int n_loc, double a0, double a1,
                                                    double b, double dt,
                                                    const double *__restrict__ R,
                                                    double *__restrict__ f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    double v = -(Y[i] - old0[i]) / dt + b * R[i];
    v -= a0 / dt * old0[i];
    if (n_loc > 1)
        v -= a1 / dt * old1[i];
    f[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_shift_var(int N, const double *__restrict__ U,
                                                      int k, double delta,
                                                      double *__restrict__ Ut)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * N)
        return;
    const double u = U[i];
    Ut[i] = (i % 5 == k) ? u + delta : u;
}

__global__ __launch_bounds__(kBlock) void k_jac_column(int N, const double *__restrict__ Rp,
                                                       const double *__restrict__ Rm, int k,
                                                       double fac, double inv_dt,
                                                       double *__restrict__ jac)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 5 * N)
        return;
    const int c = i / 5, j = i % 5;
    double v = -Rp[i];
    v += Rm[i];
    v *= fac;
    if (j == k)
        v += inv_dt;
    jac[(size_t)c * 25 + k * 5 + j] = v;
}

// Gaussian elimination with partial pivoting on [J | f]; compile-time
// indices only, so the 5x6 system stays in registers.
__global__ __launch_bounds__(kBlock) void k_block_solve(int N, const double *__restrict__ jac,
                                                        const double *__restrict__ f,
                                                        double *__restrict__ dY)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N)
        return;
    double A[5][6];
    const double *J = jac + (size_t)c * 25;
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
#pragma unroll
        for (int q = 0; q < 5; ++q)
            A[r][q] = J[q * 5 + r];
        A[r][5] = f[(size_t)c * 5 + r];
    }
#pragma unroll
    for (int col = 0; col < 5; ++col)
    {
        int p = col;
        double best = fabs(A[col][col]);
#pragma unroll
        for (int r = col + 1; r < 5; ++r)
            if (fabs(A[r][col]) > best)
            {
                best = fabs(A[r][col]);
                p = r;
            }
#pragma unroll
        for (int r = col + 1; r < 5; ++r)
            if (r == p)
#pragma unroll
                for (int q = 0; q < 6; ++q)
                {
                    const double t = A[col][q];
                    A[col][q] = A[r][q];
                    A[r][q] = t;
                }
#pragma unroll
        for (int r = col + 1; r < 5; ++r)
        {
            const double fr = A[r][col] / A[col][col];
#pragma unroll
            for (int q = col; q < 6; ++q)
                A[r][q] -= fr * A[col][q];
        }
    }
    double x[5];
#pragma unroll
    for (int r = 4; r >= 0; --r)
    {
        double v = A[r][5];
#pragma unroll
        for (int q = r + 1; q < 5; ++q)
            v -= A[r][q] * x[q];
        x[r] = v / A[r][r];
    }
#pragma unroll
    for (int r = 0; r < 5; ++r)
        dY[(size_t)c * 5 + r] = x[r];
}

__global__ __launch_bounds__(kBlock) void k_column_sumsq(int N, const double *__restrict__ U,
                                                         double *__restrict__ partials)
{
    __shared__ double sh[kBlock];
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int stride = gridDim.x * blockDim.x;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += stride)
    {
#pragma unroll
        for (int k = 0; k < 5; ++k)
        {
            const double u = U[(size_t)c * 5 + k];
            a[k] += u * u;
        }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
        dev::block_sum(sh, a[k]);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 5; ++k)
            partials[(size_t)blockIdx.x * kSumStride + k] = a[k];
}

__global__ __launch_bounds__(kBlock) void k_fold5(int nblk, const double *__restrict__ partials,
                                                  double *__restrict__ out)
{
    __shared__ double sh[kBlock];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock)
        s += partials[(size_t)b * kSumStride + k];
    dev::block_sum(sh, s);
    if (threadIdx.x == 0)
        out[k] = s;
}

__global__ __launch_bounds__(kBlock) void k_step_partials(DevMesh m, PhysDev P,
                                                          const double *__restrict__ U,
                                                          const double *__restrict__ R,
                                                          double *__restrict__ partials)
{
    __shared__ double sh[kBlock];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lc = DBL_MAX, lv = DBL_MAX;
    if (c < m.N)
    {
        const double vol = m.vol[c];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            r[j] = dev::rabs(R[(size_t)c * 5 + j]) * vol;
        double pc[10];
        dev::load_cons(U, c, pc);
        dev::con_to_prim(P, pc);
        const double dx0 = m.cdx[(size_t)c * 3], dx1 = m.cdx[(size_t)c * 3 + 1],
                     dx2 = m.cdx[(size_t)c * 3 + 2];
        const double s_rho = 1.0 / pc[0];
        const double cs = sqrt(P.kappa * pc[8] * s_rho);
        const double ds1 = (dev::rabs(pc[5]) + cs) * dx0 + (dev::rabs(pc[6]) + cs) * dx1 +
                           (dev::rabs(pc[7]) + cs) * dx2;
        lc = vol / ds1;
        const double ds2 = dev::dot3(dx0, dx1, dx2, dx0, dx1, dx2);
        lv = s_rho * ds2 / vol;
    }
#pragma unroll
    for (int j = 0; j < 5; ++j)
        dev::block_sum(sh, r[j]);
    dev::block_min(sh, lc);
    dev::block_min(sh, lv);
    if (threadIdx.x == 0)
    {
        double *p = partials + (size_t)blockIdx.x * dev::kPartialStride;
#pragma unroll
        for (int j = 0; j < 5; ++j)
            p[j] = r[j];
        p[5] = lc;
        p[6] = lv;
    }
}

} // namespace

void bdf_rhs(hipStream_t s, int n, const double *Y, const double *old0, const double *old1,
             int n_loc, double a0, double a1, double b, double dt, const double *R,
             double *f)
{
    hipLaunchKernelGGL(k_bdf_rhs, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, Y, old0,
                       old1, n_loc, a0, a1, b, dt, R, f);
}

void shift_var(hipStream_t s, int N, const double *U, int k, double delta, double *Ut)
{
    hipLaunchKernelGGL(k_shift_var, dim3(blocks_for(5L * N)), dim3(kBlock), 0, s, N, U, k,
                       delta, Ut);
}

void jac_column(hipStream_t s, int N, const double *Rp, const double *Rm, int k, double fac,
                double inv_dt, double *jac)
{
    hipLaunchKernelGGL(k_jac_column, dim3(blocks_for(5L * N)), dim3(kBlock), 0, s, N, Rp,
                       Rm, k, fac, inv_dt, jac);
}

// ---- element kernels of the linear solvers (the oracle's expressions) ----
__global__ __launch_bounds__(kBlock) void k_block_matvec(int N, const double *__restrict__ jac,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    const double *J = jac + (size_t)i * 25;
    double xi[5], bi[5];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        xi[k] = x[(size_t)i * 5 + k];
#pragma unroll
    for (int j = 0; j < 5; ++j)
        bi[j] = J[j] * xi[0];
#pragma unroll
    for (int k = 1; k < 5; ++k)
#pragma unroll
        for (int j = 0; j < 5; ++j)
            bi[j] += J[k * 5 + j] * xi[k];
#pragma unroll
    for (int j = 0; j < 5; ++j)
        b[(size_t)i * 5 + j] = bi[j];
}

__global__ __launch_bounds__(kBlock) void k_lsoe_elem(int n, int op, double a, double b,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z,
                                                     double *__restrict__ o1,
                                                     double *__restrict__ o2)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n)
        return;
    switch (op)
    {
    case 0: // o1 = x / a
        o1[q] = x[q] / a;
        break;
    case 1: // o1 = x - a y
        o1[q] = x[q] - a * y[q];
        break;
    case 2: // p = r + beta (p - omega v): o1 = p, x = r, y = v
        o1[q] = x[q] + a * (o1[q] - b * y[q]);
        break;
    case 3: // o1 += a x
        o1[q] += a * x[q];
        break;
    default: // o1 += a x + b y; o2 = y - b z  (x += alpha p + omega s; r = s - omega t)
        o1[q] += a * x[q] + b * y[q];
        o2[q] = y[q] - b * z[q];
        break;
    }
}

// x += sum_i y[i] V_i, sequentially in i (y in device memory)
__global__ __launch_bounds__(kBlock) void k_add_combination(int n, int k,
                                                           const double *__restrict__ V,
                                                           size_t ld, const double *__restrict__ y,
                                                           double *__restrict__ x)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n)
        return;
    double a = x[q];
    for (int i = 0; i < k; ++i)
        a += y[i] * V[(size_t)i * ld + q];
    x[q] = a;
}

void block_matvec(hipStream_t s, int N, const double *jac, const double *x, double *b)
{
    hipLaunchKernelGGL(k_block_matvec, dim3(blocks_for(N)), dim3(kBlock), 0, s, N, jac, x, b);
}

void lsoe_elem(hipStream_t s, int n, int op, double a, double b, const double *x,
               const double *y, const double *z, double *o1, double *o2)
{
    hipLaunchKernelGGL(k_lsoe_elem, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, op, a, b, x, y,
                       z, o1, o2);
}

void add_combination(hipStream_t s, int n, int k, const double *V, size_t ld, const double *y,
                     double *x)
{
    hipLaunchKernelGGL(k_add_combination, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, k, V, ld,
                       y, x);
}

void block_solve(hipStream_t s, int N, const double *jac, const double *f, double *dY)
{
    hipLaunchKernelGGL(k_block_solve, dim3(blocks_for(N)), dim3(kBlock), 0, s, N, jac, f,
                       dY);
}

void column_sumsq(hipStream_t s, int N, const double *U, double *partials, double *sums)
{
    hipLaunchKernelGGL(k_column_sumsq, dim3(kSumBlocks), dim3(kBlock), 0, s, N, U,
                       partials);
    hipLaunchKernelGGL(k_fold5, dim3(5), dim3(kBlock), 0, s, kSumBlocks, partials, sums);
}

void step_partials(hipStream_t s, const DevMesh &m, const PhysDev &P, const double *U,
                   const double *R, double *partials)
{
    hipLaunchKernelGGL(k_step_partials, dim3(dev::cell_blocks(m.N)), dim3(kBlock), 0, s, m,
                       P, U, R, partials);
}

} // namespace imp
} // namespace fvh
