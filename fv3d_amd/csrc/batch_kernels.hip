// Kernels of the batched mode: many independent cases stored as one mesh
// (disjoint cell ranges), advanced together.  Per-case reductions run over
// fixed chunks of each case's cell range and fold per case in chunk order
// (deterministic).
#include <cfloat>
#include <cmath>

#include "batch.h"
#include "fv_physics.h"

namespace fvh {
namespace batch {

using dev::DevMesh;
using dev::PhysDev;
using dev::kBlock;

namespace {

// Residual sums and time-step minima of one chunk: sum |R| V (5), min V/ds1,
// min ds2/(rho V).  The time step takes the density of U and the primitives
// of Uprim (the last stage's input: navier_stokes.c:171-187 on the
// reference's stale primitives; Uprim = U for a refreshed state).
__global__ __launch_bounds__(kBlock) void k_chunk_partials(DevMesh m, PhysDev P,
                                                           const int3 *__restrict__ chunks,
                                                           const double *__restrict__ U,
                                                           const double *__restrict__ Uprim,
                                                           const double *__restrict__ R,
                                                           double *__restrict__ partials)
{
    __shared__ double sh[kBlock];
    const int3 ch = chunks[blockIdx.x]; // (case, begin, end)
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    for (int c = ch.y + threadIdx.x; c < ch.z; c += kBlock)
    {
        const double vol = m.vol[c];
        if (R)
#pragma unroll
            for (int j = 0; j < 5; ++j)
                v[j] += dev::rabs(R[(size_t)c * 5 + j]) * vol;
        double pc[10];
        dev::load_cons(Uprim, c, pc);
        dev::con_to_prim(P, pc);
        const double rho = U[(size_t)c * 5];
        const double dx0 = m.cdx[(size_t)c * 3], dx1 = m.cdx[(size_t)c * 3 + 1],
                     dx2 = m.cdx[(size_t)c * 3 + 2];
        const double s_rho = 1.0 / rho;
        const double cs = sqrt(P.kappa * pc[8] * s_rho);
        const double ds1 = (dev::rabs(pc[5]) + cs) * dx0 + (dev::rabs(pc[6]) + cs) * dx1 +
                           (dev::rabs(pc[7]) + cs) * dx2;
        v[5] = dev::rmin(v[5], vol / ds1);
        const double ds2 = dev::dot3(dx0, dx1, dx2, dx0, dx1, dx2);
        v[6] = dev::rmin(v[6], s_rho * ds2 / vol);
    }
#pragma unroll
    for (int j = 0; j < 5; ++j)
        dev::block_sum(sh, v[j]);
    dev::block_min(sh, v[5]);
    dev::block_min(sh, v[6]);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < 7; ++j)
            partials[(size_t)blockIdx.x * 8 + j] = v[j];
}

// One block per case: fold its chunks chunk_begin[k] .. chunk_end[k] - 1 in
// order; residual with the step's
// dt (analyze.c:57-58), then the next dt (navier_stokes.c:182-192 with the
// case's viscosity; the viscous limit is skipped for mu = 0).
__global__ __launch_bounds__(kBlock) void k_case_fold(int mode, PhysDev P,
                                                      const int *__restrict__ chunk_begin,
                                                      const int *__restrict__ chunk_end,
                                                      const double *__restrict__ partials,
                                                      const double *__restrict__ case_mu,
                                                      const double *__restrict__ case_vol,
                                                      double *__restrict__ case_dt,
                                                      double *__restrict__ case_t,
                                                      double *__restrict__ case_res)
{
    __shared__ double sh[kBlock];
    const int k = blockIdx.x;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    for (int i = chunk_begin[k] + threadIdx.x; i < chunk_end[k]; i += kBlock)
    {
        const double *p = partials + (size_t)i * 8;
#pragma unroll
        for (int j = 0; j < 5; ++j)
            v[j] += p[j];
        v[5] = dev::rmin(v[5], p[5]);
        v[6] = dev::rmin(v[6], p[6]);
    }
#pragma unroll
    for (int j = 0; j < 5; ++j)
        dev::block_sum(sh, v[j]);
    dev::block_min(sh, v[5]);
    dev::block_min(sh, v[6]);
    if (threadIdx.x != 0)
        return;
    if (mode == 1)
    {
        const double dt = case_dt[k];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            case_res[(size_t)k * 5 + j] = v[j] * (dt / case_vol[k]);
        case_t[k] = case_t[k] + dt;
    }
    const double mu = case_mu[k];
    const double lc = v[5] * P.cfl;
    double lv = v[6];
    lv *= P.dfl * mu * P.kappa_pr * lv;
    case_dt[k] = (P.visc_dt_mode == 1 && mu == 0.0) ? lc : dev::rmin(lc, lv);
}

__global__ void k_reset_counters(unsigned int *c, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        c[i] = 0u;
}

} // namespace

void reset_counters(hipStream_t s, unsigned int *counters, int n)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters, n);
}

void chunk_partials(hipStream_t s, const DevMesh &m, const PhysDev &P, int n_chunks,
                    const int3 *chunks, const double *U, const double *Uprim, const double *R,
                    double *partials)
{
    hipLaunchKernelGGL(k_chunk_partials, dim3(n_chunks), dim3(kBlock), 0, s, m, P, chunks, U,
                       Uprim, R, partials);
}

void case_fold(hipStream_t s, int mode, const PhysDev &P, int n_cases, const int *chunk_begin,
               const int *chunk_end, const double *partials, const double *case_mu,
               const double *case_vol, double *case_dt, double *case_t, double *case_res)
{
    hipLaunchKernelGGL(k_case_fold, dim3(n_cases), dim3(kBlock), 0, s, mode, P, chunk_begin,
                       chunk_end, partials, case_mu, case_vol, case_dt, case_t, case_res);
}

} // namespace batch
} // namespace fvh
