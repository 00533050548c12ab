// Kernels of the host-mirror residual path: the reference's passes of one
// residual evaluation over device copies of its AoS arrays (layouts in
// mirror.h).  One thread per cell, face or boundary; every cell gathers its
// faces in ascending order (the reference's face-ordered scatter), with the
// same operation order as the reference, so the results equal the fast
// path's bit for bit.  A face that joins a cell to itself (a periodic
// direction one cell wide) adds and then subtracts its contribution, as the
// scatter does.
#include "fv_physics.h"
#include "mirror.h"

namespace fvh {
namespace mirror {

using dev::kBlock;

__device__ __forceinline__ void frame_of(const MirrorMesh &m, int f, dev::Frame &fr)
{
    const double *a = m.fgA + (size_t)f * dev::kFgA;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        fr.n[d] = a[d];
        fr.t1[d] = a[3 + d];
        fr.t2[d] = a[6 + d];
    }
    fr.area = a[9];
}

__global__ __launch_bounds__(kBlock) void mk_gradients(MirrorMesh m, const double *__restrict__ phi,
                                                      double *__restrict__ gx,
                                                      double *__restrict__ gy,
                                                      double *__restrict__ gz)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.N)
        return;
    double ax[10], ay[10], az[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
        ax[j] = ay[j] = az[j] = 0.0;
    for (int s = 0; s < m.ncf[c]; ++s)
    {
        const int f = m.cf[(size_t)c * m.maxf + s];
        const int c0 = m.fcell[2 * f], c1 = m.fcell[2 * f + 1];
        const double *b = m.fgB + (size_t)f * dev::kFgB;
        const double n0 = b[0], n1 = b[1], n2 = b[2], area = b[3], lam = b[4];
#pragma unroll
        for (int j = 0; j < 10; ++j)
        {
            const double p0 = phi[(size_t)c0 * 10 + j];
            const double pm = p0 + lam * (phi[(size_t)c1 * 10 + j] - p0);
            const double vx = pm * n0 * area, vy = pm * n1 * area, vz = pm * n2 * area;
            if (c0 == c)
            {
                ax[j] += vx;
                ay[j] += vy;
                az[j] += vz;
            }
            if (c1 == c)
            {
                ax[j] -= vx;
                ay[j] -= vy;
                az[j] -= vz;
            }
        }
    }
    const double sv = m.svol[c];
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        gx[(size_t)c * 10 + j] = ax[j] * sv;
        gy[(size_t)c * 10 + j] = ay[j] * sv;
        gz[(size_t)c * 10 + j] = az[j] * sv;
    }
}

// Ghost gradient rows as the reference's scatter leaves them before the
// update_gradients hook runs: zero minus the boundary face's contribution
// (reconstruction.c:33-68; its volume scaling covers the domain rows only).
__global__ __launch_bounds__(kBlock) void mk_ghost_raw(MirrorMesh m, const double *__restrict__ phi,
                                                      double *__restrict__ gx,
                                                      double *__restrict__ gy,
                                                      double *__restrict__ gz)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B)
        return;
    const int f = m.bnd_face[b];
    const int c0 = m.fcell[2 * f], c1 = m.fcell[2 * f + 1];
    const double *q = m.fgB + (size_t)f * dev::kFgB;
    const double n0 = q[0], n1 = q[1], n2 = q[2], area = q[3], lam = q[4];
    const size_t o = (size_t)(m.N + b) * 10;
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        const double p0 = phi[(size_t)c0 * 10 + j];
        const double pm = p0 + lam * (phi[(size_t)c1 * 10 + j] - p0);
        gx[o + j] = 0.0 - pm * n0 * area;
        gy[o + j] = 0.0 - pm * n1 * area;
        gz[o + j] = 0.0 - pm * n2 * area;
    }
}

__global__ __launch_bounds__(kBlock) void mk_ghost_gradients(MirrorMesh m, double *__restrict__ gx,
                                                            double *__restrict__ gy,
                                                            double *__restrict__ gz)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B)
        return;
    const int f = m.bnd_face[b];
    const int c = m.fcell[2 * f];
    const int type = m.region_type[m.bnd_region[b]];
    dev::Frame fr;
    frame_of(m, f, fr);
    const size_t o = (size_t)(m.N + b) * 10, i = (size_t)c * 10;
#pragma unroll
    for (int j = 0; j < 10; ++j)
        dev::ghost_grad(type, fr, gx[i + j], gy[i + j], gz[i + j], gx[o + j], gy[o + j],
                        gz[o + j]);
}

// neighbourhood bounds of variable j of cell c (limiter.c:57-82)
__device__ __forceinline__ void bounds(const MirrorMesh &m, const double *phi, int c, int j,
                                       double &mn, double &mx)
{
    const double p = phi[(size_t)c * 10 + j];
    mn = p;
    mx = p;
    for (int s = 0; s < m.ncf[c]; ++s)
    {
        const int f = m.cf[(size_t)c * m.maxf + s];
        const int nb = m.fcell[2 * f] == c ? m.fcell[2 * f + 1] : m.fcell[2 * f];
        const double q = phi[(size_t)nb * 10 + j];
        mn = dev::rmin(mn, q);
        mx = dev::rmax(mx, q);
    }
}

__global__ __launch_bounds__(kBlock) void mk_reconstruct(
    MirrorMesh m, dev::PhysDev P, int linear, const double *__restrict__ phi,
    const double *__restrict__ gx, const double *__restrict__ gy,
    const double *__restrict__ gz, double *__restrict__ L, double *__restrict__ R)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const int c0 = m.fcell[2 * f], c1 = m.fcell[2 * f + 1];
    const bool boundary = c1 >= m.N;
    for (int side = 0; side < 2; ++side)
    {
        const int c = side == 0 ? c0 : c1;
        double *out = (side == 0 ? L : R) + (size_t)f * 10;
        if (!linear || (side == 1 && boundary))
        {
#pragma unroll
            for (int j = 0; j < 10; ++j)
                out[j] = phi[(size_t)c * 10 + j];
            continue;
        }
        const double *d = (side == 0 ? m.d1 : m.d2) + (size_t)f * 3;
        const double r[3] = {d[0], d[1], d[2]};
        for (int j = 0; j < 10; ++j)
        {
            double mn = 0.0, mx = 0.0;
            if (P.bj)
                bounds(m, phi, c, j, mn, mx);
            const size_t q = (size_t)c * 10 + j;
            out[j] = dev::limited(phi[q], gx[q], gy[q], gz[q], r, mn, mx, P.bj);
        }
    }
}

template <int FLUX, bool VISC>
__global__ __launch_bounds__(kBlock) void mk_flux(MirrorMesh m, dev::PhysDev P,
                                                 const double *__restrict__ L,
                                                 const double *__restrict__ R,
                                                 const double *__restrict__ gx,
                                                 const double *__restrict__ gy,
                                                 const double *__restrict__ gz,
                                                 double *__restrict__ flux)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const int c0 = m.fcell[2 * f], c1 = m.fcell[2 * f + 1];
    dev::Frame fr;
    frame_of(m, f, fr);
    double l[10], r[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        l[j] = L[(size_t)f * 10 + j];
        r[j] = R[(size_t)f * 10 + j];
    }
    dev::ViscGrad gl, gr;
    if (VISC)
    {
        const double *g[3] = {gx, gy, gz};
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                gl.v[d * 4 + q] = g[d][(size_t)c0 * 10 + dev::visc_var(q)];
                gr.v[d * 4 + q] = g[d][(size_t)c1 * 10 + dev::visc_var(q)];
            }
    }
    double out[5];
    dev::face_flux<FLUX, VISC>(P, fr, l, r, gl, gr, out);
#pragma unroll
    for (int j = 0; j < 5; ++j)
        flux[(size_t)f * 5 + j] = out[j];
}

__global__ __launch_bounds__(kBlock) void mk_sum(MirrorMesh m, const double *__restrict__ flux,
                                                double *__restrict__ Rout)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.N)
        return;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < m.ncf[c]; ++s)
    {
        const int f = m.cf[(size_t)c * m.maxf + s];
        const double area = m.fgA[(size_t)f * dev::kFgA + 9];
        const bool first = m.fcell[2 * f] == c, second = m.fcell[2 * f + 1] == c;
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const double v = flux[(size_t)f * 5 + j] * area;
            if (first)
                acc[j] += v;
            if (second)
                acc[j] -= v;
        }
    }
#pragma unroll
    for (int j = 0; j < 5; ++j)
        Rout[(size_t)c * 5 + j] = -acc[j] / m.vol[c];
}

static int blocks(int n) { return (n + kBlock - 1) / kBlock; }

void gradients(hipStream_t s, const MirrorMesh &m, const double *phi, double *gx, double *gy,
               double *gz)
{
    hipLaunchKernelGGL(mk_gradients, dim3(blocks(m.N)), dim3(kBlock), 0, s, m, phi, gx, gy, gz);
}

void ghost_raw(hipStream_t s, const MirrorMesh &m, const double *phi, double *gx, double *gy,
               double *gz)
{
    if (m.B == 0)
        return;
    hipLaunchKernelGGL(mk_ghost_raw, dim3(blocks(m.B)), dim3(kBlock), 0, s, m, phi, gx, gy, gz);
}

void ghost_gradients(hipStream_t s, const MirrorMesh &m, double *gx, double *gy, double *gz)
{
    if (m.B == 0)
        return;
    hipLaunchKernelGGL(mk_ghost_gradients, dim3(blocks(m.B)), dim3(kBlock), 0, s, m, gx, gy, gz);
}

void reconstruct(hipStream_t s, const MirrorMesh &m, const dev::PhysDev &P, int linear,
                 const double *phi, const double *gx, const double *gy, const double *gz,
                 double *L, double *R)
{
    hipLaunchKernelGGL(mk_reconstruct, dim3(blocks(m.F)), dim3(kBlock), 0, s, m, P, linear, phi,
                       gx, gy, gz, L, R);
}

void flux(hipStream_t s, const MirrorMesh &m, const dev::PhysDev &P, int kind, bool viscous,
          const double *L, const double *R, const double *gx, const double *gy,
          const double *gz, double *out)
{
    const dim3 g(blocks(m.F)), b(kBlock);
    if (kind == dev::FluxRusanov)
    {
        if (viscous)
            hipLaunchKernelGGL((mk_flux<dev::FluxRusanov, true>), g, b, 0, s, m, P, L, R, gx, gy,
                               gz, out);
        else
            hipLaunchKernelGGL((mk_flux<dev::FluxRusanov, false>), g, b, 0, s, m, P, L, R, gx,
                               gy, gz, out);
    }
    else
    {
        if (viscous)
            hipLaunchKernelGGL((mk_flux<dev::FluxAUSM, true>), g, b, 0, s, m, P, L, R, gx, gy, gz,
                               out);
        else
            hipLaunchKernelGGL((mk_flux<dev::FluxAUSM, false>), g, b, 0, s, m, P, L, R, gx, gy,
                               gz, out);
    }
}

void sum_fluxes(hipStream_t s, const MirrorMesh &m, const double *flux, double *R_out)
{
    hipLaunchKernelGGL(mk_sum, dim3(blocks(m.N)), dim3(kBlock), 0, s, m, flux, R_out);
}

} // namespace mirror
} // namespace fvh
