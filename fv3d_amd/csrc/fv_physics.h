// Device physics of the residual path, shared by the stencil and the
// Newton-Krylov kernels: the reference's EOS, boundary conditions, fluxes
// and reductions restated operation by operation (compiled with
// -ffp-contract=off everywhere).
//
// The arithmetic type Real is double for every path that is compared with
// the reference (the instantiation is the plain FP64 code, constant for
// constant); Real = float serves the FP32 Jacobian-vector products of the
// mixed-precision Newton-Krylov solve.  Arrays in memory stay FP64 and are
// converted on load.
#pragma once

#include <type_traits>

#include <cfloat>
#include <cmath>

#include "fv_device.h"

namespace fvh {
namespace dev {

// reference BM_MIN / BM_MAX / BM_ABS semantics (ternaries, not fmin/fmax)
template <class Real> __device__ __forceinline__ Real rmin(Real a, Real b) { return a < b ? a : b; }
template <class Real> __device__ __forceinline__ Real rmax(Real a, Real b) { return a > b ? a : b; }
template <class Real> __device__ __forceinline__ Real rabs(Real a) { return a < 0 ? -a : a; }

// basec dot_n: sequential sum from 0.0
template <class Real>
__device__ __forceinline__ Real dot3(Real a0, Real a1, Real a2, Real b0, Real b1, Real b2)
{
    Real s = Real(0.0);
    s += a0 * b0;
    s += a1 * b1;
    s += a2 * b2;
    return s;
}

enum
{
    BC_FLOW = 0,
    BC_INFLOW = 1,
    BC_OUTFLOW = 2,
    BC_WALL_ADIABATIC = 3,
    BC_WALL_ISOTHERMAL = 4,
    BC_WALL_SLIP = 5,
    BC_SYMMETRY = 6,
    BC_STATE = 7
};

// eos.c:81-92
template <class Real>
__device__ __forceinline__ void con_to_prim(const PhysDev &P, Real *phi)
{
    phi[5] = phi[1] / phi[0];
    phi[6] = phi[2] / phi[0];
    phi[7] = phi[3] / phi[0];
    Real p = Real(P.kappa_m1) * (phi[4] - Real(0.5) * dot3(phi[1], phi[2], phi[3],
                                                             phi[5], phi[6], phi[7]));
    p = rmax(Real(1.00E-10), p);
    phi[8] = p;
    phi[9] = p / (phi[0] * Real(P.R));
}

// eos.c:120-127
template <class Real>
__device__ __forceinline__ void prim_to_con(const PhysDev &P, Real *phi)
{
    phi[1] = phi[5] * phi[0];
    phi[2] = phi[6] * phi[0];
    phi[3] = phi[7] * phi[0];
    phi[4] = Real(P.s_kappa_m1) * phi[8] +
             Real(0.5) * dot3(phi[5], phi[6], phi[7], phi[1], phi[2], phi[3]);
}

// eos.c:55-75 (compression branch with the reference's A = 2/((k+1) u_n))
template <class Real>
__device__ __forceinline__ Real riemann_p(const PhysDev &P, const Real *phi)
{
    if (phi[5] <= Real(0.0))
    {
        Real c = sqrt(Real(P.kappa) * phi[8] / phi[0]);
        return phi[8] * pow(rmax(Real(1e-4), Real(1.) + Real(0.5) * Real(P.kappa_m1) * phi[5] / c),
                            Real(2) * Real(P.kappa) * Real(P.s_kappa_m1));
    }
    Real ar = Real(2) * Real(P.s_kappa_p1) / phi[5];
    Real br = Real(P.kappa_m1) * Real(P.s_kappa_p1) * phi[8];
    return phi[8] + phi[5] / ar * Real(0.5) *
                        (phi[5] + sqrt(phi[5] * phi[5] + Real(4.) * ar * (phi[8] + br)));
}

template <class Real> struct FrameT
{
    Real n[3], t1[3], t2[3];
    Real area;
};
using Frame = FrameT<double>;

template <class Real>
__device__ __forceinline__ void load_frame(const DevMesh &m, int f, FrameT<Real> &fr)
{
    if constexpr (std::is_same<Real, float>::value)
    {
        if (m.fgA32)
        {
            const float4 *q = reinterpret_cast<const float4 *>(m.fgA32 + (size_t)f * 12);
            const float4 a = q[0], b = q[1], c = q[2];
            fr.n[0] = a.x, fr.n[1] = a.y, fr.n[2] = a.z;
            fr.t1[0] = a.w, fr.t1[1] = b.x, fr.t1[2] = b.y;
            fr.t2[0] = b.z, fr.t2[1] = b.w, fr.t2[2] = c.x;
            fr.area = c.y;
            return;
        }
    }
    const double2 *g = reinterpret_cast<const double2 *>(m.fgA + (size_t)f * kFgA);
    const double2 a = g[0], b = g[1], c = g[2], d = g[3], e = g[4];
    fr.n[0] = Real(a.x), fr.n[1] = Real(a.y), fr.n[2] = Real(b.x);
    fr.t1[0] = Real(b.y), fr.t1[1] = Real(c.x), fr.t1[2] = Real(c.y);
    fr.t2[0] = Real(d.x), fr.t2[1] = Real(d.y), fr.t2[2] = Real(e.x);
    fr.area = Real(e.y);
}

template <class Real>
__device__ __forceinline__ void rotate_back(const FrameT<Real> &fr, Real *g)
{
    Real tu = g[5] * fr.n[0] + g[6] * fr.t1[0] + g[7] * fr.t2[0];
    Real tv = g[5] * fr.n[1] + g[6] * fr.t1[1] + g[7] * fr.t2[1];
    Real tw = g[5] * fr.n[2] + g[6] * fr.t1[2] + g[7] * fr.t2[2];
    g[5] = tu;
    g[6] = tv;
    g[7] = tw;
}

// Ghost state of boundary row N+b from the interior state pc
// (boundary.c:202-319).
template <class Real>
__device__ __forceinline__ void ghost_state(const DevMesh &m, const PhysDev &P, int b,
                                            const FrameT<Real> &fr, const Real *pc, Real *g)
{
    const int region = m.bnd_region[b];
    const int type = m.region_type[region];
#pragma unroll
    for (int j = 0; j < 10; ++j)
        g[j] = pc[j];
    if (type == BC_INFLOW || type == BC_STATE)
    {
#pragma unroll
        for (int j = 0; j < 10; ++j)
            g[j] = Real(m.region_phi[region * 10 + j]);
    }
    else if (type == BC_WALL_ADIABATIC || type == BC_WALL_ISOTHERMAL)
    {
        g[5] = dot3(pc[5], pc[6], pc[7], fr.n[0], fr.n[1], fr.n[2]);
        g[6] = dot3(pc[5], pc[6], pc[7], fr.t1[0], fr.t1[1], fr.t1[2]);
        g[7] = dot3(pc[5], pc[6], pc[7], fr.t2[0], fr.t2[1], fr.t2[2]);
        g[8] = riemann_p(P, g);
        g[5] = Real(0.0);
        g[6] = Real(0.0);
        g[7] = Real(0.0);
        if (type == BC_WALL_ISOTHERMAL)
            g[9] = Real(m.region_phi[region * 10 + 9]);
        g[0] = g[8] / (Real(P.R) * g[9]);
        rotate_back(fr, g);
    }
    else if (type == BC_WALL_SLIP || type == BC_SYMMETRY)
    {
        g[5] = dot3(pc[5], pc[6], pc[7], fr.n[0], fr.n[1], fr.n[2]);
        g[6] = dot3(pc[5], pc[6], pc[7], fr.t1[0], fr.t1[1], fr.t1[2]);
        g[7] = dot3(pc[5], pc[6], pc[7], fr.t2[0], fr.t2[1], fr.t2[2]);
        g[8] = riemann_p(P, g);
        g[5] = Real(0.0);
        g[9] = g[8] / (g[0] * Real(P.R));
        rotate_back(fr, g);
    }
    prim_to_con(P, g);
}

// Ghost gradient (boundary.c:361-406): a copy of the interior gradient, with
// the normal component removed for slip walls and symmetry.
template <class Real>
__device__ __forceinline__ void ghost_grad(int type, const FrameT<Real> &fr, Real gx, Real gy,
                                           Real gz, Real &ox, Real &oy, Real &oz)
{
    if (type == BC_WALL_SLIP || type == BC_SYMMETRY)
    {
        Real ty = gx * fr.t1[0] + gy * fr.t1[1] + gz * fr.t1[2];
        Real tz = gx * fr.t2[0] + gy * fr.t2[1] + gz * fr.t2[2];
        Real tx = Real(0.0);
        ox = tx * fr.n[0] + ty * fr.t1[0] + tz * fr.t2[0];
        oy = tx * fr.n[1] + ty * fr.t1[1] + tz * fr.t2[1];
        oz = tx * fr.n[2] + ty * fr.t1[2] + tz * fr.t2[2];
    }
    else
    {
        ox = gx;
        oy = gy;
        oz = gz;
    }
}

template <class Real>
__device__ __forceinline__ void load_cons(const double *__restrict__ U, int c, Real *phi)
{
    const double *p = U + (size_t)c * 5;
#pragma unroll
    for (int j = 0; j < 5; ++j)
        phi[j] = Real(p[j]);
}

template <class Real>
__device__ __forceinline__ void load_cons(const float *__restrict__ U, int c, Real *phi)
{
    const float *p = U + (size_t)c * 5;
#pragma unroll
    for (int j = 0; j < 5; ++j)
        phi[j] = Real(p[j]);
}

// Conservative state and its primitives: read from the primitive rows
// [N][5] (u, v, w, p, T) when they are kept beside the state (the values
// con_to_prim produced from the same state, so bit-identical), else computed.
template <class Real, class UT>
__device__ __forceinline__ void load_phi(const PhysDev &P, const UT *__restrict__ U,
                                         const double *__restrict__ prims, int c, Real *phi)
{
    load_cons(U, c, phi);
    if (prims)
    {
        const double *q = prims + (size_t)c * 5;
#pragma unroll
        for (int j = 0; j < 5; ++j)
            phi[5 + j] = Real(q[j]);
    }
    else
        con_to_prim(P, phi);
}

// Workgroup -> cell tile mapping.  The hardware hands consecutive
// workgroups to the 8 XCDs in turn; this gives each XCD one contiguous range
// of cell tiles so that the neighbour data a tile gathers is mostly in that
// XCD's own L2.  A bijection on [0, nblk) for any nblk.
__device__ __forceinline__ int xcd_tile(int nblk)
{
    const int b = blockIdx.x;
    const int x = b & 7, i = b >> 3;
    const int q = nblk >> 3, r = nblk & 7;
    return x * q + (x < r ? x : r) + i;
}

// Viscous gradient subset: momentum (vars 1,2,3) and T (9) per direction,
// v[d*4 + q] = d/dx_d of var {1,2,3,9}[q].
template <class Real> struct ViscGradT
{
    Real v[12];
};
using ViscGrad = ViscGradT<double>;

__device__ __forceinline__ int visc_var(int q) { return q < 3 ? q + 1 : 9; }

template <class Real>
__device__ __forceinline__ void load_vg(const double *__restrict__ vg, int c, ViscGradT<Real> &g)
{
    const double2 *p = reinterpret_cast<const double2 *>(vg + (size_t)c * 12);
#pragma unroll
    for (int q = 0; q < 6; ++q)
    {
        const double2 v = p[q];
        g.v[2 * q] = Real(v.x);
        g.v[2 * q + 1] = Real(v.y);
    }
}

// Limited linear extrapolation to a face (reconstruction.c:193-200 with
// limiter.c:86-105: min(1, y) of the per-face-side Barth-Jespersen ratio).
template <class Real>
__device__ __forceinline__ Real limited(Real phi, Real gx, Real gy, Real gz, const Real *r,
                                        Real mn, Real mx, int bj)
{
    const Real slope = r[0] * gx + r[1] * gy + r[2] * gz;
    Real lim = Real(1.0);
    if (bj)
    {
        // y >= 1 exactly when the numerator reaches the slope (division is
        // monotonic), and then min(1, y) = 1: the divide is only needed
        // below that.  The numerator is chosen by the slope's sign first,
        // so that a wave whose lanes differ in sign runs one divide, not two
        const bool pos = slope > Real(0);
        const Real num = pos ? mx - phi : mn - phi;
        const bool below = pos ? num < slope : (slope < Real(0) && num > slope);
        Real y = Real(1.0);
        if (below)
            y = num / slope;
        lim = rmin(Real(1.0), y);
    }
    return phi + lim * slope;
}

// ---------------------------------------------------------------------------
// Fluxes (flux.c)
// ---------------------------------------------------------------------------

// flux.c:203-210
template <class Real>
__device__ __forceinline__ void euler_flux_1d(const Real *phi, Real *f)
{
    f[0] = phi[1];
    f[1] = phi[1] * phi[5] + phi[8];
    f[2] = phi[1] * phi[6];
    f[3] = phi[1] * phi[7];
    f[4] = (phi[4] + phi[8]) * phi[5];
}

// flux.c:320-336
template <class Real>
__device__ __forceinline__ void riemann_rusanov(const PhysDev &P, const Real *l, const Real *r,
                                                Real *f)
{
    const Real c_l = sqrt(Real(P.kappa) * l[8] / l[0]);
    const Real c_r = sqrt(Real(P.kappa) * r[8] / r[0]);
    const Real eig = rmax(rabs(l[5]) + c_l, rabs(r[5]) + c_r);
    Real fl[5], fr[5];
    euler_flux_1d(l, fl);
    euler_flux_1d(r, fr);
#pragma unroll
    for (int j = 0; j < 5; ++j)
        f[j] = Real(0.5) * (fl[j] + fr[j]) - Real(0.5) * eig * (r[j] - l[j]);
}

// flux.c:344-406
template <class Real>
__device__ __forceinline__ void riemann_ausm(const PhysDev &P, const Real *l, const Real *r,
                                             Real *f)
{
    const Real c_l = sqrt(Real(P.kappa) * l[8] / l[0]);
    const Real c_r = sqrt(Real(P.kappa) * r[8] / r[0]);
    const Real M_l = l[5] / c_l;
    const Real M_r = r[5] / c_r;
    const Real H_l = (l[4] + l[8]) / l[0];
    const Real H_r = (r[4] + r[8]) / r[0];
    Real M_p, P_p, M_m, P_m;
    if (M_l <= Real(-1.0))
    {
        M_p = Real(0.0);
        P_p = Real(0.0);
    }
    else if (M_l < Real(1.0))
    {
        M_p = Real(0.25) * (M_l + Real(1.0)) * (M_l + Real(1.0));
        P_p = Real(0.25) * l[8] * (Real(1.0) + M_l) * (Real(1.0) + M_l) * (Real(2.0) - M_l);
    }
    else
    {
        M_p = M_l;
        P_p = l[8];
    }
    if (M_r <= Real(-1.0))
    {
        M_m = M_r;
        P_m = r[8];
    }
    else if (M_r < Real(1.0))
    {
        M_m = Real(-0.25) * (M_r - Real(1.0)) * (M_r - Real(1.0));
        P_m = Real(0.25) * r[8] * (Real(1.0) - M_r) * (Real(1.0) - M_r) * (Real(2.0) + M_r);
    }
    else
    {
        M_m = Real(0.0);
        P_m = Real(0.0);
    }
    const Real mp = rmax(Real(0.0), M_p + M_m);
    const Real mm = rmin(Real(0.0), M_p + M_m);
    f[0] = mp * c_l * l[0];
    f[1] = mp * c_l * l[0] * l[5] + P_p;
    f[2] = mp * c_l * l[0] * l[6];
    f[3] = mp * c_l * l[0] * l[7];
    f[4] = mp * c_l * l[0] * H_l;
    f[0] += mm * c_r * r[0];
    f[1] += mm * c_r * r[0] * r[5] + P_m;
    f[2] += mm * c_r * r[0] * r[6];
    f[3] += mm * c_r * r[0] * r[7];
    f[4] += mm * c_r * r[0] * H_r;
}

// flux.c:215-263 with the reference's use of momentum gradients as velocity
// gradients; phi holds the face state (reconstructed primitives at 5..7).
template <class Real>
__device__ __forceinline__ void viscous_1d(const PhysDev &P, const Real *phi,
                                           const ViscGradT<Real> &g, Real *f, Real *gg, Real *h)
{
    const Real s_23 = Real(2.0) / Real(3.0);
    const Real s_43 = Real(4.0) / Real(3.0);
    const Real mu = Real(P.mu);
    const Real lam = Real(P.lam);
    const Real gx1 = g.v[0], gx2 = g.v[1], gx3 = g.v[2], gxT = g.v[3];
    const Real gy1 = g.v[4], gy2 = g.v[5], gy3 = g.v[6], gyT = g.v[7];
    const Real gz1 = g.v[8], gz2 = g.v[9], gz3 = g.v[10], gzT = g.v[11];
    const Real tau_xx = mu * (s_43 * gx1 - s_23 * gy2 - s_23 * gz3);
    const Real tau_yy = mu * (-s_23 * gx1 + s_43 * gy2 - s_23 * gz3);
    const Real tau_zz = mu * (-s_23 * gx1 - s_23 * gy2 + s_43 * gz3);
    const Real tau_xy = mu * (gy1 + gx2);
    const Real tau_xz = mu * (gz1 + gx3);
    const Real tau_yz = mu * (gz2 + gy3);
    f[0] = Real(0.0);
    f[1] = -tau_xx;
    f[2] = -tau_xy;
    f[3] = -tau_xz;
    f[4] = -tau_xx * phi[5] - tau_xy * phi[6] - tau_xz * phi[7] - lam * gxT;
    gg[0] = Real(0.0);
    gg[1] = -tau_xy;
    gg[2] = -tau_yy;
    gg[3] = -tau_yz;
    gg[4] = -tau_xy * phi[5] - tau_yy * phi[6] - tau_yz * phi[7] - lam * gyT;
    h[0] = Real(0.0);
    h[1] = -tau_xz;
    h[2] = -tau_yz;
    h[3] = -tau_zz;
    h[4] = -tau_xz * phi[5] - tau_yz * phi[6] - tau_zz * phi[7] - lam * gzT;
}

// One face flux in the face's own orientation (flux.c:42-111/113-197).
template <int FLUX, bool VISC, class Real>
__device__ __forceinline__ void face_flux(const PhysDev &P, const FrameT<Real> &fr,
                                          const Real *L, const Real *R,
                                          const ViscGradT<Real> &gL, const ViscGradT<Real> &gR,
                                          Real *out)
{
    Real lr[10], rr[10];
    lr[0] = L[0];
    lr[1] = dot3(L[1], L[2], L[3], fr.n[0], fr.n[1], fr.n[2]);
    lr[2] = dot3(L[1], L[2], L[3], fr.t1[0], fr.t1[1], fr.t1[2]);
    lr[3] = dot3(L[1], L[2], L[3], fr.t2[0], fr.t2[1], fr.t2[2]);
    lr[4] = L[4];
    con_to_prim(P, lr);
    rr[0] = R[0];
    rr[1] = dot3(R[1], R[2], R[3], fr.n[0], fr.n[1], fr.n[2]);
    rr[2] = dot3(R[1], R[2], R[3], fr.t1[0], fr.t1[1], fr.t1[2]);
    rr[3] = dot3(R[1], R[2], R[3], fr.t2[0], fr.t2[1], fr.t2[2]);
    rr[4] = R[4];
    con_to_prim(P, rr);

    Real fc[5];
    if (FLUX == FluxRusanov)
        riemann_rusanov(P, lr, rr, fc);
    else
        riemann_ausm(P, lr, rr, fc);

    out[0] = fc[0];
    out[1] = fc[1] * fr.n[0] + fc[2] * fr.t1[0] + fc[3] * fr.t2[0];
    out[2] = fc[1] * fr.n[1] + fc[2] * fr.t1[1] + fc[3] * fr.t2[1];
    out[3] = fc[1] * fr.n[2] + fc[2] * fr.t1[2] + fc[3] * fr.t2[2];
    out[4] = fc[4];

    if (VISC)
    {
        Real fl[5], gl[5], hl[5], fr5[5], gr[5], hr[5];
        viscous_1d(P, L, gL, fl, gl, hl);
        viscous_1d(P, R, gR, fr5, gr, hr);
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const Real fx = Real(0.5) * (fl[j] + fr5[j]);
            const Real fy = Real(0.5) * (gl[j] + gr[j]);
            const Real fz = Real(0.5) * (hl[j] + hr[j]);
            out[j] += fx * fr.n[0] + fy * fr.n[1] + fz * fr.n[2];
        }
    }
}

// Inverse of a 5x5 row-major matrix (Gauss-Jordan, partial pivoting; row
// swaps with compile-time indices so everything stays in registers).
__device__ __forceinline__ void invert5(const double *A, double *Ainv)
{
    double M[5][10];
#pragma unroll
    for (int r = 0; r < 5; ++r)
    {
#pragma unroll
        for (int q = 0; q < 5; ++q)
        {
            M[r][q] = A[r * 5 + q];
            M[r][5 + q] = r == q ? 1.0 : 0.0;
        }
    }
#pragma unroll
    for (int col = 0; col < 5; ++col)
    {
        int p = col;
        double best = fabs(M[col][col]);
#pragma unroll
        for (int r = col + 1; r < 5; ++r)
            if (fabs(M[r][col]) > best)
            {
                best = fabs(M[r][col]);
                p = r;
            }
        // row swap with compile-time indices only (no private-memory arrays)
#pragma unroll
        for (int r = col + 1; r < 5; ++r)
            if (r == p)
#pragma unroll
                for (int q = 0; q < 10; ++q)
                {
                    const double t = M[col][q];
                    M[col][q] = M[r][q];
                    M[r][q] = t;
                }
        const double inv = 1.0 / M[col][col];
#pragma unroll
        for (int q = 0; q < 10; ++q)
            M[col][q] *= inv;
#pragma unroll
        for (int r = 0; r < 5; ++r)
        {
            if (r == col)
                continue;
            const double fct = M[r][col];
#pragma unroll
            for (int q = 0; q < 10; ++q)
                M[r][q] -= fct * M[col][q];
        }
    }
#pragma unroll
    for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int q = 0; q < 5; ++q)
            Ainv[r * 5 + q] = M[r][5 + q];
}

// Block reductions: a butterfly over the 64 lanes of each wavefront (lane
// shuffles), then the waves' results combined in wave order by every thread
// from LDS.  One barrier per reduction (plus one before sh is reused); the
// order is fixed, so results are bitwise reproducible.  blockDim.x must be a
// multiple of 64; sh holds at least blockDim.x / 64 doubles.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
        v = v + __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
        v = rmin(v, __shfl_xor(v, m, 64));
    return v;
}

__device__ __forceinline__ void block_sum(double *sh, double &v)
{
    const double w = wave_sum(v);
    const int nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = w;
    __syncthreads();
    double a = sh[0];
    for (int i = 1; i < nw; ++i)
        a = a + sh[i];
    v = a;
    __syncthreads();
}

__device__ __forceinline__ void block_min(double *sh, double &v)
{
    const double w = wave_min(v);
    const int nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = w;
    __syncthreads();
    double a = sh[0];
    for (int i = 1; i < nw; ++i)
        a = rmin(a, sh[i]);
    v = a;
    __syncthreads();
}

// The 7 partials of the time loop (sums of v[0..4], minima of v[5..6]) over
// the block, with a single barrier; the result is valid in thread 0 only.
// sh holds at least 8 * blockDim.x / 64 doubles.
__device__ __forceinline__ void block_reduce7(double *sh, double *v)
{
    double w[7];
#pragma unroll
    for (int q = 0; q < 5; ++q)
        w[q] = wave_sum(v[q]);
    w[5] = wave_min(v[5]);
    w[6] = wave_min(v[6]);
    const int nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 7; ++q)
            sh[(threadIdx.x >> 6) * 8 + q] = w[q];
    __syncthreads();
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int q = 0; q < 7; ++q)
        {
            double a = sh[q];
            for (int i = 1; i < nw; ++i)
                a = q < 5 ? a + sh[i * 8 + q] : rmin(a, sh[i * 8 + q]);
            v[q] = a;
        }
    }
}

} // namespace dev
} // namespace fvh
