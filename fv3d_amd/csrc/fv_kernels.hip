// CDNA4 kernels of the residual and explicit time-stepping path.
//
// Per RK stage (or residual evaluation):
//   fv_ghost_kernel     ghost rows of the boundary faces (B threads)
//   fv_gradient_kernel  one thread per cell: Green-Gauss gradients and
//                       neighbourhood min/max, then the cell's limited state
//                       on each of its faces, written by face id and side
//   fv_residual_group_kernel
//                       LANES lanes per cell, one face per lane: the face
//                       flux from the two stored face states, summed in
//                       ascending face order, then the fused RK stage and,
//                       in the last stage, the residual-norm and time-step
//                       partials
//   fv_finalize_kernel  one block: the time loop's bookkeeping
//
// Every cell gathers over its own faces in ascending face id (the order of
// the reference's face-ordered scatter, src/Solver/src/finite_volume/
// finite_volume.c:100-112), so its sums equal the reference's exactly and
// no atomics are used: results are bitwise reproducible run to run.  Each
// face flux is evaluated by both of its cells (same inputs, same
// operations, same bits), which keeps the flux[F][5] array of the reference
// out of HBM.
//
// Arithmetic follows the reference operation by operation and is compiled
// with -ffp-contract=off, so IEEE-rounded +,-,*,/ and sqrt give the same bits
// as the reference's x86-64 build.
#include <cfloat>
#include <cmath>
#include <stdexcept>

#include "fv_device.h"

// Minimum resident waves per SIMD requested from the register allocator for
// the two stencil kernels (occupancy against spilling).
#ifndef FVH_K1_MINW
#define FVH_K1_MINW 1
#endif
#ifndef FVH_K2_MINW
#define FVH_K2_MINW 1
#endif
// Dynamic LDS reserved per gradient-kernel workgroup; a non-zero value caps
// the resident workgroups per CU (occupancy experiments).
#ifndef FVH_K1_LDS
#define FVH_K1_LDS 0
#endif

namespace fvh {
namespace dev {

// reference BM_MIN / BM_MAX / BM_ABS semantics (ternaries, not fmin/fmax)
__device__ __forceinline__ double rmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double rmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double rabs(double a) { return a < 0 ? -a : a; }

// basec dot_n: sequential sum from 0.0
__device__ __forceinline__ double dot3(double a0, double a1, double a2,
                                       double b0, double b1, double b2)
{
    double s = 0.0;
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
__device__ __forceinline__ void con_to_prim(const PhysDev &P, double *phi)
{
    phi[5] = phi[1] / phi[0];
    phi[6] = phi[2] / phi[0];
    phi[7] = phi[3] / phi[0];
    double p = P.kappa_m1 * (phi[4] - 0.5 * dot3(phi[1], phi[2], phi[3],
                                                 phi[5], phi[6], phi[7]));
    p = rmax(1.00E-10, p);
    phi[8] = p;
    phi[9] = p / (phi[0] * P.R);
}

// eos.c:120-127
__device__ __forceinline__ void prim_to_con(const PhysDev &P, double *phi)
{
    phi[1] = phi[5] * phi[0];
    phi[2] = phi[6] * phi[0];
    phi[3] = phi[7] * phi[0];
    phi[4] = P.s_kappa_m1 * phi[8] +
             0.5 * dot3(phi[5], phi[6], phi[7], phi[1], phi[2], phi[3]);
}

// eos.c:55-75 (compression branch with the reference's A = 2/((k+1) u_n))
__device__ __forceinline__ double riemann_p(const PhysDev &P, const double *phi)
{
    if (phi[5] <= 0.0)
    {
        double c = sqrt(P.kappa * phi[8] / phi[0]);
        return phi[8] * pow(rmax(1e-4, 1. + 0.5 * P.kappa_m1 * phi[5] / c),
                            2 * P.kappa * P.s_kappa_m1);
    }
    double ar = 2 * P.s_kappa_p1 / phi[5];
    double br = P.kappa_m1 * P.s_kappa_p1 * phi[8];
    return phi[8] + phi[5] / ar * 0.5 *
                        (phi[5] + sqrt(phi[5] * phi[5] + 4. * ar * (phi[8] + br)));
}

struct Frame
{
    double n[3], t1[3], t2[3];
    double area;
};

__device__ __forceinline__ void load_frame(const DevMesh &m, int f, Frame &fr)
{
    const double2 *g = reinterpret_cast<const double2 *>(m.fgA + (size_t)f * kFgA);
    const double2 a = g[0], b = g[1], c = g[2], d = g[3], e = g[4];
    fr.n[0] = a.x, fr.n[1] = a.y, fr.n[2] = b.x;
    fr.t1[0] = b.y, fr.t1[1] = c.x, fr.t1[2] = c.y;
    fr.t2[0] = d.x, fr.t2[1] = d.y, fr.t2[2] = e.x;
    fr.area = e.y;
}

__device__ __forceinline__ void rotate_back(const Frame &fr, double *g)
{
    double tu = g[5] * fr.n[0] + g[6] * fr.t1[0] + g[7] * fr.t2[0];
    double tv = g[5] * fr.n[1] + g[6] * fr.t1[1] + g[7] * fr.t2[1];
    double tw = g[5] * fr.n[2] + g[6] * fr.t1[2] + g[7] * fr.t2[2];
    g[5] = tu;
    g[6] = tv;
    g[7] = tw;
}

// Ghost state of boundary row N+b from the interior state pc
// (boundary.c:202-319).
__device__ __forceinline__ void ghost_state(const DevMesh &m, const PhysDev &P,
                                            int b, const Frame &fr,
                                            const double *pc, double *g)
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
            g[j] = m.region_phi[region * 10 + j];
    }
    else if (type == BC_WALL_ADIABATIC || type == BC_WALL_ISOTHERMAL)
    {
        g[5] = dot3(pc[5], pc[6], pc[7], fr.n[0], fr.n[1], fr.n[2]);
        g[6] = dot3(pc[5], pc[6], pc[7], fr.t1[0], fr.t1[1], fr.t1[2]);
        g[7] = dot3(pc[5], pc[6], pc[7], fr.t2[0], fr.t2[1], fr.t2[2]);
        g[8] = riemann_p(P, g);
        g[5] = 0.0;
        g[6] = 0.0;
        g[7] = 0.0;
        if (type == BC_WALL_ISOTHERMAL)
            g[9] = m.region_phi[region * 10 + 9];
        g[0] = g[8] / (P.R * g[9]);
        rotate_back(fr, g);
    }
    else if (type == BC_WALL_SLIP || type == BC_SYMMETRY)
    {
        g[5] = dot3(pc[5], pc[6], pc[7], fr.n[0], fr.n[1], fr.n[2]);
        g[6] = dot3(pc[5], pc[6], pc[7], fr.t1[0], fr.t1[1], fr.t1[2]);
        g[7] = dot3(pc[5], pc[6], pc[7], fr.t2[0], fr.t2[1], fr.t2[2]);
        g[8] = riemann_p(P, g);
        g[5] = 0.0;
        g[9] = g[8] / (g[0] * P.R);
        rotate_back(fr, g);
    }
    prim_to_con(P, g);
}

// Ghost gradient (boundary.c:361-406): a copy of the interior gradient, with
// the normal component removed for slip walls and symmetry.
__device__ __forceinline__ void ghost_grad(int type, const Frame &fr, double gx,
                                           double gy, double gz, double &ox,
                                           double &oy, double &oz)
{
    if (type == BC_WALL_SLIP || type == BC_SYMMETRY)
    {
        double ty = gx * fr.t1[0] + gy * fr.t1[1] + gz * fr.t1[2];
        double tz = gx * fr.t2[0] + gy * fr.t2[1] + gz * fr.t2[2];
        double tx = 0.0;
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

__device__ __forceinline__ void load_cons(const double *__restrict__ U, int c,
                                          double *phi)
{
    const double *p = U + (size_t)c * 5;
#pragma unroll
    for (int j = 0; j < 5; ++j)
        phi[j] = p[j];
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
struct ViscGrad
{
    double v[12];
};

__device__ __forceinline__ int visc_var(int q) { return q < 3 ? q + 1 : 9; }

__device__ __forceinline__ void load_vg(const double *__restrict__ vg, int c,
                                        ViscGrad &g)
{
    const double2 *p = reinterpret_cast<const double2 *>(vg + (size_t)c * 12);
#pragma unroll
    for (int q = 0; q < 6; ++q)
    {
        const double2 v = p[q];
        g.v[2 * q] = v.x;
        g.v[2 * q + 1] = v.y;
    }
}

// Limited linear extrapolation to a face (reconstruction.c:193-200 with
// limiter.c:86-105: min(1, y) of the per-face-side Barth-Jespersen ratio).
__device__ __forceinline__ double limited(double phi, double gx, double gy,
                                          double gz, const double *r, double mn,
                                          double mx, int bj)
{
    const double slope = r[0] * gx + r[1] * gy + r[2] * gz;
    double lim = 1.0;
    if (bj)
    {
        double y = 1.0;
        if (slope > 0)
            y = (mx - phi) / slope;
        else if (slope < 0)
            y = (mn - phi) / slope;
        lim = rmin(1.0, y);
    }
    return phi + lim * slope;
}

// ---------------------------------------------------------------------------
// Ghost rows N+b (boundary.c:188-320 and the update hook ns_update,
// navier_stokes.c:199-209): the boundary-condition state of every boundary
// face from its interior cell, primitives + conservatives [B][10]; with face
// states on, also side 1 of the boundary faces.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void fv_ghost_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U, double *__restrict__ gphi,
    double *__restrict__ fs, const DevScalars *sc)
{
    if (sc && sc->done)
        return;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B)
        return;
    const int f = m.bnd_face[b];
    const int c = m.fcell[2 * f];
    double pc[10], g[10];
    load_cons(U, c, pc);
    con_to_prim(P, pc);
    Frame fr;
    load_frame(m, f, fr);
    ghost_state(m, P, b, fr, pc, g);
#pragma unroll
    for (int j = 0; j < 10; ++j)
        gphi[(size_t)b * 10 + j] = g[j];
    if (fs)
#pragma unroll
        for (int j = 0; j < 8; ++j)
            fs[(size_t)f * 16 + 8 + j] = g[j];
}

// ---------------------------------------------------------------------------
// K1: Green-Gauss gradients of all 10 variables and neighbourhood min/max of
// one cell (reconstruction.c:20-93 in gather form, limiter.c:57-82), then
// the cell's reconstructed state on each of its faces
// (reconstruction.c:168-255 / 142-163).  Writes the face states by face id
// and side, and the gradient subset the viscous flux needs.
// ---------------------------------------------------------------------------
template <int MAXF>
__global__ __launch_bounds__(kBlock, FVH_K1_MINW) void fv_gradient_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U,
    const double *__restrict__ gphi, GradOut out, int linear,
    const DevScalars *sc)
{
    if (sc && sc->done)
        return;
    const int N = m.N;
    const int c = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    if (c >= N)
        return;

    double pc[10];
    load_cons(U, c, pc);
    con_to_prim(P, pc);

    double gx[10], gy[10], gz[10], mn[10], mx[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        gx[j] = 0.0;
        gy[j] = 0.0;
        gz[j] = 0.0;
        mn[j] = pc[j];
        mx[j] = pc[j];
    }

    const int nf = m.ncf[c];
    const int *cfc = m.cf + (size_t)c * m.maxf;
#pragma unroll
    for (int s = 0; s < MAXF; ++s)
    {
        if (s >= nf)
            break;
        const int f = cfc[s];
        const int2 fc = reinterpret_cast<const int2 *>(m.fcell)[f];
        const bool own0 = fc.x == c;
        const int nb = own0 ? fc.y : fc.x;
        const double2 *g = reinterpret_cast<const double2 *>(m.fgB + (size_t)f * kFgB);
        const double2 g0 = g[0], g1 = g[1], g2 = g[2];
        const double n0 = g0.x, n1 = g0.y, n2 = g1.x;
        const double area = g1.y, lam = g2.x;

        double pn[10];
        if (nb < N)
        {
            load_cons(U, nb, pn);
            con_to_prim(P, pn);
        }
        else
        {
            // boundary face: the ghost row written by fv_ghost_kernel
            const double *gp = gphi + (size_t)(nb - N) * 10;
#pragma unroll
            for (int j = 0; j < 10; ++j)
                pn[j] = gp[j];
        }

#pragma unroll
        for (int j = 0; j < 10; ++j)
        {
            const double p0 = own0 ? pc[j] : pn[j];
            const double p1 = own0 ? pn[j] : pc[j];
            const double pm = p0 + lam * (p1 - p0);
            const double ax = pm * n0 * area;
            const double ay = pm * n1 * area;
            const double az = pm * n2 * area;
            if (own0)
            {
                gx[j] += ax;
                gy[j] += ay;
                gz[j] += az;
            }
            else
            {
                gx[j] -= ax;
                gy[j] -= ay;
                gz[j] -= az;
            }
            mn[j] = rmin(mn[j], pn[j]);
            mx[j] = rmax(mx[j], pn[j]);
        }
    }

    const double sv = m.svol[c];
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        gx[j] *= sv;
        gy[j] *= sv;
        gz[j] *= sv;
    }

    if (out.vg)
    {
        double2 *d = reinterpret_cast<double2 *>(out.vg + (size_t)c * 12);
        d[0] = make_double2(gx[1], gx[2]);
        d[1] = make_double2(gx[3], gx[9]);
        d[2] = make_double2(gy[1], gy[2]);
        d[3] = make_double2(gy[3], gy[9]);
        d[4] = make_double2(gz[1], gz[2]);
        d[5] = make_double2(gz[3], gz[9]);
    }
    if (out.grad)
    {
        double *d = out.grad + (size_t)c * 30;
#pragma unroll
        for (int j = 0; j < 10; ++j)
        {
            d[j] = gx[j];
            d[10 + j] = gy[j];
            d[20 + j] = gz[j];
        }
    }
    if (out.mnmx)
    {
        double *d = out.mnmx + (size_t)c * 20;
#pragma unroll
        for (int j = 0; j < 10; ++j)
        {
            d[j] = mn[j];
            d[10 + j] = mx[j];
        }
    }

    if (!out.fs)
        return;
    // this cell's reconstructed state on each of its faces; the distance
    // vectors of mesh.c:44-56 (dist_cell_1 = x_f - x_c0, dist_cell_2 =
    // x_f - x_c1 through the periodic image) from the face and cell centres
    const double xc0 = m.cx[(size_t)c * 3], xc1 = m.cx[(size_t)c * 3 + 1],
                 xc2 = m.cx[(size_t)c * 3 + 2];
#pragma unroll
    for (int s = 0; s < MAXF; ++s)
    {
        if (s >= nf)
            break;
        const int f = cfc[s];
        const bool own0 = m.fcell[2 * f] == c;
        double st[8];
        if (linear)
        {
            const double2 *g = reinterpret_cast<const double2 *>(m.fgB + (size_t)f * kFgB);
            const double2 g2 = g[2], g3 = g[3];
            double r[3];
            if (own0)
            {
                r[0] = g2.y - xc0;
                r[1] = g3.x - xc1;
                r[2] = g3.y - xc2;
            }
            else
            {
                const double *sh = m.shtab + (size_t)m.fsh[f] * 3;
                r[0] = g2.y - (xc0 + sh[0]);
                r[1] = g3.x - (xc1 + sh[1]);
                r[2] = g3.y - (xc2 + sh[2]);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
                st[j] = limited(pc[j], gx[j], gy[j], gz[j], r, mn[j], mx[j], P.bj);
        }
        else
        {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                st[j] = pc[j];
        }
        double2 *d = reinterpret_cast<double2 *>(out.fs + (size_t)f * 16 + (own0 ? 0 : 8));
        d[0] = make_double2(st[0], st[1]);
        d[1] = make_double2(st[2], st[3]);
        d[2] = make_double2(st[4], st[5]);
        d[3] = make_double2(st[6], st[7]);
    }
}

// ---------------------------------------------------------------------------
// Fluxes (flux.c)
// ---------------------------------------------------------------------------

// flux.c:203-210
__device__ __forceinline__ void euler_flux_1d(const double *phi, double *f)
{
    f[0] = phi[1];
    f[1] = phi[1] * phi[5] + phi[8];
    f[2] = phi[1] * phi[6];
    f[3] = phi[1] * phi[7];
    f[4] = (phi[4] + phi[8]) * phi[5];
}

// flux.c:320-336
__device__ __forceinline__ void riemann_rusanov(const PhysDev &P, const double *l,
                                                const double *r, double *f)
{
    const double c_l = sqrt(P.kappa * l[8] / l[0]);
    const double c_r = sqrt(P.kappa * r[8] / r[0]);
    const double eig = rmax(rabs(l[5]) + c_l, rabs(r[5]) + c_r);
    double fl[5], fr[5];
    euler_flux_1d(l, fl);
    euler_flux_1d(r, fr);
#pragma unroll
    for (int j = 0; j < 5; ++j)
        f[j] = 0.5 * (fl[j] + fr[j]) - 0.5 * eig * (r[j] - l[j]);
}

// flux.c:344-406
__device__ __forceinline__ void riemann_ausm(const PhysDev &P, const double *l,
                                             const double *r, double *f)
{
    const double c_l = sqrt(P.kappa * l[8] / l[0]);
    const double c_r = sqrt(P.kappa * r[8] / r[0]);
    const double M_l = l[5] / c_l;
    const double M_r = r[5] / c_r;
    const double H_l = (l[4] + l[8]) / l[0];
    const double H_r = (r[4] + r[8]) / r[0];
    double M_p, P_p, M_m, P_m;
    if (M_l <= -1.0)
    {
        M_p = 0.0;
        P_p = 0.0;
    }
    else if (M_l < 1.0)
    {
        M_p = 0.25 * (M_l + 1.0) * (M_l + 1.0);
        P_p = 0.25 * l[8] * (1.0 + M_l) * (1.0 + M_l) * (2.0 - M_l);
    }
    else
    {
        M_p = M_l;
        P_p = l[8];
    }
    if (M_r <= -1.0)
    {
        M_m = M_r;
        P_m = r[8];
    }
    else if (M_r < 1.0)
    {
        M_m = -0.25 * (M_r - 1.0) * (M_r - 1.0);
        P_m = 0.25 * r[8] * (1.0 - M_r) * (1.0 - M_r) * (2.0 + M_r);
    }
    else
    {
        M_m = 0.0;
        P_m = 0.0;
    }
    const double mp = rmax(0.0, M_p + M_m);
    const double mm = rmin(0.0, M_p + M_m);
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
__device__ __forceinline__ void viscous_1d(const PhysDev &P, const double *phi,
                                           const ViscGrad &g, double *f,
                                           double *gg, double *h)
{
    const double s_23 = 2.0 / 3.0;
    const double s_43 = 4.0 / 3.0;
    const double mu = P.mu;
    const double gx1 = g.v[0], gx2 = g.v[1], gx3 = g.v[2], gxT = g.v[3];
    const double gy1 = g.v[4], gy2 = g.v[5], gy3 = g.v[6], gyT = g.v[7];
    const double gz1 = g.v[8], gz2 = g.v[9], gz3 = g.v[10], gzT = g.v[11];
    const double tau_xx = mu * (s_43 * gx1 - s_23 * gy2 - s_23 * gz3);
    const double tau_yy = mu * (-s_23 * gx1 + s_43 * gy2 - s_23 * gz3);
    const double tau_zz = mu * (-s_23 * gx1 - s_23 * gy2 + s_43 * gz3);
    const double tau_xy = mu * (gy1 + gx2);
    const double tau_xz = mu * (gz1 + gx3);
    const double tau_yz = mu * (gz2 + gy3);
    f[0] = 0.0;
    f[1] = -tau_xx;
    f[2] = -tau_xy;
    f[3] = -tau_xz;
    f[4] = -tau_xx * phi[5] - tau_xy * phi[6] - tau_xz * phi[7] - P.lam * gxT;
    gg[0] = 0.0;
    gg[1] = -tau_xy;
    gg[2] = -tau_yy;
    gg[3] = -tau_yz;
    gg[4] = -tau_xy * phi[5] - tau_yy * phi[6] - tau_yz * phi[7] - P.lam * gyT;
    h[0] = 0.0;
    h[1] = -tau_xz;
    h[2] = -tau_yz;
    h[3] = -tau_zz;
    h[4] = -tau_xz * phi[5] - tau_yz * phi[6] - tau_zz * phi[7] - P.lam * gzT;
}

// One face flux in the face's own orientation (flux.c:42-111/113-197).
template <int FLUX, bool VISC>
__device__ __forceinline__ void face_flux(const PhysDev &P, const Frame &fr,
                                          const double *L, const double *R,
                                          const ViscGrad &gL, const ViscGrad &gR,
                                          double *out)
{
    double lr[10], rr[10];
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

    double fc[5];
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
        double fl[5], gl[5], hl[5], fr5[5], gr[5], hr[5];
        viscous_1d(P, L, gL, fl, gl, hl);
        viscous_1d(P, R, gR, fr5, gr, hr);
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const double fx = 0.5 * (fl[j] + fr5[j]);
            const double fy = 0.5 * (gl[j] + gr[j]);
            const double fz = 0.5 * (hl[j] + hr[j]);
            out[j] += fx * fr.n[0] + fy * fr.n[1] + fz * fr.n[2];
        }
    }
}

__device__ __forceinline__ void block_sum(double *sh, double &v)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = kBlock / 2; s > 0; s >>= 1)
    {
        if (t < s)
            sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    v = sh[0];
    __syncthreads();
}

__device__ __forceinline__ void block_min(double *sh, double &v)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = kBlock / 2; s > 0; s >>= 1)
    {
        if (t < s)
            sh[t] = rmin(sh[t], sh[t + s]);
        __syncthreads();
    }
    v = sh[0];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// K2: residual dU/dt = -(sum_f +-flux_f A_f)/V (finite_volume.c:96-123) in
// lane groups: LANES consecutive lanes per cell (4 for tetrahedra, 8 for
// hexahedra), one face per lane.  The group's contributions are summed in
// ascending face order with lane shuffles (acc - x is acc + (-x) exactly,
// so the sum is bit-identical to the face-ordered scatter), and the group's
// first lane runs the cell epilogue: the 2N-storage RK stage
// (explicit.c:113-140) and, in the last stage, the residual-norm and
// time-step partial reductions (analyze.c:46-55, navier_stokes.c:171-187).
// FS: face states from K1 (reconstructed path); otherwise first order, the
// face states are the cell values of U.
// ---------------------------------------------------------------------------
template <int FLUX, bool FS, bool VISC, int LANES>
__global__ __launch_bounds__(kBlock, FVH_K2_MINW) void fv_residual_group_kernel(
    DevMesh m, PhysDev P, StageParams sp)
{
    __shared__ double sh[kBlock];
    const int N = m.N;
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int c = tid / LANES;
    const int s = tid & (LANES - 1);
    const bool cell_ok = c < N;

    if (sp.sc && sp.sc->done)
    {
        // the run has stopped: carry the state through this stage unchanged
        if (sp.do_update && cell_ok && s == 0)
#pragma unroll
            for (int j = 0; j < 5; ++j)
                sp.U_out[(size_t)c * 5 + j] = sp.U_in[(size_t)c * 5 + j];
        return;
    }

    const int nf = cell_ok ? m.ncf[c] : 0;
    double contrib[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (s < nf)
    {
        const int f = m.cf[(size_t)c * m.maxf + s];
        const int2 fc = reinterpret_cast<const int2 *>(m.fcell)[f];
        const bool own0 = fc.x == c;
        const int nb = own0 ? fc.y : fc.x;
        const bool ghost = nb >= N;
        Frame fr;
        load_frame(m, f, fr);

        double L[10], R[10];
        if (FS)
        {
            const double2 *p = reinterpret_cast<const double2 *>(sp.fs + (size_t)f * 16);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const double2 a = p[q], b = p[4 + q];
                L[2 * q] = a.x;
                L[2 * q + 1] = a.y;
                R[2 * q] = b.x;
                R[2 * q + 1] = b.y;
            }
            L[8] = L[9] = R[8] = R[9] = 0.0;
        }
        else
        {
            double pc[10], pn[10];
            load_cons(sp.U_in, c, pc);
            con_to_prim(P, pc);
            if (ghost)
            {
                const double *gp = sp.gphi + (size_t)(nb - N) * 10;
#pragma unroll
                for (int j = 0; j < 10; ++j)
                    pn[j] = gp[j];
            }
            else
            {
                load_cons(sp.U_in, nb, pn);
                con_to_prim(P, pn);
            }
#pragma unroll
            for (int j = 0; j < 10; ++j)
            {
                L[j] = own0 ? pc[j] : pn[j];
                R[j] = own0 ? pn[j] : pc[j];
            }
        }

        ViscGrad gL, gR;
        if (VISC)
        {
            ViscGrad go, gn;
            load_vg(sp.vg, c, go);
            if (ghost)
            {
                const int type = m.region_type[m.bnd_region[nb - N]];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    ghost_grad(type, fr, go.v[q], go.v[4 + q], go.v[8 + q],
                               gn.v[q], gn.v[4 + q], gn.v[8 + q]);
            }
            else
            {
                load_vg(sp.vg, nb, gn);
            }
#pragma unroll
            for (int q = 0; q < 12; ++q)
            {
                gL.v[q] = own0 ? go.v[q] : gn.v[q];
                gR.v[q] = own0 ? gn.v[q] : go.v[q];
            }
        }

        double flux[5];
        face_flux<FLUX, VISC>(P, fr, L, R, gL, gR, flux);
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const double v = flux[j] * fr.area;
            contrib[j] = own0 ? v : -v;
        }
    }

    // ordered sum over the group's faces
    double acc[5];
#pragma unroll
    for (int j = 0; j < 5; ++j)
    {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < LANES; ++k)
        {
            const double v = __shfl(contrib[j], k, LANES);
            if (k < nf)
                a = a + v;
        }
        acc[j] = a;
    }

    double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    if (cell_ok && s == 0)
    {
        const double vol = m.vol[c];
        double R5[5];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            R5[j] = -acc[j] / vol;
        if (sp.R_out)
        {
#pragma unroll
            for (int j = 0; j < 5; ++j)
                sp.R_out[(size_t)c * 5 + j] = R5[j];
        }
        if (sp.do_update)
        {
            double pc[10];
            load_cons(sp.U_in, c, pc);
            const double dt = sp.sc->dt;
            double unew[5];
#pragma unroll
            for (int j = 0; j < 5; ++j)
            {
                double kj = R5[j];
                if (!sp.stage0)
                    kj = R5[j] + sp.k[(size_t)c * 5 + j] * sp.rk_a;
                sp.k[(size_t)c * 5 + j] = kj;
                unew[j] = pc[j] + kj * dt * sp.rk_g;
                sp.U_out[(size_t)c * 5 + j] = unew[j];
            }
            if (sp.do_reduce)
            {
                // time step with the new density and the primitives of this
                // stage's input (navier_stokes.c:171-187)
                con_to_prim(P, pc);
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    part[j] = rabs(R5[j]) * vol;
                const double dx0 = m.cdx[(size_t)c * 3], dx1 = m.cdx[(size_t)c * 3 + 1],
                             dx2 = m.cdx[(size_t)c * 3 + 2];
                const double s_rho = 1.0 / unew[0];
                const double cs = sqrt(P.kappa * pc[8] * s_rho);
                const double ds1 = (rabs(pc[5]) + cs) * dx0 +
                                   (rabs(pc[6]) + cs) * dx1 +
                                   (rabs(pc[7]) + cs) * dx2;
                part[5] = vol / ds1;
                const double ds2 = dot3(dx0, dx1, dx2, dx0, dx1, dx2);
                part[6] = s_rho * ds2 / vol;
            }
        }
    }

    if (sp.do_reduce)
    {
#pragma unroll
        for (int q = 0; q < 5; ++q)
            block_sum(sh, part[q]);
        block_min(sh, part[5]);
        block_min(sh, part[6]);
        if (threadIdx.x == 0)
        {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                sp.partials[blockIdx.x * kPartialStride + q] = part[q];
        }
    }
}

// Time-step partials of a state whose primitives are current (the first
// step: the primitives were refreshed from the initial state).
__global__ __launch_bounds__(kBlock) void fv_dt_partials_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U,
    double *__restrict__ partials)
{
    __shared__ double sh[kBlock];
    const int N = m.N;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double lc = DBL_MAX, lv = DBL_MAX;
    if (c < N)
    {
        double pc[10];
        load_cons(U, c, pc);
        con_to_prim(P, pc);
        const double dx0 = m.cdx[(size_t)c * 3], dx1 = m.cdx[(size_t)c * 3 + 1],
                     dx2 = m.cdx[(size_t)c * 3 + 2];
        const double vol = m.vol[c];
        const double s_rho = 1.0 / pc[0];
        const double cs = sqrt(P.kappa * pc[8] * s_rho);
        const double ds1 = (rabs(pc[5]) + cs) * dx0 + (rabs(pc[6]) + cs) * dx1 +
                           (rabs(pc[7]) + cs) * dx2;
        lc = vol / ds1;
        const double ds2 = dot3(dx0, dx1, dx2, dx0, dx1, dx2);
        lv = s_rho * ds2 / vol;
    }
    block_min(sh, lc);
    block_min(sh, lv);
    if (threadIdx.x == 0)
    {
        partials[blockIdx.x * kPartialStride + 5] = lc;
        partials[blockIdx.x * kPartialStride + 6] = lv;
    }
}

// First level of the reduction of the per-workgroup partials: each block
// folds kFoldRows consecutive rows (sums of 0..4, minima of 5..6) in a fixed
// order into one row of `fold`.
constexpr int kFoldRows = 2048;

__global__ __launch_bounds__(kBlock) void fv_fold_partials_kernel(
    const double *__restrict__ partials, int n_partials, int mode,
    double *__restrict__ fold)
{
    __shared__ double sh[kBlock];
    const int t = threadIdx.x;
    const int lo = blockIdx.x * kFoldRows;
    const int hi = min(lo + kFoldRows, n_partials);
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    for (int i = lo + t; i < hi; i += kBlock)
    {
        const double *p = &partials[(size_t)i * kPartialStride];
        if (mode == 1)
#pragma unroll
            for (int q = 0; q < 5; ++q)
                v[q] = v[q] + p[q];
        v[5] = rmin(v[5], p[5]);
        v[6] = rmin(v[6], p[6]);
    }
    if (mode == 1)
#pragma unroll
        for (int q = 0; q < 5; ++q)
            block_sum(sh, v[q]);
    block_min(sh, v[5]);
    block_min(sh, v[6]);
    if (t == 0)
#pragma unroll
        for (int q = 0; q < 7; ++q)
            fold[(size_t)blockIdx.x * kPartialStride + q] = v[q];
}

// Time-loop bookkeeping of timedisc.c:164-207 on the device: residual norm
// (analyze.c:57-58), NaN/Inf stop, steady and iteration stops, and the next
// time step with the t_end clamp (navier_stokes.c:189-192,
// timedisc.c:167-172).  One block; fixed-order reduction.
__global__ __launch_bounds__(kBlock) void fv_finalize_kernel(
    DevScalars *sc, const double *__restrict__ partials, int n_partials,
    PhysDev P, int mode, double *log, int log_cap)
{
    __shared__ double sh[kBlock];
    const int t = threadIdx.x;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    for (int i = t; i < n_partials; i += kBlock)
    {
        const double *p = &partials[i * kPartialStride];
        if (mode == 1)
#pragma unroll
            for (int q = 0; q < 5; ++q)
                v[q] = v[q] + p[q];
        v[5] = rmin(v[5], p[5]);
        v[6] = rmin(v[6], p[6]);
    }
    if (mode == 1)
#pragma unroll
        for (int q = 0; q < 5; ++q)
            block_sum(sh, v[q]);
    block_min(sh, v[5]);
    block_min(sh, v[6]);
    if (t != 0)
        return;
    if (sc->done)
        return;

    if (mode == 1)
    {
        const double dt = sc->dt;
        double mxres = 0.0;
        bool bad = false;
        for (int q = 0; q < 5; ++q)
        {
            const double r = v[q] * (dt / sc->vglob);
            sc->res[q] = r;
            bad = bad || isnan(r) || isinf(r);
            mxres = q == 0 ? r : rmax(mxres, r);
        }
        if (bad)
        {
            sc->nan_flag = 1;
            sc->done = -1;
            return;
        }
        sc->t = sc->t + dt;
        sc->iter = sc->iter + 1;
        const int row = sc->iter - sc->log_base - 1;
        if (log && row >= 0 && row < log_cap)
        {
            double *L = &log[(size_t)row * 8];
            L[0] = sc->t;
            L[1] = dt;
            L[2] = sc->viscous;
            for (int q = 0; q < 5; ++q)
                L[3 + q] = sc->res[q];
        }
        int done = 0;
        if (!sc->transient && mxres < sc->abort_residual)
            done = 1;
        if (sc->iter >= sc->max_iter_abs)
            done = 1;
        if (sc->final_pending)
            done = 1;
        if (done)
        {
            sc->done = 1;
            return;
        }
    }

    double lc = v[5] * P.cfl;
    double lv = v[6];
    lv *= P.dfl * P.mu * P.kappa_pr * lv;
    double dtn;
    if (P.visc_dt_mode == 1 && P.mu == 0.0)
    {
        sc->viscous = 0;
        dtn = lc;
    }
    else
    {
        sc->viscous = lv < lc;
        dtn = rmin(lc, lv);
    }
    if (sc->t + dtn > sc->t_end)
    {
        dtn = sc->t_end - sc->t;
        sc->final_pending = 1;
    }
    sc->dt = dtn;
}

// Host-mirror packing: AoS rows in the reference's phi_total/phi_dt layout
// (finite_volume.c:39-52), domain rows refreshed like ns_update and ghost
// rows like update_boundaries.
__global__ __launch_bounds__(kBlock) void fv_mirror_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U,
    const double *__restrict__ R, double *__restrict__ phi_aos,
    double *__restrict__ dt_aos)
{
    const int N = m.N;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= N + m.B)
        return;
    double pc[10];
    if (row < N)
    {
        load_cons(U, row, pc);
        con_to_prim(P, pc);
        if (R && dt_aos)
#pragma unroll
            for (int j = 0; j < 5; ++j)
                dt_aos[(size_t)row * 5 + j] = R[(size_t)row * 5 + j];
    }
    else
    {
        const int b = row - N;
        const int f = m.bnd_face[b];
        const int c = m.fcell[2 * f];
        double pi[10];
        load_cons(U, c, pi);
        con_to_prim(P, pi);
        Frame fr;
        load_frame(m, f, fr);
        ghost_state(m, P, b, fr, pi, pc);
    }
#pragma unroll
    for (int j = 0; j < 10; ++j)
        phi_aos[(size_t)row * 10 + j] = pc[j];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static void check_faces(const DevMesh &m)
{
    if (m.maxf > 6)
        throw std::runtime_error("cells with more than 6 faces are not supported");
}

int residual_grid(const DevMesh &m)
{
    return cell_blocks(m.N * (m.maxf <= 4 ? 4 : 8));
}

int dt_partials_grid(const DevMesh &m) { return cell_blocks(m.N); }

void launch_ghost(hipStream_t s, const DevMesh &m, const PhysDev &P,
                  const double *U, double *gphi, double *fs,
                  const DevScalars *sc)
{
    if (m.B == 0)
        return;
    hipLaunchKernelGGL(fv_ghost_kernel, dim3(cell_blocks(m.B)), dim3(kBlock), 0,
                       s, m, P, U, gphi, fs, sc);
}

void launch_gradient(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const double *U, const double *gphi, const GradOut &out,
                     int linear, const DevScalars *sc)
{
    check_faces(m);
    if (m.maxf <= 4)
        hipLaunchKernelGGL(fv_gradient_kernel<4>, dim3(cell_blocks(m.N)),
                           dim3(kBlock), FVH_K1_LDS, s, m, P, U, gphi, out, linear,
                           sc);
    else
        hipLaunchKernelGGL(fv_gradient_kernel<6>, dim3(cell_blocks(m.N)),
                           dim3(kBlock), FVH_K1_LDS, s, m, P, U, gphi, out, linear,
                           sc);
}

template <int FLUX, bool FS, bool VISC>
static void launch_res_t(hipStream_t s, const DevMesh &m, const PhysDev &P,
                         const StageParams &sp)
{
    if (m.maxf <= 4)
        hipLaunchKernelGGL((fv_residual_group_kernel<FLUX, FS, VISC, 4>),
                           dim3(residual_grid(m)), dim3(kBlock), 0, s, m, P, sp);
    else
        hipLaunchKernelGGL((fv_residual_group_kernel<FLUX, FS, VISC, 8>),
                           dim3(residual_grid(m)), dim3(kBlock), 0, s, m, P, sp);
}

void launch_residual(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const StageParams &sp, int flux, bool face_states,
                     bool viscous)
{
    check_faces(m);
    if (sp.do_reduce && residual_grid(m) > sp.partials_cap)
        throw std::runtime_error("reduction partials buffer too small");
    if (flux == FluxRusanov)
    {
        if (face_states)
            viscous ? launch_res_t<FluxRusanov, true, true>(s, m, P, sp)
                    : launch_res_t<FluxRusanov, true, false>(s, m, P, sp);
        else
            viscous ? launch_res_t<FluxRusanov, false, true>(s, m, P, sp)
                    : launch_res_t<FluxRusanov, false, false>(s, m, P, sp);
    }
    else
    {
        if (face_states)
            viscous ? launch_res_t<FluxAUSM, true, true>(s, m, P, sp)
                    : launch_res_t<FluxAUSM, true, false>(s, m, P, sp);
        else
            viscous ? launch_res_t<FluxAUSM, false, true>(s, m, P, sp)
                    : launch_res_t<FluxAUSM, false, false>(s, m, P, sp);
    }
}

void launch_dt_partials(hipStream_t s, const DevMesh &m, const PhysDev &P,
                        const double *U, double *partials, int partials_cap)
{
    if (dt_partials_grid(m) > partials_cap)
        throw std::runtime_error("reduction partials buffer too small");
    hipLaunchKernelGGL(fv_dt_partials_kernel, dim3(dt_partials_grid(m)),
                       dim3(kBlock), 0, s, m, P, U, partials);
}

void launch_mirror(hipStream_t s, const DevMesh &m, const PhysDev &P,
                   const double *U, const double *R, double *phi_total_aos,
                   double *phi_dt_aos)
{
    hipLaunchKernelGGL(fv_mirror_kernel, dim3(cell_blocks(m.N + m.B)),
                       dim3(kBlock), 0, s, m, P, U, R, phi_total_aos,
                       phi_dt_aos);
}

int fold_rows(int n_partials) { return (n_partials + kFoldRows - 1) / kFoldRows; }

void launch_finalize(hipStream_t s, DevScalars *sc, const double *partials,
                     int n_partials, double *fold, int fold_cap,
                     const PhysDev &P, int mode, double *log, int log_cap)
{
    const int nf = fold_rows(n_partials);
    if (nf > fold_cap)
        throw std::runtime_error("reduction fold buffer too small");
    hipLaunchKernelGGL(fv_fold_partials_kernel, dim3(nf), dim3(kBlock), 0, s,
                       partials, n_partials, mode, fold);
    hipLaunchKernelGGL(fv_finalize_kernel, dim3(1), dim3(kBlock), 0, s, sc,
                       fold, nf, P, mode, log, log_cap);
}

} // namespace dev
} // namespace fvh
