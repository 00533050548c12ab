// CDNA4 kernels of the residual and explicit time-stepping path.
//
// Per RK stage (or residual evaluation):
//   fv_ghost_kernel     ghost rows of the boundary faces (B threads)
//   fv_gradient_tile_kernel (FP64 or FP32, tables built) / fv_gradient_kernel
//                       one thread per cell: Green-Gauss gradients and
//                       neighbourhood min/max (the tile kernel from its
//                       tile's states staged in LDS), then the cell's
//                       limited state on each of its faces, written to the
//                       cell's face-state rows (one per face slot)
//   fv_residual_face_kernel (tetrahedra, FP64, tables built)
//                       each face of a cell tile once: the face flux from
//                       the two sides' states, contributions per cell slot
//                       in LDS, summed in slot order, then the RK stage per
//                       (cell, variable) and, in the last stage, the
//                       residual-norm and time-step partials
//   fv_residual_group_kernel (the other cases)
//                       LANES lanes per cell, one face per lane, summed in
//                       slot order through lane shuffles, then the same
//                       stage and partials
//   fv_fold_partials_kernel, fv_finalize_kernel
//                       fixed-order folds and the time loop's bookkeeping
//
// Every cell sums its faces in ascending face id (its slot order: the order
// of the reference's face-ordered scatter, src/Solver/src/finite_volume/
// finite_volume.c:100-112), so its sums equal the reference's exactly and
// no atomics are used: results are bitwise reproducible run to run.  The
// flux[F][5] array of the reference stays out of HBM: the lane-per-slot
// kernel evaluates a face flux in both of its cells (same inputs, same
// operations, same bits), the face-centred kernel once per tile that holds
// one of its cells.
//
// Arithmetic follows the reference operation by operation and is compiled
// with -ffp-contract=off, so IEEE-rounded +,-,*,/ and sqrt give the same bits
// as the reference's x86-64 build.
#include <cfloat>
#include <cmath>
#include <stdexcept>

#include <type_traits>

#include "fv_physics.h"

// Minimum resident waves per SIMD requested from the register allocator for
// the two stencil kernels (occupancy against spilling).  The gradient kernel
// keeps a batch of neighbour rows and face records in registers (about 247
// VGPRs in FP64 with batches of 4 faces) without spills; with the face states
// stored by cell slot, a minimum of 1 wave per SIMD ran 2% faster than 2.
// Batches of 4 at 3 waves spill and ran at half the speed; batches of 1 at 3
// waves ran at the same speed with 44% more HBM reads.
#ifndef FVH_K1_MINW
#define FVH_K1_MINW 1
#endif
// Faces whose neighbour rows and records the gradient kernel loads together.
#ifndef FVH_K1_FB
#define FVH_K1_FB 4
#endif
#ifndef FVH_K2_MINW
#define FVH_K2_MINW 1
#endif
#ifndef FVH_K2_VISC_W
#define FVH_K2_VISC_W 3
#endif
// Dynamic LDS reserved per gradient-kernel workgroup; a non-zero value caps
// the resident workgroups per CU (occupancy experiments).
#ifndef FVH_K1_LDS
#define FVH_K1_LDS 0
#endif

namespace fvh {
namespace dev {

// ---------------------------------------------------------------------------
// Ghost rows N+b (boundary.c:188-320 and the update hook ns_update,
// navier_stokes.c:199-209): the boundary-condition state of every boundary
// face from its interior cell, primitives + conservatives [B][10]; with face
// states on, also side 1 of the boundary faces.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void fv_ghost_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U, double *__restrict__ gphi,
    double *__restrict__ fs, const DevScalars *sc, float *__restrict__ fs32,
    const float *__restrict__ U32)
{
    if (sc && sc->done)
        return;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.B)
        return;
    const int f = m.bnd_face[b];
    const int c = m.fcell[2 * f];
    double pc[10], g[10];
    if (U32)
        load_cons(U32, c, pc);
    else
        load_cons(U, c, pc);
    con_to_prim(P, pc);
    Frame fr;
    load_frame(m, f, fr);
    ghost_state(m, P, b, fr, pc, g);
#pragma unroll
    for (int j = 0; j < 10; ++j)
        gphi[(size_t)b * 10 + j] = g[j];
    const size_t row = (size_t)m.N * m.maxf + b;
    if (fs)
#pragma unroll
        for (int j = 0; j < 8; ++j)
            fs[row * 8 + j] = g[j];
    if (fs32)
#pragma unroll
        for (int j = 0; j < 8; ++j)
            fs32[row * 8 + j] = (float)g[j];
}

// ---------------------------------------------------------------------------
// K1: Green-Gauss gradients of all 10 variables and neighbourhood min/max of
// one cell (reconstruction.c:20-93 in gather form, limiter.c:57-82), then
// the cell's reconstructed state on each of its faces
// (reconstruction.c:168-255 / 142-163).  Writes the cell's face-state rows
// and the gradient subset the viscous flux needs.
// ---------------------------------------------------------------------------
// ALL: also the gradient of p and the bounds of p and T, which only the
// test outputs grad/mnmx read (fewer live registers without them).
template <int MAXF, class Real, bool ALL>
__global__ __launch_bounds__(kBlock, FVH_K1_MINW) void fv_gradient_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U,
    const double *__restrict__ gphi, GradOut out, int linear,
    const DevScalars *sc)
{
    if (sc && sc->done)
        return;
    const int N = m.N;
    const int c = out.c_begin + xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    if (c >= out.c_end)
        return;

    // the FP32 kernel reads the FP32 copy of the state (the same values: its
    // first operation is the rounding to FP32) and the FP32 face records
    constexpr bool f32 = std::is_same<Real, float>::value;
    const float *U32 = out.U32;
    Real pc[10];
    if (f32)
        load_phi(P, U32, nullptr, c, pc);
    else
        load_phi(P, U, nullptr, c, pc);

    Real gx[10], gy[10], gz[10], mn[10], mx[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        gx[j] = Real(0.0);
        gy[j] = Real(0.0);
        gz[j] = Real(0.0);
        mn[j] = pc[j];
        mx[j] = pc[j];
    }

    // Loads in rounds, each for several faces at once (the cell's face and
    // neighbour list; neighbour rows and face records, in batches of up to 4
    // faces), so that a wave waits for a few memory latencies rather than
    // three per face.  Slots s >= nf repeat slot 0 and are skipped in the
    // sums.
    constexpr int FB = MAXF % FVH_K1_FB == 0 ? FVH_K1_FB : (MAXF % 3 == 0 ? 3 : MAXF);
    // MAXF = 4 runs on meshes of tetrahedra only
    const int nf = MAXF == 4 ? 4 : m.ncf[c];
    const int2 *cfn = m.cfn + (size_t)c * m.maxf;
    int fid[MAXF];
    bool own[MAXF];
    int nbr[MAXF];
#pragma unroll
    for (int s = 0; s < MAXF; ++s)
    {
        const int2 e = cfn[s < m.maxf ? s : 0];
        fid[s] = e.x;
        own[s] = e.y >= 0;
        nbr[s] = own[s] ? e.y : -(e.y + 1);
    }

#pragma unroll
    for (int s0 = 0; s0 < MAXF; s0 += FB)
    {
        // neighbour rows (boundary faces: the cell's own row as a
        // placeholder, replaced by the ghost row written by fv_ghost_kernel)
        // and the face normals, areas and weights of this batch
        Real pn[FB][10];
        Real gn[FB][5];
#pragma unroll
        for (int i = 0; i < FB; ++i)
        {
            const int s = s0 + i;
            const int nb = nbr[s] < N ? nbr[s] : c;
            if (f32)
                load_cons(U32, nb, pn[i]);
            else
                load_cons(U, nb, pn[i]);
            if (f32)
            {
                const float4 *q = reinterpret_cast<const float4 *>(m.fgB32 + (size_t)fid[s] * 8);
                const float4 q0 = q[0], q1 = q[1];
                gn[i][0] = q0.x, gn[i][1] = q0.y, gn[i][2] = q0.z, gn[i][3] = q0.w;
                gn[i][4] = q1.x;
            }
            else
            {
                const double2 *g =
                    reinterpret_cast<const double2 *>(m.fgB + (size_t)fid[s] * kFgB);
                const double2 g0 = g[0], g1 = g[1], g2 = g[2];
                gn[i][0] = Real(g0.x), gn[i][1] = Real(g0.y), gn[i][2] = Real(g1.x);
                gn[i][3] = Real(g1.y), gn[i][4] = Real(g2.x);
            }
        }

#pragma unroll
        for (int i = 0; i < FB; ++i)
        {
            const int s = s0 + i;
            if (s >= nf)
                break;
            if (nbr[s] >= N)
            {
                const double *gp = gphi + (size_t)(nbr[s] - N) * 10;
#pragma unroll
                for (int j = 0; j < 10; ++j)
                    pn[i][j] = Real(gp[j]);
            }
            else
                con_to_prim(P, pn[i]);
            const bool own0 = own[s];
            const Real n0 = gn[i][0], n1 = gn[i][1], n2 = gn[i][2], area = gn[i][3],
                       lam = gn[i][4];
#pragma unroll
            for (int j = 0; j < 10; ++j)
            {
                const Real p0 = own0 ? pc[j] : pn[i][j];
                const Real p1 = own0 ? pn[i][j] : pc[j];
                const Real pm = p0 + lam * (p1 - p0);
                const Real ax = pm * n0 * area;
                const Real ay = pm * n1 * area;
                const Real az = pm * n2 * area;
                if (ALL || j != 8)
                {
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
                }
                if (ALL || j < 8)
                {
                    mn[j] = rmin(mn[j], pn[i][j]);
                    mx[j] = rmax(mx[j], pn[i][j]);
                }
            }
        }
    }

    const Real sv = Real(m.svol[c]);
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
        d[0] = make_double2(double(gx[1]), double(gx[2]));
        d[1] = make_double2(double(gx[3]), double(gx[9]));
        d[2] = make_double2(double(gy[1]), double(gy[2]));
        d[3] = make_double2(double(gy[3]), double(gy[9]));
        d[4] = make_double2(double(gz[1]), double(gz[2]));
        d[5] = make_double2(double(gz[3]), double(gz[9]));
    }
    if (ALL && out.grad)
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
    if (ALL && out.mnmx)
    {
        double *d = out.mnmx + (size_t)c * 20;
#pragma unroll
        for (int j = 0; j < 10; ++j)
        {
            d[j] = mn[j];
            d[10 + j] = mx[j];
        }
    }

    if (std::is_same<Real, float>::value ? !out.fs32 : !out.fs)
        return;
    // this cell's reconstructed state on each of its faces, one 8-value row
    // per face slot (a cell's rows are contiguous); the distance
    // vectors of mesh.c:44-56 (dist_cell_1 = x_f - x_c0, dist_cell_2 =
    // x_f - x_c1 through the periodic image), all faces' first (one round of
    // loads: the cell's own row of the slot table), then the limited states
    Real rv[MAXF][3];
    if (linear)
    {
#pragma unroll
        for (int s = 0; s < MAXF; ++s)
        {
            if (f32)
            {
                const float4 q =
                    reinterpret_cast<const float4 *>(m.crv32)[(size_t)c * m.maxf + (s < m.maxf ? s : 0)];
                rv[s][0] = q.x;
                rv[s][1] = q.y;
                rv[s][2] = q.z;
            }
            else
            {
                const double *r = m.crv + ((size_t)c * m.maxf + (s < m.maxf ? s : 0)) * 3;
                rv[s][0] = Real(r[0]);
                rv[s][1] = Real(r[1]);
                rv[s][2] = Real(r[2]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < MAXF; ++s)
    {
        if (s >= nf)
            break;
        Real st[8];
        if (linear)
        {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                st[j] = limited(pc[j], gx[j], gy[j], gz[j], rv[s], mn[j], mx[j], P.bj);
        }
        else
        {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                st[j] = pc[j];
        }
        if constexpr (std::is_same<Real, float>::value)
        {
            float4 *d = reinterpret_cast<float4 *>(out.fs32 + ((size_t)c * m.maxf + s) * 8);
            d[0] = make_float4(st[0], st[1], st[2], st[3]);
            d[1] = make_float4(st[4], st[5], st[6], st[7]);
        }
        else
        {
            double2 *d = reinterpret_cast<double2 *>(out.fs + ((size_t)c * m.maxf + s) * 8);
            d[0] = make_double2(double(st[0]), double(st[1]));
            d[1] = make_double2(double(st[2]), double(st[3]));
            d[2] = make_double2(double(st[4]), double(st[5]));
            d[3] = make_double2(double(st[6]), double(st[7]));
        }
    }
}

// ---------------------------------------------------------------------------
// K1 with the tile's states in LDS: the same gradients, bounds and face
// states as fv_gradient_kernel (FP64, without the test outputs), but the
// workgroup first stages the states of its tile of kTile consecutive cells
// (contiguous loads) and of the tile's halo (the face neighbours outside the
// tile and the ghost rows of the tile's boundary faces, gathered once per
// tile) into LDS as primitive rows, so the per-face neighbour reads are LDS
// reads and each state is converted to primitives once per tile instead of
// once per face.  LDS: 10 doubles per row, variable-major (tl[j*rows + r]).
// ---------------------------------------------------------------------------
#ifndef FVH_TILE_GROUP
#define FVH_TILE_GROUP 8
#endif
// The tile a workgroup of a tile launch takes.  The hardware hands
// consecutive workgroups to the 8 XCDs in turn; with FVH_TILE_GROUP = P > 1,
// each XCD's P workgroups of a round of 8P take P adjacent tiles, so that
// tiles that share halo rows run on one XCD, while the tiles of a round stay
// one compact run (P = 1: the identity).  On the 96^3 x 6 Taylor-Green mesh
// P = 8 and 16 were 0.5-0.8% faster than P = 1, P = 32 no faster.
__device__ __forceinline__ int tile_of_block(int b, int n)
{
    constexpr int P = FVH_TILE_GROUP;
    if (P == 1)
        return b;
    const int g = b / (8 * P), r = b - g * 8 * P;
    if ((g + 1) * 8 * P > n)
        return b;
    return g * 8 * P + (r % 8) * P + r / 8;
}
#ifndef FVH_K1T_MINW
#define FVH_K1T_MINW 2
#endif
#ifndef FVH_K1T_PREREC
#define FVH_K1T_PREREC 0
#endif
template <int MAXF, class Real>
__global__ __launch_bounds__(kTile, FVH_K1T_MINW) void fv_gradient_tile_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U,
    const double *__restrict__ gphi, GradOut out, int linear,
    const DevScalars *sc)
{
    if (sc && sc->done)
        return;
    // FP32 (the Jacobian-vector products of the Newton-Krylov solve): the
    // FP32 copy of the state and the FP32 face records and distance
    // vectors, FP32 face-state rows out (the viscous gradients stay FP64
    // storage), as fv_gradient_kernel<MAXF, float> takes and gives them
    constexpr bool f32 = std::is_same<Real, float>::value;
    extern __shared__ double smem[];
    Real *tl = reinterpret_cast<Real *>(smem);
    const int N = m.N;
    const int rows = m.tile_rows;
    // tiles in runs of 8 * FVH_TILE_GROUP, as in the face-centred residual
    // kernel
    const int t = out.c_begin / kTile + tile_of_block((int)blockIdx.x, (int)gridDim.x);
    const int base = t * kTile;
    const int h0 = m.tile_halo_off[t];
    const int nrow = kTile + (m.tile_halo_off[t + 1] - h0);
    for (int i = threadIdx.x; i < nrow; i += kTile)
    {
        Real q[10];
        const int g = i < kTile ? base + i : m.tile_halo[h0 + i - kTile];
        if (i < kTile && g >= N)
            continue;
        if (g < N)
        {
            if constexpr (f32)
                load_cons(out.U32, g, q);
            else
                load_cons(U, g, q);
            con_to_prim(P, q);
        }
        else
        {
            const double *gp = gphi + (size_t)(g - N) * 10;
#pragma unroll
            for (int j = 0; j < 10; ++j)
                q[j] = Real(gp[j]);
        }
#pragma unroll
        for (int j = 0; j < 10; ++j)
            tl[j * rows + i] = q[j];
    }
    __syncthreads();
    // cells outside the launch range (in the first and last tile of a range
    // that does not start or end on a tile boundary) take part in the
    // barriers of the output passes without computing; the output passes
    // store the rows of cells lo .. ncell - 1 of the tile only.  Their
    // states are staged like the others (on a connected mesh a cell of the
    // range may have them as neighbours)
    const int c = base + threadIdx.x;
    const bool valid = c >= out.c_begin && c < out.c_end;
    const int cc = valid ? c : base;

    Real pc[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
        pc[j] = tl[j * rows + threadIdx.x];
    Real gx[10], gy[10], gz[10], mn[10], mx[10];
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        gx[j] = Real(0.0);
        gy[j] = Real(0.0);
        gz[j] = Real(0.0);
        mn[j] = pc[j];
        mx[j] = pc[j];
    }
    // every cell of a mesh whose cells have at most 4 faces is a tetrahedron
    const int nf = valid ? (MAXF == 4 ? 4 : m.ncf[c]) : 0;
    const int2 *cfn = m.cfn + (size_t)cc * m.maxf;
    const unsigned short *nbl = m.nbl + (size_t)cc * m.maxf;
    int fid[MAXF], loc[MAXF];
    bool own[MAXF];
    if constexpr (MAXF == 4)
    {
        // tetrahedra: the cell's four slots in one 16-B and one 8-B load
        // (the face ids with their side in the sign bit, and nbl; aligned)
        const ushort4 l = *reinterpret_cast<const ushort4 *>(nbl);
        if (m.cfo)
        {
            const int4 a = *reinterpret_cast<const int4 *>(m.cfo + (size_t)cc * 4);
            fid[0] = a.x & 0x7fffffff, own[0] = a.x >= 0;
            fid[1] = a.y & 0x7fffffff, own[1] = a.y >= 0;
            fid[2] = a.z & 0x7fffffff, own[2] = a.z >= 0;
            fid[3] = a.w & 0x7fffffff, own[3] = a.w >= 0;
        }
        else
        {
            const int4 *c4 = reinterpret_cast<const int4 *>(cfn);
            const int4 a = c4[0], b = c4[1];
            fid[0] = a.x, own[0] = a.y >= 0, fid[1] = a.z, own[1] = a.w >= 0;
            fid[2] = b.x, own[2] = b.y >= 0, fid[3] = b.z, own[3] = b.w >= 0;
        }
        loc[0] = l.x, loc[1] = l.y, loc[2] = l.z, loc[3] = l.w;
    }
    else
#pragma unroll
        for (int s = 0; s < MAXF; ++s)
        {
            // slots past the mesh's face count per cell (MAXF = 6 with
            // 5-face cells) take slot 0, like the slots past a cell's own
            // count
            const int sl = s < m.maxf ? s : 0;
            const int2 e = cfn[sl];
            fid[s] = e.x;
            own[s] = e.y >= 0;
            loc[s] = nbl[sl];
        }
    // each face's record in its own iteration (FVH_K1T_PREREC 0: 158 VGPRs,
    // 3 waves per SIMD); the records of all faces first (one round of
    // gathers) need 178, 2 waves, and ran 17% slower on tetrahedra
    Real gn[MAXF][5];
    auto load_rec = [&](int s) {
        if constexpr (f32)
        {
            const float4 *q = reinterpret_cast<const float4 *>(m.fgB32 + (size_t)fid[s] * 8);
            const float4 q0 = q[0], q1 = q[1];
            gn[s][0] = q0.x, gn[s][1] = q0.y, gn[s][2] = q0.z;
            gn[s][3] = q0.w, gn[s][4] = q1.x;
        }
        else if (m.fgN)
        {
            // n and area (32 B) and lambda (8 B) instead of the 48-B record
            const double2 *g = reinterpret_cast<const double2 *>(m.fgN + (size_t)fid[s] * 4);
            const double2 g0 = g[0], g1 = g[1];
            gn[s][0] = g0.x, gn[s][1] = g0.y, gn[s][2] = g1.x;
            gn[s][3] = g1.y, gn[s][4] = m.flam[fid[s]];
        }
        else
        {
            const double2 *g = reinterpret_cast<const double2 *>(m.fgB + (size_t)fid[s] * kFgB);
            const double2 g0 = g[0], g1 = g[1], g2 = g[2];
            gn[s][0] = g0.x, gn[s][1] = g0.y, gn[s][2] = g1.x;
            gn[s][3] = g1.y, gn[s][4] = g2.x;
        }
    };
    if (FVH_K1T_PREREC)
#pragma unroll
        for (int s = 0; s < MAXF; ++s)
            load_rec(s);
#pragma unroll
    for (int s = 0; s < MAXF; ++s)
    {
        if (s >= nf)
            break;
        if (!FVH_K1T_PREREC)
            load_rec(s);
        Real pn[10];
#pragma unroll
        for (int j = 0; j < 10; ++j)
            pn[j] = tl[j * rows + loc[s]];
        const bool own0 = own[s];
        const Real n0 = gn[s][0], n1 = gn[s][1], n2 = gn[s][2], area = gn[s][3],
                   lam = gn[s][4];
#pragma unroll
        for (int j = 0; j < 10; ++j)
        {
            const Real p0 = own0 ? pc[j] : pn[j];
            const Real p1 = own0 ? pn[j] : pc[j];
            const Real pm = p0 + lam * (p1 - p0);
            const Real ax = pm * n0 * area;
            const Real ay = pm * n1 * area;
            const Real az = pm * n2 * area;
            if (j != 8)
            {
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
            }
            if (j < 8)
            {
                mn[j] = rmin(mn[j], pn[j]);
                mx[j] = rmax(mx[j], pn[j]);
            }
        }
    }
    const Real sv = Real(m.svol[cc]);
#pragma unroll
    for (int j = 0; j < 10; ++j)
    {
        gx[j] *= sv;
        gy[j] *= sv;
        gz[j] *= sv;
    }
    Real rv[MAXF][3];
    const bool fs_out = f32 ? out.fs32 != nullptr : out.fs != nullptr;
    if (linear && fs_out && MAXF == 4 && !f32)
    {
        // tetrahedra: the cell's 96 B of distance vectors in six 16-B loads
        const double2 *r2 = reinterpret_cast<const double2 *>(m.crv + (size_t)cc * 12);
        const double2 q0 = r2[0], q1 = r2[1], q2 = r2[2], q3 = r2[3], q4 = r2[4], q5 = r2[5];
        rv[0][0] = q0.x, rv[0][1] = q0.y, rv[0][2] = q1.x;
        rv[1][0] = q1.y, rv[1][1] = q2.x, rv[1][2] = q2.y;
        rv[2][0] = q3.x, rv[2][1] = q3.y, rv[2][2] = q4.x;
        rv[MAXF - 1][0] = q4.y, rv[MAXF - 1][1] = q5.x, rv[MAXF - 1][2] = q5.y;
    }
    else if (linear && fs_out)
    {
#pragma unroll
        for (int s = 0; s < MAXF; ++s)
        {
            const size_t sl = (size_t)cc * m.maxf + (s < m.maxf ? s : 0);
            if constexpr (f32)
            {
                const float4 r = reinterpret_cast<const float4 *>(m.crv32)[sl];
                rv[s][0] = r.x;
                rv[s][1] = r.y;
                rv[s][2] = r.z;
            }
            else
            {
                const double *r = m.crv + sl * 3;
                rv[s][0] = r[0];
                rv[s][1] = r[1];
                rv[s][2] = r[2];
            }
        }
    }

    // Output passes through LDS (the states table is no longer read), so
    // that each store instruction of a wave writes whole contiguous lines:
    // the tile's viscous gradient rows are one contiguous block; the face
    // states go out one pair of slots at a time (two 64-byte rows of a cell,
    // one 128-byte line).
    const int ncell = min(kTile, out.c_end - base);
    const int lo = max(out.c_begin - base, 0);
    __syncthreads();
    if (out.vg)
    {
        double *o = smem + threadIdx.x * 12;
        o[0] = gx[1], o[1] = gx[2], o[2] = gx[3], o[3] = gx[9];
        o[4] = gy[1], o[5] = gy[2], o[6] = gy[3], o[7] = gy[9];
        o[8] = gz[1], o[9] = gz[2], o[10] = gz[3], o[11] = gz[9];
        __syncthreads();
        const double2 *src = reinterpret_cast<const double2 *>(smem);
        double2 *dst = reinterpret_cast<double2 *>(out.vg + (size_t)base * 12);
        for (int i = threadIdx.x; i < (ncell - lo) * 6; i += kTile)
            dst[lo * 6 + i] = src[lo * 6 + i];
        __syncthreads();
    }
    if (!fs_out)
        return;
#pragma unroll
    for (int s0 = 0; s0 < MAXF; s0 += 2)
    {
        const int ns = min(2, m.maxf - s0);
        if (ns <= 0)
            break;
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
            const int s = s0 + h;
            if (h >= ns)
                break;
            Real *o = tl + threadIdx.x * 16 + h * 8;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                o[j] = linear ? limited(pc[j], gx[j], gy[j], gz[j], rv[s], mn[j], mx[j], P.bj)
                              : pc[j];
        }
        __syncthreads();
        if constexpr (f32)
        {
            // 2 float4 per row, ns rows per cell
            const float4 *src = reinterpret_cast<const float4 *>(tl);
            const int per = 2 * ns;
            for (int i = threadIdx.x; i < (ncell - lo) * per; i += kTile)
            {
                const int cl = lo + i / per, q = i % per;
                float4 *dst = reinterpret_cast<float4 *>(
                    out.fs32 + ((size_t)(base + cl) * m.maxf + s0) * 8);
                dst[q] = src[cl * 4 + q];
            }
        }
        else
        {
            // 4 double2 per row, ns rows per cell
            const double2 *src = reinterpret_cast<const double2 *>(tl);
            const int per = 4 * ns;
            for (int i = threadIdx.x; i < (ncell - lo) * per; i += kTile)
            {
                const int cl = lo + i / per, q = i % per;
                double2 *dst = reinterpret_cast<double2 *>(
                    out.fs + ((size_t)(base + cl) * m.maxf + s0) * 8);
                dst[q] = src[cl * 8 + q];
            }
        }
        __syncthreads();
    }
}

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
// The work of one lane of a cell group: c the cell (cell_ok: inside the
// launch range), s the lane.  sh: LDS for the block reduction (8 doubles per
// wave); partial_row: where thread 0 writes the block's partials when the
// stage reduces.  Shared by the residual kernel and the batched kernels.
template <int FLUX, bool FS, bool VISC, int LANES, class Real>
__device__ __forceinline__ void residual_group_body(const DevMesh &m, const PhysDev &P,
                                                    const StageParams &sp, int c, int s,
                                                    bool cell_ok, double *sh,
                                                    double *partial_row)
{
    const int N = m.N;
    // four lanes per cell only on meshes of tetrahedra: four faces, no load
    // for the count before the face work
    const int nf = cell_ok ? (LANES == 4 ? 4 : m.ncf[c]) : 0;
    // the epilogue's operands (the cell's state and RK register, element j in
    // lane j % LANES of the group, and its volume), loaded before the face
    // work so that their latency overlaps it
    constexpr int NQ = (5 + LANES - 1) / LANES;
    const int cc = cell_ok ? c : sp.c_begin;
    double pu[NQ], pk[NQ];
    if (sp.do_update)
    {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
        {
            const int j = q * LANES + s < 5 ? q * LANES + s : 0;
            pu[q] = sp.U_in[(size_t)cc * 5 + j];
            pk[q] = sp.stage0 ? 0.0 : sp.k[(size_t)cc * 5 + j];
        }
    }
    const double vol = m.vol[cc];
    Real contrib[5] = {Real(0), Real(0), Real(0), Real(0), Real(0)};
    if (s < nf)
    {
        const int2 e = m.cfn[(size_t)c * m.maxf + s];
        const int f = e.x;
        const bool own0 = e.y >= 0;
        const int nb = own0 ? e.y : -(e.y + 1);
        const bool ghost = nb >= N;
        FrameT<Real> fr;
        load_frame(m, f, fr);

        Real L[10], R[10];
        if (FS)
        {
            // the face's two states: this cell's row of the slot, and the
            // neighbour's row of the same face (or the ghost row)
            const size_t ro = (size_t)c * m.maxf + s;
            const size_t rn = ghost ? (size_t)N * m.maxf + (nb - N)
                                    : (size_t)nb * m.maxf + m.nsl[ro];
            const size_t rl = own0 ? ro : rn, rr = own0 ? rn : ro;
            if constexpr (std::is_same<Real, float>::value)
            {
                const float4 *pl = reinterpret_cast<const float4 *>(sp.fs32 + rl * 8);
                const float4 *pr = reinterpret_cast<const float4 *>(sp.fs32 + rr * 8);
#pragma unroll
                for (int q = 0; q < 2; ++q)
                {
                    const float4 a = pl[q], b = pr[q];
                    L[4 * q] = a.x, L[4 * q + 1] = a.y, L[4 * q + 2] = a.z, L[4 * q + 3] = a.w;
                    R[4 * q] = b.x, R[4 * q + 1] = b.y, R[4 * q + 2] = b.z, R[4 * q + 3] = b.w;
                }
            }
            else
            {
                const double2 *pl = reinterpret_cast<const double2 *>(sp.fs + rl * 8);
                const double2 *pr = reinterpret_cast<const double2 *>(sp.fs + rr * 8);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    const double2 a = pl[q], b = pr[q];
                    L[2 * q] = Real(a.x);
                    L[2 * q + 1] = Real(a.y);
                    R[2 * q] = Real(b.x);
                    R[2 * q + 1] = Real(b.y);
                }
            }
            L[8] = L[9] = R[8] = R[9] = Real(0.0);
        }
        else
        {
            constexpr bool f32 = std::is_same<Real, float>::value;
            Real pc[10], pn[10];
            if (f32 && sp.U32)
                load_phi(P, sp.U32, nullptr, c, pc);
            else
                load_phi(P, sp.U_in, nullptr, c, pc);
            if (ghost)
            {
                const double *gp = sp.gphi + (size_t)(nb - N) * 10;
#pragma unroll
                for (int j = 0; j < 10; ++j)
                    pn[j] = Real(gp[j]);
            }
            else if (f32 && sp.U32)
            {
                load_phi(P, sp.U32, nullptr, nb, pn);
            }
            else
            {
                load_phi(P, sp.U_in, nullptr, nb, pn);
            }
#pragma unroll
            for (int j = 0; j < 10; ++j)
            {
                L[j] = own0 ? pc[j] : pn[j];
                R[j] = own0 ? pn[j] : pc[j];
            }
        }

        ViscGradT<Real> gL, gR;
        if (VISC)
        {
            // the neighbour's row is loaded unconditionally (the cell's own
            // row stands in on boundary faces), so that its load is issued
            // with the face's other loads
            ViscGradT<Real> go, gn;
            load_vg(sp.vg, c, go);
            load_vg(sp.vg, ghost ? c : nb, gn);
            if (ghost)
            {
                const int type = m.region_type[m.bnd_region[nb - N]];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    ghost_grad(type, fr, go.v[q], go.v[4 + q], go.v[8 + q],
                               gn.v[q], gn.v[4 + q], gn.v[8 + q]);
            }
#pragma unroll
            for (int q = 0; q < 12; ++q)
            {
                gL.v[q] = own0 ? go.v[q] : gn.v[q];
                gR.v[q] = own0 ? gn.v[q] : go.v[q];
            }
        }

        Real flux[5];
        face_flux<FLUX, VISC>(P, fr, L, R, gL, gR, flux);
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const Real v = flux[j] * fr.area;
            contrib[j] = own0 ? v : -v;
        }
    }

    // ordered sum over the group's faces
    Real acc[5];
#pragma unroll
    for (int j = 0; j < 5; ++j)
    {
        Real a = Real(0.0);
#pragma unroll
        for (int k = 0; k < LANES; ++k)
        {
            const Real v = __shfl(contrib[j], k, LANES);
            if (k < nf)
                a = a + v;
        }
        acc[j] = a;
    }

    double u5[5], k5[5];
    if (sp.do_update)
    {
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            u5[j] = __shfl(pu[j / LANES], j % LANES, LANES);
            k5[j] = __shfl(pk[j / LANES], j % LANES, LANES);
        }
    }

    double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    if (cell_ok)
    {
        // every lane of the group has the sums; each computes the update of
        // all five variables (the same instructions as one lane) and stores
        // the variables j = s, s + LANES, ..., so that a store instruction
        // writes the group's row, not one value per group
        double R5[5];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            R5[j] = -double(acc[j]) / vol;
        double kk[5], unew[5];
        if (sp.do_update)
        {
            const double dt = sp.case_dt ? sp.case_dt[sp.cell_case[c]] : sp.sc->dt;
#pragma unroll
            for (int j = 0; j < 5; ++j)
            {
                double kj = R5[j];
                if (!sp.stage0)
                    kj = R5[j] + k5[j] * sp.rk_a;
                kk[j] = kj;
                unew[j] = u5[j] + kj * dt * sp.rk_g;
            }
        }
        auto pick = [](const double *a, int j) {
            return j == 0 ? a[0] : j == 1 ? a[1] : j == 2 ? a[2] : j == 3 ? a[3] : a[4];
        };
#pragma unroll
        for (int q = 0; q < NQ; ++q)
        {
            const int j = q * LANES + s;
            if (j < 5)
            {
                if (sp.R_out)
                    sp.R_out[(size_t)c * 5 + j] = pick(R5, j);
                if (sp.do_update)
                {
                    if (!sp.k_dead)
                        sp.k[(size_t)c * 5 + j] = pick(kk, j);
                    sp.U_out[(size_t)c * 5 + j] = pick(unew, j);
                }
            }
        }
        if (s == 0 && sp.do_update)
        {
            double pc[10];
#pragma unroll
            for (int j = 0; j < 5; ++j)
                pc[j] = u5[j];
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
        block_reduce7(sh, part);
        if (threadIdx.x == 0)
        {
#pragma unroll
            for (int q = 0; q < 7; ++q)
                partial_row[q] = part[q];
        }
    }
}

// The FP64 viscous kernel on face states is held to 3 waves per SIMD
// (FVH_K2_VISC_W): with the epilogue operands loaded up front it needs about
// 146 VGPRs, and at 4 waves its spills made it 1.6x slower; the other
// instantiations keep the compiler's choice.
template <int FLUX, bool FS, bool VISC, int LANES, class Real>
__global__ __launch_bounds__(kBlock, (FS && VISC && std::is_same<Real, double>::value)
                                         ? FVH_K2_VISC_W
                                         : FVH_K2_MINW) void fv_residual_group_kernel(
    DevMesh m, PhysDev P, StageParams sp)
{
    __shared__ double sh[8 * kBlock / 64];
    const int tid = xcd_tile(gridDim.x) * kBlock + threadIdx.x;
    const int c = sp.c_begin + tid / LANES;
    const int s = tid & (LANES - 1);
    const bool cell_ok = c < sp.c_end;
    // the run has stopped: this stage does nothing (both state buffers keep
    // their contents; the host finds the final one from the step count)
    if (sp.sc && sp.sc->done)
        return;
    residual_group_body<FLUX, FS, VISC, LANES, Real>(
        m, P, sp, c, s, cell_ok, sh, sp.partials + (size_t)blockIdx.x * kPartialStride);
}



// ---------------------------------------------------------------------------
// Batched cases: per stage one persistent launch for each (flux, viscous)
// class of cases (one flux and viscosity variant per kernel keeps its
// register count, and so its occupancy, that of the single-case kernel).
// Each workgroup takes tiles from the launch's work queue until it is
// empty; a tile is bt.sub_tiles runs of BS / LANES consecutive cells of one
// case, computed one run after the other with that case's viscosity and
// time step by the residual kernel's own lane-group code (first order).
// With do_reduce, run u of tile t writes its partials to row
// t * sub_tiles + u of sp.partials (a case's tiles are consecutive and its
// last tile ends with it, so its rows are consecutive too and the per-case
// fold reads them in a fixed order).  The launch bounds ask for the waves
// per SIMD of the single-case kernel of the same variant (the tile loop
// would otherwise keep enough loop-invariant values live to lose one).
// ---------------------------------------------------------------------------
// Tile t of a batched launch: its runs of BS / LANES cells one after the
// other, with its case's viscosity.
template <int BS, int LANES, int FLUX, bool VISC>
__device__ __forceinline__ void batch_tile_body(const DevMesh &m, const PhysDev &P,
                                                const StageParams &sp, const BatchTiles &bt,
                                                int t, double *sh)
{
    constexpr int GC = BS / LANES;
    const int k = bt.tile_case[t];
    const int begin = bt.tile_begin[t], end = bt.case_end[k];
    PhysDev Pk = P;
    Pk.mu = bt.case_mu[k];
    Pk.lam = bt.case_lam[k];
    const int s = threadIdx.x & (LANES - 1);
    for (int u = 0; u < bt.sub_tiles; ++u)
    {
        StageParams q = sp;
        q.c_begin = begin + u * GC;
        if (q.c_begin >= end)
            break;
        q.c_end = min(q.c_begin + GC, end);
        const int c = q.c_begin + (int)threadIdx.x / LANES;
        double *row = sp.partials + ((size_t)t * bt.sub_tiles + u) * kPartialStride;
        residual_group_body<FLUX, false, VISC, LANES, double>(m, Pk, q, c, s, c < q.c_end, sh,
                                                             row);
    }
}

template <int BS, int LANES, int FLUX, bool VISC>
__global__ __launch_bounds__(BS, LANES == 4 ? (VISC ? 3 : 5) : (VISC ? 2 : 3)) void
fv_batch_stage_kernel(DevMesh m, PhysDev P, StageParams sp,
                                                           BatchTiles bt)
{
    __shared__ double sh[8 * BS / 64];
    __shared__ int s_tile;
    // One lane-0 section per iteration (end stamp of the previous tile, the
    // next tile, its start stamp) and a wave-uniform tile index: the loop
    // exit is a scalar branch and every barrier sits in code that all lanes
    // run.  Lane 0 keeps the following queue index in flight while the tile
    // is computed, so the atomic's latency overlaps the tile's loads.
    int next = 0;
    if (threadIdx.x == 0)
        next = (int)atomicAdd(bt.counter, 1u);
    int t = -1;
    for (;;)
    {
        __syncthreads();
        if (threadIdx.x == 0)
        {
            if (bt.stamps && t >= 0)
                bt.stamps[(size_t)t * 2 + 1] = wall_clock64();
            const int nt = next < bt.n_tiles ? bt.tile_base + next : -1;
            s_tile = nt;
            if (nt >= 0)
            {
                next = (int)atomicAdd(bt.counter, 1u);
                if (bt.stamps)
                    bt.stamps[(size_t)nt * 2] = wall_clock64();
            }
        }
        __syncthreads();
        t = __builtin_amdgcn_readfirstlane(s_tile);
        if (t < 0)
            return;
        batch_tile_body<BS, LANES, FLUX, VISC>(m, P, sp, bt, t, sh);
    }
}

// The same tiles, one workgroup each (no work queue: the hardware
// dispatcher hands out the tiles), t = bt.tile_base + blockIdx.x.
template <int BS, int LANES, int FLUX, bool VISC>
__global__ __launch_bounds__(BS, LANES == 4 ? (VISC ? 3 : 5) : (VISC ? 2 : 3)) void
fv_batch_tile_kernel(DevMesh m, PhysDev P, StageParams sp,
                                                           BatchTiles bt)
{
    __shared__ double sh[8 * BS / 64];
    const int t = bt.tile_base + (int)blockIdx.x;
    if (bt.stamps && threadIdx.x == 0)
        bt.stamps[(size_t)t * 2] = wall_clock64();
    batch_tile_body<BS, LANES, FLUX, VISC>(m, P, sp, bt, t, sh);
    if (bt.stamps)
    {
        __syncthreads();
        if (threadIdx.x == 0)
            bt.stamps[(size_t)t * 2 + 1] = wall_clock64();
    }
}

template <int LANES, int FLUX, bool VISC>
static void launch_batch_lanes(hipStream_t s, const DevMesh &m, const PhysDev &P,
                               const StageParams &sp, const BatchTiles &bt, int block, int grid)
{
    if (!bt.counter)
    {
        // one workgroup per tile
        if (block == 128)
            hipLaunchKernelGGL((fv_batch_tile_kernel<128, LANES, FLUX, VISC>), dim3(bt.n_tiles),
                               dim3(128), 0, s, m, P, sp, bt);
        else if (block == 256)
            hipLaunchKernelGGL((fv_batch_tile_kernel<256, LANES, FLUX, VISC>), dim3(bt.n_tiles),
                               dim3(256), 0, s, m, P, sp, bt);
        else if (block == 512)
            hipLaunchKernelGGL((fv_batch_tile_kernel<512, LANES, FLUX, VISC>), dim3(bt.n_tiles),
                               dim3(512), 0, s, m, P, sp, bt);
        else
            throw std::runtime_error("launch_batch_stage: block must be 128, 256 or 512");
        return;
    }
    if (block == 128)
        hipLaunchKernelGGL((fv_batch_stage_kernel<128, LANES, FLUX, VISC>), dim3(grid),
                           dim3(128), 0, s, m, P, sp, bt);
    else if (block == 256)
        hipLaunchKernelGGL((fv_batch_stage_kernel<256, LANES, FLUX, VISC>), dim3(grid),
                           dim3(256), 0, s, m, P, sp, bt);
    else if (block == 512)
        hipLaunchKernelGGL((fv_batch_stage_kernel<512, LANES, FLUX, VISC>), dim3(grid),
                           dim3(512), 0, s, m, P, sp, bt);
    else
        throw std::runtime_error("launch_batch_stage: block must be 128, 256 or 512");
}

template <int LANES>
static void launch_batch_class(hipStream_t s, const DevMesh &m, const PhysDev &P,
                               const StageParams &sp, const BatchTiles &bt, int flux,
                               bool viscous, int block, int grid)
{
    if (flux == FluxRusanov)
    {
        if (viscous)
            launch_batch_lanes<LANES, FluxRusanov, true>(s, m, P, sp, bt, block, grid);
        else
            launch_batch_lanes<LANES, FluxRusanov, false>(s, m, P, sp, bt, block, grid);
    }
    else if (flux == FluxAUSM)
    {
        if (viscous)
            launch_batch_lanes<LANES, FluxAUSM, true>(s, m, P, sp, bt, block, grid);
        else
            launch_batch_lanes<LANES, FluxAUSM, false>(s, m, P, sp, bt, block, grid);
    }
    else
        throw std::runtime_error("launch_batch_stage: unknown flux");
}

void launch_batch_stage(hipStream_t s, const DevMesh &m, const PhysDev &P,
                        const StageParams &sp, const BatchTiles &bt, int flux, bool viscous,
                        int block, int grid)
{
    if (m.maxf > 6)
        throw std::runtime_error("cells with more than 6 faces are not supported");
    if (!sp.case_dt || !sp.cell_case || !sp.do_update)
        throw std::runtime_error("launch_batch_stage: needs per-case time steps");
    if (grid < 1 || bt.n_tiles < 1 || bt.tile_base < 0 || bt.sub_tiles < 1)
        throw std::runtime_error("launch_batch_stage: empty launch");
    if (viscous && !sp.vg)
        throw std::runtime_error("launch_batch_stage: viscous tiles need the gradients");
    if (m.maxf <= 4)
        launch_batch_class<4>(s, m, P, sp, bt, flux, viscous, block, grid);
    else
        launch_batch_class<8>(s, m, P, sp, bt, flux, viscous, block, grid);
}

// The 2N-storage RK stage (explicit.c:113-140) from a residual R computed
// elsewhere (the host-mirror path): one thread per cell, the same arithmetic
// as the residual kernel's epilogue, partials of the last stage per
// workgroup of kBlock cells.
__global__ __launch_bounds__(kBlock) void fv_stage_update_kernel(DevMesh m, PhysDev P,
                                                                StageParams sp,
                                                                const double *__restrict__ R)
{
    __shared__ double sh[8 * kBlock / 64];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = c < m.N;
    double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    if (sp.sc->done)
        return;
    if (ok)
    {
        const double vol = m.vol[c];
        const double dt = sp.sc->dt;
        double pc[10], R5[5], unew[5];
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            pc[j] = sp.U_in[(size_t)c * 5 + j];
            R5[j] = R[(size_t)c * 5 + j];
            double kj = R5[j];
            if (!sp.stage0)
                kj = R5[j] + sp.k[(size_t)c * 5 + j] * sp.rk_a;
            sp.k[(size_t)c * 5 + j] = kj;
            unew[j] = pc[j] + kj * dt * sp.rk_g;
            sp.U_out[(size_t)c * 5 + j] = unew[j];
        }
        if (sp.do_reduce)
        {
            con_to_prim(P, pc);
#pragma unroll
            for (int j = 0; j < 5; ++j)
                part[j] = rabs(R5[j]) * vol;
            const double dx0 = m.cdx[(size_t)c * 3], dx1 = m.cdx[(size_t)c * 3 + 1],
                         dx2 = m.cdx[(size_t)c * 3 + 2];
            const double s_rho = 1.0 / unew[0];
            const double cs = sqrt(P.kappa * pc[8] * s_rho);
            const double ds1 = (rabs(pc[5]) + cs) * dx0 + (rabs(pc[6]) + cs) * dx1 +
                               (rabs(pc[7]) + cs) * dx2;
            part[5] = vol / ds1;
            const double ds2 = dot3(dx0, dx1, dx2, dx0, dx1, dx2);
            part[6] = s_rho * ds2 / vol;
        }
    }
    if (sp.do_reduce)
    {
        block_reduce7(sh, part);
        if (threadIdx.x == 0)
#pragma unroll
            for (int q = 0; q < 7; ++q)
                sp.partials[blockIdx.x * kPartialStride + q] = part[q];
    }
}

void launch_stage_update(hipStream_t s, const DevMesh &m, const PhysDev &P,
                         const StageParams &sp, const double *R)
{
    if (!sp.sc)
        throw std::runtime_error("launch_stage_update: no time step");
    if (sp.do_reduce && cell_blocks(m.N) > sp.partials_cap)
        throw std::runtime_error("reduction partials buffer too small");
    hipLaunchKernelGGL(fv_stage_update_kernel, dim3(cell_blocks(m.N)), dim3(kBlock), 0, s, m, P,
                       sp, R);
}

// Time-step partials of a state whose primitives are current (the first
// step: the primitives were refreshed from the initial state).
__global__ __launch_bounds__(kBlock) void fv_dt_partials_kernel(
    DevMesh m, PhysDev P, const double *__restrict__ U,
    double *__restrict__ partials, const double *__restrict__ prims)
{
    __shared__ double sh[kBlock];
    const int N = m.N;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double lc = DBL_MAX, lv = DBL_MAX;
    if (c < N)
    {
        double pc[10];
        load_phi(P, U, prims, c, pc);
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
    block_reduce7(sh, v);
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
    block_reduce7(sh, v);
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
        if (sc->host_dt)
            return;
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
    double *__restrict__ dt_aos, const double *__restrict__ Up)
{
    const int N = m.N;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= N + m.B)
        return;
    double pc[10];
    if (row < N)
    {
        load_cons(Up, row, pc);
        con_to_prim(P, pc);
        load_cons(U, row, pc);
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
        load_cons(Up, c, pi);
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
                  const DevScalars *sc, float *fs32, const float *U32)
{
    if (m.B == 0)
        return;
    hipLaunchKernelGGL(fv_ghost_kernel, dim3(cell_blocks(m.B)), dim3(kBlock), 0,
                       s, m, P, U, gphi, fs, sc, fs32, U32);
}

// The tiles of the tile gradient kernel that cover the cells [c_begin,
// c_end): from the tile of c_begin to the tile of c_end - 1.
static int range_tiles(const GradOut &out)
{
    return (out.c_end + kTile - 1) / kTile - out.c_begin / kTile;
}

void launch_gradient(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const double *U, const double *gphi, const GradOut &out_in,
                     int linear, const DevScalars *sc, bool fp32)
{
    check_faces(m);
    GradOut out = out_in;
    if (out.c_end <= 0)
    {
        out.c_begin = 0;
        out.c_end = m.N;
    }
    if (out.c_begin < 0 || out.c_end > m.N || out.c_begin >= out.c_end)
        throw std::runtime_error("launch_gradient: bad cell range");
    const int grid = cell_blocks(out.c_end - out.c_begin);
    if (fp32)
    {
        if (out.grad || out.mnmx)
            throw std::runtime_error("launch_gradient: the test outputs are FP64 only");
        if (!out.U32 || !m.fgB32 || !m.crv32)
            throw std::runtime_error("launch_gradient: FP32 needs the FP32 state and face records");
        if (m.tile_halo_off)
        {
            const size_t lds = max((size_t)10 * m.tile_rows * sizeof(float),
                                   (size_t)12 * kTile * sizeof(double));
            const int tiles = range_tiles(out);
            if (m.maxf <= 4)
                hipLaunchKernelGGL((fv_gradient_tile_kernel<4, float>), dim3(tiles), dim3(kTile),
                                   lds, s, m, P, U, gphi, out, linear, sc);
            else
                hipLaunchKernelGGL((fv_gradient_tile_kernel<6, float>), dim3(tiles), dim3(kTile),
                                   lds, s, m, P, U, gphi, out, linear, sc);
            return;
        }
        if (m.maxf <= 4)
            hipLaunchKernelGGL((fv_gradient_kernel<4, float, false>), dim3(grid), dim3(kBlock),
                               FVH_K1_LDS, s, m, P, U, gphi, out, linear, sc);
        else
            hipLaunchKernelGGL((fv_gradient_kernel<6, float, false>), dim3(grid), dim3(kBlock),
                               FVH_K1_LDS, s, m, P, U, gphi, out, linear, sc);
        return;
    }
    if (out.grad || out.mnmx)
    {
        if (m.maxf <= 4)
            hipLaunchKernelGGL((fv_gradient_kernel<4, double, true>), dim3(grid),
                               dim3(kBlock), FVH_K1_LDS, s, m, P, U, gphi, out, linear,
                               sc);
        else
            hipLaunchKernelGGL((fv_gradient_kernel<6, double, true>), dim3(grid),
                               dim3(kBlock), FVH_K1_LDS, s, m, P, U, gphi, out, linear,
                               sc);
        return;
    }
    if (m.tile_halo_off)
    {
        const size_t lds = (size_t)max(10 * m.tile_rows, 16 * kTile) * sizeof(double);
        const int tiles = range_tiles(out);
        if (m.maxf <= 4)
            hipLaunchKernelGGL((fv_gradient_tile_kernel<4, double>), dim3(tiles), dim3(kTile),
                               lds, s, m, P, U, gphi, out, linear, sc);
        else
            hipLaunchKernelGGL((fv_gradient_tile_kernel<6, double>), dim3(tiles), dim3(kTile),
                               lds, s, m, P, U, gphi, out, linear, sc);
        return;
    }
    if (m.maxf <= 4)
        hipLaunchKernelGGL((fv_gradient_kernel<4, double, false>), dim3(grid),
                           dim3(kBlock), FVH_K1_LDS, s, m, P, U, gphi, out, linear,
                           sc);
    else
        hipLaunchKernelGGL((fv_gradient_kernel<6, double, false>), dim3(grid),
                           dim3(kBlock), FVH_K1_LDS, s, m, P, U, gphi, out, linear,
                           sc);
}

// ---------------------------------------------------------------------------
// K2 face-centred (tetrahedral meshes): the workgroup of tile t evaluates
// each face of its cells once (a face between two of the tile's cells gives
// both their contributions; a face to a cell outside the tile is also
// evaluated by that cell's tile), with the same states, gradients and flux
// as the lane-per-slot kernel, and keeps the contributions +-flux*area per
// cell slot in LDS.  Then one thread per (cell, variable) sums the cell's
// slots in slot order (the lane-per-slot kernel's order) and applies the
// 2N-storage RK stage, so that the stores of a wave are contiguous; in the
// last stage one thread per cell forms the residual-norm and time-step
// partials, one row per 64 cells (over the whole mesh: the rows of the
// lane-per-slot kernel's grid; with tile_c0 tiles: kTile / 64 rows per
// tile).  FS: the face states written by the gradient kernel; otherwise
// (first order) the two cells' values.  Tiles are kTile consecutive cells,
// or DevMesh::tile_c0's (batched cases: no tile spans two cases); a launch
// covers tiles sp.tile_begin .. + gridDim.x - 1.  LDS: kTile * 4 * 5
// doubles.
// ---------------------------------------------------------------------------
#ifndef FVH_K2F_MINW
#define FVH_K2F_MINW 3
#endif
// One entry of DevMesh::tile_faces as {face, left row, right row, left slot |
// right slot << 16} (0xffff for a side outside the tile); base: the tile's
// first cell.
__device__ __forceinline__ int4 unpack_tile_face(unsigned long long q, int base)
{
    const int f = (int)(q & 0x3ffffffULL);
    const int a = (int)((q >> 26) & 0x1ffULL);
    const int b4 = base * 4;
    if (q >> 63)
    {
        const bool right = (q >> 35) & 1ULL;
        const int other = (int)((q >> 36) & 0x7ffffffULL);
        return right ? make_int4(f, other, b4 + a, 0xffff | (a << 16))
                     : make_int4(f, b4 + a, other, a | (0xffff << 16));
    }
    const int b = (int)((q >> 35) & 0x1ffULL);
    return make_int4(f, b4 + a, b4 + b, a | (b << 16));
}
// 1: the update's operands are loaded before the face work (20 more VGPRs);
// 2: only the RK register; 0: none (with FVH_K2F_LATE the RK register is
// then loaded with the others after the face work: 12 VGPRs fewer, the same
// time on the 96^3 x 6 Taylor-Green mesh)
#ifndef FVH_K2F_PRE
#define FVH_K2F_PRE 2
#endif
#ifndef FVH_K2F_LATE
#define FVH_K2F_LATE 1
#endif
template <int FLUX, bool FS, bool VISC>
__global__ __launch_bounds__(kTile, FVH_K2F_MINW) void fv_residual_face_kernel(DevMesh m,
                                                                              PhysDev P,
                                                                              StageParams sp)
{
    if (sp.sc && sp.sc->done)
        return;
    extern __shared__ double ct[]; // [kTile * 4][5]
    const int N = m.N;
    // tiles in runs: the hardware hands consecutive workgroups to the 8 XCDs
    // in turn, so all XCDs work on one compact region of the mesh at a time
    // (5% faster than XCD-contiguous tile ranges on the 96^3 x 6 Taylor-Green
    // mesh; batched cases: each case done in a short window), each XCD on
    // FVH_TILE_GROUP adjacent tiles of a run (tile_of_block)
    const int t = sp.tile_begin + tile_of_block((int)blockIdx.x, (int)gridDim.x);
    const int base = m.tile_c0 ? m.tile_c0[t] : t * kTile;
    const int ncell = m.tile_c0 ? m.tile_c0[t + 1] - base : min(kTile, N - base);
    // per-tile timing (tile_c0 tiles only: every thread reaches the end)
    const bool stamp = sp.stamps && m.tile_c0;
    if (stamp && threadIdx.x == 0)
        sp.stamps[(size_t)t * 2] = wall_clock64();
    const int e0 = m.tile_face_off[t], e1 = m.tile_face_off[t + 1];
    // element p = lc * 5 + j of the tile's rows; with FVH_K2F_PRE the
    // update's operands are loaded before the face work so that their
    // latency overlaps it
    constexpr int kE = 5; // elements per thread: kTile * 5 / kTile
    double u_in[kE], k_in[kE];
    if (FVH_K2F_PRE)
#pragma unroll
        for (int i = 0; i < kE; ++i)
        {
            const int p = (int)threadIdx.x + i * kTile;
            const size_t g = (size_t)base * 5 + (p < ncell * 5 ? p : 0);
            u_in[i] = (sp.do_update && FVH_K2F_PRE == 1) ? sp.U_in[g] : 0.0;
            k_in[i] = (sp.do_update && !sp.stage0) ? sp.k[g] : 0.0;
        }
    // the next face's record is loaded while this face is computed
    int e = e0 + (int)threadIdx.x;
    unsigned long long q_next = m.tile_faces[e < e1 ? e : e0];
    for (; e < e1; e += kTile)
    {
        const int4 r = unpack_tile_face(q_next, base);
        if (e + kTile < e1)
            q_next = m.tile_faces[e + kTile];
        const int f = r.x;
        const bool ghost = r.z >= N * 4;
        Frame fr;
        if (m.tangents_derived)
        {
            // 32 B per face instead of 80: the tangents follow from n
            const double2 *g = reinterpret_cast<const double2 *>(m.fgN + (size_t)f * 4);
            const double2 a = g[0], b = g[1];
            fr.n[0] = a.x, fr.n[1] = a.y, fr.n[2] = b.x;
            fr.area = b.y;
            tangents_from_normal(fr.n, fr.t1, fr.t2);
        }
        else
            load_frame(m, f, fr);
        double L[10], R[10];
        if (FS)
        {
            const double2 *pl = reinterpret_cast<const double2 *>(sp.fs + (size_t)r.y * 8);
            const double2 *pr = reinterpret_cast<const double2 *>(sp.fs + (size_t)r.z * 8);
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const double2 a = pl[q], b = pr[q];
                L[2 * q] = a.x;
                L[2 * q + 1] = a.y;
                R[2 * q] = b.x;
                R[2 * q + 1] = b.y;
            }
            L[8] = L[9] = R[8] = R[9] = 0.0;
        }
        else
        {
            // first order: the two cells' values (a ghost row on a boundary
            // face), as the lane-per-slot kernel takes them
            load_phi(P, sp.U_in, nullptr, r.y >> 2, L);
            if (ghost)
            {
                const double *gp = sp.gphi + (size_t)(r.z - N * 4) * 10;
#pragma unroll
                for (int j = 0; j < 10; ++j)
                    R[j] = gp[j];
            }
            else
                load_phi(P, sp.U_in, nullptr, r.z >> 2, R);
        }
        ViscGrad gL, gR;
        if (VISC)
        {
            // the left cell's row, and the right cell's (the left one's
            // stands in for a ghost, whose gradient follows from it)
            load_vg(sp.vg, r.y >> 2, gL);
            load_vg(sp.vg, ghost ? (r.y >> 2) : (r.z >> 2), gR);
            if (ghost)
            {
                const int type = m.region_type[m.bnd_region[r.z - N * 4]];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    ghost_grad(type, fr, gL.v[q], gL.v[4 + q], gL.v[8 + q], gR.v[q],
                               gR.v[4 + q], gR.v[8 + q]);
            }
        }
        double flux[5];
        face_flux<FLUX, VISC>(P, fr, L, R, gL, gR, flux);
        const int dA = r.w & 0xffff, dB = (r.w >> 16) & 0xffff;
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            const double v = flux[j] * fr.area;
            if (dA != 0xffff)
                ct[dA * 5 + j] = v;
            if (dB != 0xffff)
                ct[dB * 5 + j] = -v;
        }
    }
    // the update's remaining operands (and the last stage's per-cell ones)
    // are loaded here, after the face work, so that their latency overlaps
    // the barrier's wait (FVH_K2F_LATE 0: loaded where they are used, each
    // wait after the barrier)
    double vol_el[kE], uin_el[kE], k_late[kE];
    double lastU[5], lastV = 1.0, lastD[3] = {0.0, 0.0, 0.0};
    const double dt_all = (sp.do_update && !sp.case_dt) ? sp.sc->dt : 0.0;
    if (FVH_K2F_LATE)
    {
#pragma unroll
        for (int i = 0; i < kE; ++i)
        {
            const int p = (int)threadIdx.x + i * kTile;
            const int lc = p < ncell * 5 ? p / 5 : 0;
            const size_t g = (size_t)base * 5 + (p < ncell * 5 ? p : 0);
            vol_el[i] = m.vol[base + lc];
            uin_el[i] = (sp.do_update && FVH_K2F_PRE != 1) ? sp.U_in[g] : 0.0;
            k_late[i] = (sp.do_update && !sp.stage0 && FVH_K2F_PRE == 0) ? sp.k[g] : 0.0;
        }
        if (sp.do_reduce && sp.do_update)
        {
            const int c = base + ((int)threadIdx.x < ncell ? (int)threadIdx.x : 0);
#pragma unroll
            for (int j = 0; j < 5; ++j)
                lastU[j] = sp.U_in[(size_t)c * 5 + j];
            lastV = m.vol[c];
            lastD[0] = m.cdx[(size_t)c * 3];
            lastD[1] = m.cdx[(size_t)c * 3 + 1];
            lastD[2] = m.cdx[(size_t)c * 3 + 2];
        }
    }
    __syncthreads();

    // per (cell, variable)
    double r_el[kE], u_el[kE];
#pragma unroll
    for (int i = 0; i < kE; ++i)
    {
        const int p = (int)threadIdx.x + i * kTile;
        const int lc = p / 5, j = p - lc * 5;
        r_el[i] = 0.0;
        u_el[i] = 0.0;
        if (lc >= ncell)
            continue;
        const int c = base + lc;
        // tetrahedra (the face tables exist for tetrahedral meshes only):
        // four slots per cell
        double a = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            a = a + ct[(lc * 4 + q) * 5 + j];
        const double Rj = -a / (FVH_K2F_LATE ? vol_el[i] : m.vol[c]);
        const size_t g = (size_t)base * 5 + p;
        r_el[i] = Rj;
        if (sp.R_out)
            sp.R_out[g] = Rj;
        if (sp.do_update)
        {
            const double dt = sp.case_dt ? sp.case_dt[sp.cell_case[c]] : dt_all;
            double kj = Rj;
            if (!sp.stage0)
                kj = Rj + (FVH_K2F_PRE ? k_in[i] : FVH_K2F_LATE ? k_late[i] : sp.k[g]) * sp.rk_a;
            const double uin = FVH_K2F_PRE == 1 ? u_in[i] : FVH_K2F_LATE ? uin_el[i] : sp.U_in[g];
            const double un = uin + kj * dt * sp.rk_g;
            u_el[i] = un;
            if (!sp.k_dead)
                sp.k[g] = kj;
            sp.U_out[g] = un;
        }
    }
    if (!sp.do_reduce)
    {
        if (stamp)
        {
            __syncthreads();
            if (threadIdx.x == 0)
                sp.stamps[(size_t)t * 2 + 1] = wall_clock64();
        }
        return;
    }

    // per cell, last stage: the new state and residual rows through LDS (the
    // contributions are no longer read), then the partials
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kE; ++i)
    {
        const int p = (int)threadIdx.x + i * kTile;
        if (p < ncell * 5)
        {
            ct[p] = u_el[i];
            ct[kTile * 5 + p] = r_el[i];
        }
    }
    __syncthreads();
    const int lc = threadIdx.x;
    const int c = base + lc;
    double part[7] = {0.0, 0.0, 0.0, 0.0, 0.0, DBL_MAX, DBL_MAX};
    if (lc < ncell && sp.do_update)
    {
        double unew[5], R5[5], pc[10];
#pragma unroll
        for (int j = 0; j < 5; ++j)
        {
            unew[j] = ct[lc * 5 + j];
            R5[j] = ct[kTile * 5 + lc * 5 + j];
            pc[j] = FVH_K2F_LATE ? lastU[j] : sp.U_in[(size_t)c * 5 + j];
        }
        // time step with the new density and the primitives of this
        // stage's input (navier_stokes.c:171-187)
        con_to_prim(P, pc);
        const double vol = FVH_K2F_LATE ? lastV : m.vol[c];
#pragma unroll
        for (int j = 0; j < 5; ++j)
            part[j] = rabs(R5[j]) * vol;
        const double dx0 = FVH_K2F_LATE ? lastD[0] : m.cdx[(size_t)c * 3],
                     dx1 = FVH_K2F_LATE ? lastD[1] : m.cdx[(size_t)c * 3 + 1],
                     dx2 = FVH_K2F_LATE ? lastD[2] : m.cdx[(size_t)c * 3 + 2];
        const double s_rho = 1.0 / unew[0];
        const double cs = sqrt(P.kappa * pc[8] * s_rho);
        const double ds1 = (rabs(pc[5]) + cs) * dx0 + (rabs(pc[6]) + cs) * dx1 +
                           (rabs(pc[7]) + cs) * dx2;
        part[5] = vol / ds1;
        const double ds2 = dot3(dx0, dx1, dx2, dx0, dx1, dx2);
        part[6] = s_rho * ds2 / vol;
    }
    // one row per wave of 64 cells: by cell over the whole mesh, or by tile
    // (every wave of the tile writes its row)
    int wrow;
    if (m.tile_c0)
        wrow = t * (kTile / 64) + (lc >> 6);
    else
    {
        wrow = (base + (lc & ~63)) / 64;
        if (base + (lc & ~63) >= N)
            return;
    }
    double w[7];
#pragma unroll
    for (int q = 0; q < 5; ++q)
        w[q] = wave_sum(part[q]);
    w[5] = wave_min(part[5]);
    w[6] = wave_min(part[6]);
    if ((lc & 63) == 0)
#pragma unroll
        for (int q = 0; q < 7; ++q)
            sp.partials[(size_t)wrow * kPartialStride + q] = w[q];
    if (stamp)
    {
        __syncthreads();
        if (threadIdx.x == 0)
            sp.stamps[(size_t)t * 2 + 1] = wall_clock64();
    }
}

// Whether the face-centred kernel runs a residual launch: FP64 face states
// over the whole tetrahedral mesh with the face tables built, the stage's
// dt from sc.
static bool use_face_kernel(const DevMesh &m, const StageParams &sp, bool fs, bool fp32)
{
    return fs && !fp32 && m.tile_faces && m.maxf == 4 && sp.c_begin == 0 && sp.c_end == m.N &&
           !sp.case_dt && sp.fs && (!sp.do_update || sp.sc);
}

template <int FLUX, bool FS, bool VISC>
static void launch_face_t(hipStream_t s, const DevMesh &m, const PhysDev &P,
                          const StageParams &sp_in, int tile_begin, int n_tiles)
{
    StageParams sp = sp_in;
    sp.tile_begin = tile_begin;
    const size_t lds = (size_t)kTile * 4 * 5 * sizeof(double);
    hipLaunchKernelGGL((fv_residual_face_kernel<FLUX, FS, VISC>), dim3(n_tiles), dim3(kTile),
                       lds, s, m, P, sp);
}

template <int FLUX, bool FS, bool VISC>
static void launch_res_t(hipStream_t s, const DevMesh &m, const PhysDev &P,
                         const StageParams &sp, bool fp32)
{
    if (FS && use_face_kernel(m, sp, FS, fp32))
    {
        launch_face_t<FLUX, true, VISC>(s, m, P, sp, 0, (m.N + kTile - 1) / kTile);
        return;
    }
    const int lanes = m.maxf <= 4 ? 4 : 8;
    const int grid = cell_blocks((sp.c_end - sp.c_begin) * lanes);
    if (fp32)
    {
        if (m.maxf <= 4)
            hipLaunchKernelGGL((fv_residual_group_kernel<FLUX, FS, VISC, 4, float>),
                               dim3(grid), dim3(kBlock), 0, s, m, P, sp);
        else
            hipLaunchKernelGGL((fv_residual_group_kernel<FLUX, FS, VISC, 8, float>),
                               dim3(grid), dim3(kBlock), 0, s, m, P, sp);
        return;
    }
    if (m.maxf <= 4)
        hipLaunchKernelGGL((fv_residual_group_kernel<FLUX, FS, VISC, 4, double>),
                           dim3(grid), dim3(kBlock), 0, s, m, P, sp);
    else
        hipLaunchKernelGGL((fv_residual_group_kernel<FLUX, FS, VISC, 8, double>),
                           dim3(grid), dim3(kBlock), 0, s, m, P, sp);
}

void launch_residual(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const StageParams &sp_in, int flux, bool face_states,
                     bool viscous, bool fp32)
{
    check_faces(m);
    StageParams sp = sp_in;
    if (sp.c_end <= 0)
    {
        sp.c_begin = 0;
        sp.c_end = m.N;
    }
    if (sp.c_begin < 0 || sp.c_end > m.N || sp.c_begin >= sp.c_end)
        throw std::runtime_error("launch_residual: bad cell range");
    if (sp.do_update && !sp.case_dt && !sp.sc)
        throw std::runtime_error("launch_residual: stage update without a time step");
    if (fp32 && (sp.do_update || sp.do_reduce))
        throw std::runtime_error("launch_residual: FP32 arithmetic is for residuals only");
    if (sp.do_reduce && cell_blocks((sp.c_end - sp.c_begin) * (m.maxf <= 4 ? 4 : 8)) >
                            sp.partials_cap)
        throw std::runtime_error("reduction partials buffer too small");
    if (flux == FluxRusanov)
    {
        if (face_states)
            viscous ? launch_res_t<FluxRusanov, true, true>(s, m, P, sp, fp32)
                    : launch_res_t<FluxRusanov, true, false>(s, m, P, sp, fp32);
        else
            viscous ? launch_res_t<FluxRusanov, false, true>(s, m, P, sp, fp32)
                    : launch_res_t<FluxRusanov, false, false>(s, m, P, sp, fp32);
    }
    else
    {
        if (face_states)
            viscous ? launch_res_t<FluxAUSM, true, true>(s, m, P, sp, fp32)
                    : launch_res_t<FluxAUSM, true, false>(s, m, P, sp, fp32);
        else
            viscous ? launch_res_t<FluxAUSM, false, true>(s, m, P, sp, fp32)
                    : launch_res_t<FluxAUSM, false, false>(s, m, P, sp, fp32);
    }
}

void launch_residual_tiles(hipStream_t s, const DevMesh &m, const PhysDev &P,
                           const StageParams &sp, int flux, bool face_states, bool viscous,
                           int tile_begin, int n_tiles)
{
    check_faces(m);
    if (m.maxf != 4 || !m.tile_faces || !m.tile_face_off)
        throw std::runtime_error("launch_residual_tiles: needs the face tables of a "
                                 "tetrahedral mesh");
    if (tile_begin < 0 || n_tiles < 1)
        throw std::runtime_error("launch_residual_tiles: bad tile range");
    if (face_states && !sp.fs)
        throw std::runtime_error("launch_residual_tiles: no face states");
    if (viscous && !sp.vg)
        throw std::runtime_error("launch_residual_tiles: no viscous gradients");
    if (sp.do_update && !sp.case_dt && !sp.sc)
        throw std::runtime_error("launch_residual_tiles: stage update without a time step");
    if (sp.stamps && !m.tile_c0)
        throw std::runtime_error("launch_residual_tiles: per-tile stamps need tile_c0 tiles");
    if (sp.do_reduce && (!m.tile_c0 || (size_t)(tile_begin + n_tiles) * (kTile / 64) >
                                           (size_t)sp.partials_cap))
        throw std::runtime_error("launch_residual_tiles: reduction needs tile rows "
                                 "(tile_c0) and a partials buffer of kTile / 64 rows per tile");
    if (flux == FluxRusanov)
    {
        if (face_states)
            viscous ? launch_face_t<FluxRusanov, true, true>(s, m, P, sp, tile_begin, n_tiles)
                    : launch_face_t<FluxRusanov, true, false>(s, m, P, sp, tile_begin, n_tiles);
        else
            viscous ? launch_face_t<FluxRusanov, false, true>(s, m, P, sp, tile_begin, n_tiles)
                    : launch_face_t<FluxRusanov, false, false>(s, m, P, sp, tile_begin, n_tiles);
    }
    else
    {
        if (face_states)
            viscous ? launch_face_t<FluxAUSM, true, true>(s, m, P, sp, tile_begin, n_tiles)
                    : launch_face_t<FluxAUSM, true, false>(s, m, P, sp, tile_begin, n_tiles);
        else
            viscous ? launch_face_t<FluxAUSM, false, true>(s, m, P, sp, tile_begin, n_tiles)
                    : launch_face_t<FluxAUSM, false, false>(s, m, P, sp, tile_begin, n_tiles);
    }
}

void launch_dt_partials(hipStream_t s, const DevMesh &m, const PhysDev &P,
                        const double *U, double *partials, int partials_cap,
                        const double *prims)
{
    if (dt_partials_grid(m) > partials_cap)
        throw std::runtime_error("reduction partials buffer too small");
    hipLaunchKernelGGL(fv_dt_partials_kernel, dim3(dt_partials_grid(m)),
                       dim3(kBlock), 0, s, m, P, U, partials, prims);
}


void launch_mirror(hipStream_t s, const DevMesh &m, const PhysDev &P,
                   const double *U, const double *R, double *phi_total_aos,
                   double *phi_dt_aos, const double *U_prims)
{
    hipLaunchKernelGGL(fv_mirror_kernel, dim3(cell_blocks(m.N + m.B)),
                       dim3(kBlock), 0, s, m, P, U, R, phi_total_aos,
                       phi_dt_aos, U_prims ? U_prims : U);
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
