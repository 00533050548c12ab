// Device-side data structures and kernel launchers of libfv3d_hip.so
// (internal; the public C-ABI is include/fv3d_solver.h + fv3d_hip.h).
//
// HBM layout.  Everything a stencil kernel gathers through an index is a
// per-element record, so one gather touches one or two cache lines; FP64
// values, int32 indices.
//   cell records      : U[N][5] (ping-pong pair), k[N][5] (RK register),
//                       R[N][5] (dU/dt), vg[N][12] (gradients of rho*u,
//                       rho*v, rho*w and T per direction, d*4 + q),
//                       cdx[N][3], cx[N][3] (centroids), cf[N][maxf]
//                       (face ids, ascending, -1
//                       padded); vol[N], svol[N] = 1/vol, ncf[N]
//   face-state rows   : fs[N*maxf + B][8] (row c*maxf + s: cell c's
//                       reconstructed state of variables 0..7 on its face
//                       slot s; row N*maxf + b: the ghost state of boundary
//                       b); nsl[N*maxf] gives the slot of the same face in
//                       the neighbour's list
//   face records      : fgA[F][10] (n, t1, t2, area: flux kernels),
//                       fgB[F][6] (n, area, lambda, pad: gradient
//                       kernels), fgN[F][4] (n, area) with flam[F]
//                       (lambda) for the tile gradient kernel, and fgN
//                       alone for the face-centred residual kernel when
//                       every face's stored t1 and t2 follow from n by the
//                       builders' formula; fsh[F] (index into shtab[][3],
//                       the exact periodic shifts; 0 = none), fcell[F][2]
//                       (cells[1] >= N is the ghost row N + boundary)
//   boundary records  : gphi[B][10] (ghost rows), bnd_region[B],
//                       bnd_face[B]; region_type[R], region_phi[R][10]
//   test-only outputs : grad[N][30] (d*10 + var), mnmx[N][20]
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace fvh {
namespace dev {

constexpr int kFgA = 10; // doubles per flux-kernel face record
constexpr int kFgB = 6;  // doubles per gradient-kernel face record

// A face's tangents from its unit normal n, as the mesh builders form them:
// t1 = a / |a| with a = (-n_z, 0, n_x) when |n_x| > |n_y|, else
// (0, -n_z, n_y); t2 = b / |b| with b = t1 x n; |a| = sqrt((a_x a_x +
// a_y a_y) + a_z a_z), each component divided by it.  One definition for
// the host, which checks a mesh's stored tangents against it, and for the
// kernels that form the tangents instead of loading them.
__host__ __device__ inline void tangents_from_normal(const double *n, double *t1, double *t2)
{
    const double ax = n[0] < 0 ? -n[0] : n[0], ay = n[1] < 0 ? -n[1] : n[1];
    const bool xs = ax > ay;
    const double x = xs ? -n[2] : 0.0, y = xs ? 0.0 : -n[2], z = xs ? n[0] : n[1];
    const double l1 = sqrt(x * x + y * y + z * z);
    t1[0] = x / l1;
    t1[1] = y / l1;
    t1[2] = z / l1;
    const double bx = t1[1] * n[2] - t1[2] * n[1];
    const double by = t1[2] * n[0] - t1[0] * n[2];
    const double bz = t1[0] * n[1] - t1[1] * n[0];
    const double l2 = sqrt(bx * bx + by * by + bz * bz);
    t2[0] = bx / l2;
    t2[1] = by / l2;
    t2[2] = bz / l2;
}

struct DevMesh
{
    int N, F, B, maxf;
    const int *__restrict__ cf;     // [N*maxf]
    const int *__restrict__ ncf;    // [N]
    const int *__restrict__ fcell;  // [F*2]
    // per cell face slot: {face, v} with v = neighbour when the cell is the
    // face's first cell, else -(neighbour + 1); slots past the cell's face
    // count repeat slot 0 [N*maxf]
    const int2 *__restrict__ cfn;
    // tetrahedral meshes: per slot the face id, with the sign bit set when
    // the cell is the face's second cell (cfn's side), [N*4] (null: other
    // meshes)
    const int *__restrict__ cfo;
    // per cell face slot: the slot that the same face has in the neighbour's
    // list (0 on boundary faces and past the cell's face count) [N*maxf]
    const unsigned char *__restrict__ nsl;
    // per cell face slot: the distance vector from the cell centre to the
    // face centre (x_f - x_c; on side 1 to the periodic image of the cell,
    // x_f - (x_c + shift)), as the reconstruction uses it [N*maxf][3]
    const double *__restrict__ crv;
    const double *__restrict__ fgA; // [F*kFgA]
    const double *__restrict__ fgB; // [F*kFgB]
    // [F][4] n, area and [F] lambda (null: fgA / fgB only)
    const double *__restrict__ fgN;
    const double *__restrict__ flam;
    // 1 when t1 and t2 of every face are what tangents_from_normal forms
    // from n (checked bit for bit at set-up): the face-centred residual
    // kernel then reads fgN instead of fgA
    int tangents_derived;
    const int *__restrict__ fsh;    // [F]
    const double *__restrict__ shtab; // [nsh*3]
    const double *__restrict__ cx;  // [N*3]
    const double *__restrict__ vol;
    const double *__restrict__ svol;
    const double *__restrict__ cdx; // [N*3]
    const int *__restrict__ bnd_region;
    const int *__restrict__ bnd_face;
    const int *__restrict__ region_type;
    const double *__restrict__ region_phi;
    // FP32 face records of the FP32 residual (built on first use):
    // fgA32 [F][12] = n, t1, t2, area, 2 pad; fgB32 [F][8] = n, area,
    // lambda, 3 pad; crv32 [N*maxf][4] = the slot distance vectors, 1 pad;
    // each value the FP32 rounding of the FP64 one
    const float *__restrict__ fgA32;
    const float *__restrict__ fgB32;
    const float *__restrict__ crv32;
    // cell tiles of the LDS-staged gradient kernel (null: not built): tile t
    // is cells [t*kTile, (t+1)*kTile); its halo, the face neighbours outside
    // the tile, is tile_halo[tile_halo_off[t] .. tile_halo_off[t+1]); nbl
    // [N*maxf] gives each slot's neighbour as a row of the tile's LDS table
    // (0..kTile-1 the tile's own cells, kTile + k halo cell k, 0xffff a
    // boundary face); tile_rows = kTile + the largest halo
    const int *__restrict__ tile_halo_off;
    const int *__restrict__ tile_halo;
    const unsigned short *__restrict__ nbl;
    int tile_rows;
    // faces of each cell tile for the face-centred residual kernel (null: not
    // built; tetrahedral meshes of fewer than 2^26 faces and 2^27 face-state
    // rows): tile t's faces are tile_faces[tile_face_off[t] ..
    // tile_face_off[t+1]), each face once, in 8 bytes: bits 0-25 the face,
    // 26-34 a tile-local slot a = (cell - tile's first cell) * 4 + s; bit 63
    // clear: both cells in the tile, bits 35-43 the right cell's slot b, the
    // left state's row tile base * 4 + a and the right one's base * 4 + b;
    // bit 63 set: one side in the tile (bit 35: 0 left, 1 right), its row
    // base * 4 + a, the other side's row in bits 36-62 (N*4 + b: ghost b).
    // The left slot receives +flux*area, the right one -flux*area.
    const int *__restrict__ tile_face_off;
    const unsigned long long *__restrict__ tile_faces;
    // first cell of each face-table tile, [tiles + 1] (null: tile t is cells
    // [t*kTile, (t+1)*kTile)); set for tiles that follow other boundaries
    // (batched cases: no tile spans two cases), each at most kTile cells
    const int *__restrict__ tile_c0;
};

// cells per tile of the LDS-staged gradient kernel and the face-centred
// residual kernel (a multiple of 64)
#ifndef FVH_KTILE
#define FVH_KTILE 128
#endif
constexpr int kTile = FVH_KTILE;
static_assert(kTile % 64 == 0 && kTile <= 1024, "cell tile: whole waves, one workgroup");

struct PhysDev
{
    double kappa, kappa_m1, s_kappa_m1, s_kappa_p1, R, mu, lam;
    double cfl, dfl, kappa_pr;
    int visc_dt_mode;
    int bj; // Barth-Jespersen limiter on (linear reconstruction)
};

// Device scalars of the time loop (one instance in device memory).
struct DevScalars
{
    double t, dt, t_end, abort_residual, vglob;
    double res[5];
    int iter, max_iter_abs, done, final_pending;
    int transient, viscous, nan_flag, log_base;
    int host_dt; // the host sets dt before each step (a time-step hook
                 // that is not the built-in one); the finalize keeps it
};

// Outputs of the gradient/reconstruction kernel (null pointers are skipped).
struct GradOut
{
    double *vg;   // [N*12] viscous gradient subset
    double *fs;   // [N*maxf + B][8] face-state rows
    float *fs32;  // the same rows in FP32 (the FP32 kernels write these)
    double *grad; // [N*30] all gradients (tests)
    double *mnmx; // [N*20] neighbourhood min/max (tests)
    int c_begin, c_end; // cells [c_begin, c_end); c_end = 0: all cells
    const float *U32;    // input of the FP32 kernel: the state in FP32 (null: U)
};

struct StageParams
{
    const double *U_in;
    double *U_out;
    double *k;
    double *R_out;         // may be null
    const double *vg;      // [N*12] when viscous
    const double *fs;      // [N*maxf + B][8] face-state rows (reconstructed path)
    const float *fs32;     // the same rows in FP32 (the FP32 kernels read these)
    const double *gphi;    // [B*10] ghost rows (first-order path)
    double rk_a, rk_g;
    int stage0;            // first stage: k = R
    int do_update;         // RK stage update into U_out / k
    int do_reduce;         // write residual and dt partials (last stage)
    int k_dead;            // k is not read after this stage: not stored
    DevScalars *sc;        // dt, done flag
    double *partials;      // [partials_cap * 8]
    int partials_cap;      // rows available in partials
    int c_begin, c_end;    // cells [c_begin, c_end); c_end = 0: all cells
    const double *case_dt; // batched cases: dt of case cell_case[c] (sc unused)
    const int *cell_case;
    const float *U32;       // the FP32 kernel's input state in FP32 (null: U_in)
    int tile_begin;         // face-centred kernel: first tile of the launch
    // face-centred kernel over tile_c0 tiles: wall-clock start and end of
    // tile t at stamps[2t], stamps[2t + 1] (null: not recorded)
    unsigned long long *stamps;
};

constexpr int kBlock = 256;
constexpr int kPartialStride = 8;

// Work queue of the batched stage kernel: tiles of consecutive cells, each
// inside one case, taken by persistent workgroups through an atomic counter
// (the counter only schedules: each tile's results do not depend on which
// workgroup computes it).  Per case: flux kind, viscous flag, mu, lambda.
struct BatchTiles
{
    int tile_base, n_tiles; // the launch takes tiles tile_base .. tile_base + n_tiles - 1
    int sub_tiles;          // runs of block / lanes cells per tile
    const int *tile_case;  // [all tiles]
    const int *tile_begin; // [all tiles] first cell of the tile (tiles end with their case)
    const int *case_end;   // [K] one past the case's last cell
    const double *case_mu, *case_lam; // [K]
    unsigned int *counter; // zero before the launch; null: one workgroup per tile
    unsigned long long *stamps; // [all tiles][2] wall-clock start/end per tile (null: off)
};

inline int cell_blocks(int N) { return (N + kBlock - 1) / kBlock; }

// Workgroups of a launch_residual grid (one partials row each when the
// launch reduces) and of launch_dt_partials.
int residual_grid(const DevMesh &m);
int dt_partials_grid(const DevMesh &m);

enum FluxKind
{
    FluxAUSM = 0,
    FluxRusanov = 1
};

// Ghost rows gphi[B][10] of the current state (and side 1 of the boundary
// faces in fs when fs is not null).  No-op without boundaries.
void launch_ghost(hipStream_t s, const DevMesh &m, const PhysDev &P,
                  const double *U, double *gphi, double *fs,
                  const DevScalars *sc, float *fs32 = nullptr,
                  const float *U32 = nullptr);
// linear: limited linear face states; otherwise the cell values.
// fp32: the arithmetic in FP32 (inputs and outputs stay FP64 arrays), for
// the FP32 Jacobian-vector products of the Newton-Krylov solve.
void launch_gradient(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const double *U, const double *gphi, const GradOut &out,
                     int linear, const DevScalars *sc, bool fp32 = false);
// face_states: read the face states written by launch_gradient; otherwise
// (first order) take the cell values of U directly.
void launch_residual(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const StageParams &sp, int flux, bool face_states,
                     bool viscous, bool fp32 = false);
// The face-centred residual kernel over face-table tiles tile_begin ..
// tile_begin + n_tiles - 1 of m (tetrahedral meshes; m.tile_c0 gives the
// tiles' cells).  Without face_states the face states are the cell values
// (first order); sp.case_dt / sp.cell_case give per-case time steps; with
// do_reduce, tile t writes partials rows t * (kTile / 64) .. + kTile / 64 - 1,
// one per 64 of its cells (identity rows past its last cell).
void launch_residual_tiles(hipStream_t s, const DevMesh &m, const PhysDev &P,
                           const StageParams &sp, int flux, bool face_states, bool viscous,
                           int tile_begin, int n_tiles);
// The RK stage update of StageParams (U_out, k and, with do_reduce, the
// partials of cell_blocks(N) workgroups) from the residual R [N][5].
void launch_stage_update(hipStream_t s, const DevMesh &m, const PhysDev &P,
                         const StageParams &sp, const double *R);
// prims: the primitive rows [N][5] to use with U's density (null: from U).
void launch_dt_partials(hipStream_t s, const DevMesh &m, const PhysDev &P,
                        const double *U, double *partials, int partials_cap,
                        const double *prims = nullptr);
// One RK stage of batched first-order cases (StageParams with case_dt /
// cell_case; partials per tile when do_reduce) over the tiles of bt, all of
// cases with flux `flux` and viscosity on or off (`viscous`): `block`
// threads per workgroup (128, 256 or 512; a tile is bt.sub_tiles runs of
// block / lanes cells), `grid` persistent workgroups taking tiles from the
// work queue bt.counter, or with bt.counter null one workgroup per tile.
void launch_batch_stage(hipStream_t s, const DevMesh &m, const PhysDev &P,
                        const StageParams &sp, const BatchTiles &bt, int flux, bool viscous,
                        int block, int grid);
// Host-mirror rows: AoS phi_total[(N+B)*10] (conservative columns of U,
// primitives by con_to_prim of U_prims and ghost rows by the boundary
// conditions from U_prims; U_prims null: U) and phi_dt[N*5] (R may be null).
void launch_mirror(hipStream_t s, const DevMesh &m, const PhysDev &P,
                   const double *U, const double *R, double *phi_total_aos,
                   double *phi_dt_aos, const double *U_prims = nullptr);
// Rows of the first-level fold of n_partials partials rows.
int fold_rows(int n_partials);
// Folds the partials (into fold, fold_cap rows) and runs the time-loop
// bookkeeping.  mode 0: initial dt only; mode 1: end of step + next dt.
void launch_finalize(hipStream_t s, DevScalars *sc, const double *partials,
                     int n_partials, double *fold, int fold_cap,
                     const PhysDev &P, int mode, double *log, int log_cap);

} // namespace dev
} // namespace fvh
