// Device-side data structures and kernel launchers of libfv3d_hip.so
// (internal; the public C-ABI is include/fv3d_solver.h + fv3d_hip.h).
//
// HBM layout (structure of arrays, FP64 values, int32 indices):
//   cell arrays [N]   : U[5][N] (ping-pong pair), k[5][N] (RK register),
//                       vg[12][N] (gradients of rho*u, rho*v, rho*w and T
//                       per direction, row d*4 + q), vol, svol = 1/vol,
//                       dx[3][N], ncf, cf[maxf][N] (face ids per slot,
//                       ascending, -1 padded)
//   face-state array  : fs[2][8][F]: the reconstructed (limited) state of
//                       variables 0..7 on side 0 (cells[0]) and side 1
//                       (cells[1] or the ghost) of every face
//   ghost rows        : gphi[10][B] (boundary-condition states)
//   test-only outputs : grad[30][N] ((dim*10 + var) rows), mnmx[20][N]
//   face arrays [F]   : fc0, fc1 (fc1 >= N is ghost row N + boundary),
//                       n, t1, t2, d1 = x_f - x_c0, d2 = x_f - x_c1 image
//                       ([3][F] each), area, lambda
//   boundary arrays   : bnd_region[B]; region_type[R], region_phi[R*10]
#pragma once

#include <hip/hip_runtime.h>

namespace fvh {
namespace dev {

struct DevMesh
{
    int N, F, B, maxf;
    const int *__restrict__ cf;
    const int *__restrict__ ncf;
    const int *__restrict__ fc0;
    const int *__restrict__ fc1;
    const double *__restrict__ fn;  // [3*F]
    const double *__restrict__ ft1; // [3*F]
    const double *__restrict__ ft2; // [3*F]
    const double *__restrict__ fd1; // [3*F]
    const double *__restrict__ fd2; // [3*F]
    const double *__restrict__ farea;
    const double *__restrict__ flam;
    const double *__restrict__ vol;
    const double *__restrict__ svol;
    const double *__restrict__ cdx; // [3*N]
    const int *__restrict__ bnd_region;
    const int *__restrict__ bnd_face;
    const int *__restrict__ region_type;
    const double *__restrict__ region_phi;
};

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
};

// Outputs of the gradient/reconstruction kernel (null pointers are skipped).
struct GradOut
{
    double *vg;   // [12*N] viscous gradient subset
    double *fs;   // [16*F] face states
    double *grad; // [30*N] all gradients (tests)
    double *mnmx; // [20*N] neighbourhood min/max (tests)
};

struct StageParams
{
    const double *U_in;
    double *U_out;
    double *k;
    double *R_out;         // may be null
    const double *vg;      // [12*N] when viscous
    const double *fs;      // [16*F] face states (reconstructed path)
    const double *gphi;    // [10*B] ghost rows (first-order path)
    double rk_a, rk_g;
    int stage0;            // first stage: k = R
    int do_update;         // RK stage update into U_out / k
    int do_reduce;         // write residual and dt partials (last stage)
    DevScalars *sc;        // dt, done flag
    double *partials;      // [partials_cap * 8]
    int partials_cap;      // rows available in partials
};

constexpr int kBlock = 256;
constexpr int kPartialStride = 8;

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

// Ghost rows gphi[10][B] of the current state (and side 1 of the boundary
// faces in fs when fs is not null).  No-op without boundaries.
void launch_ghost(hipStream_t s, const DevMesh &m, const PhysDev &P,
                  const double *U, double *gphi, double *fs,
                  const DevScalars *sc);
// linear: limited linear face states; otherwise the cell values.
void launch_gradient(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const double *U, const double *gphi, const GradOut &out,
                     int linear, const DevScalars *sc);
// face_states: read the face states written by launch_gradient; otherwise
// (first order) take the cell values of U directly.
void launch_residual(hipStream_t s, const DevMesh &m, const PhysDev &P,
                     const StageParams &sp, int flux, bool face_states,
                     bool viscous);
void launch_dt_partials(hipStream_t s, const DevMesh &m, const PhysDev &P,
                        const double *U, double *partials, int partials_cap);
// Host-mirror rows: AoS phi_total[(N+B)*10] (primitives from con_to_prim,
// ghost rows from the boundary conditions) and phi_dt[N*5] (R may be null).
void launch_mirror(hipStream_t s, const DevMesh &m, const PhysDev &P,
                   const double *U, const double *R, double *phi_total_aos,
                   double *phi_dt_aos);
// mode 0: initial dt only; mode 1: end of step bookkeeping + next dt
void launch_finalize(hipStream_t s, DevScalars *sc, const double *partials,
                     int n_partials, const PhysDev &P, int mode,
                     double *log, int log_cap);

} // namespace dev
} // namespace fvh
