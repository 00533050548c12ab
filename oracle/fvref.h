/*
 * fvref — CPU oracle for the fv3d residual and time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C99 restatement of the
 * reference solver's arithmetic (src/Solver/src/... and src/FV3D/src/...),
 * kept in the reference's own data model (array-of-structs phi_total[N x 10],
 * face-state arrays, function-pointer limiter/Riemann/reconstruction hooks and
 * separate face loops) so that it is both the parity checker and a faithful
 * CPU cost baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product (fv3d_amd/) never does.
 *
 * Pinning status of each part is listed in fvref.c's header comment.
 */
#ifndef FVREF_H
#define FVREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fvref fvref_t;

/* Region (boundary) types, same numbering as the reference's BoundaryType
 * enum (src/FV3D/src/navier_stokes/navier_stokes_module.h:47-59). */
enum
{
    FVREF_BC_FLOW = 0,
    FVREF_BC_INFLOW = 1,
    FVREF_BC_OUTFLOW = 2,
    FVREF_BC_WALL_ADIABATIC = 3,
    FVREF_BC_WALL_ISOTHERMAL = 4,
    FVREF_BC_WALL_SLIP = 5,
    FVREF_BC_SYMMETRY = 6,
    FVREF_BC_STATE = 7,
    FVREF_BC_FUNCTION = 8
};

/* Mesh as stored in the reference's mesh file (Py/write.py layout):
 * face_cells[F*2] holds -1 as second cell of boundary faces; face_shift[F*3]
 * (may be NULL) is the periodic translation of cells[1]'s image. */
fvref_t *fvref_create(int n_cells, int max_cell_faces, int n_faces,
                      int n_boundaries, int n_regions,
                      const int *cell_n_faces, const int *cell_faces,
                      const double *cell_x, const double *cell_volume,
                      const double *cell_dx,
                      const int *face_cells, const int *face_boundary,
                      const double *face_x, const double *face_n,
                      const double *face_t1, const double *face_t2,
                      const double *face_area, const double *face_lambda,
                      const double *face_shift,
                      const int *boundary_id, const int *boundary_face,
                      const int *region_is_boundary);
void fvref_destroy(fvref_t *s);

/* Physics/scheme parameters (reference keys Equation/Navier-Stokes/...,
 * FV/Reconstruction, FV/Limiter).  flux: 0 = AUSM, 1 = Rusanov.
 * recon: 0 = first order, 1 = linear.  limiter: 0 = Barth-Jespersenn,
 * 1 = none.  visc_dt_mode: 0 = reference formula verbatim, 1 = skip the
 * viscous limit when mu == 0 (otherwise the reference gives dt = 0). */
void fvref_set_physics(fvref_t *s, double cfl_scale, double dfl_scale,
                       double mu_mix, double R_mix, double Pr, double kappa,
                       int flux, int recon, int limiter, int visc_dt_mode);

/* Region type and primitive state (u, v, w, p, T); the conservative part is
 * derived exactly as boundary_initialize does. */
void fvref_set_region(fvref_t *s, int region, int type, const double *prim5);

/* Set conservative state of the domain cells (U[N*5]) and run the update hook
 * (con_to_prim + ghost cells), as set_solution does after the IC. */
void fvref_set_state(fvref_t *s, const double *U);
/* The time step of fvref_timedisc_explicit from a time-step hook that
 * returns dt (> 0), instead of calc_ns_timestep (0). */
void fvref_set_fixed_dt(fvref_t *s, double dt);

/* One residual evaluation (finite_volume_time_derivative). */
void fvref_time_derivative(fvref_t *s, double t);

/* The timestep hook (calc_ns_timestep): returns dt; *viscous gets the flag. */
double fvref_calc_timestep(fvref_t *s, int *viscous);

/* Global residual (calc_global_residual) into res[5]. */
void fvref_global_residual(fvref_t *s, double dt, double *res);

/* Explicit timedisc loop (timedisc + timestep_lserkw2).  scheme: 0 Euler,
 * 1 RK3-3, 2 RK4-5.  Writes per-step log rows (t, dt, viscous, res[5]) into
 * log[max_iter*8] if non-NULL.  Returns the number of steps done, or -1 on a
 * NaN/Inf residual (the reference aborts there). */
int fvref_timedisc_explicit(fvref_t *s, int scheme, int max_iter,
                            int transient, double abort_residual,
                            double t_start, double t_end, double *log);

/* Amplification factor of the LSERK loop on y' = lambda*y, z = lambda*dt. */
double fvref_lserk_amplification(int scheme, double z);

/* Reference-semantics implicit step (timestep_newton): BDF1/BDF2 with the
 * lumped finite-difference block Jacobian; the block-diagonal linear system
 * is solved by the Krylov method set with fvref_set_lsoe (the reference's
 * external basec solvers, restated in fvref.c).  old_levels [2*N*5] is the BDF history (phi_old[0],
 * phi_old[1]) the caller keeps across steps; phi_dt must hold R of the
 * current state.  Returns the inner iterations (-2: max_iter_inner
 * reached, where the reference aborts). */
int fvref_timestep_newton(fvref_t *s, int scheme, int iter, double t,
                          double dt, int transient, int max_iter_inner,
                          double *old_levels, double *jac_out);

/* The implicit step's linear solve (implicit.c:285-300): solver 0
 * BiCGStab, 1 restarted GMRES(dims) with `restarts` restarts (the default,
 * with tol 1e-12, max_iter 100, dims 15, restarts 2), -1 the exact per-cell
 * block solve.  The solve stops at |f - J dY| <= tol |f| or max_iter
 * iterations; its previous solution is the initial guess of the next one.
 * fvref_lsoe_iterations: the iterations of the last solve. */
void fvref_set_lsoe(fvref_t *s, int solver, double tol, int max_iter, int dims,
                    int restarts);
int fvref_lsoe_iterations(const fvref_t *s);

/* timedisc() with the implicit step (scheme 0 Euler, 1 BDF2); log rows as
 * fvref_timedisc_explicit.  Returns the steps done, -1 on NaN/Inf, -2 when
 * the inner Newton loop ran out. */
int fvref_timedisc_implicit(fvref_t *s, int scheme, int max_iter,
                            int transient, double abort_residual,
                            double t_start, double t_end, int max_iter_inner,
                            double *log);

/* Field access.  what: 0 phi_total[(N+B)*10], 1 phi_dt[(N+B)*5],
 * 2 grad_x, 3 grad_y, 4 grad_z [(N+B)*10], 5 flux[F*5], 6 left[F*10],
 * 7 right[F*10], 8 dist_cell_1[F*3], 9 dist_cell_2[F*3],
 * 10 face cells after metrics [F*2] (as double). */
void fvref_get(const fvref_t *s, int what, double *out);
int fvref_n_elements(const fvref_t *s);
double fvref_global_volume(const fvref_t *s);

/* Stand-alone physics helpers for unit tests (use the parameters set by
 * fvref_set_physics). */
void fvref_con_to_prim(fvref_t *s, double *phi10);
void fvref_prim_to_con(fvref_t *s, double *phi10);
double fvref_riemann_p(fvref_t *s, const double *phi10);
void fvref_riemann(fvref_t *s, int flux, const double *phi_l10,
                   const double *phi_r10, double *f5);

#ifdef __cplusplus
}
#endif

#endif /* FVREF_H */
