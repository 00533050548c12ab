/*
 * fv3d_hip.h — extension API (fvh_*) of libfv3d_hip.so.
 *
 * The Solver API of fv3d_solver.h is the drop-in boundary; the functions
 * here have no counterpart in the reference and serve the things the
 * reference does not have: mesh generation with periodic pairing, direct
 * device state access for tests, device-side stepping control and the
 * measurement hooks.  All return 0 on success and a negative status on
 * error (fvh_last_error() gives the message), unlike the reference-named
 * functions, which abort on error as the reference does.
 */
#ifndef FV3D_HIP_H
#define FV3D_HIP_H

#include "fv3d_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FVH_OK 0
#define FVH_ERR_ARG -1
#define FVH_ERR_STATE -2
#define FVH_ERR_HIP -3
#define FVH_ERR_IO -4
#define FVH_ERR_NAN -5

const char *fvh_last_error(void);

/* ---- mesh generation (replaces reading a pyMeshFV3D file) ------------- */
/* Structured box [x0,x0+lx]x[y0,y0+ly]x[z0,z0+lz] with nx*ny*nz hexahedra,
 * optionally Kuhn-split into 6 tetrahedra each (tets != 0).  periodic is a
 * bit mask (1 = x, 2 = y, 4 = z); non-periodic sides become boundary
 * regions "xmin","xmax","ymin","ymax","zmin","zmax" (in that order, after
 * the flow region "flow").  tile > 0 orders the hexahedra in tile^3 blocks
 * for cache locality (0 = lexicographic x-fastest).  Builds solver_mesh
 * exactly as init_mesh does (reader layout + calc_mesh_metrics). */
int fvh_mesh_box(int nx, int ny, int nz, double x0, double y0, double z0,
                 double lx, double ly, double lz, int tets, int periodic,
                 int tile);

/* Bump channel x in [-2,2], y in [h(x),1], h = 0.1*exp(-25 x^2),
 * z in [0,0.5]; nx*ny*nz hexahedra Kuhn-split into tetrahedra; interior
 * nodes jittered uniformly by +-jitter*spacing per axis (splitmix64 with
 * the given seed).  Regions: "flow","inflow","outflow","bottom","top",
 * "zmin","zmax". */
int fvh_mesh_channel(int nx, int ny, int nz, double jitter,
                     unsigned long long seed);

/* Read a Gmsh 4.1 ASCII mesh (tetrahedra, hexahedra, prisms, pyramids as
 * cells; triangles and quadrangles as boundaries; physical names as
 * regions), scale the vertices by (sx, sy, sz) and preprocess it as the
 * reference's pyMeshFV3D does (replaces tools/pyMeshFV3D/pymeshfv3d/reader/
 * gmsh.py:35-139 + process.py:23-214 followed by init_mesh of the written
 * file).  Builds solver_mesh like fvh_mesh_box; fvh_mesh_write saves it. */
int fvh_mesh_gmsh(cstring_t path, double sx, double sy, double sz);

/* Write the mesh in the reference's HDF5 layout (path ending in .h5; libhdf5
 * 1.10 is loaded at run time) or in a raw little-endian container with the
 * same groups and datasets (any other name).  init_mesh() reads either,
 * recognised by the file signature.  Both add FACES/periodic_shift [F*3]. */
int fvh_mesh_write(cstring_t path);
/* 1 when libhdf5 can be loaded (FV3D_HDF5_LIB, the loader path or the conda
 * copy), else 0. */
int fvh_hdf5_available(void);
/* Format of write_output: 0 HDF5 when available, else raw (default);
 * 1 raw container (<title>.%09d.fvs); 2 HDF5 (<title>.%09d.h5). */
int fvh_set_output_format(int format);
/* Periodic translation of each face's cells[1] image, [F*3] (zeros on
 * non-periodic faces).  Faces_t has no field for it. */
int fvh_get_face_shift(double *shift);

/* ---- physics parameters not carried by the reference's globals -------- */
/* visc_dt_mode 0: reference dt formula verbatim (dt = 0 when mu_mix = 0);
 * 1: skip the viscous limit when mu_mix == 0. */
int fvh_set_visc_dt_mode(int mode);
/* Built-in initial conditions usable with set_exact_function:
 * kind 0 = density wave rho = a + b*sin(2*pi*(x+y+z)), u=v=w=c, p = d
 *          (params a, b, c, d);
 * kind 1 = Taylor-Green vortex on [0,2pi]^3 (params p0, rho0, U0) with
 *          rho = p/(R_mix*T0), T0 = p0/(rho0*R_mix);
 * kind 2 = uniform "flow" region state (the reference's exact function). */
int fvh_set_ic(int kind, const double *params, int n_params);
void fvh_exact_function_builtin(int id, double t, double *x, double *phi);

/* ---- state access ------------------------------------------------------ */
/* Conservative state of the domain cells, AoS [N*5]. */
int fvh_get_state(double *U);
int fvh_set_state(const double *U);
/* Last residual (phi_dt) of the domain cells, AoS [N*5]. */
int fvh_get_phi_dt(double *R);
/* Device fields for tests: what 0 = gradients [N*30] as (dim, var) =
 * g[c*30 + d*10 + j]; 1 = neighbourhood min/max [N*20] (min j, max j). */
int fvh_get_field(int what, double *out);
/* Refresh the host mirror solver_data->phi_total/phi_dt from the device. */
int fvh_sync_host(void);

/* ---- stepping ---------------------------------------------------------- */
/* Print per-step residual lines from timedisc() (default 1). */
int fvh_set_verbose(int verbose);
/* Steps per host synchronisation inside timedisc() (default 16). */
int fvh_set_sync_interval(int steps);
/* timedisc() returning a status instead of aborting: FVH_OK, or
 * FVH_ERR_NAN when the residual became NaN/Inf (the reference aborts). */
int fvh_timedisc(void);
/* Per-step log of the last timedisc(): rows of (t, dt, viscous, res[5]);
 * returns the number of rows copied (<= max_rows). */
int fvh_get_log(double *rows, int max_rows);
/* Status of the last timedisc(): iterations done and 0 / FVH_ERR_NAN. */
int fvh_last_status(int *iterations);

/* Time the explicit step loop: runs n_steps after n_warmup, returns the
 * elapsed seconds of the timed steps (device events on the library's
 * stream) and the per-kernel average durations in microseconds
 * (k_us[0] gradient kernel, k_us[1] residual/stage kernel, k_us[2]
 * finalize), measured with HIP events in a separate, instrumented pass. */
int fvh_bench_explicit(int n_steps, int n_warmup, double *seconds,
                       double *k_us);

/* ---- allocation-free residual ------------------------------------------- */
/* dU/dt of the state U (device memory, [N*5] AoS, the domain cells) into R
 * (device memory, [N*5]): finite_volume_time_derivative's arithmetic on the
 * current mesh and physics, with the built-in hooks, enqueued on `stream`
 * (a hipStream_t; NULL = the library's stream) without synchronising.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops).  It
 * allocates nothing: the intermediate ghost rows, face states and viscous
 * gradients live in a fixed 64 MiB device array (meshes up to about 190,000
 * viscous second-order tetrahedra; larger meshes are rejected with
 * FVH_ERR_STATE and a message).  Every call on a device shares that
 * device's array: calls that overlap on one GPU race. */
int fvh_time_derivative_noalloc(const double *U, double *R, void *stream);

/* ---- device information ----------------------------------------------- */
int fvh_device_info(char *name, int name_len, int *cu_count,
                    double *hbm_gbps_peak);

/* ---- steady Newton-Krylov (JFNK) solve --------------------------------- */
/* Drives the current device state to the steady state R(U) = 0 of the
 * configured residual (finite_volume_time_derivative semantics) with
 * pseudo-transient Newton steps (I/dtau - dR/dU) dU = R(U), dtau the local
 * convective step at the pseudo CFL, each solved by right-preconditioned
 * GMRES(krylov_dims) with a matrix-free finite-difference operator
 * J v = [R(U + eps z) - R(U)] / eps, eps = fd_eps (1 + |U|_2) / |z|_2.
 * No counterpart in the reference (its implicit mode is the lumped-Jacobian
 * block solve of init_implicit / timedisc). */
#define FVH_JFNK_FP64 0      /* Krylov vectors and J v in FP64 */
#define FVH_JFNK_FP32 1      /* Krylov vectors in FP32 and J v from two FP32-
                                arithmetic residuals (step fd_eps32); Newton
                                residuals, dots and the update in FP64 */
#define FVH_PRECOND_SCALAR 0 /* M = I / dtau */
#define FVH_PRECOND_BLOCK 1  /* block Jacobi: exact 5x5 blocks of the
                                first-order inviscid residual */
#define FVH_PRECOND_LEARNED 2 /* 5x5 blocks predicted by the trained cell
                                 network; block Jacobi where rejected */
#define FVH_COUPLING_NONE 0  /* the preconditioner is block-diagonal */
#define FVH_COUPLING_DILU 1  /* multicolour block ILU(0) of the first-order
                                inviscid Jacobian A1 = I/dtau - dR1/dU: the
                                precond blocks on the diagonal, face-
                                neighbour blocks by differences of the shared
                                face flux, cells ordered by a distance-1
                                colouring; M agrees with A1 on every diagonal
                                and face-neighbour block.  The solve stores
                                the factors in FP32 (in both precisions) */

typedef struct
{
    int max_newton;      /* Newton steps (default 300) */
    double tol;          /* stop when max_j (1/V) sum_i |R_ij| V_i <= tol
                            (default 1e-8) */
    double cfl0;         /* initial pseudo CFL (default 10) */
    double cfl_growth;   /* CFL factor after a successful step (default 1.5) */
    double cfl_cut;      /* CFL factor after a step that raised the stop norm
                            above 3x the previous one (default 0.3) */
    double cfl_stall;    /* GMRES relative residual above which a step counts
                            as stalled (default 0.5; > 1 disables) */
    double cfl_stall_cut; /* CFL factor after a stalled step (default 0.7) */
    double cfl_max;      /* (default 1e6) */
    int krylov_dims;     /* GMRES restart length m (default 30, <= 60) */
    int max_restarts;    /* GMRES cycles per Newton step minus one (default 3) */
    double rtol;         /* GMRES relative tolerance (default 1e-2) */
    double fd_eps;       /* finite-difference scale (default 1e-7) */
    double fd_eps32;     /* the same for FP32 J v (default 3e-4 ~ sqrt of the
                            FP32 unit roundoff) */
    double max_change;   /* damping: largest relative change of rho and
                            rho*E per step (default 0.2) */
    int precision;       /* FVH_JFNK_FP64 / FVH_JFNK_FP32 */
    int precond;         /* FVH_PRECOND_* (default FVH_PRECOND_BLOCK) */
    int verbose;         /* print one line per Newton step */
    int learned_check;   /* FVH_PRECOND_LEARNED: compare the predicted blocks
                            with the block-Jacobi blocks every learned_check
                            Newton steps (and at the first); 0 = first only */
    double learned_tol;  /* mean per-cell relative Frobenius error above which
                            the solve falls back to block Jacobi (default 0.2) */
    int coupling;        /* FVH_COUPLING_* (default FVH_COUPLING_NONE); DILU
                            needs precond BLOCK or LEARNED */
    double reorth;       /* a second Gram-Schmidt pass runs when the new
                            Krylov vector kept at most this fraction of its
                            square norm (default 0.1; 0.5 is the 1/sqrt(2)
                            test of Daniel, Gragg, Kaufman and Stewart; 0
                            never) */
} fvh_jfnk_options_t;

typedef struct
{
    int status;           /* 0 converged, 1 Newton step limit reached */
    int newton_steps;
    int gmres_iterations; /* total Arnoldi steps */
    int residual_evals;   /* full residual evaluations */
    double seconds;       /* wall time of the Newton loop */
    double norm0, norm;   /* stop norm at the start and at the end */
    double l2_drop;       /* |R|_2 end / start */
    int precond_fallbacks; /* learned -> block-Jacobi fallbacks (0 or 1) */
    double learned_error;  /* last measured mean block error of the network */
    double precond_seconds; /* wall time spent building preconditioners */
} fvh_jfnk_stats_t;

int fvh_jfnk_default_options(fvh_jfnk_options_t *opt);
int fvh_jfnk_solve(const fvh_jfnk_options_t *opt, fvh_jfnk_stats_t *stats);
/* Per Newton step rows {norm, cfl, gmres iterations, gmres relative
 * residual, damping factor, elapsed seconds}; returns the row count. */
int fvh_jfnk_history(double *rows, int max_rows);
/* The preconditioner of one Newton step at the current state: the local
 * pseudo time steps dtau[N] at the given CFL and the inverse blocks
 * minv[N][5][5] (row-major) of M_c = I/dtau_c - dR1_c/dU_c (precond
 * FVH_PRECOND_BLOCK) or of I/dtau_c (FVH_PRECOND_SCALAR). */
int fvh_jfnk_blocks(double cfl, int precond, double *dtau, double *minv);
/* z = M^-1 v (host arrays of 5N values) for the preconditioner that a Newton
 * step at the current state and the given CFL builds with precond / coupling
 * (FVH_PRECOND_LEARNED needs a network; it is applied without the
 * block-error check), with the factors kept in FP64. */
int fvh_jfnk_precond_apply(double cfl, int precond, int coupling, const double *v, double *z);

/* ---- batched cases ------------------------------------------------------ */
/* Many independent cases in one device state, advanced by one launch per
 * stage (see fvh_batch_set_schedule): n_cases periodic boxes [0,length]^3 of n_hex[k]^3 hexahedra,
 * each Kuhn-split into 6 tetrahedra, stored as one mesh with the cases'
 * cells in order (case k owns cells offsets[k] .. offsets[k+1]).  Builds
 * solver_mesh like fvh_mesh_box; the usual init sequence follows
 * (navier_stokes_initialize ... init_finite_volume, first-order
 * reconstruction, init_explicit), then fvh_batch_setup. */
int fvh_mesh_batch(int n_cases, const int *n_hex, double length, int tile);
/* Per-case flux (0 AUSM, 1 Rusanov), viscous flag and viscosity; cases
 * with equal (flux, viscous, mu) must be adjacent (they form one launch
 * group).  Computes every case's first time step from the current state. */
int fvh_batch_setup(int n_cases, const int *flux, const int *viscous,
                    const double *mu);
/* Advances every case by `steps` steps of the explicit scheme, each case
 * with its own time step (cfl_scale, dfl_scale of the globals); use_graph
 * replays a captured HIP graph of one step.  seconds: device time. */
int fvh_batch_run(int steps, int use_graph, double *seconds);
int fvh_batch_info(int *n_cases, int *offsets, int max_offsets);
/* Step schedule: persistent = 0 (the default) launches the stage kernel per
 * (flux, viscous) group of cases and reduces the per-case norms in a pass
 * of its own; persistent = 2 (tetrahedral cases) launches the face-centred
 * stage kernel per group over tiles of at most 128 cells of one case (each
 * face's flux evaluated once per tile) and reduces in its last stage
 * (block, grid and sub_tiles unused); persistent = 1 runs each stage as one persistent launch per
 * (flux, viscous) class of cases (workgroups of `block` = 128, 256 or 512
 * threads take tiles of sub_tiles * block / lanes cells of one case from a
 * work queue, lanes = 4 on tetrahedra; `grid` workgroups, 0 = min(tiles,
 * 4 x CUs), -1 = one workgroup per tile and no queue) and reduces in its
 * last stage.  Both run one gradient launch
 * per run of viscous cases and give the same states and time steps. */
int fvh_batch_set_schedule(int persistent, int block, int grid, int sub_tiles);
/* Per-case step time inside the batched launches (persistent or face
 * schedule; the stage kernel's launches, not the gradient launches):
 * runs `steps` steps and returns in ms[K] each case's mean per step of the
 * time from the start of its first tile to the end of its last, summed over
 * the stages (device wall clock). */
int fvh_batch_case_step_times(int steps, double *ms);
/* Per case: next dt [K], time [K], residual norms of the last step [K*5]
 * (analyze.c norms with the case's volume and dt). */
int fvh_batch_case_scalars(double *dt, double *t, double *res);

/* ---- learned block preconditioner (network of FVH_PRECOND_LEARNED) ---- */
/* A per-cell MLP 28 -> 64 -> 64 -> 25 (tanh) maps local features of the
 * state (normalised primitives, their scaled Green-Gauss gradients, face
 * geometry moments, boundary face count, Mach number) to the dimensionless
 * first-order diagonal Jacobian block K_c = -(V_c/ds1_c) dR1_c/dU_c
 * (normalised by the flow-region state), from which the block-Jacobi block
 * (I + cfl K_c)/dtau_c is inverted exactly for any pseudo CFL.  BF16 MFMA kernels with FP32 accumulation run inference and
 * training; master weights and Adam are FP32.  The parameters are one FP32
 * vector of fvh_nn_param_count() values (W1[32][64], b1[64], W2[64][64],
 * b2[64], W3[64][32], b3[32], row-major [in][out]; input rows 28..31 and
 * output columns 25..31 are padding). */
int fvh_nn_param_count(void);
int fvh_nn_init(unsigned long long seed);
int fvh_nn_get_params(float *theta);
int fvh_nn_set_params(const float *theta);
/* Features [N][32] of the current state at the given pseudo CFL. */
int fvh_nn_features(double cfl, float *feat);
/* Appends the features and block-Jacobi targets of the current state at
 * the given CFL to the device training set; n_total receives its size. */
int fvh_nn_collect(double cfl, long long *n_total);
int fvh_nn_clear(void);
/* Adam training over the training set, mini-batches in shuffled order;
 * loss[epochs] receives the mean batch loss per epoch. */
int fvh_nn_train(int epochs, int batch, double lr, unsigned long long seed,
                 double *loss);
/* Network output [n][32] for host features [n][32] (inference kernel). */
int fvh_nn_forward(const float *feat, int n, float *out);
/* Loss and its gradient [param_count] for host features and targets
 * [n][32] (the training kernel's forward and backward pass). */
int fvh_nn_loss_grad(const float *feat, const float *target, int n,
                     float *grad, double *loss);
/* Max and mean per-cell relative Frobenius error of the network's inverse
 * blocks against the block-Jacobi blocks at the current state. */
int fvh_nn_validate(double cfl, double *err);

#ifdef __cplusplus
}
#endif

#endif /* FV3D_HIP_H */
