/*
 * fv3d_solver.h — drop-in Solver API of the MI355X residual/time-stepping
 * library (libfv3d_hip.so).
 *
 * Every symbol here keeps the name, signature and meaning of the reference's
 * Solver library (src/Solver/include/solver/solver_module.h, types in
 * src/Solver/include/solver/solver_type.h), so an FV3D-style application
 * links against this library unchanged.  The host structs below are the
 * host mirror: the device SoA copies are authoritative while timedisc() or
 * finite_volume_time_derivative() run, and the mirror is refreshed before a
 * host hook or write_output needs it.
 *
 * The reference's external base library (basec) supplies string_t,
 * cstring_t, bool_t and hsize_t; they are defined here directly.
 */
#ifndef FV3D_SOLVER_H
#define FV3D_SOLVER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef char *string_t;
typedef const char *cstring_t;
typedef int bool_t;
typedef unsigned long long hsize_t;

#define BC_TRUE 1
#define BC_FALSE 0

#define SMALL 1e-16  /* solver_type.h:15 */
#define FACE_CELLS 2 /* solver_type.h:16 */
#define DIM 3        /* solver_type.h:17 */

/* solver_type.h:22-39 */
typedef struct Boundaries
{
    int n_global_boundaries;
    int n_boundaries;
    int max_boundary_vertices;

    int *id;
    int *type;
    int *n_vertices;
    int *vertices;
    int *face;
    double *distance;
    double *n;
    double *t1;
    double *t2;

    hsize_t *stride;
} Boundaries_t;

/* solver_type.h:44-63 */
typedef struct Cells
{
    int n_global_cells;
    int n_local_cells;
    int n_domain_cells;
    int max_cell_vertices;
    int max_cell_faces;

    int *id;
    int *type;
    int *n_vertices;
    int *vertices;
    int *n_faces;
    int *faces;
    double *x;
    double *volume;
    double *dx;

    hsize_t *stride;
} Cells_t;

/* solver_type.h:68-95 */
typedef struct Faces
{
    int n_global_faces;
    int n_faces;
    int max_face_vertices;

    int *type;
    int *n_vertices;
    int *vertices;
    int *cells;
    int *boundary;
    double *area;
    double *lambda;
    double *x;
    double *n;
    double *t1;
    double *t2;

    hsize_t *stride;

    double *dist_cell_1;
    double *dist_cell_2;

    int n_internal_faces;
    int *internal_faces;
    int n_boundary_faces;
    int *boundary_faces;
} Faces_t;

/* solver_type.h:100-121 (multi-rank partitions; unused on one GPU) */
typedef struct Partition
{
    int n_partitions;
    int n_partition_cells;
    int n_partition_boundaries;
    int n_partition_faces;
    int n_partition_sends;
    int n_partition_receives;

    int *partition_cells;
    int *partition_boundaries;
    int *partition_faces;
    int *partition_sends;
    int *partition_sends_pid;
    int *partition_receives;
    int *partition_receives_pid;

    int *n_partition_sends_to;
    int *partition_sends_to;
    int *n_partition_receives_from;
    int *partition_receives_from;
} Partition_t;

/* solver_type.h:126-139 */
typedef struct Regions
{
    int n_regions;
    int max_name_length;

    string_t *name;
    int *is_boundary;

    int flow_region;

    int *type;
    int *function_id;
    double *phi_total;
} Regions_t;

/* solver_type.h:144-149 */
typedef struct Vertices
{
    int n_vertices;

    double *x;
} Vertices_t;

/* solver_type.h:154-168 */
typedef struct Mesh
{
    int dimension;
    int is_partitioned;

    Partition_t *partition;
    Vertices_t *vertices;
    Cells_t *cells;
    Boundaries_t *boundaries;
    Faces_t *faces;
    Regions_t *regions;

    double local_volume;
    double global_volume;
} Mesh_t;

/* solver_type.h:173-191 */
typedef struct Variable
{
    string_t name;
} Variable_t;

typedef struct Variables
{
    int n_sol_variables;
    Variable_t *sol_variables;

    int n_dep_variables;
    Variable_t *dep_variables;

    int n_tot_variables;
    Variable_t **tot_variables;
} Variables_t;

/* solver_type.h:196-207 */
typedef struct Data
{
    double *flux;
    double *phi_total_left;
    double *phi_total_right;

    double *grad_phi_total_x;
    double *grad_phi_total_y;
    double *grad_phi_total_z;
    double *phi_dt;
    double *phi_total;
} data_t;

/* solver_type.h:212-262 */
typedef enum LimiterType
{
    BarthJespersenn,
    NoLimter
} limiter_type_t;

typedef enum ReconstructionType
{
    FirstOrder,
    Linear
} reconstruction_type_t;

typedef enum TimediscType
{
    Explicit,
    Implicit
} timedisc_type_t;

typedef enum ExplicitScheme
{
    EulerExplicit,
    RungeKutta33,
    RungeKutta45
} explicit_scheme_t;

typedef enum ImplicitScheme
{
    EulerImplicit,
    BDF2,
} implicit_scheme_t;

typedef enum ImplicitSolver
{
    BiCGStab,
    GMRes,
} implicit_solver_t;

/* solver_type.h:264-268 */
typedef void (*void_calc_exact_ft)(int id, double t, double *x, double *phi);
typedef void (*void_calc_flux_ft)(double time);
typedef double (*double_calc_timestep_ft)(double time);
typedef void (*void_update_ft)(double time);
typedef void (*void_update_gradients_ft)(double time);

/* solver_module.h:16-18 — host mirror globals */
extern Mesh_t *solver_mesh;
extern Variables_t *solver_variables;
extern data_t *solver_data;

/* solver_module.h:25-32 */
int add_sol_variable(cstring_t name);
int add_dep_variable(cstring_t name);
/* solver_module.h:38 — residual of the last stage, into the host mirror */
void calc_global_residual(double dt);
/* solver_module.h:44 */
void create_file_header(cstring_t file_name);
/* solver_module.h:50 — one residual evaluation on the GPU; refreshes
 * solver_data->phi_total (primitives) and solver_data->phi_dt */
void finite_volume_time_derivative(double time);
/* solver_module.h:55-110 */
void free_analyze(void);
void free_equation(void);
void free_explicit(void);
void free_finite_volume(void);
void free_implicit(void);
void free_limiter(void);
void free_mesh(void);
void free_reconstruction(void);
void free_output(void);
void free_restart(void);
void free_solver(void);
void free_timedisc(void);
/* solver_module.h:115-202 */
cstring_t get_simulation_title(void);
void init_analyze(void);
void init_equation(void);
void init_explicit(explicit_scheme_t explicit_scheme);
void init_finite_volume(void);
void init_implicit(implicit_scheme_t implicit_scheme,
                   implicit_solver_t implicit_solver,
                   int max_iter_inner,
                   double tolerance_lsoe, int max_iter_lsoe,
                   int max_krylov_dims, int max_krylov_restarts);
void init_limiter(limiter_type_t limiter_type);
void init_mesh(cstring_t mesh_file);
void init_reconstruction(reconstruction_type_t reconstruction_type);
void init_output(int i_output_data);
void init_restart(bool_t use_restart, cstring_t restart_file);
void init_solver(cstring_t title);
void init_timedisc(timedisc_type_t timedisc_type, int max_iter,
                   bool_t is_transient, double abort_residual,
                   double t_start, double t_end);
/* solver_module.h:208-273 */
bool_t is_explicit(void);
void print_mesh_info(void);
void print_variables(void);
void read_restart_data(cstring_t restart_file);
void set_calc_flux(void_calc_flux_ft fun_ptr);
void set_calc_timestep(double_calc_timestep_ft fun_ptr);
void set_exact_function(void_calc_exact_ft fun_ptr);
void set_update(void_update_ft fun_ptr);
void set_update_gradients(void_update_gradients_ft fun_ptr);
void set_viscous_dt(bool_t is_viscous_dt);
void timedisc(void);
void write_output(int iter, double t);

/*
 * Navier–Stokes physics plug-in (the reference keeps it in the FV3D
 * application, src/FV3D/src/navier_stokes/).  This library ships it as the
 * built-in device hooks: registering these function pointers with the
 * set_* calls above selects the GPU kernels.  Any other flux, timestep,
 * update or gradient-update hook is rejected (the library has no host
 * fallback for the hot path); a custom exact function is accepted, since
 * it is evaluated once on the host for the initial condition.
 */
/* parameters, navier_stokes.c:16-24 (defaults as there) */
extern double cfl_scale;
extern double dfl_scale;
extern double mu_mix;
extern double R_mix;
extern double Pr;
extern double kappa;
/* derived constants, navier_stokes.c:96-104 */
extern double kappa_m1;
extern double kappa_p1;
extern double s_kappa;
extern double s_kappa_m1;
extern double s_kappa_p1;
extern double cv;
extern double cp;
extern double kappa_pr;
extern double lambda;
/* variable indices, navier_stokes.c:234-244 */
extern int ic_rho, ic_rho_u, ic_rho_v, ic_rho_w, ic_rho_e;
extern int ip_u, ip_v, ip_w, ip_p, ip_T;

/* navier_stokes_module.h:47-59 */
enum BoundaryType
{
    BoundaryFlow,
    BoundaryInflow,
    BoundaryOutflow,
    BoundaryAdiabaticWall,
    BoundaryIsothermalWall,
    BoundarySlipWall,
    BoundarySymmetry,
    BoundaryState,
    BoundaryFunction,
    BoundaryTypeMax
};

/* navier_stokes.c:85-124 without the JSON glue: derives the constants from
 * the parameter globals, registers the 5 + 5 variables and the hooks */
void navier_stokes_initialize(void);
/* boundary.c:55-139 without the JSON glue: one region's type and primitive
 * state (u, v, w, p, T); only T is used for isothermal walls */
void boundary_set_region(int region, int type, const double *prim5);
/* flux.c:295-312 without the JSON glue: "AUSM" or "Rusanov" */
void flux_initialize_scheme(cstring_t flux_scheme_name);

/* navier_stokes_module.h:83-243 hooks (built-ins) */
void calc_ns_flux(double time);
double calc_ns_timestep(double time);
void calc_ns_exact_function(int id, double t, double *x, double *phi);
void ns_update(double time);
void ns_update_gradients(double time);

#ifdef __cplusplus
}
#endif

#endif /* FV3D_SOLVER_H */
