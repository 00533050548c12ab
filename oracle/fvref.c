/*
 * fvref — CPU oracle for the fv3d residual and time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY (see fvref.h).  A C99 restatement of the
 * reference's arithmetic in the reference's data model: AoS phi_total with
 * 10 doubles per row (5 conservative + 5 primitive), ghost rows after the
 * domain cells, face-state arrays, function-pointer hooks and one loop per
 * face pass.  Every function cites the reference file:line it follows
 * (paths relative to the reference root).
 *
 * Compiled like the reference's Release build: gcc -O3 -std=c99, no -march,
 * no OpenMP, with -ffp-contract=off stated explicitly.
 *
 * Pinning (what checks this restatement independently of itself; the tests
 * in tests/test_oracle.py carry them out):
 *   con_to_prim/prim_to_con  round trip to 1 ulp, textbook ideal gas values
 *   riemann (Rusanov, AUSM)  consistency F(U,U) = F_n(U); supersonic AUSM
 *                            upwinding; L<->R mirror symmetry
 *   riemann_p                isentropic closed form at u_n <= 0
 *   calc_gradients           exact for linear fields on interior cells of a
 *                            uniform periodic hex box; zero for constants
 *   limiter_barth_jespersenn face value inside the neighbourhood [min,max];
 *                            brute force on hand stencils
 *   time derivative          free-stream preservation; exact conservation of
 *                            sum V*U on periodic meshes; hand-summed fluxes
 *   viscous flux             zero for mu = 0; hand-computed tau for linear
 *                            momentum fields
 *   calc_ns_timestep         closed form on a uniform hex box
 *   LSERK tables             stability polynomial on y' = lambda*y (RK3-3:
 *                            1+z+z^2/2+z^3/6; RK4-5 order 4)
 *   global residual          hand sums
 *   lumped Jacobian          brute-force finite differences on small meshes
 * The external dependency basec (not vendored, version unpinned) provides
 * dot_n/len_n/sum_n; they are assumed to sum sequentially from 0.0, left to
 * right.  basec's Krylov solvers are not restated: the reference's implicit
 * system is block diagonal, so it is solved here exactly per cell
 * (parity unpinned for the Krylov iterates themselves).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#include "fvref.h"

#define FACE_CELLS 2
#define DIM 3
#define SMALL 1e-16
#define NTOT 10
#define NSOL 5

/* variable indices, registration order of navier_stokes.c:234-244 */
#define IC_RHO 0
#define IC_RHO_U 1
#define IC_RHO_V 2
#define IC_RHO_W 3
#define IC_RHO_E 4
#define IP_U 5
#define IP_V 6
#define IP_W 7
#define IP_P 8
#define IP_T 9

#define REF_MIN(a, b) ((a) < (b) ? (a) : (b))
#define REF_MAX(a, b) ((a) > (b) ? (a) : (b))
#define REF_ABS(a) ((a) < 0 ? -(a) : (a))

typedef double (*limiter_ft)(fvref_t *s, int i_cell, int i_var, double slope);
typedef void (*riemann_ft)(fvref_t *s, double *phi_l, double *phi_r, double *f);
typedef void (*reconstruction_ft)(fvref_t *s, double t);

struct fvref
{
    /* mesh (Si/solver_type.h:22-168 layout) */
    int n_cells, max_cell_faces, n_faces, n_boundaries, n_regions;
    int *cell_n_faces, *cell_faces;
    double *cell_x, *cell_volume, *cell_dx;
    int *face_cells, *face_boundary;
    double *face_x, *face_n, *face_t1, *face_t2, *face_area, *face_lambda;
    double *face_shift;
    double *dist_cell_1, *dist_cell_2;
    int n_internal_faces, n_boundary_faces;
    int *internal_faces, *boundary_faces;
    int *boundary_id, *boundary_face;
    int *region_is_boundary;
    int flow_region;
    double global_volume;

    /* regions (boundary.c:61-64) */
    int *region_type;
    double *region_phi;

    /* data (finite_volume.c:37-52) */
    double *flux, *phi_left, *phi_right;
    double *grad_x, *grad_y, *grad_z;
    double *phi_dt, *phi_total;

    /* physics (navier_stokes.c:16-24, 96-104) */
    double cfl_scale, dfl_scale, mu_mix, R_mix, Pr, kappa;
    double kappa_m1, kappa_p1, s_kappa, s_kappa_m1, s_kappa_p1;
    double cv, cp, kappa_pr, lambda;
    double fixed_dt; /* > 0: a time-step hook returning this constant */
    /* the implicit step's linear solve (implicit.c:285-300): solver 0
     * BiCGStab, 1 GMRES(dims) with restarts, -1 the exact block solve;
     * tolerance relative to |f|, iteration limit; dY is kept across calls
     * (the reference's dY_n buffer is the next solve's initial guess) */
    int lsoe_solver, lsoe_max_iter, lsoe_dims, lsoe_restarts;
    double lsoe_tol;
    double *lsoe_dY;
    int lsoe_last_iter;
    int visc_dt_mode;
    int is_viscous_dt;

    limiter_ft limiter;
    riemann_ft riemann;
    reconstruction_ft reconstruction;
};

/* basec vector helpers, assumed semantics (sequential from 0.0) */
static double dot_n(const double *a, const double *b, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        s += a[i] * b[i];
    return s;
}

static void copy_n(const double *a, int n, double *b)
{
    for (int i = 0; i < n; ++i)
        b[i] = a[i];
}

static void set_value_n(double v, int n, double *a)
{
    for (int i = 0; i < n; ++i)
        a[i] = v;
}

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p)
        abort();
    return p;
}

/* ------------------------------------------------------------------------ */
/* EOS: src/FV3D/src/navier_stokes/eos.c                                      */
/* ------------------------------------------------------------------------ */

/* eos.c:33-36 */
static double calc_ig_rho(double p, double T, double R)
{
    return p / (R * T);
}

/* eos.c:45-48 */
static double calc_ig_T(double p, double rho, double R)
{
    return p / (rho * R);
}

/* eos.c:55-75.  The compression branch uses A = 2/((kappa+1) u_n) where
 * the textbook shock relation has 2/((kappa+1) rho); kept for parity. */
static double calc_riemann_p(fvref_t *s, const double *phi)
{
    if (phi[IP_U] <= 0.0)
    {
        double c = sqrt(s->kappa * phi[IP_P] / phi[IC_RHO]);
        return phi[IP_P] *
               pow(REF_MAX(1e-4, 1. + 0.5 * s->kappa_m1 * phi[IP_U] / c),
                   2 * s->kappa * s->s_kappa_m1);
    }
    else
    {
        double ar = 2 * s->s_kappa_p1 / phi[IP_U];
        double br = s->kappa_m1 * s->s_kappa_p1 * phi[IP_P];
        return phi[IP_P] +
               phi[IP_U] / ar * 0.5 *
                   (phi[IP_U] + sqrt(phi[IP_U] * phi[IP_U] +
                                     4. * ar * (phi[IP_P] + br)));
    }
}

/* eos.c:81-92 */
static void con_to_prim(fvref_t *s, double *phi)
{
    phi[IP_U] = phi[IC_RHO_U] / phi[IC_RHO];
    phi[IP_V] = phi[IC_RHO_V] / phi[IC_RHO];
    phi[IP_W] = phi[IC_RHO_W] / phi[IC_RHO];
    phi[IP_P] = s->kappa_m1 *
                (phi[IC_RHO_E] - 0.5 * dot_n(&phi[IC_RHO_U], &phi[IP_U], DIM));
    phi[IP_P] = REF_MAX(1.00E-10, phi[IP_P]);
    phi[IP_T] = calc_ig_T(phi[IP_P], phi[IC_RHO], s->R_mix);
}

/* eos.c:120-127 */
static void prim_to_con(fvref_t *s, double *phi)
{
    phi[IC_RHO_U] = phi[IP_U] * phi[IC_RHO];
    phi[IC_RHO_V] = phi[IP_V] * phi[IC_RHO];
    phi[IC_RHO_W] = phi[IP_W] * phi[IC_RHO];
    phi[IC_RHO_E] = s->s_kappa_m1 * phi[IP_P] +
                    0.5 * dot_n(&phi[IP_U], &phi[IC_RHO_U], DIM);
}

/* ------------------------------------------------------------------------ */
/* Mesh metrics: src/Solver/src/mesh/mesh.c:19-83                           */
/* ------------------------------------------------------------------------ */
static void calc_mesh_metrics(fvref_t *s)
{
    /* sum_n over domain cells (mesh.c:25) */
    double v = 0.0;
    for (int i = 0; i < s->n_cells; ++i)
        v += s->cell_volume[i];
    s->global_volume = v;

    s->n_internal_faces = 0;
    for (int i = 0; i < s->n_faces; ++i)
        if (s->face_boundary[i] < 0)
            s->n_internal_faces += 1;
    s->n_boundary_faces = s->n_faces - s->n_internal_faces;

    s->dist_cell_1 = xcalloc((size_t)s->n_faces * DIM, sizeof(double));
    s->dist_cell_2 = xcalloc((size_t)s->n_faces * DIM, sizeof(double));
    s->internal_faces = xcalloc(s->n_internal_faces, sizeof(int));
    s->boundary_faces = xcalloc(s->n_boundary_faces, sizeof(int));

    int k = 0, l = 0;
    for (int i = 0; i < s->n_faces; ++i)
    {
        int *fc = &s->face_cells[i * FACE_CELLS];
        for (int j = 0; j < DIM; ++j)
            s->dist_cell_1[i * DIM + j] =
                s->face_x[i * DIM + j] - s->cell_x[fc[0] * DIM + j];

        if (fc[1] >= 0)
        {
            /* periodic extension: the image of cells[1] sits at x + shift */
            for (int j = 0; j < DIM; ++j)
            {
                double xc1 = s->cell_x[fc[1] * DIM + j];
                if (s->face_shift)
                    xc1 = xc1 + s->face_shift[i * DIM + j];
                s->dist_cell_2[i * DIM + j] = s->face_x[i * DIM + j] - xc1;
            }
            s->internal_faces[k++] = i;
        }
        else
        {
            for (int j = 0; j < DIM; ++j)
                s->dist_cell_2[i * DIM + j] = 0.0;
            fc[1] = s->n_cells + s->face_boundary[i];
            s->boundary_faces[l++] = i;
        }
    }

    s->flow_region = -1;
    for (int i = 0; i < s->n_regions; ++i)
        if (s->region_is_boundary[i] == 0)
        {
            s->flow_region = i;
            break;
        }
}

/* ------------------------------------------------------------------------ */
/* Boundaries: src/FV3D/src/navier_stokes/boundary.c                          */
/* ------------------------------------------------------------------------ */

static void rotate_back(const double *n, const double *t1, const double *t2,
                        double *phi)
{
    double tmp_u = phi[IP_U] * n[0] + phi[IP_V] * t1[0] + phi[IP_W] * t2[0];
    double tmp_v = phi[IP_U] * n[1] + phi[IP_V] * t1[1] + phi[IP_W] * t2[1];
    double tmp_w = phi[IP_U] * n[2] + phi[IP_V] * t1[2] + phi[IP_W] * t2[2];
    phi[IP_U] = tmp_u;
    phi[IP_V] = tmp_v;
    phi[IP_W] = tmp_w;
}

/* boundary.c:188-320 */
static void update_boundaries(fvref_t *s, double t)
{
    (void)t;
    for (int i = 0; i < s->n_boundaries; ++i)
    {
        int bf = s->boundary_face[i];
        int bc = s->face_cells[bf * FACE_CELLS];
        int id = s->boundary_id[i];
        const double *n = &s->face_n[bf * DIM];
        const double *t1 = &s->face_t1[bf * DIM];
        const double *t2 = &s->face_t2[bf * DIM];
        double *phi = &s->phi_total[(s->n_cells + i) * NTOT];
        const double *phi_c = &s->phi_total[bc * NTOT];

        copy_n(phi_c, NTOT, phi);

        switch (s->region_type[id])
        {
        case FVREF_BC_INFLOW:
        case FVREF_BC_STATE:
            copy_n(&s->region_phi[id * NTOT], NTOT, phi);
            break;
        case FVREF_BC_OUTFLOW:
            break;
        case FVREF_BC_WALL_ADIABATIC:
            phi[IP_U] = dot_n(&phi_c[IP_U], n, DIM);
            phi[IP_V] = dot_n(&phi_c[IP_U], t1, DIM);
            phi[IP_W] = dot_n(&phi_c[IP_U], t2, DIM);
            phi[IP_P] = calc_riemann_p(s, phi);
            set_value_n(0.0, DIM, &phi[IP_U]);
            phi[IC_RHO] = calc_ig_rho(phi[IP_P], phi[IP_T], s->R_mix);
            rotate_back(n, t1, t2, phi);
            break;
        case FVREF_BC_WALL_ISOTHERMAL:
            phi[IP_U] = dot_n(&phi_c[IP_U], n, DIM);
            phi[IP_V] = dot_n(&phi_c[IP_U], t1, DIM);
            phi[IP_W] = dot_n(&phi_c[IP_U], t2, DIM);
            phi[IP_P] = calc_riemann_p(s, phi);
            set_value_n(0.0, DIM, &phi[IP_U]);
            phi[IP_T] = s->region_phi[id * NTOT + IP_T];
            phi[IC_RHO] = calc_ig_rho(phi[IP_P], phi[IP_T], s->R_mix);
            rotate_back(n, t1, t2, phi);
            break;
        case FVREF_BC_WALL_SLIP:
        case FVREF_BC_SYMMETRY:
            phi[IP_U] = dot_n(&phi_c[IP_U], n, DIM);
            phi[IP_V] = dot_n(&phi_c[IP_U], t1, DIM);
            phi[IP_W] = dot_n(&phi_c[IP_U], t2, DIM);
            phi[IP_P] = calc_riemann_p(s, phi);
            phi[IP_U] = 0.0;
            phi[IP_T] = calc_ig_T(phi[IP_P], phi[IC_RHO], s->R_mix);
            rotate_back(n, t1, t2, phi);
            break;
        default:
            /* "function" is unreachable in the reference (boundary.c:127 vs
             * navier_stokes.c:142-150) */
            abort();
        }

        prim_to_con(s, phi);
    }
}

/* boundary.c:325-412 */
static void update_gradients_boundaries(fvref_t *s, double t)
{
    (void)t;
    double tmp_x[NTOT], tmp_y[NTOT], tmp_z[NTOT];
    for (int i = 0; i < s->n_boundaries; ++i)
    {
        int bf = s->boundary_face[i];
        int bc = s->face_cells[bf * FACE_CELLS];
        int id = s->boundary_id[i];
        const double *n = &s->face_n[bf * DIM];
        const double *t1 = &s->face_t1[bf * DIM];
        const double *t2 = &s->face_t2[bf * DIM];
        double *gx = &s->grad_x[(s->n_cells + i) * NTOT];
        double *gy = &s->grad_y[(s->n_cells + i) * NTOT];
        double *gz = &s->grad_z[(s->n_cells + i) * NTOT];

        copy_n(&s->grad_x[bc * NTOT], NTOT, gx);
        copy_n(&s->grad_y[bc * NTOT], NTOT, gy);
        copy_n(&s->grad_z[bc * NTOT], NTOT, gz);

        switch (s->region_type[id])
        {
        case FVREF_BC_WALL_SLIP:
        case FVREF_BC_SYMMETRY:
            for (int j = 0; j < NTOT; ++j)
            {
                tmp_x[j] = gx[j] * n[0] + gy[j] * n[1] + gz[j] * n[2];
                tmp_y[j] = gx[j] * t1[0] + gy[j] * t1[1] + gz[j] * t1[2];
                tmp_z[j] = gx[j] * t2[0] + gy[j] * t2[1] + gz[j] * t2[2];
            }
            set_value_n(0.0, NTOT, tmp_x);
            for (int j = 0; j < NTOT; ++j)
            {
                gx[j] = tmp_x[j] * n[0] + tmp_y[j] * t1[0] + tmp_z[j] * t2[0];
                gy[j] = tmp_x[j] * n[1] + tmp_y[j] * t1[1] + tmp_z[j] * t2[1];
                gz[j] = tmp_x[j] * n[2] + tmp_y[j] * t1[2] + tmp_z[j] * t2[2];
            }
            break;
        default:
            break;
        }
    }
}

/* navier_stokes.c:199-209 */
static void ns_update(fvref_t *s, double t)
{
    for (int i = 0; i < s->n_cells; ++i)
        con_to_prim(s, &s->phi_total[i * NTOT]);
    update_boundaries(s, t);
}

/* ------------------------------------------------------------------------ */
/* Gradients and reconstruction: src/Solver/src/finite_volume/reconstruction.c */
/* ------------------------------------------------------------------------ */

/* reconstruction.c:20-93 (MPI halo exchanges omitted: one rank) */
static void calc_gradients(fvref_t *s, double t)
{
    int n_el = s->n_cells + s->n_boundaries;
    set_value_n(0.0, NTOT * n_el, s->grad_x);
    set_value_n(0.0, NTOT * n_el, s->grad_y);
    set_value_n(0.0, NTOT * n_el, s->grad_z);

    for (int i = 0; i < s->n_faces; ++i)
    {
        int *fc = &s->face_cells[i * FACE_CELLS];
        const double *n = &s->face_n[i * DIM];
        double area = s->face_area[i];

        for (int j = 0; j < NTOT; ++j)
        {
            double phi_mean =
                s->phi_total[fc[0] * NTOT + j] +
                s->face_lambda[i] *
                    (s->phi_total[fc[1] * NTOT + j] -
                     s->phi_total[fc[0] * NTOT + j]);

            s->grad_x[fc[0] * NTOT + j] += phi_mean * n[0] * area;
            s->grad_y[fc[0] * NTOT + j] += phi_mean * n[1] * area;
            s->grad_z[fc[0] * NTOT + j] += phi_mean * n[2] * area;

            s->grad_x[fc[1] * NTOT + j] -= phi_mean * n[0] * area;
            s->grad_y[fc[1] * NTOT + j] -= phi_mean * n[1] * area;
            s->grad_z[fc[1] * NTOT + j] -= phi_mean * n[2] * area;
        }
    }

    for (int i = 0; i < s->n_cells; ++i)
    {
        double s_volume = 1.0 / s->cell_volume[i];
        for (int j = 0; j < NTOT; ++j)
        {
            s->grad_x[i * NTOT + j] *= s_volume;
            s->grad_y[i * NTOT + j] *= s_volume;
            s->grad_z[i * NTOT + j] *= s_volume;
        }
    }

    /* update_gradients hook = ns_update_gradients (navier_stokes.c:215-218) */
    update_gradients_boundaries(s, t);
}

/* limiter.c:49-106 */
static double limiter_barth_jespersenn(fvref_t *s, int i_cell, int i_var,
                                       double slope)
{
    const int *cf = &s->cell_faces[i_cell * s->max_cell_faces];
    double phi_i = s->phi_total[i_cell * NTOT + i_var];
    double phi_min = phi_i;
    double phi_max = phi_i;

    for (int i = 0; i < s->cell_n_faces[i_cell]; ++i)
    {
        const int *fc = &s->face_cells[cf[i] * FACE_CELLS];
        if (fc[0] == i_cell)
        {
            phi_min = REF_MIN(phi_min, s->phi_total[fc[1] * NTOT + i_var]);
            phi_max = REF_MAX(phi_max, s->phi_total[fc[1] * NTOT + i_var]);
        }
        else
        {
            phi_min = REF_MIN(phi_min, s->phi_total[fc[0] * NTOT + i_var]);
            phi_max = REF_MAX(phi_max, s->phi_total[fc[0] * NTOT + i_var]);
        }
    }

    double tmp = 1.0;
    for (int i = 0; i < s->cell_n_faces[i_cell]; ++i)
    {
        double y = 1.0;
        if (slope > 0)
            y = (phi_max - s->phi_total[i_cell * NTOT + i_var]) / slope;
        else if (slope < 0)
            y = (phi_min - s->phi_total[i_cell * NTOT + i_var]) / slope;
        tmp = REF_MIN(tmp, y);
    }
    return tmp;
}

/* limiter.c:115-123 */
static double limiter_none(fvref_t *s, int i_cell, int i_var, double slope)
{
    (void)s;
    (void)i_cell;
    (void)i_var;
    (void)slope;
    return 1.0;
}

/* reconstruction.c:142-163 */
static void reconstruction_first_order(fvref_t *s, double t)
{
    ns_update(s, t);
    calc_gradients(s, t);

    for (int i = 0; i < s->n_faces; ++i)
    {
        const int *fc = &s->face_cells[i * FACE_CELLS];
        for (int j = 0; j < NTOT; ++j)
        {
            s->phi_left[i * NTOT + j] = s->phi_total[fc[0] * NTOT + j];
            s->phi_right[i * NTOT + j] = s->phi_total[fc[1] * NTOT + j];
        }
    }
}

/* reconstruction.c:168-255 */
static void reconstruction_linear(fvref_t *s, double t)
{
    ns_update(s, t);
    calc_gradients(s, t);

    for (int ii = 0; ii < s->n_internal_faces; ++ii)
    {
        int i = s->internal_faces[ii];
        const int *fc = &s->face_cells[i * FACE_CELLS];

        const double *r = &s->dist_cell_1[i * DIM];
        const double *gx = &s->grad_x[fc[0] * NTOT];
        const double *gy = &s->grad_y[fc[0] * NTOT];
        const double *gz = &s->grad_z[fc[0] * NTOT];
        for (int j = 0; j < NTOT; ++j)
        {
            double slope = r[0] * gx[j] + r[1] * gy[j] + r[2] * gz[j];
            double lim = s->limiter(s, fc[0], j, slope);
            s->phi_left[i * NTOT + j] =
                s->phi_total[fc[0] * NTOT + j] + lim * slope;
        }

        r = &s->dist_cell_2[i * DIM];
        gx = &s->grad_x[fc[1] * NTOT];
        gy = &s->grad_y[fc[1] * NTOT];
        gz = &s->grad_z[fc[1] * NTOT];
        for (int j = 0; j < NTOT; ++j)
        {
            double slope = r[0] * gx[j] + r[1] * gy[j] + r[2] * gz[j];
            double lim = s->limiter(s, fc[1], j, slope);
            s->phi_right[i * NTOT + j] =
                s->phi_total[fc[1] * NTOT + j] + lim * slope;
        }
    }

    for (int ii = 0; ii < s->n_boundary_faces; ++ii)
    {
        int i = s->boundary_faces[ii];
        const int *fc = &s->face_cells[i * FACE_CELLS];

        const double *r = &s->dist_cell_1[i * DIM];
        const double *gx = &s->grad_x[fc[0] * NTOT];
        const double *gy = &s->grad_y[fc[0] * NTOT];
        const double *gz = &s->grad_z[fc[0] * NTOT];
        for (int j = 0; j < NTOT; ++j)
        {
            double slope = r[0] * gx[j] + r[1] * gy[j] + r[2] * gz[j];
            double lim = s->limiter(s, fc[0], j, slope);
            s->phi_left[i * NTOT + j] =
                s->phi_total[fc[0] * NTOT + j] + lim * slope;
        }
        for (int j = 0; j < NTOT; ++j)
            s->phi_right[i * NTOT + j] = s->phi_total[fc[1] * NTOT + j];
    }
}

/* ------------------------------------------------------------------------ */
/* Fluxes: src/FV3D/src/navier_stokes/flux.c                                  */
/* ------------------------------------------------------------------------ */

/* flux.c:203-210 */
static void eval_euler_flux_1d(const double *phi, double *f)
{
    f[0] = phi[IC_RHO_U];
    f[1] = phi[IC_RHO_U] * phi[IP_U] + phi[IP_P];
    f[2] = phi[IC_RHO_U] * phi[IP_V];
    f[3] = phi[IC_RHO_U] * phi[IP_W];
    f[4] = (phi[IC_RHO_E] + phi[IP_P]) * phi[IP_U];
}

/* flux.c:320-336 */
static void riemann_rusanov(fvref_t *s, double *phi_l, double *phi_r, double *f)
{
    double c_l = sqrt(s->kappa * phi_l[IP_P] / phi_l[IC_RHO]);
    double c_r = sqrt(s->kappa * phi_r[IP_P] / phi_r[IC_RHO]);
    double eigval = REF_MAX(REF_ABS(phi_l[IP_U]) + c_l,
                            REF_ABS(phi_r[IP_U]) + c_r);
    double f_l[NSOL], f_r[NSOL];
    eval_euler_flux_1d(phi_l, f_l);
    eval_euler_flux_1d(phi_r, f_r);
    for (int j = 0; j < NSOL; ++j)
        f[j] = 0.5 * (f_l[j] + f_r[j]) - 0.5 * eigval * (phi_r[j] - phi_l[j]);
}

/* flux.c:344-406 */
static void riemann_ausm(fvref_t *s, double *phi_l, double *phi_r, double *f)
{
    double c_l = sqrt(s->kappa * phi_l[IP_P] / phi_l[IC_RHO]);
    double c_r = sqrt(s->kappa * phi_r[IP_P] / phi_r[IC_RHO]);
    double M_l = phi_l[IP_U] / c_l;
    double M_r = phi_r[IP_U] / c_r;
    double H_l = (phi_l[IC_RHO_E] + phi_l[IP_P]) / phi_l[IC_RHO];
    double H_r = (phi_r[IC_RHO_E] + phi_r[IP_P]) / phi_r[IC_RHO];
    double M_m, M_p, P_m, P_p;

    if (M_l <= -1.0)
    {
        M_p = 0.0;
        P_p = 0.0;
    }
    else if (M_l < 1.0)
    {
        M_p = 0.25 * (M_l + 1.0) * (M_l + 1.0);
        P_p = 0.25 * phi_l[IP_P] * (1.0 + M_l) * (1.0 + M_l) * (2.0 - M_l);
    }
    else
    {
        M_p = M_l;
        P_p = phi_l[IP_P];
    }

    if (M_r <= -1.0)
    {
        M_m = M_r;
        P_m = phi_r[IP_P];
    }
    else if (M_r < 1.0)
    {
        M_m = -0.25 * (M_r - 1.0) * (M_r - 1.0);
        P_m = 0.25 * phi_r[IP_P] * (1.0 - M_r) * (1.0 - M_r) * (2.0 + M_r);
    }
    else
    {
        M_m = 0.0;
        P_m = 0.0;
    }

    f[0] = REF_MAX(0.0, M_p + M_m) * c_l * phi_l[IC_RHO];
    f[1] = REF_MAX(0.0, M_p + M_m) * c_l * phi_l[IC_RHO] * phi_l[IP_U] + P_p;
    f[2] = REF_MAX(0.0, M_p + M_m) * c_l * phi_l[IC_RHO] * phi_l[IP_V];
    f[3] = REF_MAX(0.0, M_p + M_m) * c_l * phi_l[IC_RHO] * phi_l[IP_W];
    f[4] = REF_MAX(0.0, M_p + M_m) * c_l * phi_l[IC_RHO] * H_l;

    f[0] += REF_MIN(0.0, M_p + M_m) * c_r * phi_r[IC_RHO];
    f[1] += REF_MIN(0.0, M_p + M_m) * c_r * phi_r[IC_RHO] * phi_r[IP_U] + P_m;
    f[2] += REF_MIN(0.0, M_p + M_m) * c_r * phi_r[IC_RHO] * phi_r[IP_V];
    f[3] += REF_MIN(0.0, M_p + M_m) * c_r * phi_r[IC_RHO] * phi_r[IP_W];
    f[4] += REF_MIN(0.0, M_p + M_m) * c_r * phi_r[IC_RHO] * H_r;
}

/* flux.c:215-263 (gradient indices 1..3 are the momentum gradients, as in
 * the reference) */
static void eval_viscous_flux_1d(fvref_t *s, const double *phi,
                                 const double *gx, const double *gy,
                                 const double *gz,
                                 double *f, double *g, double *h)
{
    const double s_23 = 2.0 / 3.0;
    const double s_43 = 4.0 / 3.0;
    double mu = s->mu_mix;
    double tau_xx = mu * (s_43 * gx[1] - s_23 * gy[2] - s_23 * gz[3]);
    double tau_yy = mu * (-s_23 * gx[1] + s_43 * gy[2] - s_23 * gz[3]);
    double tau_zz = mu * (-s_23 * gx[1] - s_23 * gy[2] + s_43 * gz[3]);
    double tau_xy = mu * (gy[1] + gx[2]);
    double tau_xz = mu * (gz[1] + gx[3]);
    double tau_yz = mu * (gz[2] + gy[3]);

    f[0] = 0.0;
    f[1] = -tau_xx;
    f[2] = -tau_xy;
    f[3] = -tau_xz;
    f[4] = -tau_xx * phi[IP_U] - tau_xy * phi[IP_V] - tau_xz * phi[IP_W] -
           s->lambda * gx[IP_T];

    g[0] = 0.0;
    g[1] = -tau_xy;
    g[2] = -tau_yy;
    g[3] = -tau_yz;
    g[4] = -tau_xy * phi[IP_U] - tau_yy * phi[IP_V] - tau_yz * phi[IP_W] -
           s->lambda * gy[IP_T];

    h[0] = 0.0;
    h[1] = -tau_xz;
    h[2] = -tau_yz;
    h[3] = -tau_zz;
    h[4] = -tau_xz * phi[IP_U] - tau_yz * phi[IP_V] - tau_zz * phi[IP_W] -
           s->lambda * gz[IP_T];
}

/* flux.c:422-447 */
static void viscous_flux(fvref_t *s, const double *phi_l, const double *gx_l,
                         const double *gy_l, const double *gz_l,
                         const double *phi_r, const double *gx_r,
                         const double *gy_r, const double *gz_r,
                         double *f, double *g, double *h)
{
    double f_l[NSOL], g_l[NSOL], h_l[NSOL];
    double f_r[NSOL], g_r[NSOL], h_r[NSOL];
    eval_viscous_flux_1d(s, phi_l, gx_l, gy_l, gz_l, f_l, g_l, h_l);
    eval_viscous_flux_1d(s, phi_r, gx_r, gy_r, gz_r, f_r, g_r, h_r);
    for (int j = 0; j < NSOL; ++j)
    {
        f[j] = 0.5 * (f_l[j] + f_r[j]);
        g[j] = 0.5 * (g_l[j] + g_r[j]);
        h[j] = 0.5 * (h_l[j] + h_r[j]);
    }
}

/* flux.c:42-111 (internal) and :113-197 (boundary): the same arithmetic */
static void ns_face_flux(fvref_t *s, int i)
{
    const int *fc = &s->face_cells[i * FACE_CELLS];
    double *flux_i = &s->flux[i * NSOL];
    const double *n = &s->face_n[i * DIM];
    const double *t1 = &s->face_t1[i * DIM];
    const double *t2 = &s->face_t2[i * DIM];
    const double *pl = &s->phi_left[i * NTOT];
    const double *pr = &s->phi_right[i * NTOT];
    double lr[NTOT], rr[NTOT], flux_c[NSOL];
    double fdx[NSOL], fdy[NSOL], fdz[NSOL];

    lr[IC_RHO] = pl[IC_RHO];
    lr[IC_RHO_U] = dot_n(&pl[IC_RHO_U], n, DIM);
    lr[IC_RHO_V] = dot_n(&pl[IC_RHO_U], t1, DIM);
    lr[IC_RHO_W] = dot_n(&pl[IC_RHO_U], t2, DIM);
    lr[IC_RHO_E] = pl[IC_RHO_E];
    con_to_prim(s, lr);

    rr[IC_RHO] = pr[IC_RHO];
    rr[IC_RHO_U] = dot_n(&pr[IC_RHO_U], n, DIM);
    rr[IC_RHO_V] = dot_n(&pr[IC_RHO_U], t1, DIM);
    rr[IC_RHO_W] = dot_n(&pr[IC_RHO_U], t2, DIM);
    rr[IC_RHO_E] = pr[IC_RHO_E];
    con_to_prim(s, rr);

    s->riemann(s, lr, rr, flux_c);

    flux_i[0] = flux_c[0];
    flux_i[1] = flux_c[1] * n[0] + flux_c[2] * t1[0] + flux_c[3] * t2[0];
    flux_i[2] = flux_c[1] * n[1] + flux_c[2] * t1[1] + flux_c[3] * t2[1];
    flux_i[3] = flux_c[1] * n[2] + flux_c[2] * t1[2] + flux_c[3] * t2[2];
    flux_i[4] = flux_c[4];

    viscous_flux(s, pl, &s->grad_x[fc[0] * NTOT], &s->grad_y[fc[0] * NTOT],
                 &s->grad_z[fc[0] * NTOT], pr, &s->grad_x[fc[1] * NTOT],
                 &s->grad_y[fc[1] * NTOT], &s->grad_z[fc[1] * NTOT],
                 fdx, fdy, fdz);

    for (int j = 0; j < NSOL; ++j)
        flux_i[j] += fdx[j] * n[0] + fdy[j] * n[1] + fdz[j] * n[2];
}

/* flux.c:23-198 */
static void calc_ns_flux(fvref_t *s, double t)
{
    (void)t;
    for (int ii = 0; ii < s->n_internal_faces; ++ii)
        ns_face_flux(s, s->internal_faces[ii]);
    for (int ii = 0; ii < s->n_boundary_faces; ++ii)
        ns_face_flux(s, s->boundary_faces[ii]);
}

/* ------------------------------------------------------------------------ */
/* Residual: src/Solver/src/finite_volume/finite_volume.c:81-124              */
/* ------------------------------------------------------------------------ */
static void time_derivative(fvref_t *s, double t)
{
    s->reconstruction(s, t);
    calc_ns_flux(s, t);

    set_value_n(0.0, NSOL * (s->n_cells + s->n_boundaries), s->phi_dt);

    for (int i = 0; i < s->n_faces; ++i)
    {
        const int *fc = &s->face_cells[i * FACE_CELLS];
        double area = s->face_area[i];
        for (int j = 0; j < NSOL; ++j)
        {
            s->phi_dt[fc[0] * NSOL + j] += s->flux[i * NSOL + j] * area;
            s->phi_dt[fc[1] * NSOL + j] -= s->flux[i * NSOL + j] * area;
        }
    }

    for (int i = 0; i < s->n_cells; ++i)
    {
        double volume = s->cell_volume[i];
        for (int j = 0; j < NSOL; ++j)
            s->phi_dt[i * NSOL + j] = -s->phi_dt[i * NSOL + j] / volume;
    }
}

/* ------------------------------------------------------------------------ */
/* Timestep: src/FV3D/src/navier_stokes/navier_stokes.c:159-193               */
/* ------------------------------------------------------------------------ */
static double calc_ns_timestep(fvref_t *s)
{
    double lambda_conv = DBL_MAX;
    double lambda_visc = DBL_MAX;

    for (int i = 0; i < s->n_cells; ++i)
    {
        const double *phi = &s->phi_total[i * NTOT];
        const double *dx = &s->cell_dx[i * DIM];
        double s_rho = 1.0 / phi[IC_RHO];
        double c = sqrt(s->kappa * phi[IP_P] * s_rho);

        double ds1 = (REF_ABS(phi[IP_U]) + c) * dx[0] +
                     (REF_ABS(phi[IP_V]) + c) * dx[1] +
                     (REF_ABS(phi[IP_W]) + c) * dx[2];
        lambda_conv = REF_MIN(lambda_conv, s->cell_volume[i] / ds1);

        double ds2 = dot_n(dx, dx, DIM);
        lambda_visc = REF_MIN(lambda_visc, s_rho * ds2 / s->cell_volume[i]);
    }

    lambda_conv *= s->cfl_scale;
    lambda_visc *= s->dfl_scale * s->mu_mix * s->kappa_pr * lambda_visc;

    if (s->visc_dt_mode == 1 && s->mu_mix == 0.0)
    {
        s->is_viscous_dt = 0;
        return lambda_conv;
    }

    s->is_viscous_dt = lambda_visc < lambda_conv;
    return REF_MIN(lambda_conv, lambda_visc);
}

/* ------------------------------------------------------------------------ */
/* Analyze: src/Solver/src/analyze/analyze.c:37-61                            */
/* ------------------------------------------------------------------------ */
void fvref_global_residual(fvref_t *s, double dt, double *res)
{
    double tmp[NSOL];
    set_value_n(0.0, NSOL, tmp);
    for (int i = 0; i < s->n_cells; ++i)
    {
        double volume = s->cell_volume[i];
        for (int j = 0; j < NSOL; ++j)
            tmp[j] += REF_ABS(s->phi_dt[i * NSOL + j]) * volume;
    }
    for (int j = 0; j < NSOL; ++j)
        tmp[j] *= dt / s->global_volume;
    copy_n(tmp, NSOL, res);
}

/* ------------------------------------------------------------------------ */
/* Explicit LSERK: src/Solver/src/timedisc/explicit.c:19-39, 96-141           */
/* ------------------------------------------------------------------------ */
static const double rk_a_euler[1] = {0.0};
static const double rk_b_euler[1] = {0.0};
static const double rk_g_euler[1] = {1.0};
static const double rk_a_rk33[3] = {0., -5. / 9., -153. / 128.};
static const double rk_b_rk33[3] = {0., 1. / 3., 3. / 4.};
static const double rk_g_rk33[3] = {1. / 3., 15. / 16., 8. / 15.};
static const double rk_a_rk45[5] = {0., -567301805773.0 / 1357537059087.0,
                                    -2404267990393.0 / 2016746695238.0,
                                    -3550918686646.0 / 2091501179385.0,
                                    -1275806237668.0 / 842570457699.0};
static const double rk_b_rk45[5] = {0., 1432997174477.0 / 9575080441755.0,
                                    2526269341429.0 / 6820363962896.0,
                                    2006345519317.0 / 3224310063776.0,
                                    2802321613138.0 / 2924317926251.0};
static const double rk_g_rk45[5] = {1432997174477.0 / 9575080441755.0,
                                    5161836677717.0 / 13612068292357.0,
                                    1720146321549.0 / 2090206949498.0,
                                    3134564353537.0 / 4481467310338.0,
                                    2277821191437.0 / 14882151754819.0};

static void timestep_lserkw2(fvref_t *s, double *k, int n_stages,
                             const double *rk_a, const double *rk_b,
                             const double *rk_g, double t, double dt)
{
    int n = s->n_cells;
    double t_stage = t;
    time_derivative(s, t_stage);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < NSOL; ++j)
            k[i * NSOL + j] = s->phi_dt[i * NSOL + j];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < NSOL; ++j)
            s->phi_total[i * NTOT + j] += k[i * NSOL + j] * dt * rk_g[0];

    for (int st = 1; st < n_stages; ++st)
    {
        t_stage = t + dt * rk_b[st];
        time_derivative(s, t_stage);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < NSOL; ++j)
                k[i * NSOL + j] =
                    s->phi_dt[i * NSOL + j] + k[i * NSOL + j] * rk_a[st];
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < NSOL; ++j)
                s->phi_total[i * NTOT + j] += k[i * NSOL + j] * dt * rk_g[st];
    }
}

/* Amplification factor of the 2N-storage loop of timestep_lserkw2 on the
 * scalar ODE y' = lambda*y with z = lambda*dt (pins the tables and the loop
 * form k = R + a*k, y += k*dt*g). */
double fvref_lserk_amplification(int scheme, double z)
{
    const double *a, *g;
    int ns;
    switch (scheme)
    {
    case 0:
        a = rk_a_euler, g = rk_g_euler, ns = 1;
        break;
    case 1:
        a = rk_a_rk33, g = rk_g_rk33, ns = 3;
        break;
    default:
        a = rk_a_rk45, g = rk_g_rk45, ns = 5;
        break;
    }
    double y = 1.0, k = 0.0;
    for (int st = 0; st < ns; ++st)
    {
        double r = z * y; /* dt * R(y) with dt = 1 */
        k = st == 0 ? r : r + k * a[st];
        y += k * 1.0 * g[st];
    }
    return y;
}

/* src/Solver/src/timedisc/timedisc.c:134-222 (explicit branch) */
int fvref_timedisc_explicit(fvref_t *s, int scheme, int max_iter,
                            int transient, double abort_residual,
                            double t_start, double t_end, double *log)
{
    const double *a, *b, *g;
    int ns;
    switch (scheme)
    {
    case 0:
        a = rk_a_euler, b = rk_b_euler, g = rk_g_euler, ns = 1;
        break;
    case 1:
        a = rk_a_rk33, b = rk_b_rk33, g = rk_g_rk33, ns = 3;
        break;
    default:
        a = rk_a_rk45, b = rk_b_rk45, g = rk_g_rk45, ns = 5;
        break;
    }
    double t_end_loc = transient ? t_end : DBL_MAX;
    double *k = xcalloc((size_t)s->n_cells * NSOL, sizeof(double));
    int iter = 0;
    double t = t_start;
    double res[NSOL];
    int finalize = 0;

    time_derivative(s, t);

    while (1)
    {
        /* timedisc.c:164: the time-step hook (calc_ns_timestep, or a hook
         * returning a constant) */
        double dt = s->fixed_dt > 0.0 ? s->fixed_dt : calc_ns_timestep(s);
        if (t + dt > t_end_loc)
        {
            dt = t_end_loc - t;
            finalize = 1;
        }
        timestep_lserkw2(s, k, ns, a, b, g, t, dt);
        fvref_global_residual(s, dt, res);

        for (int j = 0; j < NSOL; ++j)
            if (isnan(res[j]) || isinf(res[j]))
            {
                free(k);
                return -1;
            }

        t = t + dt;
        iter = iter + 1;

        if (log)
        {
            double *row = &log[(iter - 1) * 8];
            row[0] = t;
            row[1] = dt;
            row[2] = s->is_viscous_dt;
            for (int j = 0; j < NSOL; ++j)
                row[3 + j] = res[j];
        }

        if (!transient)
        {
            double mx = res[0];
            for (int j = 1; j < NSOL; ++j)
                mx = REF_MAX(mx, res[j]);
            if (mx < abort_residual)
                finalize = 1;
        }
        if (iter >= max_iter)
            finalize = 1;
        if (finalize)
            break;
    }
    free(k);
    return iter;
}

/* ------------------------------------------------------------------------ */
/* Implicit (reference semantics): src/Solver/src/timedisc/implicit.c         */
/* ------------------------------------------------------------------------ */

/* Gaussian elimination with partial pivoting on one 5x5 block stored
 * column-major as the reference stores jac (jac[k*5 + j] = J[j][k]). */
static void solve_block5(const double *jac, const double *rhs, double *x)
{
    double A[5][6];
    for (int r = 0; r < 5; ++r)
    {
        for (int c = 0; c < 5; ++c)
            A[r][c] = jac[c * 5 + r];
        A[r][5] = rhs[r];
    }
    for (int c = 0; c < 5; ++c)
    {
        int p = c;
        for (int r = c + 1; r < 5; ++r)
            if (fabs(A[r][c]) > fabs(A[p][c]))
                p = r;
        if (p != c)
            for (int q = 0; q < 6; ++q)
            {
                double tmp = A[c][q];
                A[c][q] = A[p][q];
                A[p][q] = tmp;
            }
        for (int r = c + 1; r < 5; ++r)
        {
            double f = A[r][c] / A[c][c];
            for (int q = c; q < 6; ++q)
                A[r][q] -= f * A[c][q];
        }
    }
    for (int r = 4; r >= 0; --r)
    {
        double v = A[r][5];
        for (int c = r + 1; c < 5; ++c)
            v -= A[r][c] * x[c];
        x[r] = v / A[r][r];
    }
}

/* b_i = J_i x_i, J_i column-major (implicit.c:362-385, its order) */
static void block_matvec(const double *jac, int n, const double *x, double *b)
{
    for (int i = 0; i < n; ++i)
    {
        const double *J = &jac[i * 25];
        const double *xi = &x[i * NSOL];
        double *bi = &b[i * NSOL];
        for (int j = 0; j < NSOL; ++j)
            bi[j] = J[j] * xi[0];
        for (int k = 1; k < NSOL; ++k)
            for (int j = 0; j < NSOL; ++j)
                bi[j] += J[k * NSOL + j] * xi[k];
    }
}

/* The linear solvers behind basec's solve_gmres_n_m / solve_bicgstab_n_m
 * (polynomialchaos/basec, version unpinned, not in the reference tree), as
 * the reference calls them (implicit.c:285-300): b = f, x = dY (its previous
 * content is the initial guess), *res in = the absolute tolerance
 * tolerance_lsoe * |f|, *n_iter in = max_iter_lsoe; out: the iterations
 * done and the final residual norm.  basec's own iteration is not
 * available here, so these are the textbook methods:
 *   GMRES(m) with restarts: Saad, Iterative Methods for Sparse Linear
 *     Systems, 2nd ed., Algorithm 6.9 with Givens rotations (Section 6.5.3),
 *     Arnoldi by classical Gram-Schmidt applied twice; at most restarts + 1
 *     cycles of m steps and *n_iter steps in total; the residual is
 *     recomputed at the start of each cycle;
 *   BiCGStab: van der Vorst (1992), SIAM J. Sci. Stat. Comput. 13:631-644,
 *     Algorithm 1, r_hat = r_0, stopping on |r| or |s| <= tolerance. */
static void gmres_nm(const double *jac, int n, const double *b, double *x, int m,
                     int restarts, int *n_iter, double *res)
{
    const int nv = n * NSOL;
    const double tol = *res;
    const int max_it = *n_iter;
    double *V = xcalloc((size_t)(m + 1) * nv, sizeof(double));
    double *w = xcalloc((size_t)nv, sizeof(double));
    double *H = xcalloc((size_t)(m + 1) * m, sizeof(double));
    double *g = xcalloc((size_t)m + 1, sizeof(double));
    double *cs = xcalloc((size_t)m, sizeof(double));
    double *sn = xcalloc((size_t)m, sizeof(double));
    double *y = xcalloc((size_t)m, sizeof(double));
    double h1[64], h2[64];
    int it = 0;
    double resid = 0.0;
    for (int cycle = 0; cycle <= restarts; ++cycle)
    {
        block_matvec(jac, n, x, w);
        for (int q = 0; q < nv; ++q)
            w[q] = b[q] - w[q];
        const double beta = sqrt(dot_n(w, w, nv));
        resid = beta;
        if (beta <= tol || it >= max_it)
            break;
        for (int q = 0; q < nv; ++q)
            V[q] = w[q] / beta;
        for (int i = 0; i <= m; ++i)
            g[i] = 0.0;
        g[0] = beta;
        int k = 0;
        for (int j = 0; j < m && it < max_it; ++j)
        {
            double *vj1 = &V[(size_t)(j + 1) * nv];
            block_matvec(jac, n, &V[(size_t)j * nv], vj1);
            ++it;
            for (int pass = 0; pass < 2; ++pass)
            {
                double *h = pass == 0 ? h1 : h2;
                for (int i = 0; i <= j; ++i)
                    h[i] = dot_n(&V[(size_t)i * nv], vj1, nv);
                for (int i = 0; i <= j; ++i)
                    for (int q = 0; q < nv; ++q)
                        vj1[q] -= h[i] * V[(size_t)i * nv + q];
            }
            for (int i = 0; i <= j; ++i)
                H[i * m + j] = h1[i] + h2[i];
            const double hn = sqrt(dot_n(vj1, vj1, nv));
            H[(j + 1) * m + j] = hn;
            if (hn > 0.0)
                for (int q = 0; q < nv; ++q)
                    vj1[q] = vj1[q] / hn;
            for (int i = 0; i < j; ++i)
            {
                const double a = H[i * m + j], c = H[(i + 1) * m + j];
                H[i * m + j] = cs[i] * a + sn[i] * c;
                H[(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
            }
            const double a = H[j * m + j], c = H[(j + 1) * m + j];
            const double r = sqrt(a * a + c * c);
            cs[j] = r > 0.0 ? a / r : 1.0;
            sn[j] = r > 0.0 ? c / r : 0.0;
            H[j * m + j] = r;
            H[(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            resid = fabs(g[j + 1]);
            k = j + 1;
            if (resid <= tol)
                break;
        }
        for (int i = k - 1; i >= 0; --i)
        {
            double v = g[i];
            for (int q = i + 1; q < k; ++q)
                v -= H[i * m + q] * y[q];
            y[i] = v / H[i * m + i];
        }
        for (int i = 0; i < k; ++i)
            for (int q = 0; q < nv; ++q)
                x[q] += y[i] * V[(size_t)i * nv + q];
        if (resid <= tol || it >= max_it)
            break;
    }
    *n_iter = it;
    *res = resid;
    free(V), free(w), free(H), free(g), free(cs), free(sn), free(y);
}

static void bicgstab_nm(const double *jac, int n, const double *b, double *x, int *n_iter,
                        double *res)
{
    const int nv = n * NSOL;
    const double tol = *res;
    const int max_it = *n_iter;
    double *r = xcalloc((size_t)nv, sizeof(double)), *rh = xcalloc((size_t)nv, sizeof(double));
    double *p = xcalloc((size_t)nv, sizeof(double)), *v = xcalloc((size_t)nv, sizeof(double));
    double *sv = xcalloc((size_t)nv, sizeof(double)), *t = xcalloc((size_t)nv, sizeof(double));
    block_matvec(jac, n, x, r);
    for (int q = 0; q < nv; ++q)
    {
        r[q] = b[q] - r[q];
        rh[q] = r[q];
    }
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    double rn = sqrt(dot_n(r, r, nv));
    int it = 0;
    while (rn > tol && it < max_it)
    {
        ++it;
        const double rho1 = dot_n(rh, r, nv);
        const double beta = (rho1 / rho) * (alpha / omega);
        for (int q = 0; q < nv; ++q)
            p[q] = r[q] + beta * (p[q] - omega * v[q]);
        block_matvec(jac, n, p, v);
        alpha = rho1 / dot_n(rh, v, nv);
        for (int q = 0; q < nv; ++q)
            sv[q] = r[q] - alpha * v[q];
        const double sn = sqrt(dot_n(sv, sv, nv));
        if (sn <= tol)
        {
            for (int q = 0; q < nv; ++q)
                x[q] += alpha * p[q];
            rn = sn;
            break;
        }
        block_matvec(jac, n, sv, t);
        omega = dot_n(t, sv, nv) / dot_n(t, t, nv);
        for (int q = 0; q < nv; ++q)
        {
            x[q] += alpha * p[q] + omega * sv[q];
            r[q] = sv[q] - omega * t[q];
        }
        rn = sqrt(dot_n(r, r, nv));
        rho = rho1;
    }
    *n_iter = it;
    *res = rn;
    free(r), free(rh), free(p), free(v), free(sv), free(t);
}

/* f = -(Y - old_0)/dt + b R - sum_s a_s/dt old_s  (implicit.c:254-270) */
static void bdf_rhs(int nv, const double *Y, double *const *old, int n_loc,
                    const double *a, double b, double dt, const double *R,
                    double *f)
{
    for (int idx = 0; idx < nv; ++idx)
        f[idx] = -(Y[idx] - old[0][idx]) / dt + b * R[idx];
    for (int st = 0; st < n_loc; ++st)
        for (int idx = 0; idx < nv; ++idx)
            f[idx] -= a[st] / dt * old[st][idx];
}

/* One implicit step, implicit.c:224-352 (timestep_newton) with
 * calc_jacobian_numerical (implicit.c:61-120).  old[0..n_bdf-1] is the BDF
 * history kept by the caller across steps (phi_old); on entry phi_dt holds
 * R of the current state.  Returns the inner iterations, or -2 when the
 * inner loop ran to max_iter_inner (the reference aborts there). */
static int timestep_newton(fvref_t *s, int scheme, int iter, double t,
                           double dt, int transient, int max_iter_inner,
                           double **old, double *jac_out)
{
    static const double bdf_a_euler[1] = {0.0};
    static const double bdf_a_bdf2[2] = {-1. / 3., 1 / 3.};
    const int n = s->n_cells;
    const int n_bdf = scheme == 1 ? 2 : 1;
    const int n_loc = (iter < n_bdf) ? 1 : n_bdf;
    const double *a = (iter < n_bdf) ? bdf_a_euler
                                     : (scheme == 1 ? bdf_a_bdf2 : bdf_a_euler);
    const double b = (iter < n_bdf) ? 1.0 : (scheme == 1 ? 2. / 3. : 1.0);
    const double tpdt = t + dt;
    const int nv = n * NSOL;

    double *Y = xcalloc((size_t)nv, sizeof(double));
    double *f = xcalloc((size_t)nv, sizeof(double));
    double *dY = xcalloc((size_t)nv, sizeof(double));
    double *dYdt = xcalloc((size_t)nv, sizeof(double));
    double *jac = xcalloc((size_t)nv * NSOL, sizeof(double));

    for (int i = n_loc - 1; i > 0; --i)
        copy_n(old[i - 1], nv, old[i]);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < NSOL; ++j)
            old[0][i * NSOL + j] = s->phi_total[i * NTOT + j];
    copy_n(old[0], nv, Y);
    copy_n(s->phi_dt, nv, dYdt);
    bdf_rhs(nv, Y, old, n_loc, a, b, dt, dYdt, f);

    const double e0 = sqrt(dot_n(f, f, nv));
    double e_old = e0;
    int inner;
    for (inner = 1; inner <= max_iter_inner; ++inner)
    {
        for (int iv = 0; iv < NSOL; ++iv)
        {
            double eps = 0.0;
            for (int i = 0; i < n; ++i)
                eps += Y[i * NSOL + iv] * Y[i * NSOL + iv];
            eps = sqrt(eps) * 1e-4;

            for (int i = 0; i < n; ++i)
            {
                for (int j = 0; j < NSOL; ++j)
                    s->phi_total[i * NTOT + j] = Y[i * NSOL + j];
                s->phi_total[i * NTOT + iv] += 0.5 * eps;
            }
            time_derivative(s, tpdt);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < NSOL; ++j)
                    jac[i * 25 + iv * 5 + j] = -s->phi_dt[i * NSOL + j];

            for (int i = 0; i < n; ++i)
            {
                for (int j = 0; j < NSOL; ++j)
                    s->phi_total[i * NTOT + j] = Y[i * NSOL + j];
                s->phi_total[i * NTOT + iv] -= 0.5 * eps;
            }
            time_derivative(s, tpdt);
            for (int i = 0; i < n; ++i)
            {
                int idx = i * 25 + iv * 5;
                for (int j = 0; j < NSOL; ++j)
                {
                    jac[idx + j] += s->phi_dt[i * NSOL + j];
                    jac[idx + j] *= b / (eps + SMALL);
                }
                jac[idx + iv] += 1.0 / dt;
            }
        }
        if (jac_out)
            copy_n(jac, nv * NSOL, jac_out);

        if (e_old >= e0)
        {
            if (s->lsoe_solver < 0)
                for (int i = 0; i < n; ++i)
                    solve_block5(&jac[i * 25], &f[i * NSOL], &dY[i * NSOL]);
            else
            {
                /* implicit.c:285-300: initial guess the previous dY */
                if (!s->lsoe_dY)
                    s->lsoe_dY = xcalloc((size_t)nv, sizeof(double));
                copy_n(s->lsoe_dY, nv, dY);
                int n_it = s->lsoe_max_iter;
                double res = s->lsoe_tol * e_old;
                if (s->lsoe_solver == 0)
                    bicgstab_nm(jac, n, f, dY, &n_it, &res);
                else
                    gmres_nm(jac, n, f, dY, s->lsoe_dims, s->lsoe_restarts, &n_it, &res);
                copy_n(dY, nv, s->lsoe_dY);
                s->lsoe_last_iter = n_it;
            }
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < NSOL; ++j)
                {
                    Y[i * NSOL + j] += dY[i * NSOL + j];
                    s->phi_total[i * NTOT + j] = Y[i * NSOL + j];
                }
            time_derivative(s, tpdt);
            copy_n(s->phi_dt, nv, dYdt);
            bdf_rhs(nv, Y, old, n_loc, a, b, dt, dYdt, f);
            e_old = sqrt(dot_n(f, f, nv));
        }
        else
        {
            time_derivative(s, tpdt);
        }
        if (e_old < e0)
            break;
        if (!transient)
            break;
    }
    free(Y);
    free(f);
    free(dY);
    free(dYdt);
    free(jac);
    if (inner >= max_iter_inner)
        return -2;
    return inner;
}

int fvref_timestep_newton(fvref_t *s, int scheme, int iter, double t,
                          double dt, int transient, int max_iter_inner,
                          double *old_levels, double *jac_out)
{
    const int nv = s->n_cells * NSOL;
    double *old[2] = {old_levels, old_levels + nv};
    return timestep_newton(s, scheme, iter, t, dt, transient, max_iter_inner,
                           old, jac_out);
}

/* timedisc.c:134-222 with the implicit step: the same loop as the explicit
 * driver (dt from calc_ns_timestep on the refreshed state, t_end clamp,
 * residual, NaN stop, steady and iteration stops); log rows as there. */
int fvref_timedisc_implicit(fvref_t *s, int scheme, int max_iter,
                            int transient, double abort_residual,
                            double t_start, double t_end, int max_iter_inner,
                            double *log)
{
    const int nv = s->n_cells * NSOL;
    double *levels = xcalloc((size_t)nv * 2, sizeof(double));
    double *old[2] = {levels, levels + nv};
    double t_end_loc = transient ? t_end : DBL_MAX;
    int iter = 0;
    double t = t_start;
    double res[NSOL];
    int finalize = 0;

    time_derivative(s, t);

    while (1)
    {
        double dt = calc_ns_timestep(s);
        if (t + dt > t_end_loc)
        {
            dt = t_end_loc - t;
            finalize = 1;
        }
        if (timestep_newton(s, scheme, iter, t, dt, transient, max_iter_inner,
                            old, NULL) < 0)
        {
            free(levels);
            return -2;
        }
        fvref_global_residual(s, dt, res);
        for (int j = 0; j < NSOL; ++j)
            if (isnan(res[j]) || isinf(res[j]))
            {
                free(levels);
                return -1;
            }
        t = t + dt;
        iter = iter + 1;
        if (log)
        {
            double *row = &log[(iter - 1) * 8];
            row[0] = t;
            row[1] = dt;
            row[2] = s->is_viscous_dt;
            for (int j = 0; j < NSOL; ++j)
                row[3 + j] = res[j];
        }
        if (!transient)
        {
            double mx = res[0];
            for (int j = 1; j < NSOL; ++j)
                mx = REF_MAX(mx, res[j]);
            if (mx < abort_residual)
                finalize = 1;
        }
        if (iter >= max_iter)
            finalize = 1;
        if (finalize)
            break;
    }
    free(levels);
    return iter;
}

/* ------------------------------------------------------------------------ */
/* Construction and access                                                    */
/* ------------------------------------------------------------------------ */
#define DUP(dst, src, n, T)                                                    \
    do                                                                         \
    {                                                                          \
        (dst) = xcalloc((size_t)(n), sizeof(T));                               \
        if (src)                                                               \
            memcpy((dst), (src), (size_t)(n) * sizeof(T));                     \
    } while (0)

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
                      const int *region_is_boundary)
{
    fvref_t *s = xcalloc(1, sizeof(fvref_t));
    /* the reference's TimeDisc/Implicit defaults (src/FV3D/src/timedisc/
     * implicit.c:12-18): GMRES(15), 2 restarts, 1e-12, 100 iterations */
    s->lsoe_solver = 1;
    s->lsoe_tol = 1e-12;
    s->lsoe_max_iter = 100;
    s->lsoe_dims = 15;
    s->lsoe_restarts = 2;
    s->n_cells = n_cells;
    s->max_cell_faces = max_cell_faces;
    s->n_faces = n_faces;
    s->n_boundaries = n_boundaries;
    s->n_regions = n_regions;
    DUP(s->cell_n_faces, cell_n_faces, n_cells, int);
    DUP(s->cell_faces, cell_faces, (size_t)n_cells * max_cell_faces, int);
    DUP(s->cell_x, cell_x, (size_t)n_cells * DIM, double);
    DUP(s->cell_volume, cell_volume, n_cells, double);
    DUP(s->cell_dx, cell_dx, (size_t)n_cells * DIM, double);
    DUP(s->face_cells, face_cells, (size_t)n_faces * 2, int);
    DUP(s->face_boundary, face_boundary, n_faces, int);
    DUP(s->face_x, face_x, (size_t)n_faces * DIM, double);
    DUP(s->face_n, face_n, (size_t)n_faces * DIM, double);
    DUP(s->face_t1, face_t1, (size_t)n_faces * DIM, double);
    DUP(s->face_t2, face_t2, (size_t)n_faces * DIM, double);
    DUP(s->face_area, face_area, n_faces, double);
    DUP(s->face_lambda, face_lambda, n_faces, double);
    if (face_shift)
        DUP(s->face_shift, face_shift, (size_t)n_faces * DIM, double);
    DUP(s->boundary_id, boundary_id, n_boundaries, int);
    DUP(s->boundary_face, boundary_face, n_boundaries, int);
    DUP(s->region_is_boundary, region_is_boundary, n_regions, int);
    s->region_type = xcalloc(n_regions, sizeof(int));
    s->region_phi = xcalloc((size_t)n_regions * NTOT, sizeof(double));

    calc_mesh_metrics(s);

    size_t ne = (size_t)n_cells + n_boundaries;
    s->flux = xcalloc((size_t)n_faces * NSOL, sizeof(double));
    s->phi_left = xcalloc((size_t)n_faces * NTOT, sizeof(double));
    s->phi_right = xcalloc((size_t)n_faces * NTOT, sizeof(double));
    s->grad_x = xcalloc(ne * NTOT, sizeof(double));
    s->grad_y = xcalloc(ne * NTOT, sizeof(double));
    s->grad_z = xcalloc(ne * NTOT, sizeof(double));
    s->phi_dt = xcalloc(ne * NSOL, sizeof(double));
    s->phi_total = xcalloc(ne * NTOT, sizeof(double));

    fvref_set_physics(s, 1.0, 1.0, 18.13e-6, 8.31446261815324 / 28.96e-3,
                      0.718, 1.4, 0, 0, 1, 0);
    return s;
}

void fvref_destroy(fvref_t *s)
{
    if (!s)
        return;
    void *ptrs[] = {s->cell_n_faces, s->cell_faces, s->cell_x,
                    s->cell_volume, s->cell_dx, s->face_cells,
                    s->face_boundary, s->face_x, s->face_n, s->face_t1,
                    s->face_t2, s->face_area, s->face_lambda, s->face_shift,
                    s->dist_cell_1, s->dist_cell_2, s->internal_faces,
                    s->boundary_faces, s->boundary_id, s->boundary_face,
                    s->region_is_boundary, s->region_type, s->region_phi,
                    s->flux, s->phi_left, s->phi_right, s->grad_x, s->grad_y,
                    s->grad_z, s->phi_dt, s->phi_total, s->lsoe_dY};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
        free(ptrs[i]);
    free(s);
}

/* navier_stokes.c:96-104 derived constants; flux.c:295-312 and
 * limiter.c:26-40 / reconstruction.c:110-137 hook selection */
void fvref_set_physics(fvref_t *s, double cfl_scale, double dfl_scale,
                       double mu_mix, double R_mix, double Pr, double kappa,
                       int flux, int recon, int limiter, int visc_dt_mode)
{
    s->cfl_scale = cfl_scale;
    s->dfl_scale = dfl_scale;
    s->mu_mix = mu_mix;
    s->R_mix = R_mix;
    s->Pr = Pr;
    s->kappa = kappa;
    s->kappa_m1 = kappa - 1.0;
    s->kappa_p1 = kappa + 1.0;
    s->s_kappa = 1.0 / kappa;
    s->s_kappa_m1 = 1.0 / s->kappa_m1;
    s->s_kappa_p1 = 1.0 / s->kappa_p1;
    s->cv = R_mix * s->s_kappa_m1;
    s->cp = kappa * s->cv;
    s->kappa_pr = kappa / Pr;
    s->lambda = mu_mix * s->cp / Pr;
    s->visc_dt_mode = visc_dt_mode;
    s->riemann = flux == 1 ? riemann_rusanov : riemann_ausm;
    s->reconstruction = recon == 1 ? reconstruction_linear
                                   : reconstruction_first_order;
    s->limiter = limiter == 0 ? limiter_barth_jespersenn : limiter_none;
}

/* boundary.c:76-121 and parse_primitive_state boundary.c:144-183 */
void fvref_set_region(fvref_t *s, int region, int type, const double *prim5)
{
    double *phi = &s->region_phi[region * NTOT];
    s->region_type[region] = type;
    if (type == FVREF_BC_FLOW || type == FVREF_BC_INFLOW ||
        type == FVREF_BC_STATE)
    {
        set_value_n(0.0, NTOT, phi);
        phi[IP_U] = prim5[0];
        phi[IP_V] = prim5[1];
        phi[IP_W] = prim5[2];
        phi[IP_P] = prim5[3];
        phi[IP_T] = prim5[4];
        phi[IC_RHO] = calc_ig_rho(phi[IP_P], phi[IP_T], s->R_mix);
        prim_to_con(s, phi);
    }
    else if (type == FVREF_BC_WALL_ISOTHERMAL)
    {
        phi[IP_T] = prim5[4];
    }
}

void fvref_set_fixed_dt(fvref_t *s, double dt) { s->fixed_dt = dt; }

void fvref_set_lsoe(fvref_t *s, int solver, double tol, int max_iter, int dims,
                    int restarts)
{
    s->lsoe_solver = solver;
    s->lsoe_tol = tol;
    s->lsoe_max_iter = max_iter;
    s->lsoe_dims = dims < 1 ? 1 : (dims > 64 ? 64 : dims);
    s->lsoe_restarts = restarts;
    free(s->lsoe_dY);
    s->lsoe_dY = NULL;
}

int fvref_lsoe_iterations(const fvref_t *s) { return s->lsoe_last_iter; }

void fvref_set_state(fvref_t *s, const double *U)
{
    for (int i = 0; i < s->n_cells; ++i)
        for (int j = 0; j < NSOL; ++j)
            s->phi_total[i * NTOT + j] = U[i * NSOL + j];
    ns_update(s, 0.0);
}

void fvref_time_derivative(fvref_t *s, double t)
{
    time_derivative(s, t);
}

double fvref_calc_timestep(fvref_t *s, int *viscous)
{
    double dt = calc_ns_timestep(s);
    if (viscous)
        *viscous = s->is_viscous_dt;
    return dt;
}

int fvref_n_elements(const fvref_t *s)
{
    return s->n_cells + s->n_boundaries;
}

double fvref_global_volume(const fvref_t *s)
{
    return s->global_volume;
}

void fvref_get(const fvref_t *s, int what, double *out)
{
    size_t ne = (size_t)s->n_cells + s->n_boundaries;
    size_t nf = (size_t)s->n_faces;
    switch (what)
    {
    case 0:
        memcpy(out, s->phi_total, ne * NTOT * sizeof(double));
        break;
    case 1:
        memcpy(out, s->phi_dt, ne * NSOL * sizeof(double));
        break;
    case 2:
        memcpy(out, s->grad_x, ne * NTOT * sizeof(double));
        break;
    case 3:
        memcpy(out, s->grad_y, ne * NTOT * sizeof(double));
        break;
    case 4:
        memcpy(out, s->grad_z, ne * NTOT * sizeof(double));
        break;
    case 5:
        memcpy(out, s->flux, nf * NSOL * sizeof(double));
        break;
    case 6:
        memcpy(out, s->phi_left, nf * NTOT * sizeof(double));
        break;
    case 7:
        memcpy(out, s->phi_right, nf * NTOT * sizeof(double));
        break;
    case 8:
        memcpy(out, s->dist_cell_1, nf * DIM * sizeof(double));
        break;
    case 9:
        memcpy(out, s->dist_cell_2, nf * DIM * sizeof(double));
        break;
    case 10:
        for (size_t i = 0; i < nf * 2; ++i)
            out[i] = (double)s->face_cells[i];
        break;
    default:
        break;
    }
}

void fvref_con_to_prim(fvref_t *s, double *phi10)
{
    con_to_prim(s, phi10);
}

void fvref_prim_to_con(fvref_t *s, double *phi10)
{
    prim_to_con(s, phi10);
}

double fvref_riemann_p(fvref_t *s, const double *phi10)
{
    return calc_riemann_p(s, phi10);
}

void fvref_riemann(fvref_t *s, int flux, const double *phi_l10,
                   const double *phi_r10, double *f5)
{
    double l[NTOT], r[NTOT];
    copy_n(phi_l10, NTOT, l);
    copy_n(phi_r10, NTOT, r);
    if (flux == 1)
        riemann_rusanov(s, l, r, f5);
    else
        riemann_ausm(s, l, r, f5);
}
