// Kernels of the reference-semantics implicit time step (timestep_newton,
// src/Solver/src/timedisc/implicit.c:61-385): BDF right-hand side, lumped
// finite-difference block Jacobian, per-cell block solve.  Internal.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_device.h"

namespace fvh {
namespace imp {

// f = -(Y - old0)/dt + b R, then f -= a_s/dt old_s for s < n_loc
// (implicit.c:254-270 / 315-331)
void bdf_rhs(hipStream_t s, int n, const double *Y, const double *old0,
             const double *old1, int n_loc, double a0, double a1, double b,
             double dt, const double *R, double *f);

// Ut = U with column k shifted by delta in every cell (implicit.c:74-81 /
// 96-103: delta = +0.5 eps, then -0.5 eps)
void shift_var(hipStream_t s, int N, const double *U, int k, double delta,
               double *Ut);

// jac[i*25 + k*5 + j] = (-Rp_ij + Rm_ij) * fac, plus inv_dt when j == k
// (implicit.c:85-118, fac = b/(eps + SMALL))
void jac_column(hipStream_t s, int N, const double *Rp, const double *Rm,
                int k, double fac, double inv_dt, double *jac);

// dY_i = J_i^-1 f_i per cell, J_i column-major as jac stores it; Gaussian
// elimination with partial pivoting.
void block_solve(hipStream_t s, int N, const double *jac, const double *f,
                 double *dY);

// b_i = J_i x_i per cell (implicit.c:362-385, the reference's order)
void block_matvec(hipStream_t s, int N, const double *jac, const double *x, double *b);
// Element operations of the linear solvers: op 0 o1 = x / a; 1 o1 = x - a y;
// 2 o1 = x + a (o1 - b y); 3 o1 += a x; 4 o1 += a x + b y and o2 = y - b z.
void lsoe_elem(hipStream_t s, int n, int op, double a, double b, const double *x,
               const double *y, const double *z, double *o1, double *o2);
// x += sum_{i<k} y[i] V_i in order of i (y in device memory)
void add_combination(hipStream_t s, int n, int k, const double *V, size_t ld, const double *y,
                     double *x);

// sums[k] = sum_i U_ik^2, k < 5 (fixed order); partials: kDotBlocks*64
void column_sumsq(hipStream_t s, int N, const double *U, double *partials,
                  double *sums);

// One row per workgroup of cell_blocks(N) in the time loop's partials
// layout: [0..4] sum |R_ij| V_i, [5] min V/ds1, [6] min ds2/(rho V) of U
// with primitives refreshed from U.
void step_partials(hipStream_t s, const dev::DevMesh &m, const dev::PhysDev &P,
                   const double *U, const double *R, double *partials);

} // namespace imp
} // namespace fvh
