// Kernels of the Newton-Krylov (JFNK) path (internal).
//
// Vectors have the state's layout [N][5] (n = 5N values).  Krylov vectors are
// stored as T = double (plain FP64) or T = float (mixed precision: FP32
// storage and FP32 Jacobian-vector results; every dot product accumulates in
// FP64 and the residuals and the state update stay FP64).
#pragma once

#include <hip/hip_runtime.h>

#include "fv_device.h"

namespace fvh {
namespace kry {

constexpr int kMaxDots = 64;       // Krylov vectors per multi-dot / multi-axpy
constexpr int kDotBlocks = 1024;   // fixed grid of the deterministic dots

// Local pseudo time step dtau_c = cfl * V / sum_d (|u_d| + c) dx_d (the
// per-cell convective limit of navier_stokes.c:171-180).
void local_dt(hipStream_t s, const dev::DevMesh &m, const dev::PhysDev &P,
              const double *U, double cfl, double *dtau);

// Inverse of the block-Jacobi block M_c = I/dtau_c - dR1_c/dU_c, R1 the
// first-order inviscid residual; dR1_c/dU_c by one-sided differences of the
// face fluxes in the cell's own state (ghost states follow the perturbed
// interior state).  Minv[c*25 + r*5 + k] row-major.  With blocks = 0 the
// block is I/dtau_c (scalar pseudo-time preconditioner).
void precond_blocks(hipStream_t s, const dev::DevMesh &m, const dev::PhysDev &P,
                    int flux, const double *U, const double *dtau, int blocks,
                    double *Minv);

template <typename T>
void precond_apply(hipStream_t s, int N, const double *Minv, const T *v, T *z);

// z = Minv v and zz[0] = |z|_2^2 (fixed-order FP64 sum of the stored values)
template <typename T>
void precond_apply_norm(hipStream_t s, int N, const double *Minv, const T *v, T *z,
                        double *partials, double *zz);

// eps[0] = num / sqrt(zz[0]) (num / 1e-300 for zz = 0)
void fd_eps(hipStream_t s, const double *zz, double num, double *eps);

// Ut = U + eps[0] * z  (eps in device memory)
template <typename T>
void perturb(hipStream_t s, int n, const double *U, const double *eps,
             const T *z, double *Ut);

// w = z / dtau - (Rt - R0) / eps[0]   (A z with A = I/dtau - J, J z by
// finite difference)
template <typename T>
void op_finish(hipStream_t s, int N, const T *z, const double *dtau,
               const double *Rt, const double *R0, const double *eps, T *w);

// w = b - w
template <typename T>
void rhs_minus(hipStream_t s, int n, const double *b, T *w);

// dots[j] = <V_j, w> for j < k, V_j at V + j*ld; FP64 accumulation in a fixed
// order.  partials holds kDotBlocks*kMaxDots doubles.
template <typename T>
void multidot(hipStream_t s, int n, int k, const T *V, size_t ld, const T *w,
              double *partials, double *dots);

// w -= sum_j h[j] V_j  (h in device memory)
template <typename T>
void multiaxpy(hipStream_t s, int n, int k, const T *V, size_t ld,
               const double *h, T *w);

// out = w * scale
template <typename T>
void scale(hipStream_t s, int n, const T *w, double scale, T *out);

// x = sum_j y[j] V_j (y in device memory)
template <typename T>
void combine(hipStream_t s, int n, int k, const T *V, size_t ld,
             const double *y, T *x);

// x (FP64) += z
template <typename T>
void accumulate(hipStream_t s, int n, const T *z, double *x);

template <typename T>
void convert(hipStream_t s, int n, const double *src, T *dst);

// sums[j] = sum_c |R_cj| V_c (the reference norm before the dt / V_glob
// factor), fixed order.
void residual_sums(hipStream_t s, int N, const double *R, const double *vol,
                   double *partials, double *sums);

// out[0] = max_c max(|dU_c0 / U_c0|, |dU_c4 / U_c4|)
void update_ratio(hipStream_t s, int N, const double *U, const double *dU,
                  double *partials, double *out);

// U += alpha * dU
void update(hipStream_t s, int n, double *U, const double *dU, double alpha);

} // namespace kry
} // namespace fvh
