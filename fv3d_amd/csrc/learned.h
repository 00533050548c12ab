// Learned block preconditioner: a per-cell MLP (28 -> 64 -> 64 -> 25, tanh)
// that predicts the 5x5 block-Jacobi block of the Newton-Krylov solve from
// local features.  BF16 MFMA kernels with FP32 accumulation for inference
// and training; FP32 master weights and Adam.  Internal.
#pragma once

#include <hip/hip_runtime.h>

#include "fv_device.h"

namespace fvh {
namespace nn {

constexpr int kFeat = 28;  // features used
constexpr int kIn = 32;    // padded input width
constexpr int kHid = 64;
constexpr int kOut = 32;   // padded output width (25 used)
constexpr int kBlk = 25;

// Parameter vector theta: W1 [kIn][kHid], b1, W2 [kHid][kHid], b2,
// W3 [kHid][kOut], b3 (row-major [in][out]).
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kIn * kHid;
constexpr int oW2 = oB1 + kHid;
constexpr int oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid;
constexpr int oB3 = oW3 + kHid * kOut;
constexpr int kParams = oB3 + kOut;

constexpr int kTrainBlocks = 128; // workgroups (one wave each) of a gradient pass

// Normalisation of the conservative variables: S = diag(1/rho_r,
// 1/(rho_r a_r) x3, 1/(rho_r a_r^2)), a_r = sqrt(kappa p_r / rho_r).
struct Refs
{
    double rho, p, a;
};

// Features [N][kIn] of the current state U (cfl is not used):
//  0      rho/rho_r - 1
//  1..3   u/c, v/c, w/c (c the cell's sound speed)
//  4      p/p_r - 1
//  5..19  h * Green-Gauss gradient of (rho/rho_r, u/a_r, v/a_r, w/a_r,
//         p/p_r), h = V^(1/3), face values the mean of the two sides
//         (ghost state on boundary faces); order q*3 + d
//  20..25 sum_f A_f n_f n_f^T / V^(2/3): xx, yy, zz, xy, xz, yz
//  26     boundary faces / 4
//  27     |u| / c
//  28..31 zero
void features(hipStream_t s, const dev::DevMesh &m, const dev::PhysDev &P,
              const double *U, double cfl, Refs r, float *feat);

// Targets [N][kOut]: S K_c S^-1 (25 entries row-major, then zeros), where
// the block-Jacobi block is M_c = Minv_c^-1 = (I + cfl K_c) / dtau_c; K_c is
// the dimensionless first-order diagonal Jacobian block -(V/ds1) J_c,
// independent of the CFL number.
void targets(hipStream_t s, int N, const double *Minv, const double *dtau,
             double cfl, Refs r, float *T);

// out [N][kOut] = MLP(feat) with the parameters theta (FP32 master values,
// rounded to BF16 for the MFMA operands).
void forward(hipStream_t s, int N, const float *theta, const float *feat,
             float *out);

// Minv_c = dtau_c (I + cfl S^-1 out_c S)^-1 (FP64 blocks, row-major).
void apply_blocks(hipStream_t s, int N, const float *out, const double *dtau,
                  double cfl, Refs r, double *Minv);

// Loss sum_c sum_{j<25} (y_cj - t_cj)^2 / (n * 25) over samples [0, n) and
// its gradient grad[kParams]; partials holds kTrainBlocks * (kParams + 1)
// floats.  Fixed-order reductions (deterministic).
void loss_grad(hipStream_t s, int n, const float *theta, const float *feat,
               const float *target, float *partials, float *grad, float *loss);

// Adam step t (1-based) on theta with the moments m, v.
void adam(hipStream_t s, float *theta, const float *grad, float *m, float *v,
          int t, float lr, float beta1, float beta2, float eps);

// Per-cell relative Frobenius error |A - B|_F / |B|_F of two block sets,
// reduced to the maximum and the mean (out[0], out[1]).
void block_error(hipStream_t s, int N, const double *A, const double *B,
                 double *partials, double *out);

} // namespace nn
} // namespace fvh
