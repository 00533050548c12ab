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

// ---- coupled preconditioner: multicolour block ILU(0) ("DILU") -------------
// M = (D + L) D^-1 (D + U) over the cells in colour order (a distance-1
// colouring: no two cells of one colour share a face), L and U the
// off-diagonal 5x5 blocks of A1 = I/dtau - dR1/dU (R1 the first-order
// inviscid residual) below and above the diagonal in that order, and
// D_i = A1_ii - sum_{k lower neighbour of i} A1_ik D_k^-1 A1_ki, so that M
// agrees with A1 on the diagonal blocks and on every face-neighbour block.
//
// Blocks are built by colour-sorted position p (cell ccell[p]) and face slot
// t (the slot order of the cell's face list): off[(p*maxf + t)*25 + r*5 + k]
// = A1_{i,nbr} row r, column k.  The sweeps read tiled copies: entry e of E
// per position at ((p/64)*E + e)*64 + p%64 (offB: E = maxf*25, e = t*25 +
// r*5 + k; dinvB: E = 25; nbr: E = maxf, e = t).  nbr holds the neighbour
// cell or -1 (boundary face, or t beyond the cell's face count), allocated
// for whole tiles; back[p*maxf + t] the flat slot index (pk*maxf + t') of
// the transposed coupling A1_{nbr,i}; flags[p] has bit t set for a
// neighbour of a lower colour and bit 8 + t for one of a higher colour.
// bfloat16 storage (the upper 16 bits of an FP32 pattern)
struct bf16_t
{
    unsigned short v;
};

struct DiluMesh
{
    int N, maxf, ncol;
    const int *ccell, *nbr, *back, *flags;
};

// off = the off-diagonal blocks of A1 at U, by one-sided differences of the
// shared face flux in the column cell's state (h = 1e-7 max(|U|, 1), the
// steps of precond_blocks).
void dilu_offdiag(hipStream_t s, const dev::DevMesh &m, const dev::PhysDev &P, int flux,
                  const DiluMesh &d, const double *U, double *off);

// Factorisation of the positions [p0, p1) of one colour (colours in
// ascending order): dinv[p] = D_i^-1 with A1_ii = Minv[i]^-1 (Minv by cell,
// the block-Jacobi or learned inverse blocks); the position's blocks are
// also written into the tiled sweep copies: dinvB in precision B, offB in O
// (double, float or bfloat16).
template <typename B, typename O>
void dilu_factor(hipStream_t s, const DiluMesh &d, int p0, int p1, const double *Minv,
                 const double *off, double *dinv, O *offB, B *dinvB);

// z = M^-1 v: the forward sweep over the colours in ascending order, then
// the backward sweep in descending order (z in place); coff[q] is the first
// position of colour q (host array of ncol + 1 entries).
template <typename T, typename B, typename O>
void dilu_apply(hipStream_t s, const DiluMesh &d, const int *coff, const O *off,
                const B *dinv, const T *v, T *z);

// Ut = U + eps * z (computed in FP64, stored as O) with eps = num /
// sqrt(zz[0]) (num / 1e-300 for zz = 0), zz in device memory; eps also
// stored to eps[0]
template <typename T, typename O>
void perturb(hipStream_t s, int n, const double *U, const double *zz, double num,
             const T *z, O *Ut, double *eps);

// w = z / dtau - (Rt - R0) / eps[0]   (A z with A = I/dtau - J, J z by
// finite difference)
template <typename T>
void op_finish(hipStream_t s, int N, const T *z, const double *dtau,
               const double *Rt, const double *R0, const double *eps, T *w);

// w = b - w
template <typename T>
void rhs_minus(hipStream_t s, int n, const double *b, T *w);

// dots[j] = <V_j, w> for j < k, V_j at V + j*ld; FP64 accumulation in a fixed
// order.  partials holds kDotBlocks*kMaxDots doubles.  With a device flag
// `on`, the kernels do nothing when *on == 0 (a pass decided on the device).
template <typename T>
void multidot(hipStream_t s, int n, int k, const T *V, size_t ld, const T *w,
              double *partials, double *dots, const int *on = nullptr);

// w -= sum_j h[j] V_j  (h in device memory; `on` as for multidot)
template <typename T>
void multiaxpy(hipStream_t s, int n, int k, const T *V, size_t ld,
               const double *h, T *w, const int *on = nullptr);

// Arnoldi step bookkeeping on the device (one thread each):
// arnoldi_norm: ar[0] = ww[0] - sum_{i<k} h[i]^2, ar[1] = *again = whether a
// second Gram-Schmidt pass is needed (ar[0] <= eta ww); arnoldi_finish: after
// that pass (run or skipped), with hn2 = ww2[0] if it ran, else ar[0]:
// ar[2] = hn = sqrt(max(hn2, 0)), ar[3] = 1/hn (1 when hn <= tiny), ar[4] =
// hn2, and w *= ar[3].
void arnoldi_norm(hipStream_t s, int k, const double *h, const double *ww, double *ar,
                  int *again, double eta);
template <typename T>
void arnoldi_finish(hipStream_t s, int n, T *w, const int *again, const double *ww2, double *ar,
                    double tiny);

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
