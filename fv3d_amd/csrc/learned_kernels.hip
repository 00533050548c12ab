// Kernels of the learned block preconditioner: features, BF16-MFMA MLP
// inference and training (forward + backward, FP32 accumulation), Adam,
// and the block conversions.
//
// MFMA tiles on one 64-wide wavefront, D[16][16] += A[16][K] B[K][16]:
// - the dense layers (K a multiple of 32): gfx950's v_mfma_f32_16x16x32_bf16.
//   Lane l holds A[l % 16][8 (l / 16) + q] and B[8 (l / 16) + q][l % 16]
//   (q = 0..7); the activation fragments of a 16-row tile are loaded once
//   and reused for every 16-column output tile;
// - the weight-gradient products over the 16 rows of a tile (K = 16):
//   v_mfma_f32_16x16x16_bf16, lane l holding A[l % 16][4 (l / 16) + q] and
//   B[4 (l / 16) + q][l % 16] (q = 0..3).
// Both give D[4 (l / 16) + q][l % 16] (q = 0..3) in lane l.  Matrices are
// staged row-major in LDS in FP32; fragments are rounded to BF16 (round to
// nearest even) when they are loaded.
#include <cfloat>
#include <cmath>

#include "fv_physics.h"
#include "learned.h"

namespace fvh {
namespace nn {

using dev::DevMesh;
using dev::PhysDev;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kB = 256;

__device__ __forceinline__ short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ __bf16 f2bf8(float f)
{
    return __builtin_bit_cast(__bf16, f2bf(f));
}

__device__ __forceinline__ int lane() { return threadIdx.x & 63; }

// A[i][k] of the 16x16 tile (row0, k0) of M (row-major, leading dim ld);
// TRANS: A[i][k] = M[k][i].
template <bool TRANS>
__device__ __forceinline__ bf16x4 frag_a(const float *M, int ld, int row0, int k0)
{
    const int l = lane();
    const int i = row0 + (l & 15), kb = k0 + 4 * (l >> 4);
    bf16x4 r;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        r[q] = f2bf(TRANS ? M[(kb + q) * ld + i] : M[i * ld + kb + q]);
    return r;
}

// B[k][n] of the 16x16 tile (k0, col0) of M (row-major [K][N]); TRANS:
// B[k][n] = M[n][k].
template <bool TRANS>
__device__ __forceinline__ bf16x4 frag_b(const float *M, int ld, int k0, int col0)
{
    const int l = lane();
    const int n = col0 + (l & 15), kb = k0 + 4 * (l >> 4);
    bf16x4 r;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        r[q] = f2bf(TRANS ? M[n * ld + kb + q] : M[(kb + q) * ld + n]);
    return r;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// The K = 32 fragments: A[i][k] of the 16x32 tile (row0, k0) of M
// (row-major, leading dim ld) and B[k][n] of the 32x16 tile (k0, col0) of M
// (row-major [K][N]; TRANS: B[k][n] = M[n][k]).
__device__ __forceinline__ bf16x8 frag_a8(const float *M, int ld, int row0, int k0)
{
    const int l = lane();
    const float *p = M + (row0 + (l & 15)) * ld + k0 + 8 * (l >> 4);
    bf16x8 r;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        r[q] = f2bf8(p[q]);
    return r;
}

template <bool TRANS>
__device__ __forceinline__ bf16x8 frag_b8(const float *M, int ld, int k0, int col0)
{
    const int l = lane();
    const int n = col0 + (l & 15), kb = k0 + 8 * (l >> 4);
    bf16x8 r;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        r[q] = f2bf8(TRANS ? M[n * ld + kb + q] : M[(kb + q) * ld + n]);
    return r;
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// out[16][N] = act(A[16][K] W[K][N] + bias), A and W row-major in LDS.
// ACT 0: identity, 1: tanh.
template <int K, int N, int ACT>
__device__ __forceinline__ void dense(const float *A, const float *W, const float *bias,
                                      float *out)
{
    static_assert(K % 32 == 0 && N % 16 == 0, "dense: K a multiple of 32");
    const int l = lane();
    bf16x8 a[K / 32];
#pragma unroll
    for (int kt = 0; kt < K / 32; ++kt)
        a[kt] = frag_a8(A, K, 0, 32 * kt);
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt)
    {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < K / 32; ++kt)
            acc = mfma32(a[kt], frag_b8<false>(W, N, 32 * kt, 16 * nt), acc);
        const int j = 16 * nt + (l & 15);
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            float v = acc[q] + bias[j];
            if (ACT == 1)
                v = tanhf(v);
            out[(4 * (l >> 4) + q) * N + j] = v;
        }
    }
}

// out[16][N] = A[16][K] B^T, B row-major [N][K] (B^T[k][n] = B[n][k]).
template <int K, int N>
__device__ __forceinline__ void dense_bt(const float *A, const float *B, float *out)
{
    static_assert(K % 32 == 0 && N % 16 == 0, "dense_bt: K a multiple of 32");
    const int l = lane();
    bf16x8 a[K / 32];
#pragma unroll
    for (int kt = 0; kt < K / 32; ++kt)
        a[kt] = frag_a8(A, K, 0, 32 * kt);
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt)
    {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < K / 32; ++kt)
            acc = mfma32(a[kt], frag_b8<true>(B, K, 32 * kt, 16 * nt), acc);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            out[(4 * (l >> 4) + q) * N + 16 * nt + (l & 15)] = acc[q];
    }
}

// acc[M/16][N/16] += X^T D over the 16 rows of a tile: X [16][M], D [16][N].
template <int M, int N>
__device__ __forceinline__ void outer_acc(const float *X, const float *D, f32x4 *acc)
{
#pragma unroll
    for (int mt = 0; mt < M / 16; ++mt)
    {
        const bf16x4 a = frag_a<true>(X, M, 16 * mt, 0);
#pragma unroll
        for (int nt = 0; nt < N / 16; ++nt)
            acc[mt * (N / 16) + nt] = mfma(a, frag_b<false>(D, N, 0, 16 * nt),
                                           acc[mt * (N / 16) + nt]);
    }
}

template <int M, int N>
__device__ __forceinline__ void store_acc(const f32x4 *acc, float *dst)
{
    const int l = lane();
#pragma unroll
    for (int mt = 0; mt < M / 16; ++mt)
#pragma unroll
        for (int nt = 0; nt < N / 16; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                dst[(16 * mt + 4 * (l >> 4) + q) * N + 16 * nt + (l & 15)] =
                    acc[mt * (N / 16) + nt][q];
}

__device__ __forceinline__ void load_theta(const float *__restrict__ theta, float *sT)
{
    for (int i = threadIdx.x; i < kParams; i += blockDim.x)
        sT[i] = theta[i];
}

// ---- features ---------------------------------------------------------------
__global__ __launch_bounds__(kB) void k_features(DevMesh m, PhysDev P,
                                                 const double *__restrict__ U, double cfl,
                                                 Refs r, float *__restrict__ feat)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.N)
        return;
    double pc[10];
    dev::load_cons(U, c, pc);
    dev::con_to_prim(P, pc);
    const double vol = m.vol[c];
    const double h = cbrt(vol);
    auto q5 = [&](const double *p, double *q) {
        q[0] = p[0] / r.rho;
        q[1] = p[5] / r.a;
        q[2] = p[6] / r.a;
        q[3] = p[7] / r.a;
        q[4] = p[8] / r.p;
    };
    double qc[5];
    q5(pc, qc);
    double g[15] = {0.0};
    double mom[6] = {0.0};
    int nb_bnd = 0;
    const int nf = m.ncf[c];
    for (int s = 0; s < nf; ++s)
    {
        const int f = m.cf[(size_t)c * m.maxf + s];
        const int c0 = m.fcell[2 * f], c1 = m.fcell[2 * f + 1];
        const bool own0 = c0 == c;
        const int nb = own0 ? c1 : c0;
        dev::Frame fr;
        dev::load_frame(m, f, fr);
        double pn[10];
        if (nb < m.N)
        {
            dev::load_cons(U, nb, pn);
            dev::con_to_prim(P, pn);
        }
        else
        {
            dev::ghost_state(m, P, nb - m.N, fr, pc, pn);
            dev::con_to_prim(P, pn);
            ++nb_bnd;
        }
        double qn[5];
        q5(pn, qn);
        const double sg = own0 ? fr.area : -fr.area; // outward area vector scale
        for (int q = 0; q < 5; ++q)
        {
            const double qf = 0.5 * (qc[q] + qn[q]);
            for (int d = 0; d < 3; ++d)
                g[q * 3 + d] += qf * fr.n[d] * sg;
        }
        const double a = fr.area;
        mom[0] += a * fr.n[0] * fr.n[0];
        mom[1] += a * fr.n[1] * fr.n[1];
        mom[2] += a * fr.n[2] * fr.n[2];
        mom[3] += a * fr.n[0] * fr.n[1];
        mom[4] += a * fr.n[0] * fr.n[2];
        mom[5] += a * fr.n[1] * fr.n[2];
    }
    const double cs = sqrt(P.kappa * pc[8] / pc[0]);
    float *o = feat + (size_t)c * kIn;
    o[0] = (float)(qc[0] - 1.0);
    o[1] = (float)(pc[5] / cs);
    o[2] = (float)(pc[6] / cs);
    o[3] = (float)(pc[7] / cs);
    o[4] = (float)(qc[4] - 1.0);
    for (int i = 0; i < 15; ++i)
        o[5 + i] = (float)(g[i] * h / vol);
    const double h2 = h * h;
    for (int i = 0; i < 6; ++i)
        o[20 + i] = (float)(mom[i] / h2);
    o[26] = (float)(nb_bnd * 0.25);
    o[27] = (float)(sqrt(pc[5] * pc[5] + pc[6] * pc[6] + pc[7] * pc[7]) / cs);
    for (int i = kFeat; i < kIn; ++i)
        o[i] = 0.f;
}

__device__ __forceinline__ void scales(const Refs &r, double *s)
{
    s[0] = 1.0 / r.rho;
    s[1] = s[2] = s[3] = 1.0 / (r.rho * r.a);
    s[4] = 1.0 / (r.rho * r.a * r.a);
}

// The block M_c = I/dtau_c - J_c (J_c the first-order diagonal Jacobian
// block) written as M_c = (I + cfl K_c) / dtau_c: K_c = -(V_c/ds1_c) J_c is
// dimensionless and independent of the CFL number.  The network predicts
// S K_c S^-1.
__global__ __launch_bounds__(kB) void k_targets(int N, const double *__restrict__ Minv,
                                                const double *__restrict__ dtau, double cfl,
                                                Refs r, float *__restrict__ T)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N)
        return;
    double s[5];
    scales(r, s);
    double M[25];
    dev::invert5(Minv + (size_t)c * 25, M);
    const double dt = dtau[c];
    float *o = T + (size_t)c * kOut;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
        {
            const double k = (dt * M[i * 5 + j] - (i == j ? 1.0 : 0.0)) / cfl;
            o[i * 5 + j] = (float)(s[i] * k / s[j]);
        }
    for (int i = kBlk; i < kOut; ++i)
        o[i] = 0.f;
}

// Minv_c = dtau_c (I + cfl S^-1 out_c S)^-1
__global__ __launch_bounds__(kB) void k_apply_blocks(int N, const float *__restrict__ out,
                                                     const double *__restrict__ dtau,
                                                     double cfl, Refs r,
                                                     double *__restrict__ Minv)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N)
        return;
    double s[5];
    scales(r, s);
    const float *y = out + (size_t)c * kOut;
    double A[25], Ai[25];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j)
            A[i * 5 + j] = (i == j ? 1.0 : 0.0) + cfl * ((double)y[i * 5 + j] * s[j] / s[i]);
    dev::invert5(A, Ai);
    const double dt = dtau[c];
    for (int i = 0; i < 25; ++i)
        Minv[(size_t)c * 25 + i] = dt * Ai[i];
}

// ---- inference -------------------------------------------------------------
// 4 waves per block, 16 cells per wave and iteration; the block's loop trip
// count is uniform so that every wave reaches each barrier.
constexpr int kInferWaves = 4;

__global__ __launch_bounds__(64 * kInferWaves) void k_forward(int N,
                                                              const float *__restrict__ theta,
                                                              const float *__restrict__ feat,
                                                              float *__restrict__ out)
{
    __shared__ float sT[kParams];
    __shared__ float sX[kInferWaves][16 * kIn];
    __shared__ float sH1[kInferWaves][16 * kHid];
    __shared__ float sH2[kInferWaves][16 * kHid];
    __shared__ float sY[kInferWaves][16 * kOut];
    load_theta(theta, sT);
    __syncthreads();
    const int w = threadIdx.x >> 6, l = lane();
    const int per_iter = 16 * kInferWaves;
    for (int base = blockIdx.x * per_iter; base < N; base += gridDim.x * per_iter)
    {
        const int c0 = base + 16 * w;
        for (int i = l; i < 16 * kIn; i += 64)
        {
            const int c = c0 + i / kIn;
            sX[w][i] = c < N ? feat[(size_t)c0 * kIn + i] : 0.f;
        }
        __syncthreads();
        dense<kIn, kHid, 1>(sX[w], sT + oW1, sT + oB1, sH1[w]);
        __syncthreads();
        dense<kHid, kHid, 1>(sH1[w], sT + oW2, sT + oB2, sH2[w]);
        __syncthreads();
        dense<kHid, kOut, 0>(sH2[w], sT + oW3, sT + oB3, sY[w]);
        __syncthreads();
        for (int i = l; i < 16 * kOut; i += 64)
        {
            const int c = c0 + i / kOut;
            if (c < N)
                out[(size_t)c0 * kOut + i] = sY[w][i];
        }
        __syncthreads();
    }
}

// ---- training --------------------------------------------------------------
// One wave per block; grid-stride over 16-sample tiles; the weight
// gradients accumulate in MFMA registers and are written once per block.
__global__ __launch_bounds__(64) void k_loss_grad(int n, const float *__restrict__ theta,
                                                  const float *__restrict__ feat,
                                                  const float *__restrict__ target,
                                                  float *__restrict__ partials)
{
    __shared__ float sT[kParams];
    __shared__ float sX[16 * kIn];
    __shared__ float sA1[16 * kHid];
    __shared__ float sA2[16 * kHid];
    __shared__ float sY[16 * kOut];
    __shared__ float sD3[16 * kOut];
    __shared__ float sD2[16 * kHid];
    __shared__ float sD1[16 * kHid];
    __shared__ float sDb[kHid + kHid + kOut];
    __shared__ float sLoss[64];
    load_theta(theta, sT);
    const int l = lane();
    for (int i = l; i < kHid + kHid + kOut; i += 64)
        sDb[i] = 0.f;
    f32x4 g1[(kIn / 16) * (kHid / 16)], g2[(kHid / 16) * (kHid / 16)],
        g3[(kHid / 16) * (kOut / 16)];
    for (auto &a : g1)
        a = f32x4{0.f, 0.f, 0.f, 0.f};
    for (auto &a : g2)
        a = f32x4{0.f, 0.f, 0.f, 0.f};
    for (auto &a : g3)
        a = f32x4{0.f, 0.f, 0.f, 0.f};
    float loss = 0.f;
    const float gscale = 2.f / ((float)n * (float)kBlk);
    __syncthreads();
    for (int c0 = 16 * blockIdx.x; c0 < n; c0 += 16 * gridDim.x)
    {
        for (int i = l; i < 16 * kIn; i += 64)
        {
            const int c = c0 + i / kIn;
            sX[i] = c < n ? feat[(size_t)c0 * kIn + i] : 0.f;
        }
        __syncthreads();
        dense<kIn, kHid, 1>(sX, sT + oW1, sT + oB1, sA1);
        __syncthreads();
        dense<kHid, kHid, 1>(sA1, sT + oW2, sT + oB2, sA2);
        __syncthreads();
        dense<kHid, kOut, 0>(sA2, sT + oW3, sT + oB3, sY);
        __syncthreads();
        for (int i = l; i < 16 * kOut; i += 64)
        {
            const int c = c0 + i / kOut, o = i % kOut;
            float d = 0.f;
            if (c < n && o < kBlk)
            {
                const float e = sY[i] - target[(size_t)c0 * kOut + i];
                loss += e * e;
                d = gscale * e;
            }
            sD3[i] = d;
        }
        __syncthreads();
        // layer 3: dW3 += A2^T dZ3, db3, dA2 = dZ3 W3^T
        outer_acc<kHid, kOut>(sA2, sD3, g3);
        if (l < kOut)
        {
            float s = 0.f;
            for (int c = 0; c < 16; ++c)
                s += sD3[c * kOut + l];
            sDb[2 * kHid + l] += s;
        }
        dense_bt<kOut, kHid>(sD3, sT + oW3, sD2);
        __syncthreads();
        for (int i = l; i < 16 * kHid; i += 64)
            sD2[i] *= 1.f - sA2[i] * sA2[i];
        __syncthreads();
        // layer 2
        outer_acc<kHid, kHid>(sA1, sD2, g2);
        {
            float s = 0.f;
            for (int c = 0; c < 16; ++c)
                s += sD2[c * kHid + l];
            sDb[kHid + l] += s;
        }
        dense_bt<kHid, kHid>(sD2, sT + oW2, sD1);
        __syncthreads();
        for (int i = l; i < 16 * kHid; i += 64)
            sD1[i] *= 1.f - sA1[i] * sA1[i];
        __syncthreads();
        // layer 1
        outer_acc<kIn, kHid>(sX, sD1, g1);
        {
            float s = 0.f;
            for (int c = 0; c < 16; ++c)
                s += sD1[c * kHid + l];
            sDb[l] += s;
        }
        __syncthreads();
    }
    float *dst = partials + (size_t)blockIdx.x * (kParams + 1);
    store_acc<kIn, kHid>(g1, dst + oW1);
    store_acc<kHid, kHid>(g2, dst + oW2);
    store_acc<kHid, kOut>(g3, dst + oW3);
    dst[oB1 + l] = sDb[l];
    dst[oB2 + l] = sDb[kHid + l];
    if (l < kOut)
        dst[oB3 + l] = sDb[2 * kHid + l];
    sLoss[l] = loss;
    __syncthreads();
    if (l == 0)
    {
        float s = 0.f;
        for (int i = 0; i < 64; ++i)
            s += sLoss[i];
        dst[kParams] = s / ((float)n * (float)kBlk);
    }
}

__global__ __launch_bounds__(kB) void k_fold_grad(int nblk, const float *__restrict__ partials,
                                                  float *__restrict__ grad,
                                                  float *__restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kParams)
        return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b)
        s += partials[(size_t)b * (kParams + 1) + p];
    if (p < kParams)
        grad[p] = s;
    else
        loss[0] = s;
}

__global__ __launch_bounds__(kB) void k_adam(float *__restrict__ theta,
                                             const float *__restrict__ grad,
                                             float *__restrict__ m, float *__restrict__ v,
                                             float c1, float c2, float lr, float b1, float b2,
                                             float eps)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams)
        return;
    const float g = grad[p];
    const float mp = b1 * m[p] + (1.f - b1) * g;
    const float vp = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mp;
    v[p] = vp;
    theta[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

__global__ __launch_bounds__(kB) void k_block_error(int N, const double *__restrict__ A,
                                                    const double *__restrict__ B,
                                                    double *__restrict__ partials)
{
    __shared__ double sh[kB];
    double mx = 0.0, sum = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x)
    {
        double d2 = 0.0, b2 = 0.0;
        for (int i = 0; i < 25; ++i)
        {
            const double a = A[(size_t)c * 25 + i], b = B[(size_t)c * 25 + i];
            d2 += (a - b) * (a - b);
            b2 += b * b;
        }
        const double e = sqrt(d2 / (b2 > 0.0 ? b2 : 1e-300));
        mx = fmax(mx, e);
        sum += e;
    }
    dev::block_sum(sh, sum);
    double nmx = -mx;
    dev::block_min(sh, nmx);
    if (threadIdx.x == 0)
    {
        partials[2 * blockIdx.x] = -nmx;
        partials[2 * blockIdx.x + 1] = sum;
    }
}

__global__ __launch_bounds__(kB) void k_fold_error(int nblk, int N,
                                                   const double *__restrict__ partials,
                                                   double *__restrict__ out)
{
    __shared__ double sh[kB];
    double mx = 0.0, sum = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kB)
    {
        mx = fmax(mx, partials[2 * b]);
        sum += partials[2 * b + 1];
    }
    dev::block_sum(sh, sum);
    double nmx = -mx;
    dev::block_min(sh, nmx);
    if (threadIdx.x == 0)
    {
        out[0] = -nmx;
        out[1] = sum / (N > 0 ? N : 1);
    }
}

inline int blocks_for(long n) { return (int)((n + kB - 1) / kB); }

} // namespace

void features(hipStream_t s, const DevMesh &m, const PhysDev &P, const double *U, double cfl,
              Refs r, float *feat)
{
    hipLaunchKernelGGL(k_features, dim3(blocks_for(m.N)), dim3(kB), 0, s, m, P, U, cfl, r,
                       feat);
}

void targets(hipStream_t s, int N, const double *Minv, const double *dtau, double cfl, Refs r,
             float *T)
{
    hipLaunchKernelGGL(k_targets, dim3(blocks_for(N)), dim3(kB), 0, s, N, Minv, dtau, cfl, r,
                       T);
}

void forward(hipStream_t s, int N, const float *theta, const float *feat, float *out)
{
    const int per_block = 16 * kInferWaves;
    int grid = (N + per_block - 1) / per_block;
    grid = grid < 2048 ? grid : 2048;
    if (grid < 1)
        return;
    hipLaunchKernelGGL(k_forward, dim3(grid), dim3(64 * kInferWaves), 0, s, N, theta, feat,
                       out);
}

void apply_blocks(hipStream_t s, int N, const float *out, const double *dtau, double cfl,
                  Refs r, double *Minv)
{
    hipLaunchKernelGGL(k_apply_blocks, dim3(blocks_for(N)), dim3(kB), 0, s, N, out, dtau, cfl,
                       r, Minv);
}

void loss_grad(hipStream_t s, int n, const float *theta, const float *feat,
               const float *target, float *partials, float *grad, float *loss)
{
    hipLaunchKernelGGL(k_loss_grad, dim3(kTrainBlocks), dim3(64), 0, s, n, theta, feat,
                       target, partials);
    hipLaunchKernelGGL(k_fold_grad, dim3(blocks_for(kParams + 1)), dim3(kB), 0, s,
                       kTrainBlocks, partials, grad, loss);
}

void adam(hipStream_t s, float *theta, const float *grad, float *m, float *v, int t, float lr,
          float beta1, float beta2, float eps)
{
    const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(kParams)), dim3(kB), 0, s, theta, grad, m, v,
                       c1, c2, lr, beta1, beta2, eps);
}

void block_error(hipStream_t s, int N, const double *A, const double *B, double *partials,
                 double *out)
{
    const int grid = 512;
    hipLaunchKernelGGL(k_block_error, dim3(grid), dim3(kB), 0, s, N, A, B, partials);
    hipLaunchKernelGGL(k_fold_error, dim3(1), dim3(kB), 0, s, grid, N, partials, out);
}

} // namespace nn
} // namespace fvh
