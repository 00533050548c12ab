// Device runtime shared by the host translation units that drive kernels
// (solver.cpp owns the state; the Newton-Krylov driver borrows it).
// Internal; not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "fv_device.h"

namespace fvh {

#define FVH_HIP_CHECK(expr)                                                    \
    do                                                                         \
    {                                                                          \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            throw std::runtime_error(std::string("HIP error: ") +             \
                                     hipGetErrorString(e_) + " at " +          \
                                     __FILE__ + ":" + std::to_string(__LINE__)); \
    } while (0)

// The device state set up by init_finite_volume with the built-in hooks.
struct RuntimeView
{
    hipStream_t stream;
    dev::DevMesh m;
    dev::PhysDev P;
    int flux;      // dev::FluxKind
    int linear;    // linear reconstruction
    int viscous;   // mu != 0
    double *U;     // current conservative state [N][5]
    double vglob;  // total volume
    double rho_ref, p_ref; // density and pressure of the flow region state
};

// Throws unless init_finite_volume ran with the built-in hooks and the
// reconstruction and limiter are set.
// allocate = false: nothing is allocated (the gradient and face-state
// buffers may then be null).
RuntimeView runtime_view(bool allocate = true);

// Set-up of the allocation-free residual on the current device (the device
// address of its scratch array); called by init_finite_volume.
void noalloc_setup();

// The explicit stepping buffers: the ping-pong state pair (U[*cur] is
// current), the RK register k, the residual R and the LSERK tables of the
// scheme set by init_explicit.
struct StageBuffers
{
    double *U[2];
    int *cur;
    double *k, *R;
    int stages;
    const double *rk_a, *rk_g;
};
StageBuffers stage_buffers();

// Face tables of the face-centred residual kernel for cell tiles
// [c0[t], c0[t+1]) of at most dev::kTile cells (tetrahedral meshes):
// per tile every face of its cells once (off[t] .. off[t+1] - 1 of ent) as
// {face, left row, right row, left slot | right slot << 16}, 0xffff for a
// side outside the tile.  Reads the host mesh.
void build_tile_faces(const std::vector<int> &c0, std::vector<int> &off,
                      std::vector<int4> &ent);
// The same table in the 8-byte DevMesh tile_faces layout, each entry checked
// against its tile's slots; false when the mesh is too large for it.
bool pack_tile_faces_table(const std::vector<int> &c0, const std::vector<int> &off,
                           const std::vector<int4> &ent, std::vector<unsigned long long> &out);

// Enqueues R = dU/dt(U) (finite_volume_time_derivative semantics, the
// reference's ghost rows recomputed from U) on the runtime stream.  fp32:
// the gradient and flux arithmetic in FP32 (for Jacobian-vector products),
// with the state read from U32 when given (an FP32 copy; U is then unused
// on the linear-reconstruction path).
void eval_residual(const double *U, double *R, bool fp32 = false, const float *U32 = nullptr);

// The current state was changed outside the time loop: the cached residual
// and host mirror are stale.
void state_changed();

// Releases device work buffers of the Newton-Krylov driver and the
// per-mesh buffers of the learned preconditioner (called when the device
// state is freed).
void jfnk_release();
void learned_release_mesh();
void batch_release();

// Learned preconditioner (learned.cpp): whether a network is loaded, its
// inverse blocks for the current state at the pseudo steps dtau, and the
// max / mean per-cell relative Frobenius error of blocks A against B.
bool learned_ready();
void learned_blocks(const RuntimeView &v, double cfl, const double *dtau, double *Minv);
void learned_block_error(const RuntimeView &v, const double *A, const double *B,
                         double *err);

} // namespace fvh
