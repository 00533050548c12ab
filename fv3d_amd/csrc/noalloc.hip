// Allocation-free residual: fvh_time_derivative_noalloc, the residual of
// finite_volume_time_derivative for a state the caller passes in device
// memory, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops).  The ghost rows, face states and viscous
// gradients that the residual's kernels pass between them live in a
// file-scope device array of fixed size instead of library allocations.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>

#include "fv3d_hip.h"
#include "fv_device.h"
#include "fvh_internal.h"
#include "runtime.h"

namespace fvh {
namespace {

// 64 MiB: meshes up to about 190,000 tetrahedra (44 doubles per cell) or
// 130,000 hexahedra (60 doubles per cell) with viscous second-order fluxes.
constexpr size_t kScratchDoubles = (size_t)8 << 20;
__device__ __align__(256) double scratch[kScratchDoubles];

// The array's device address per device (hipGetSymbolAddress with that
// device current); the array's name is not a device address in host code.
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};

double *scratch_of_current_device()
{
    int d = 0;
    FVH_HIP_CHECK(hipGetDevice(&d));
    if (d < 0 || d >= kMaxDevices)
        throw std::runtime_error("fvh_time_derivative_noalloc: device index out of range");
    if (!g_scratch[d])
    {
        void *p = nullptr;
        FVH_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[d] = static_cast<double *>(p);
    }
    return g_scratch[d];
}

size_t round256(size_t n_doubles) { return (n_doubles + 31) / 32 * 32; }

} // namespace

void noalloc_setup() { (void)scratch_of_current_device(); }

} // namespace fvh

using namespace fvh;

extern "C" int fvh_time_derivative_noalloc(const double *U, double *R, void *stream)
{
    try
    {
        if (!U || !R)
            throw std::runtime_error("fvh_time_derivative_noalloc: null state or residual");
        const RuntimeView v = runtime_view(false);
        const dev::DevMesh &m = v.m;
        const size_t N = m.N, B = m.B;
        const size_t n_gphi = round256(10 * B);
        const size_t n_fs = v.linear ? round256((N * m.maxf + B) * 8) : 0;
        const size_t n_vg = v.viscous ? round256(12 * N) : 0;
        if (n_gphi + n_fs + n_vg > kScratchDoubles)
        {
            char msg[200];
            std::snprintf(msg, sizeof(msg),
                          "fvh_time_derivative_noalloc: the mesh needs %zu doubles of scratch, "
                          "the maximum is %zu (64 MiB)",
                          n_gphi + n_fs + n_vg, kScratchDoubles);
            throw std::runtime_error(msg);
        }
        double *base = scratch_of_current_device();
        double *gphi = base;
        double *fs = v.linear ? base + n_gphi : nullptr;
        double *vg = v.viscous ? base + n_gphi + n_fs : nullptr;
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : v.stream;
        dev::launch_ghost(s, m, v.P, U, gphi, fs, nullptr);
        if (v.linear || v.viscous)
        {
            dev::GradOut go{};
            go.vg = vg;
            go.fs = fs;
            dev::launch_gradient(s, m, v.P, U, gphi, go, v.linear, nullptr);
        }
        dev::StageParams sp{};
        sp.U_in = U;
        sp.R_out = R;
        sp.vg = vg;
        sp.fs = fs;
        sp.gphi = gphi;
        dev::launch_residual(s, m, v.P, sp, v.flux, v.linear, v.viscous);
        FVH_HIP_CHECK(hipGetLastError());
    }
    catch (const std::exception &e)
    {
        set_error(e.what());
        return FVH_ERR_STATE;
    }
    return FVH_OK;
}
