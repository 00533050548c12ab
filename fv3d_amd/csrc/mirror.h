// Host-mirror residual path (internal): the reference's face-ordered data
// flow of finite_volume_time_derivative (src/Solver/src/finite_volume/
// finite_volume.c:81-124) over device copies of the reference's AoS arrays,
// so that a hook registered with set_update / set_update_gradients /
// set_calc_flux that is not a built-in can run on the host between the
// device passes, reading and writing solver_data exactly as in the
// reference:
//   phi  [(N+B)][10]  phi_total (domain rows + ghost rows)
//   g[3] [(N+B)][10]  grad_phi_total_x/y/z
//   L, R [F][10]      phi_total_left / phi_total_right
//   flux [F][5]       flux
#pragma once

#include <hip/hip_runtime.h>

#include "fv_device.h"

namespace fvh {
namespace mirror {

struct MirrorMesh
{
    int N, F, B, maxf;
    const int *cf;        // [N*maxf] cell faces, ascending
    const int *ncf;       // [N]
    const int *fcell;     // [F*2] (cells[1] = N + boundary on boundary faces)
    const double *fgA;    // [F][10] n, t1, t2, area
    const double *fgB;    // [F][6] n, area, lambda
    const double *d1, *d2; // [F*3] dist_cell_1 / dist_cell_2
    const double *vol, *svol;
    const int *bnd_face, *bnd_region, *region_type;
};

// Green-Gauss gradients of all 10 variables of the domain cells
// (reconstruction.c:20-93): rows 0..N-1 of gx/gy/gz; ghost rows untouched.
void gradients(hipStream_t s, const MirrorMesh &m, const double *phi, double *gx,
               double *gy, double *gz);
// Ghost gradient rows N+b as the gradient scatter leaves them (minus the
// boundary face's contribution), before the update_gradients hook.
void ghost_raw(hipStream_t s, const MirrorMesh &m, const double *phi, double *gx, double *gy,
               double *gz);
// Ghost gradient rows N+b (boundary.c:325-412, the built-in
// ns_update_gradients).
void ghost_gradients(hipStream_t s, const MirrorMesh &m, double *gx, double *gy, double *gz);
// Face states (reconstruction.c:142-255): linear = limited linear
// reconstruction (Barth-Jespersen when bj, limiter.c:49-106), otherwise the
// cell values; the right state of a boundary face is its ghost row.
void reconstruct(hipStream_t s, const MirrorMesh &m, const dev::PhysDev &P, int linear,
                 const double *phi, const double *gx, const double *gy, const double *gz,
                 double *L, double *R);
// The built-in calc_ns_flux (flux.c:23-198): flux[F][5] from the face states
// and the gradients of the face's two rows.
void flux(hipStream_t s, const MirrorMesh &m, const dev::PhysDev &P, int kind, bool viscous,
          const double *L, const double *R, const double *gx, const double *gy,
          const double *gz, double *flux);
// dU/dt = -(sum_f +-flux_f A_f)/V of the domain cells (finite_volume.c:96-123),
// each cell's faces in ascending order; R_out [N][5].
void sum_fluxes(hipStream_t s, const MirrorMesh &m, const double *flux, double *R_out);

} // namespace mirror
} // namespace fvh
