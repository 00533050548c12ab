// Internal declarations shared by the host runtime translation units of
// libfv3d_hip.so.  Not part of the public C-ABI.
#pragma once

#include <string>
#include <vector>

#include "fv3d_solver.h"
#include "fv3d_hip.h"

namespace fvh {

// Error reporting: fvh_* return codes carry the message here; functions with
// reference names abort (the reference's BM_CHECK_EXPRESSION behaviour).
void set_error(const std::string &msg);
[[noreturn]] void fatal(const std::string &msg);

// Mesh as written by the reference's preprocessor (Py/write.py) before the
// solver's calc_mesh_metrics: second face cell -1 on boundary faces.
struct RawMesh
{
    int dimension = 3;
    std::vector<double> vertices;              // [nv*3]
    int max_cell_vertices = 0, max_cell_faces = 0, max_face_vertices = 0;
    std::vector<int> cell_region, cell_type, cell_nv, cell_vertices;
    std::vector<int> cell_nf, cell_faces;      // faces ascending, -1 padded
    std::vector<double> cell_volume, cell_x, cell_dx;
    std::vector<int> face_type, face_nv, face_vertices, face_cells,
        face_boundary;
    std::vector<double> face_x, face_n, face_t1, face_t2, face_area,
        face_lambda, face_shift;               // shift: periodic extension
    int max_boundary_vertices = 0;
    std::vector<int> bnd_region, bnd_type, bnd_nv, bnd_vertices, bnd_face;
    std::vector<double> bnd_n, bnd_t1, bnd_t2, bnd_distance;
    std::vector<std::string> region_names;
    std::vector<int> region_is_boundary;

    int n_cells() const { return (int)cell_volume.size(); }
    int n_faces() const { return (int)face_area.size(); }
    int n_boundaries() const { return (int)bnd_face.size(); }
};

// Generators (mesh_gen.cpp)
int generate_box(RawMesh &m, int nx, int ny, int nz, double x0, double y0,
                 double z0, double lx, double ly, double lz, bool tets,
                 int periodic, int tile);
int generate_channel(RawMesh &m, int nx, int ny, int nz, double jitter,
                     unsigned long long seed);

// Gmsh 4.1 ASCII mesh file (gmsh_reader.cpp): parse, scale the vertices per
// axis and preprocess into the reference layout; returns the cell count.
int read_gmsh(RawMesh &m, const char *path, const double scale[3]);

// Mesh file I/O (mesh_io.cpp): HDF5 (.h5, the reference layout) or the raw
// container, chosen by extension when writing and by signature when
// reading.  Throw on error.
int write_mesh_file(const RawMesh &m, const char *path);
int read_mesh_file(RawMesh &m, const char *path);

// Install a raw mesh as solver_mesh (reader layout) and run the metrics
// (solver.cpp).
void install_mesh(RawMesh &&m);

// Periodic shift of each face (translation of cells[1]'s image), kept beside
// solver_mesh because Faces_t has no field for it.
const std::vector<double> &face_shift();

} // namespace fvh
