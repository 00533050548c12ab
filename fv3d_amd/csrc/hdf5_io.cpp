// HDF5 mesh reader: the reference stores meshes as HDF5
// (src/Solver/src/mesh/mesh.c:121-630).  libhdf5 is optional at run time.
#include "fvh_internal.h"

namespace fvh {

int read_hdf5_mesh(RawMesh &m, const char *path)
{
    (void)m;
    set_error(std::string("HDF5 mesh files are not supported by this build: ") +
              path + " (convert to .fvm)");
    return -1;
}

} // namespace fvh
