"""fv3d_amd — MI355X-native residual and time-stepping path of fv3d.

The product is the C-ABI shared library ``fv3d_amd/libfv3d_hip.so`` (HIP
kernels for gfx950 + C++ host runtime) exporting the reference's Solver API
(include/fv3d_solver.h) and the fvh_* extension API (include/fv3d_hip.h).
This package is the Python host mirror over that boundary: ctypes bindings
with the reference's names (:mod:`fv3d_amd.solver`) and a JSON driver that
accepts the reference's configuration keys (:mod:`fv3d_amd.driver`).

Importing the package does not load the library; :func:`fv3d_amd.lib`
does, and raises if it has not been built — there is no fallback path.
"""
from ._lib import lib, lib_path  # noqa: F401

__all__ = ["lib", "lib_path"]
