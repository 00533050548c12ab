"""Mesh preprocessing from the reference's preprocessing parameter file.

The keys are those of pyMeshFV3D's parameter file
(tools/pyMeshFV3D/pymeshfv3d/bin/pyMeshFV3D.py:62-87, e.g.
misc/gmsh/naca0012/naca0012.json): General/title, Reader/mesh_reader
("Gmsh"), Reader/mesh_file (a Gmsh 4.1 ASCII file), Process/mesh_scale.
The mesh is read and preprocessed by the library (fvh_mesh_gmsh) and written
as <title>.mesh.h5 in the reference layout (or <title>.mesh.fvm, the raw
container, when libhdf5 cannot be loaded).  Process/n_partitions is read
and ignored: the whole mesh runs on one GPU.  The 1D "Lin1D" reader is not
part of this build (the solver path is three-dimensional).

    python -m fv3d_amd.meshprep naca0012.json
"""
import argparse
import json
import os
import sys

from . import solver as S
from ._lib import lib


def _v(node, *path, default=None):
    for k in path:
        if not isinstance(node, dict) or k not in node:
            return default
        node = node[k]
    if isinstance(node, dict) and "value" in node:
        return node["value"]
    return node


def preprocess(params, base_dir=".", out_dir=None):
    """Read, preprocess and write the mesh; returns the written path."""
    reader = _v(params, "Reader", "mesh_reader", default="Gmsh")
    if reader != "Gmsh":
        raise ValueError("mesh reader %r is not supported (Gmsh only)" % reader)
    path = _v(params, "Reader", "mesh_file")
    if path is None:
        raise KeyError("Reader/mesh_file is required")
    if not os.path.isabs(path):
        path = os.path.join(base_dir, path)
    scale = [float(x) for x in _v(params, "Process", "mesh_scale", default=[1.0, 1.0, 1.0])]
    title = _v(params, "General", "title", default="untitled")
    L = lib()
    L.init_solver(title.encode())
    S.check(L.fvh_mesh_gmsh(path.encode(), *scale))
    ext = ".mesh.h5" if L.fvh_hdf5_available() else ".mesh.fvm"
    out = os.path.join(out_dir if out_dir is not None else base_dir, title + ext)
    S.check(L.fvh_mesh_write(out.encode()))
    L.free_mesh()
    L.free_solver()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Preprocess a Gmsh mesh for the solver.")
    ap.add_argument("parameter_file")
    a = ap.parse_args(argv)
    with open(a.parameter_file) as f:
        params = json.load(f)
    base = os.path.dirname(os.path.abspath(a.parameter_file))
    print(preprocess(params, base, os.getcwd()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
