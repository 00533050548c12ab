"""Mesh files in the reference's HDF5 layout (libhdf5 loaded at run time):
a written mesh has the groups, datasets and attributes of the reference's
writer (tools/pyMeshFV3D/pymeshfv3d/write.py:19-202, checked with the
independent h5dump tool when present) and reads back through init_mesh
into the identical solver mesh; the raw container round-trips the same
way."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from fv3d_amd import solver as S
from fv3d_amd._lib import lib

KEYS = ["cell_volume", "cell_x", "cell_dx", "cell_faces", "cell_n_faces", "face_cells",
        "face_boundary", "face_x", "face_n", "face_t1", "face_t2", "face_area", "face_lambda",
        "dist_cell_1", "dist_cell_2", "boundary_id", "boundary_face", "region_is_boundary",
        "face_shift"]


def _channel():
    L = lib()
    L.init_solver(b"hdf5-test")
    S.check(L.fvh_mesh_channel(6, 3, 2, 0.05, 7))
    return S.mesh_arrays()


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert a["region_names"] == b["region_names"]


@pytest.mark.parametrize("ext", [".h5", ".fvm"])
def test_mesh_round_trip(tmp_path, ext):
    L = lib()
    if ext == ".h5" and not L.fvh_hdf5_available():
        pytest.skip("libhdf5 not loadable")
    m0 = _channel()
    path = os.path.join(str(tmp_path), "mesh" + ext)
    S.check(L.fvh_mesh_write(path.encode()))
    with open(path, "rb") as f:
        sig = f.read(8)
    assert (sig == b"\x89HDF\r\n\x1a\n") == (ext == ".h5")
    L.init_mesh(path.encode())
    _same(S.mesh_arrays(), m0)


def test_hdf5_layout_matches_reference_writer(tmp_path):
    L = lib()
    dump = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                       if os.path.exists("/opt/conda/bin/h5dump") else None)
    if not L.fvh_hdf5_available() or dump is None:
        pytest.skip("libhdf5 or h5dump not available")
    _channel()
    path = os.path.join(str(tmp_path), "mesh.h5")
    S.check(L.fvh_mesh_write(path.encode()))
    out = subprocess.run([dump, "-H", path], capture_output=True, text=True, check=True).stdout
    want = {
        "/": (["dimension", "is_partitioned"], []),
        "VERTICES": (["n_vertices"], ["x"]),
        "CELLS": (["n_cells", "max_cell_vertices", "max_cell_faces"],
                  ["id", "type", "n_vertices", "vertices", "n_faces", "faces", "volume", "x", "dx"]),
        "FACES": (["n_faces", "max_face_vertices"],
                  ["type", "cells", "boundary", "n_vertices", "vertices", "x", "n", "t1", "t2",
                   "area", "lambda"]),
        "BOUNDARIES": (["n_boundaries", "max_boundary_vertices"],
                       ["id", "type", "n_vertices", "vertices", "face", "n", "t1", "t2",
                        "distance"]),
        "REGIONS": (["n_regions"], ["name", "is_boundary"]),
    }
    for group, (attrs, dsets) in want.items():
        if group != "/":
            assert 'GROUP "%s"' % group in out, group
        for a in attrs:
            assert 'ATTRIBUTE "%s"' % a in out, (group, a)
        for d in dsets:
            assert 'DATASET "%s"' % d in out, (group, d)
    # shapes of two datasets the reader takes as [n][3] and [n][2]
    assert "H5T_STD_I32LE" in out or "H5T_STD_I64LE" in out
    assert "H5T_IEEE_F64LE" in out
    assert "H5T_STRING" in out
