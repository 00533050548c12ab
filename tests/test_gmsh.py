"""The Gmsh 4.1 reader and mesh preprocessing (fvh_mesh_gmsh), host code only.

- Structured boxes written as .msh (hexahedra, Kuhn tetrahedra) give exactly
  the mesh that the box generator builds from the same vertices and cells
  (same face numbering, cell face lists, boundaries and geometry); the
  generator is pinned by tests/test_mesh.py.
- A jittered mixed mesh (hexahedra, prisms, pyramids) matches the oracle's
  restatement of pyMeshFV3D (oracle/meshproc.py): integer arrays exactly,
  geometry to 1e-13.
- The oracle restatement is pinned by closed forms: unit-element volumes
  (tetrahedron 1/6, prism 1/2, pyramid 1/3, hexahedron 1), the total volume
  of the mixed box, closed cells (sum of outward n A = 0), lambda = 1/2 on
  uniform hexahedra, boundary distance = h/2, per-axis scaling.
- Malformed files are rejected with an error status.
"""
import os

import numpy as np
import pytest

from fv3d_amd import solver as S
from fv3d_amd._lib import lib
from oracle import meshproc as MP
from tests import gmsh_fixtures as G


def _boundary(n):
    b = S.mesh_struct().boundaries.contents
    return dict(n=np.ctypeslib.as_array(b.n, (3 * n,)).copy().reshape(n, 3),
                t1=np.ctypeslib.as_array(b.t1, (3 * n,)).copy().reshape(n, 3),
                t2=np.ctypeslib.as_array(b.t2, (3 * n,)).copy().reshape(n, 3),
                distance=np.ctypeslib.as_array(b.distance, (n,)).copy())


def _cells_extra(n):
    c = S.mesh_struct().cells.contents
    return dict(type=np.ctypeslib.as_array(c.type, (n,)).copy(),
                id=np.ctypeslib.as_array(c.id, (n,)).copy())


def _load(path, scale=(1.0, 1.0, 1.0)):
    L = lib()
    L.init_solver(b"gmsh-test")
    S.check(L.fvh_mesh_gmsh(str(path).encode(), *scale))
    m = S.mesh_arrays()
    m.update({"bnd_" + k: v for k, v in _boundary(m["n_boundaries"]).items()})
    m.update({"cell_" + k: v for k, v in _cells_extra(m["n_cells"]).items()})
    return m


_INT = ["n_cells", "n_faces", "n_boundaries", "cell_n_faces", "cell_faces", "face_cells",
        "face_boundary", "boundary_id", "boundary_face", "region_is_boundary"]
_FLOAT = ["cell_x", "cell_volume", "cell_dx", "face_x", "face_n", "face_t1", "face_t2",
          "face_area", "face_lambda", "dist_cell_1", "dist_cell_2"]


@pytest.mark.parametrize("tets", [False, True])
def test_box_matches_generator(tmp_path, tets):
    n = (3, 2, 2)
    X, cells = G.box_cells(n, tets=tets)
    path = tmp_path / "box.msh"
    G.write_msh(path, X, cells)
    a = _load(path)
    L = lib()
    L.init_solver(b"gmsh-test")
    S.check(L.fvh_mesh_box(n[0], n[1], n[2], 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, int(tets), 0, 0))
    b = S.mesh_arrays()
    b.update({"bnd_" + k: v for k, v in _boundary(b["n_boundaries"]).items()})
    for k in _INT:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    for k in _FLOAT + ["bnd_n", "bnd_t1", "bnd_t2", "bnd_distance"]:
        assert np.allclose(a[k], b[k], rtol=1e-14, atol=1e-15), k
    assert a["region_names"] == b["region_names"]


def test_mixed_matches_oracle(tmp_path):
    X, cells = G.mixed_cells()
    path = tmp_path / "mixed.msh"
    G.write_msh(path, X, cells, jitter_seed=5)
    scale = (2.0, 1.0, 0.5)
    a = _load(path, scale)
    o = MP.process(path, scale)
    N, F, B = o["n_cells"], o["n_faces"], o["n_boundaries"]
    assert (a["n_cells"], a["n_faces"], a["n_boundaries"]) == (N, F, B)
    assert sorted(set(o["cell_type"].tolist())) == [MP.HEXAEDER, MP.PRISM, MP.PYRAMID]
    assert np.array_equal(a["cell_type"], o["cell_type"])
    assert np.array_equal(a["cell_id"], o["cell_region"])
    maxf = a["max_cell_faces"]
    assert np.array_equal(a["cell_faces"].reshape(N, maxf), o["cell_faces"])
    assert np.array_equal(a["cell_n_faces"], o["cell_n_faces"])
    assert np.array_equal(a["face_cells_file"].reshape(F, 2), o["face_cells"])
    assert np.array_equal(a["face_boundary"], o["face_boundary"])
    assert np.array_equal(a["boundary_face"], o["boundary_face"])
    assert np.array_equal(a["boundary_id"], o["boundary_region"])
    assert np.array_equal(a["region_is_boundary"], o["region_is_boundary"])
    for k, ko, shape in [("cell_x", "cell_x", (N, 3)), ("cell_volume", "cell_volume", (N,)),
                         ("cell_dx", "cell_dx", (N, 3)), ("face_x", "face_x", (F, 3)),
                         ("face_n", "face_n", (F, 3)), ("face_t1", "face_t1", (F, 3)),
                         ("face_t2", "face_t2", (F, 3)), ("face_area", "face_area", (F,)),
                         ("face_lambda", "face_lambda", (F,)),
                         ("bnd_n", "boundary_n", (B, 3)), ("bnd_t1", "boundary_t1", (B, 3)),
                         ("bnd_t2", "boundary_t2", (B, 3)),
                         ("bnd_distance", "boundary_distance", (B,))]:
        assert np.allclose(np.asarray(a[k]).reshape(shape), o[ko], rtol=1e-13, atol=1e-14), k


def test_oracle_closed_forms(tmp_path):
    I = np.eye(3)
    o = np.zeros(3)
    tet = np.array([o, I[0], I[1], I[2]])
    assert abs(MP.volume(MP.TETRAEDER, tet) - 1 / 6) < 1e-15
    prism = np.array([o, I[0], I[1], I[2], I[0] + I[2], I[1] + I[2]])
    assert abs(MP.volume(MP.PRISM, prism) - 0.5) < 1e-15
    pyr = np.array([o, I[0], I[0] + I[1], I[1], [0.5, 0.5, 1.0]])
    assert abs(MP.volume(MP.PYRAMID, pyr) - 1 / 3) < 1e-15
    hexa = np.array([o, I[0], I[0] + I[1], I[1], I[2], I[0] + I[2], I[0] + I[1] + I[2],
                     I[1] + I[2]])
    assert abs(MP.volume(MP.HEXAEDER, hexa) - 1.0) < 1e-15
    assert abs(MP.area(MP.QUADRANGLE, hexa[:4]) - 1.0) < 1e-15
    assert abs(MP.area(MP.TRIANGLE, tet[:3]) - 0.5) < 1e-15

    # planar faces (no jitter): the cell volumes tile the box exactly
    X, cells = G.mixed_cells()
    path = tmp_path / "mixed.msh"
    G.write_msh(path, X, cells)
    m = MP.process(path, (2.0, 1.0, 0.5))
    assert abs(m["cell_volume"].sum() - 1.0) < 1e-13
    N = m["n_cells"]
    fc = m["face_cells"]
    s = np.zeros((N, 3))
    for f in range(m["n_faces"]):
        v = m["face_n"][f] * m["face_area"][f]
        s[fc[f, 0]] += v
        if fc[f, 1] >= 0:
            s[fc[f, 1]] -= v
    assert np.abs(s).max() < 1e-14
    # every face normal points out of its first cell
    for f in range(m["n_faces"]):
        assert np.dot(m["face_n"][f], m["face_x"][f] - m["cell_x"][fc[f, 0]]) > 0

    X, cells = G.box_cells((3, 3, 3))
    path = tmp_path / "hex.msh"
    G.write_msh(path, X, cells)
    m = MP.process(path)
    inner = m["face_cells"][:, 1] >= 0
    assert np.allclose(m["face_lambda"][inner], 0.5, rtol=0, atol=1e-15)
    assert np.allclose(m["boundary_distance"], 0.5 / 3, rtol=0, atol=1e-15)
    assert np.allclose(m["cell_dx"], 2.0 / 9.0, rtol=0, atol=1e-15)


def test_malformed_files_rejected(tmp_path):
    X, cells = G.box_cells((2, 2, 2))
    good = tmp_path / "good.msh"
    G.write_msh(good, X, cells)
    text = good.read_text()
    L = lib()
    L.init_solver(b"gmsh-test")
    bad = tmp_path / "version.msh"
    bad.write_text(text.replace("4.1 0 8", "2.2 0 8"))
    assert L.fvh_mesh_gmsh(str(bad).encode(), 1.0, 1.0, 1.0) != 0
    # drop the last boundary element (tags stay 1..n after renumbering):
    # a boundary face without a boundary element
    lines = text.split("\n")
    i = lines.index("$Elements")
    blocks, ne, _, _ = [int(x) for x in lines[i + 1].split()]
    j, last_bnd = i + 2, None
    for _ in range(blocks):
        d, ent, g, m = [int(x) for x in lines[j].split()]
        if d == 2:
            last_bnd = (j, m)
        j += 1 + m
    jb, m = last_bnd
    hdr = lines[jb].split()
    lines[jb] = " ".join(hdr[:3] + [str(m - 1)])
    del lines[jb + m]
    tag = 1
    j = i + 2
    for _ in range(blocks):
        d, ent, g, m = [int(x) for x in lines[j].split()]
        for q in range(m):
            parts = lines[j + 1 + q].split()
            lines[j + 1 + q] = " ".join([str(tag)] + parts[1:])
            tag += 1
        j += 1 + m
    lines[i + 1] = "%d %d 1 %d" % (blocks, ne - 1, ne - 1)
    bad2 = tmp_path / "missing.msh"
    bad2.write_text("\n".join(lines))
    assert L.fvh_mesh_gmsh(str(bad2).encode(), 1.0, 1.0, 1.0) != 0
    assert b"has no boundary element" in L.fvh_last_error()
    assert L.fvh_mesh_gmsh(str(tmp_path / "none.msh").encode(), 1.0, 1.0, 1.0) != 0
    assert L.fvh_mesh_gmsh(str(good).encode(), 1.0, 1.0, 1.0) == 0


def test_meshprep_writes_reference_layout(tmp_path):
    from fv3d_amd import meshprep

    X, cells = G.mixed_cells()
    G.write_msh(tmp_path / "m.msh", X, cells, jitter_seed=3)
    params = {"General": {"title": {"value": "mixed"}},
              "Reader": {"mesh_reader": {"value": "Gmsh"}, "mesh_file": {"value": "m.msh"}},
              "Process": {"mesh_scale": {"value": [1.0, 2.0, 1.0]},
                          "n_partitions": {"value": 4}}}
    out = meshprep.preprocess(params, str(tmp_path))
    assert os.path.exists(out)
    a = _load(tmp_path / "m.msh", (1.0, 2.0, 1.0))
    L = lib()
    L.init_solver(b"gmsh-test")
    L.init_mesh(out.encode())
    b = S.mesh_arrays()
    for k in _INT:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    for k in _FLOAT:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert a["region_names"] == b["region_names"]
