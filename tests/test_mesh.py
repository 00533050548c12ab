"""Mesh generator (host code of libfv3d_hip.so): geometry and connectivity
checked against closed forms and invariants (CPU only)."""
import math
import os

import numpy as np
import pytest

from fv3d_amd import solver as S
from fv3d_amd._lib import lib


def _box(n, tets, periodic, length=(1.0, 1.0, 1.0), tile=0):
    L = lib()
    L.init_solver(b"mesh-test")
    S.check(L.fvh_mesh_box(n[0], n[1], n[2], 0.0, 0.0, 0.0, length[0], length[1], length[2],
                           1 if tets else 0, periodic, tile))
    return S.mesh_arrays()


def _checks(m):
    N, F = m["n_cells"], m["n_faces"]
    fc = m["face_cells_file"].reshape(F, 2)
    n = m["face_n"].reshape(F, 3)
    A = m["face_area"]
    # closed cells: sum of outward n*A is zero
    s = np.zeros((N, 3))
    np.add.at(s, fc[:, 0], n * A[:, None])
    internal = fc[:, 1] >= 0
    np.add.at(s, fc[internal, 1], -(n * A[:, None])[internal])
    assert np.abs(s).max() < 1e-12 * A.max()
    # unit normals, orthonormal frames
    t1 = m["face_t1"].reshape(F, 3)
    t2 = m["face_t2"].reshape(F, 3)
    for v in (n, t1, t2):
        assert np.allclose((v * v).sum(1), 1.0, atol=1e-14)
    assert np.abs((n * t1).sum(1)).max() < 1e-14
    assert np.abs((n * t2).sum(1)).max() < 1e-14
    assert np.abs((t1 * t2).sum(1)).max() < 1e-14
    # cell face lists ascending, each face listed by its cells
    cf = m["cell_faces"].reshape(N, -1)
    nf = m["cell_n_faces"]
    for c in range(0, N, max(1, N // 200)):
        lst = cf[c, : nf[c]]
        assert np.all(np.diff(lst) > 0)
        for f in lst:
            assert c in fc[f]
    # normals point out of cells[0] (towards the image of cells[1])
    xc = m["cell_x"].reshape(N, 3)
    xf = m["face_x"].reshape(F, 3)
    assert np.all(((xf - xc[fc[:, 0]]) * n).sum(1) > 0)
    d2 = m["dist_cell_2"].reshape(F, 3)
    assert np.all(((-d2[internal]) * n[internal]).sum(1) > 0)
    # 0 < lambda < 1 inside, 1 on boundaries
    lam = m["face_lambda"]
    assert np.all((lam[internal] > 0) & (lam[internal] < 1))
    assert np.all(lam[~internal] == 1.0)
    # face owners: the first cell of a face has the lower index
    assert np.all(fc[internal, 0] < fc[internal, 1])
    return fc, internal


@pytest.mark.parametrize("tets", [False, True])
def test_periodic_box(tets):
    n = 5
    m = _box((n, n, n), tets, 7, length=(2.0, 2.0, 2.0))
    N, F = m["n_cells"], m["n_faces"]
    assert N == n ** 3 * (6 if tets else 1)
    assert F == (2 * N if tets else 3 * N)
    assert m["n_boundaries"] == 0
    assert m["global_volume"] == pytest.approx(8.0, rel=1e-14)
    h = 2.0 / n
    expect_v = h ** 3 / 6 if tets else h ** 3
    assert np.allclose(m["cell_volume"], expect_v, rtol=1e-13)
    fc, internal = _checks(m)
    assert internal.all()
    # periodic shifts are lattice vectors of the box
    sh = m["face_shift"].reshape(F, 3)
    assert set(np.unique(np.abs(sh))) <= {0.0, 2.0}
    assert np.count_nonzero(np.abs(sh).sum(1)) > 0
    if not tets:
        assert np.allclose(m["face_lambda"], 0.5)
        dx = m["cell_dx"].reshape(N, 3)
        assert np.allclose(dx, 2 * h * h, rtol=1e-14)


def test_box_boundaries_and_regions():
    m = _box((3, 4, 2), True, 1)  # periodic in x only
    assert m["region_names"] == ["flow", "ymin", "ymax", "zmin", "zmax"]
    assert list(m["region_is_boundary"]) == [0, 1, 1, 1, 1]
    B = m["n_boundaries"]
    assert B == 2 * 2 * (3 * 2 + 3 * 4)  # 2 triangles per boundary quad
    fc, internal = _checks(m)
    bf = m["boundary_face"]
    assert np.all(m["face_boundary"][bf] == np.arange(B))
    xf = m["face_x"].reshape(-1, 3)
    reg = m["boundary_id"]
    assert np.allclose(xf[bf[reg == 1], 1], 0.0)
    assert np.allclose(xf[bf[reg == 2], 1], 1.0)
    assert np.allclose(xf[bf[reg == 3], 2], 0.0)
    assert np.allclose(xf[bf[reg == 4], 2], 1.0)


def test_tiled_order_is_a_permutation_of_the_same_mesh():
    a = _box((4, 4, 4), True, 7, tile=2)
    b = _box((4, 4, 4), True, 7, tile=0)
    assert a["n_cells"] == b["n_cells"] and a["n_faces"] == b["n_faces"]
    ka = np.sort(np.round(a["cell_x"].reshape(-1, 3), 12).view("f8,f8,f8").ravel())
    kb = np.sort(np.round(b["cell_x"].reshape(-1, 3), 12).view("f8,f8,f8").ravel())
    assert np.array_equal(ka, kb)


def test_channel_mesh():
    L = lib()
    L.init_solver(b"mesh-test")
    S.check(L.fvh_mesh_channel(16, 6, 3, 0.05, 7))
    m = S.mesh_arrays()
    assert m["n_cells"] == 16 * 6 * 3 * 6
    assert m["region_names"] == ["flow", "inflow", "outflow", "bottom", "top", "zmin", "zmax"]
    fc, internal = _checks(m)
    # volume: integral of (1 - h(x)) over x in [-2,2] times depth 0.5 (the
    # bottom wall is piecewise linear between nodes)
    x = np.linspace(-2, 2, 17)
    hb = 0.1 * np.exp(-25 * x ** 2)
    vol = 0.5 * np.trapezoid(1 - hb, x)
    assert m["global_volume"] == pytest.approx(vol, rel=1e-12)
    reg = m["boundary_id"]
    counts = np.bincount(reg, minlength=7)
    assert list(counts[1:]) == [2 * 6 * 3, 2 * 6 * 3, 2 * 16 * 3, 2 * 16 * 3, 2 * 16 * 6, 2 * 16 * 6]
    # jitter is deterministic in the seed
    v1 = m["cell_volume"].copy()
    S.check(L.fvh_mesh_channel(16, 6, 3, 0.05, 7))
    assert np.array_equal(S.mesh_arrays()["cell_volume"], v1)
    S.check(L.fvh_mesh_channel(16, 6, 3, 0.05, 8))
    assert not np.array_equal(S.mesh_arrays()["cell_volume"], v1)


def _tangents_from_normal(n):
    """t1, t2 from unit normals n [F,3]: t1 = a/|a| with a = (-n_z, 0, n_x)
    where |n_x| > |n_y|, else (0, -n_z, n_y); t2 = b/|b| with b = t1 x n;
    |a| = sqrt((a_x a_x + a_y a_y) + a_z a_z) -- the order of operations the
    face-centred kernel uses to form them from the 32-B face records."""
    xs = np.abs(n[:, 0]) > np.abs(n[:, 1])
    a = np.stack([np.where(xs, -n[:, 2], 0.0), np.where(xs, 0.0, -n[:, 2]),
                  np.where(xs, n[:, 0], n[:, 1])], axis=1)
    l1 = np.sqrt(a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1] + a[:, 2] * a[:, 2])
    t1 = a / l1[:, None]
    b = np.stack([t1[:, 1] * n[:, 2] - t1[:, 2] * n[:, 1],
                  t1[:, 2] * n[:, 0] - t1[:, 0] * n[:, 2],
                  t1[:, 0] * n[:, 1] - t1[:, 1] * n[:, 0]], axis=1)
    l2 = np.sqrt(b[:, 0] * b[:, 0] + b[:, 1] * b[:, 1] + b[:, 2] * b[:, 2])
    return t1, b / l2[:, None]


@pytest.mark.parametrize("kind", ["box-tets", "box-hexes", "channel"])
def test_stored_tangents_follow_the_normal(kind):
    """The generated meshes store t1 and t2 bit for bit as formed from the
    normal, so the face-centred kernel reads 32-B records (n, area) and
    forms the tangents instead of loading 80-B frames.  On the jittered
    channel a changed summation order in |b| changes t2 on 461 of 3,780
    faces, and a multiply by the reciprocal changes most of t1 and t2; the
    boxes' regular normals exercise both branches of the formula."""
    if kind == "channel":
        L = lib()
        L.init_solver(b"mesh-test")
        S.check(L.fvh_mesh_channel(16, 6, 3, 0.05, 7))
        m = S.mesh_arrays()
    else:
        m = _box((5, 4, 3), kind == "box-tets", 7, length=(1.0, 0.8, 0.6), tile=2)
    F = m["n_faces"]
    n = m["face_n"].reshape(F, 3)
    t1, t2 = _tangents_from_normal(n)
    for got, want in ((t1, m["face_t1"]), (t2, m["face_t2"])):
        want = np.ascontiguousarray(want.reshape(F, 3))
        assert np.array_equal(np.ascontiguousarray(got).view(np.int64), want.view(np.int64))
    # both branches of the formula occur
    xs = np.abs(n[:, 0]) > np.abs(n[:, 1])
    assert xs.any() and (~xs).any()


def test_mesh_file_round_trip(tmp_path):
    L = lib()
    L.init_solver(b"mesh-test")
    S.check(L.fvh_mesh_channel(6, 4, 2, 0.05, 3))
    a = S.mesh_arrays()
    path = os.path.join(str(tmp_path), "c.fvm")
    S.check(L.fvh_mesh_write(path.encode()))
    L.init_mesh(path.encode())
    b = S.mesh_arrays()
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, b[k]), k
        else:
            assert v == b[k], k
