"""CPU checks of the batched mesh (fvh_mesh_batch): the union of the cases'
boxes holds each case's mesh unchanged up to index offsets, and no face
joins two cases."""
import ctypes

import numpy as np

from fv3d_amd import solver as S
from fv3d_amd._lib import lib


def _box(n, tile):
    L = lib()
    S.check(L.fvh_mesh_box(n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1, 7, tile))
    return S.mesh_arrays()


def test_union_is_the_cases_with_offsets():
    L = lib()
    L.init_solver(b"batch-mesh")
    sizes, tile = [3, 5, 4], 2
    singles = [_box(n, tile) for n in sizes]
    nh = (ctypes.c_int * 3)(*sizes)
    S.check(L.fvh_mesh_batch(3, nh, 1.0, tile))
    u = S.mesh_arrays()
    co = np.cumsum([0] + [m["n_cells"] for m in singles])
    fo = np.cumsum([0] + [m["n_faces"] for m in singles])
    assert u["n_cells"] == co[-1] and u["n_faces"] == fo[-1] and u["n_boundaries"] == 0
    fc = u["face_cells"].reshape(-1, 2)
    for k, m in enumerate(singles):
        cs, fs = slice(co[k], co[k + 1]), slice(fo[k], fo[k + 1])
        assert np.array_equal(u["cell_volume"][cs], m["cell_volume"])
        assert np.array_equal(u["cell_x"].reshape(-1, 3)[cs], m["cell_x"].reshape(-1, 3))
        assert np.array_equal(fc[fs], m["face_cells"].reshape(-1, 2) + co[k])
        cf = u["cell_faces"].reshape(u["n_cells"], -1)[cs]
        assert np.array_equal(cf, m["cell_faces"].reshape(m["n_cells"], -1) + fo[k])
        assert np.array_equal(u["face_shift"].reshape(-1, 3)[fs], m["face_shift"].reshape(-1, 3))
        # every face of case k joins two cells of case k
        assert np.all((fc[fs] >= co[k]) & (fc[fs] < co[k + 1]))
