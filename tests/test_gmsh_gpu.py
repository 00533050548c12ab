"""GPU parity on a Gmsh mesh of mixed cells (hexahedra, prisms, pyramids;
up to 6 faces per cell, 5-face cells among them), read by fvh_mesh_gmsh:
dU/dt and the state after 10 RK3-3 steps against the oracle on the same
mesh arrays, with all six boundary types, viscosity, linear reconstruction
and the Barth-Jespersen limiter.  Bar: 1e-12 relative per component."""
import numpy as np
import pytest

from fv3d_amd import configs
from tests import gmsh_fixtures as G
from tests.helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu


def test_mixed_gmsh_mesh_parity(tmp_path):
    X, cells = G.mixed_cells(n=(5, 4, 3))
    path = tmp_path / "mixed.msh"
    G.write_msh(path, X, cells, jitter_seed=11)
    cfg = configs.all_boundaries_box(steps=10)
    cfg["mesh"] = ("gmsh", dict(path=str(path), scale=(1.0, 0.8, 0.6)))
    print("mixed gmsh parity", len(cells), "cells", flush=True)
    case, o, mesh = make_pair(cfg)
    assert mesh["max_cell_faces"] == 6
    Rg = case.time_derivative()
    Ro = o.time_derivative()
    er = rel_err(Rg, Ro)
    rc, _ = case.run()
    n, _ = o.timedisc("RK3-3", 10, True, 1e-10, 0.0, 1e300)
    eu = rel_err(case.state(), o.state())
    print("mixed gmsh parity res_rel=%g u_rel=%g" % (er, eu), flush=True)
    assert rc == 0 and n == 10
    assert er <= 1e-12 and eu <= 1e-12
    case.close()
