"""GPU tests of the coupled (multicolour block ILU(0)) JFNK preconditioner,
through the C-ABI (fvh_jfnk_precond_apply, fvh_jfnk_solve).

- M^-1 v on the device against the oracle's Dilu on the same state, pseudo
  CFL and vector: rel <= 1e-6 (both build the blocks by one-sided
  differences with the same step, by different routes: face-flux
  differences on the device, coloured residual differences in the oracle;
  they differ by rounding, ~1e-16/1e-7, amplified by the factorisation).
  The block-diagonal apply is checked the same way.
- The FP64 JFNK with the coupled preconditioner reaches the oracle's coupled
  JFNK solution (<= 1e-10 relative at stop norm 1e-12).
- On a larger channel the coupled FP64 and FP32 solves reach the block-Jacobi
  steady state (<= 1e-10 relative at stop norm 1e-12) with fewer Krylov
  iterations.
- M^-1 v on Gmsh meshes of prisms only (5 faces per cell, tables 5 slots
  wide under the 6-slot kernels) and of mixed cells (hexahedra, prisms,
  pyramids), with walls, inflow and outflow boundaries, against the oracle's
  Dilu (same bar).
"""
import numpy as np
import pytest

from fv3d_amd import configs
from oracle.jfnk import OracleJfnk
from tests import gmsh_fixtures as G
from tests.helpers import make_pair

pytestmark = pytest.mark.gpu


def _print(tag, **kv):
    print(tag, " ".join("%s=%s" % (k, v) for k, v in kv.items()), flush=True)


def _perturbed(U, seed):
    rng = np.random.default_rng(seed)
    V = U.copy()
    V[:, 0] *= 1.0 + 0.02 * rng.standard_normal(len(V))
    V[:, 1:4] += 0.05 * rng.standard_normal((len(V), 3))
    V[:, 4] *= 1.0 + 0.02 * rng.standard_normal(len(V))
    return V


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_dilu_apply_matches_oracle():
    cfg = configs.bump_channel_steady(nx=16, ny=4, nz=3)
    case, _, mesh = make_pair(cfg)
    U = _perturbed(case.state(), 3)
    case.set_state(U)
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    rng = np.random.default_rng(4)
    v = rng.standard_normal((case.n_cells, 5))
    for cfl in (10.0, 1e3, 1e6):
        _print("dilu apply", cfl=cfl)
        dtau = oj.local_dt(U, cfl)
        zd = case.jfnk_precond_apply(cfl, v, "block", "dilu")
        zo = oj.dilu(U, dtau).apply(v).reshape(-1, 5)
        e = _rel(zd, zo)
        zb = case.jfnk_precond_apply(cfl, v, "block", "none")
        zbo = np.einsum("nij,nj->ni", np.linalg.inv(oj.blocks(U, dtau)), v)
        eb = _rel(zb, zbo)
        _print("dilu apply", cfl=cfl, rel=e, rel_block=eb, differ=_rel(zd, zb))
        assert e <= 1e-6 and eb <= 1e-6
        assert _rel(zd, zb) > 1e-3  # the coupling acts
    case.close()


@pytest.mark.parametrize("kind", ["prisms", "mixed"])
def test_dilu_apply_on_gmsh_cells(tmp_path, kind):
    if kind == "prisms":
        X, cells = G.prism_cells((4, 3, 3))
    else:
        X, cells = G.mixed_cells((5, 4, 3))
    path = tmp_path / ("%s.msh" % kind)
    G.write_msh(path, X, cells, jitter_seed=7)
    cfg = configs.all_boundaries_box(steps=10)
    cfg["mesh"] = ("gmsh", dict(path=str(path), scale=(1.0, 1.0, 1.0)))
    _print("dilu apply gmsh", kind=kind, cells=len(cells))
    case, _, mesh = make_pair(cfg)
    assert mesh["max_cell_faces"] == (5 if kind == "prisms" else 6)
    U = _perturbed(case.state(), 5)
    case.set_state(U)
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    v = np.random.default_rng(6).standard_normal((case.n_cells, 5))
    for cfl in (10.0, 1e3):
        dtau = oj.local_dt(U, cfl)
        zd = case.jfnk_precond_apply(cfl, v, "block", "dilu")
        zo = oj.dilu(U, dtau).apply(v).reshape(-1, 5)
        e = _rel(zd, zo)
        _print("dilu apply gmsh", kind=kind, cfl=cfl, rel=e)
        assert e <= 1e-6
    case.close()


def test_jfnk_dilu_matches_oracle_jfnk():
    cfg = configs.bump_channel_steady(nx=16, ny=4, nz=3)
    case, _, mesh = make_pair(cfg)
    U0 = case.state()
    st, _ = case.jfnk(tol=1e-12, precision="fp64", coupling="dilu")
    Ug = case.state()
    _print("jfnk dilu fp64", **st)
    assert st["status"] == 0
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    Uo, st_o, _ = oj.solve(U0, tol=1e-12, coupling="dilu")
    e = _rel(Ug, Uo)
    _print("device vs oracle coupled JFNK", rel=e, newton_dev=st["newton_steps"],
           newton_oracle=st_o["newton_steps"], gmres_dev=st["gmres_iterations"],
           gmres_oracle=st_o["gmres_iterations"])
    assert st_o["status"] == 0
    assert e <= 1e-10
    case.close()


def test_jfnk_dilu_same_solution_fewer_iterations():
    cfg = configs.bump_channel_steady(nx=32, ny=8, nz=3)
    case, _, _ = make_pair(cfg)
    U0 = case.state()
    stb, _ = case.jfnk(tol=1e-12, precision="fp64")
    Ub = case.state()
    res = {}
    for prec in ("fp64", "fp32"):
        case.set_state(U0)
        st, _ = case.jfnk(tol=1e-12, precision=prec, coupling="dilu")
        res[prec] = (st, _rel(case.state(), Ub))
        _print("jfnk dilu", precision=prec, rel_vs_block=res[prec][1],
               newton=st["newton_steps"], gmres=st["gmres_iterations"],
               seconds=st["seconds"], block_newton=stb["newton_steps"],
               block_gmres=stb["gmres_iterations"], block_seconds=stb["seconds"])
    assert stb["status"] == 0
    for prec, (st, e) in res.items():
        assert st["status"] == 0
        assert e <= 1e-10
        assert st["gmres_iterations"] < stb["gmres_iterations"]
    case.close()
