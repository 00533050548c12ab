"""GPU tests of the steady Newton-Krylov path (fvh_jfnk_*), through the C-ABI.

- The device's block-Jacobi blocks (per-cell one-sided differences of the
  first-order inviscid residual) against the oracle's blocks, computed by
  coloured global differences of the oracle residual: rel <= 1e-6 (both
  are one-sided differences with the same step; they differ by rounding in
  the differenced residuals, ~1e-16/1e-7).
- The FP64 JFNK solution is a fixed point of the oracle's residual (stop
  norm of the oracle residual <= 1e-11 when the device stops at 1e-12) and
  agrees with the oracle's own JFNK solution to <= 1e-10 relative.
- The FP32-Krylov variant converges to the same solution as FP64
  (<= 1e-10 relative).
- Every variant (FP32 Krylov with block Jacobi, with the coupled DILU
  preconditioner, and with the learned blocks in DILU) reaches the oracle's
  FP64 JFNK solution to <= 1e-10 relative at stop norm 1e-12; each error is
  printed.
"""
import math

import numpy as np
import pytest

from fv3d_amd import configs
from oracle.jfnk import OracleJfnk
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


def test_blocks_match_oracle():
    cfg = configs.bump_channel_steady(nx=16, ny=4, nz=3)
    case, _, mesh = make_pair(cfg)
    U = _perturbed(case.state(), 1)
    case.set_state(U)
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    N = case.n_cells
    for cfl in (5.0, 1e4):
        dtau = np.zeros(N)
        minv = np.zeros((N, 5, 5))
        case.jfnk_blocks(cfl, "block", dtau, minv)
        dt_o = oj.local_dt(U, cfl)
        assert np.allclose(dtau, dt_o, rtol=1e-14, atol=0)
        M_o = oj.blocks(U, dt_o)
        M_g = np.linalg.inv(minv)
        e = float(np.abs(M_g - M_o).max() / np.abs(M_o).max())
        _print("blocks", cfl=cfl, rel=e)
        assert e <= 1e-6
        case.jfnk_blocks(cfl, "scalar", dtau, minv)
        assert np.allclose(minv, np.eye(5)[None] * dtau[:, None, None], rtol=0, atol=0)
    case.close()


def test_jfnk_fp64_matches_oracle_jfnk():
    cfg = configs.bump_channel_steady(nx=16, ny=4, nz=3)
    case, o, mesh = make_pair(cfg)
    U0 = case.state()
    st, hist = case.jfnk(tol=1e-12, precision="fp64")
    Ug = case.state()
    _print("jfnk fp64", **st)
    assert st["status"] == 0 and st["norm"] <= 1e-12
    assert len(hist) == st["newton_steps"]
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    n_o = oj.stop_norm(oj.residual(Ug))
    _print("oracle stop norm of the device solution", norm=n_o)
    assert n_o <= 1e-11
    Uo, st_o, _ = oj.solve(U0, tol=1e-12)
    e = _rel(Ug, Uo)
    _print("device vs oracle JFNK solution", rel=e, newton_dev=st["newton_steps"],
           newton_oracle=st_o["newton_steps"])
    assert st_o["status"] == 0
    assert e <= 1e-10
    case.close()


def test_jfnk_fp32_same_solution():
    cfg = configs.bump_channel_steady(nx=32, ny=8, nz=3)
    case, _, _ = make_pair(cfg)
    U0 = case.state()
    st64, _ = case.jfnk(tol=1e-12, precision="fp64")
    U64 = case.state()
    case.set_state(U0)
    st32, _ = case.jfnk(tol=1e-12, precision="fp32")
    U32 = case.state()
    e = _rel(U32, U64)
    _print("jfnk fp32 vs fp64", rel=e, newton64=st64["newton_steps"],
           newton32=st32["newton_steps"], gmres64=st64["gmres_iterations"],
           gmres32=st32["gmres_iterations"], s64=st64["seconds"], s32=st32["seconds"])
    assert st64["status"] == 0 and st32["status"] == 0
    assert e <= 1e-10
    case.close()


def test_jfnk_block_beats_scalar():
    cfg = configs.bump_channel_steady(nx=16, ny=4, nz=3)
    case, _, _ = make_pair(cfg)
    U0 = case.state()
    stb, _ = case.jfnk(tol=1e-6, precond="block")
    case.set_state(U0)
    sts, _ = case.jfnk(tol=1e-6, precond="scalar", max_newton=400)
    _print("block vs scalar", gmres_block=stb["gmres_iterations"],
           gmres_scalar=sts["gmres_iterations"], status_scalar=sts["status"])
    assert stb["status"] == 0
    assert stb["gmres_iterations"] < sts["gmres_iterations"]
    case.close()


def test_jfnk_variants_vs_oracle_fp64():
    from fv3d_amd import learned
    cfg = configs.bump_channel_steady(nx=16, ny=4, nz=3)
    case, o, mesh = make_pair(cfg)
    U0 = case.state()
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    Uo, st_o, _ = oj.solve(U0, tol=1e-12)
    assert st_o["status"] == 0
    learned.load_pretrained()
    coupled = dict(coupling="dilu", krylov_dims=40, max_restarts=2)
    variants = {"fp64_block": dict(precision="fp64", precond="block"),
                "fp32_block": dict(precision="fp32", precond="block"),
                "fp32_dilu": dict(precision="fp32", precond="block", **coupled),
                "learned_fp32_dilu": dict(precision="fp32", precond="learned", **coupled)}
    for name, kw in variants.items():
        case.set_state(U0)
        st, _ = case.jfnk(tol=1e-12, max_newton=400, **kw)
        e = _rel(case.state(), Uo)
        _print("variant vs oracle FP64 JFNK", variant=name, rel=e, newton=st["newton_steps"],
               fallback=st["precond_fallbacks"])
        assert st["status"] == 0
        assert e <= 1e-10, name
    case.close()
