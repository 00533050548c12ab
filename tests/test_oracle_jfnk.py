"""CPU checks that pin the oracle's Newton-Krylov pieces (no GPU).

- gmres: on explicit nonsymmetric matrices it returns numpy's solve when the
  tolerance is tight, and its Givens estimate of the residual equals the
  true residual |b - A x| / |b| at a loose tolerance (a wrong rotation sign
  or a transposed Hessenberg index breaks both).
- coloured block differences equal single-cell differences (a wrong
  colouring lets a neighbour's perturbation leak into a block).
- a tiny channel converges to the steady state under the oracle JFNK, and
  the block-Jacobi preconditioner needs fewer Krylov steps than I/dtau.
"""
import math

import numpy as np
import pytest

from fv3d_amd import configs
from fv3d_amd import solver as S
from fv3d_amd._lib import lib
from oracle.jfnk import OracleJfnk, colouring, gmres


def _channel(nx, ny, nz):
    cfg = configs.bump_channel_steady(nx=nx, ny=ny, nz=nz)
    L = lib()
    L.init_solver(b"oracle-jfnk")
    S.check(L.fvh_mesh_channel(nx, ny, nz, 0.05, 20261203))
    mesh = S.mesh_arrays()
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    u0 = 0.5 * math.sqrt(1.4)
    U = np.zeros((mesh["n_cells"], 5))
    U[:, 0] = 1.0
    U[:, 1] = u0
    U[:, 4] = 1 / 0.4 + 0.5 * u0 * u0
    return oj, mesh, U


@pytest.mark.parametrize("n,m,restarts", [(40, 40, 0), (60, 15, 60)])
def test_gmres_solves_dense_systems(n, m, restarts):
    rng = np.random.default_rng(n)
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.5
    b = rng.standard_normal(n)
    x, its, rel = gmres(lambda v: A @ v, b, m, restarts, 1e-13)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-11)
    # loose tolerance: the Givens estimate is the true residual
    x, its, rel = gmres(lambda v: A @ v, b, m, restarts, 1e-3)
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rel <= 1e-3
    assert abs(true - rel) <= 1e-9 + 1e-6 * rel


def test_coloured_blocks_equal_single_cell_blocks():
    oj, mesh, U = _channel(6, 2, 2)
    rng = np.random.default_rng(3)
    U = U * (1.0 + 0.02 * rng.standard_normal(U.shape))
    colour = colouring(mesh)
    fc = np.asarray(mesh["face_cells_file"]).reshape(-1, 2)
    inner = fc[(fc[:, 1] >= 0) & (fc[:, 1] < mesh["n_cells"])]
    assert np.all(colour[inner[:, 0]] != colour[inner[:, 1]])
    dtau = oj.local_dt(U, 10.0)
    M = oj.blocks(U, dtau)
    base = oj.residual(U, first=True)
    for c in (0, 7, mesh["n_cells"] - 1):
        J = np.zeros((5, 5))
        for k in range(5):
            Up = U.copy()
            h = 1e-7 * max(abs(U[c, k]), 1.0)
            Up[c, k] += h
            J[:, k] = (oj.residual(Up, first=True)[c] - base[c]) / (Up[c, k] - U[c, k])
        assert np.allclose(M[c], np.eye(5) / dtau[c] - J, rtol=1e-12, atol=1e-12 * np.abs(J).max())


def test_oracle_jfnk_converges_on_tiny_channel():
    oj, mesh, U = _channel(8, 2, 2)
    Ub, stb, hist = oj.solve(U, tol=1e-9)
    assert stb["status"] == 0
    assert oj.stop_norm(oj.residual(Ub)) <= 1e-9
    # the steady state is not the initial free stream (the bump deflects it)
    assert np.abs(Ub[:, 2]).max() > 1e-3
    Us, sts, _ = oj.solve(U, tol=1e-9, precond="scalar", max_newton=400)
    assert stb["gmres_iterations"] < sts["gmres_iterations"]
    if sts["status"] == 0:
        assert np.linalg.norm(Us - Ub) / np.linalg.norm(Ub) <= 1e-7
