"""CPU checks that pin the oracle's multicolour block ILU(0) preconditioner
(coupling "dilu") without a GPU.

- The distance-2 coloured differences give the same off-diagonal blocks as
  perturbing one cell at a time (bit-for-bit: a first-order residual row sees
  exactly one perturbed neighbour), and a single-cell perturbation changes
  no residual row beyond the cell and its face neighbours (a wrong
  neighbour table or colouring breaks either).
- The factorisation has the ILU(0) property: M = (D + L) D^-1 (D + U)
  equals A1 on every diagonal and face-neighbour block (a dropped or
  transposed coupling term, or a wrong colour order, leaves an error of the
  size of the blocks).  M^-1 applied by the sweeps is the inverse of that M.
- The oracle JFNK with the coupled preconditioner reaches the same steady
  state as with block Jacobi (<= 1e-10 relative at stop norm 1e-12) in fewer
  Krylov iterations.
"""
import math

import numpy as np

from fv3d_amd import configs
from fv3d_amd import solver as S
from fv3d_amd._lib import lib
from oracle.jfnk import Dilu, OracleJfnk, colouring2


def _channel(nx, ny, nz, seed=20261203):
    cfg = configs.bump_channel_steady(nx=nx, ny=ny, nz=nz)
    L = lib()
    L.init_solver(b"oracle-dilu")
    S.check(L.fvh_mesh_channel(nx, ny, nz, 0.05, seed))
    mesh = S.mesh_arrays()
    oj = OracleJfnk(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    u0 = 0.5 * math.sqrt(1.4)
    U = np.zeros((mesh["n_cells"], 5))
    U[:, 0] = 1.0
    U[:, 1] = u0
    U[:, 4] = 1 / 0.4 + 0.5 * u0 * u0
    return oj, U


def _perturbed(U, seed):
    rng = np.random.default_rng(seed)
    V = U.copy()
    N = len(V)
    V[:, 0] *= 1.0 + 0.1 * rng.random(N)
    V[:, 1] *= 1.0 + 0.2 * rng.random(N)
    V[:, 2] = 0.1 * rng.standard_normal(N)
    V[:, 4] = 1 / 0.4 + 0.5 * (V[:, 1] ** 2 + V[:, 2] ** 2) / V[:, 0]
    return V


def test_colouring_and_offdiag_match_single_cell_differences():
    oj, U = _channel(8, 3, 3, seed=7)
    U = _perturbed(U, 1)
    nb, N = oj.nb, oj.N
    # distance-1 colouring: no face neighbours share a colour
    for c in range(N):
        for k in nb[c][nb[c] >= 0]:
            assert oj.colour[k] != oj.colour[c]
    col2 = colouring2(nb)
    for c in range(N):
        ring = set(nb[c][nb[c] >= 0])
        for k in list(ring):
            ring |= set(nb[k][nb[k] >= 0])
        ring.discard(c)
        assert all(col2[k] != col2[c] for k in ring)
    off = oj.offdiag(U)
    base = oj.residual(U, first=True)
    for c in range(0, N, 7):
        for k in range(5):
            h = 1e-7 * max(abs(U[c, k]), 1.0)
            Up = U.copy()
            Up[c, k] += h
            d = (oj.residual(Up, first=True) - base) / (Up[c, k] - U[c, k])
            near = np.concatenate([[c], nb[c][nb[c] >= 0]])
            far = np.setdiff1d(np.arange(N), near)
            assert np.abs(d[far]).max() == 0.0
            for i, t in zip(*np.where(nb == c)):
                assert np.array_equal(off[i, t, :, k], -d[i])


def test_factorisation_has_ilu0_property():
    oj, U = _channel(8, 3, 3, seed=7)
    U = _perturbed(U, 2)
    N, nb = oj.N, oj.nb
    dtau = oj.local_dt(U, 1e3)
    diag = oj.blocks(U, dtau)
    off = oj.offdiag(U)
    M = Dilu(nb, oj.colour, diag, off)
    n = 5 * N
    A = np.zeros((n, n))
    pat = np.zeros((N, N), bool)
    for c in range(N):
        A[5 * c:5 * c + 5, 5 * c:5 * c + 5] = diag[c]
        pat[c, c] = True
        for t in range(nb.shape[1]):
            k = nb[c, t]
            if k >= 0:
                A[5 * c:5 * c + 5, 5 * k:5 * k + 5] += off[c, t]
                pat[c, k] = True
    Minv = np.stack([M.apply(e) for e in np.eye(n)], axis=1)
    Md = np.linalg.inv(Minv)
    P = np.kron(pat, np.ones((5, 5), bool))
    scale = np.abs(A).max()
    assert np.abs((Md - A)[P]).max() <= 1e-12 * scale
    # the product has fill outside the pattern (it is not A1 itself)
    assert np.abs((Md - A)[~P]).max() > 1e-6 * scale
    # explicit (D + L) D^-1 (D + U) in the colour order
    order = np.argsort(oj.colour, kind="stable")
    pos = np.empty(N, int)
    pos[order] = np.arange(N)
    Dblk = np.zeros((n, n))
    Lw = np.zeros((n, n))
    Up = np.zeros((n, n))
    for c in range(N):
        Dblk[5 * c:5 * c + 5, 5 * c:5 * c + 5] = np.linalg.inv(M.Dinv[c])
        for t in range(nb.shape[1]):
            k = nb[c, t]
            if k >= 0:
                tgt = Lw if pos[k] < pos[c] else Up
                tgt[5 * c:5 * c + 5, 5 * k:5 * k + 5] = off[c, t]
    M2 = (Dblk + Lw) @ np.linalg.inv(Dblk) @ (Dblk + Up)
    assert np.abs(M2 - Md).max() <= 1e-10 * scale


def test_oracle_jfnk_dilu_same_steady_state():
    oj, U = _channel(16, 4, 3)
    Ub, stb, _ = oj.solve(U, tol=1e-12)
    Ud, std, _ = oj.solve(U, tol=1e-12, coupling="dilu")
    assert stb["status"] == 0 and std["status"] == 0
    assert std["gmres_iterations"] < stb["gmres_iterations"]
    assert np.linalg.norm(Ud - Ub) / np.linalg.norm(Ub) <= 1e-10
