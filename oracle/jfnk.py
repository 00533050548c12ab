"""CPU restatement of the steady Newton-Krylov (JFNK) solve on the oracle's
residual.

TEST INFRASTRUCTURE ONLY: used by tests/ and the CPU-baseline leg of
bench.py as the checker, never by the product.

The reference has no JFNK; its residual is `finite_volume_time_derivative`
(src/Solver/src/finite_volume/finite_volume.c:81-124), which the oracle
restates.  The algorithm here is the one the product's fvh_jfnk_solve runs
(same options, same pseudo-time step, same operator and the same GMRES),
written plainly in FP64 numpy:
  - dtau_c = cfl * V_c / sum_d (|u_d| + c) dx_d  (the convective limit of
    src/FV3D/src/navier_stokes/navier_stokes.c:171-180, per cell);
  - block-Jacobi blocks M_c = I/dtau_c - dR1_c/dU_c of the first-order
    inviscid residual R1, by one-sided differences with h = 1e-7 max(|U|, 1)
    over a distance-1 colouring of the cells (all cells of one colour
    perturbed at once: a first-order residual sees no two of them), a
    different route to the same derivatives as the device's per-cell
    differences;
  - optionally (coupling "dilu") the multicolour block ILU(0) of
    A1 = I/dtau - dR1/dU: cells ordered by the greedy distance-1 colouring
    (ascending cell index within a colour), off-diagonal face-neighbour
    blocks by one-sided differences over a distance-2 colouring (every
    neighbour of a perturbed cell sees exactly one perturbed cell), the
    block-Jacobi blocks on the diagonal, D_i = A1_ii - sum_{k lower}
    A1_ik D_k^-1 A1_ki, M = (D + L) D^-1 (D + U);
  - right-preconditioned GMRES(m) with restarts on
    A z = z/dtau - [R(U + eps z) - R(U)]/eps, eps = fd_eps (1+|U|)/|z|,
    classical Gram-Schmidt with selective reorthogonalisation, Givens
    rotations, relative tolerance rtol;
  - update U += alpha M^-1 x with alpha = min(1, max_change / max relative
    change of rho and rho*E);
  - CFL times cfl_cut (not below 1) when the new stop norm exceeds 3x the
    old one, else times cfl_stall_cut when GMRES ended above cfl_stall
    relative residual, else times cfl_growth (capped at cfl_max).
Stop norm: max_j (1/V_glob) sum_i |R_ij| V_i (src/Solver/src/analyze/
analyze.c:37-61 without the dt factor).
"""
import math
import time

import numpy as np

from . import fvref

DEFAULTS = dict(max_newton=300, tol=1e-8, cfl0=10.0, cfl_growth=1.5, cfl_cut=0.3,
                cfl_stall=0.5, cfl_stall_cut=0.7, cfl_max=1e6, krylov_dims=30, max_restarts=3, rtol=1e-2, fd_eps=1e-7,
                max_change=0.2, precond="block", coupling="none", reorth=0.1)


def colouring(mesh):
    """Greedy distance-1 colouring of the cell graph (cells sharing a face)."""
    N, F = mesh["n_cells"], mesh["n_faces"]
    fc = np.asarray(mesh["face_cells_file"]).reshape(F, 2)
    adj = [[] for _ in range(N)]
    for a, b in fc:
        if 0 <= b < N:
            adj[a].append(b)
            adj[b].append(a)
    colour = -np.ones(N, dtype=np.int64)
    for c in range(N):
        used = {colour[x] for x in adj[c]}
        k = 0
        while k in used:
            k += 1
        colour[c] = k
    return colour


def gmres(op, b, m, max_restarts, rtol, reorth=0.5):
    """Restarted GMRES(m) for op(x) = b from x = 0: classical Gram-Schmidt in
    one pass with the new vector's norm by Pythagoras (|w|^2 - |h|^2), and a
    second pass when that kept at most the fraction reorth of |w|^2 (0.5:
    the 1/sqrt(2) test of Daniel, Gragg, Kaufman and Stewart); Givens
    rotations; stops when the estimated residual is at most rtol |b|.
    Returns (x, Arnoldi steps, estimated relative residual)."""
    x = np.zeros(b.size)
    beta0 = None
    its = 0
    rel = 1.0
    for cycle in range(max_restarts + 1):
        r = b.copy() if cycle == 0 else b - op(x)
        beta = np.linalg.norm(r)
        if beta0 is None:
            beta0 = beta
        if not beta > 0:
            rel = beta / beta0 if beta0 > 0 else 0.0
            break
        Vk = np.zeros((m + 1, b.size))
        H = np.zeros((m + 1, m))
        cs, sn = np.zeros(m), np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        Vk[0] = r / beta
        k = 0
        done = False
        for j in range(m):
            w = op(Vk[j])
            h1 = Vk[: j + 1] @ w
            ww = w @ w
            w = w - h1 @ Vk[: j + 1]
            hn2 = ww - h1 @ h1
            h2 = np.zeros_like(h1)
            if not hn2 > reorth * ww:
                h2 = Vk[: j + 1] @ w
                w = w - h2 @ Vk[: j + 1]
                hn2 = w @ w
            its += 1
            H[: j + 1, j] = h1 + h2
            hn = math.sqrt(max(hn2, 0.0))
            H[j + 1, j] = hn
            for i in range(j):
                a_, b_ = H[i, j], H[i + 1, j]
                H[i, j] = cs[i] * a_ + sn[i] * b_
                H[i + 1, j] = -sn[i] * a_ + cs[i] * b_
            rr = np.hypot(H[j, j], H[j + 1, j])
            cs[j] = H[j, j] / rr if rr > 0 else 1.0
            sn[j] = H[j + 1, j] / rr if rr > 0 else 0.0
            H[j, j] = rr
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            k = j + 1
            rel = abs(g[j + 1]) / beta0
            if rel <= rtol or not hn > 1e-300 * beta0:
                done = True
                break
            Vk[j + 1] = w / hn
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
        x = x + y @ Vk[:k]
        if done:
            break
    return x, its, rel


def neighbours(mesh):
    """nb[c, t]: the interior face neighbour of cell c across its face slot t
    (cell_faces order), -1 for boundary faces and unused slots."""
    N, maxf = mesh["n_cells"], mesh["max_cell_faces"]
    cf = np.asarray(mesh["cell_faces"]).reshape(N, maxf)
    fc = np.asarray(mesh["face_cells"]).reshape(-1, 2)
    nb = -np.ones((N, maxf), dtype=np.int64)
    for c in range(N):
        for t in range(maxf):
            f = cf[c, t]
            if f < 0:
                continue
            o = fc[f, 1] if fc[f, 0] == c else fc[f, 0]
            if 0 <= o < N:
                nb[c, t] = o
    return nb


def colouring2(nb):
    """Greedy distance-2 colouring (no two cells within two face steps share
    a colour)."""
    N = nb.shape[0]
    colour = -np.ones(N, dtype=np.int64)
    for c in range(N):
        used = set()
        for x in nb[c]:
            if x < 0:
                continue
            used.add(colour[x])
            for y in nb[x]:
                if y >= 0:
                    used.add(colour[y])
        k = 0
        while k in used:
            k += 1
        colour[c] = k
    return colour


class Dilu:
    """M = (D + L) D^-1 (D + U) in the colour order; diag [N,5,5] the A1
    diagonal blocks, off [N,maxf,5,5] the blocks A1_{c, nb[c,t]}."""

    def __init__(self, nb, colour, diag, off):
        self.nb, self.colour, self.off = nb, colour, off
        N, maxf = nb.shape
        valid = nb >= 0
        cn = np.where(valid, colour[np.maximum(nb, 0)], -1)
        self.low = valid & (cn < colour[:, None])
        self.high = valid & (cn > colour[:, None])
        self.groups = [np.where(colour == q)[0] for q in range(colour.max() + 1)]
        self.Dinv = np.zeros((N, 5, 5))
        for g in self.groups:
            for i in g:
                D = diag[i].copy()
                for t in np.where(self.low[i])[0]:
                    k = nb[i, t]
                    tk = int(np.where(nb[k] == i)[0][0])
                    D -= off[i, t] @ self.Dinv[k] @ off[k, tk]
                self.Dinv[i] = np.linalg.inv(D)

    def apply(self, v):
        N = self.nb.shape[0]
        v = np.asarray(v).reshape(N, 5)
        z = np.zeros((N, 5))
        for g in self.groups:
            r = v[g].copy()
            for t in range(self.nb.shape[1]):
                m = self.low[g, t]
                if m.any():
                    i = g[m]
                    r[m] -= np.einsum("nij,nj->ni", self.off[i, t], z[self.nb[i, t]])
            z[g] = np.einsum("nij,nj->ni", self.Dinv[g], r)
        for g in self.groups[::-1]:
            r = np.zeros((len(g), 5))
            for t in range(self.nb.shape[1]):
                m = self.high[g, t]
                if m.any():
                    i = g[m]
                    r[m] += np.einsum("nij,nj->ni", self.off[i, t], z[self.nb[i, t]])
            z[g] -= np.einsum("nij,nj->ni", self.Dinv[g], r)
        return z.reshape(-1)


class OracleJfnk:
    def __init__(self, mesh, physics, regions, numerics):
        self.mesh = mesh
        self.N = mesh["n_cells"]
        self.V = np.asarray(mesh["cell_volume"], dtype=np.float64)
        self.Vg = self.V.sum()
        self.dx = np.asarray(mesh["cell_dx"], dtype=np.float64).reshape(self.N, 3)
        self.kappa = physics.get("kappa", 1.4)
        self.full = fvref.from_case(mesh, physics, regions, numerics)
        # the preconditioners' residual R1: first order and inviscid, whatever
        # the case's viscosity
        self.first = fvref.from_case(mesh, dict(physics, mu_mix=0.0), regions,
                                     dict(numerics, reconstruction="First-Order",
                                          limiter="None"))
        self.colour = colouring(mesh)
        self.nb = neighbours(mesh)
        self._colour2 = None
        self.evals = 0

    def residual(self, U, first=False):
        o = self.first if first else self.full
        o.set_state(U.reshape(-1))
        if not first:
            self.evals += 1
        return o.time_derivative()

    def stop_norm(self, R):
        return (np.abs(R) * self.V[:, None]).sum(0).max() / self.Vg

    def local_dt(self, U, cfl):
        rho = U[:, 0]
        u = U[:, 1:4] / rho[:, None]
        p = (self.kappa - 1.0) * (U[:, 4] - 0.5 * (U[:, 1:4] * u).sum(1))
        p = np.maximum(p, 1e-10)
        c = np.sqrt(self.kappa * p / rho)
        return cfl * (self.V / ((np.abs(u) + c[:, None]) * self.dx).sum(1))

    def blocks(self, U, dtau):
        """M_c = I/dtau_c - dR1_c/dU_c, [N,5,5] (row j, column k)."""
        base = self.residual(U, first=True)
        J = np.zeros((self.N, 5, 5))
        for col in range(self.colour.max() + 1):
            mask = self.colour == col
            for k in range(5):
                h = 1e-7 * np.maximum(np.abs(U[mask, k]), 1.0)
                Up = U.copy()
                Up[mask, k] += h
                hk = Up[mask, k] - U[mask, k]
                d = self.residual(Up, first=True) - base
                J[mask, :, k] = d[mask] / hk[:, None]
        return np.eye(5)[None] / dtau[:, None, None] - J

    def offdiag(self, U):
        """A1_{c, nb[c,t]} = -dR1_c/dU_nb, [N, maxf, 5, 5], by one-sided
        differences (h = 1e-7 max(|U|, 1)) over a distance-2 colouring."""
        if self._colour2 is None:
            self._colour2 = colouring2(self.nb)
        col2, nb = self._colour2, self.nb
        N, maxf = nb.shape
        base = self.residual(U, first=True)
        off = np.zeros((N, maxf, 5, 5))
        valid = nb >= 0
        ncol = np.where(valid, col2[np.maximum(nb, 0)], -1)
        for q in range(col2.max() + 1):
            mask = col2 == q
            sel = valid & (ncol == q)
            rows, slots = np.where(sel)
            src = nb[rows, slots]
            for k in range(5):
                h = 1e-7 * np.maximum(np.abs(U[mask, k]), 1.0)
                Up = U.copy()
                Up[mask, k] += h
                hk = np.zeros(N)
                hk[mask] = Up[mask, k] - U[mask, k]
                d = self.residual(Up, first=True) - base
                off[rows, slots, :, k] = -d[rows] / hk[src][:, None]
        return off

    def dilu(self, U, dtau):
        return Dilu(self.nb, self.colour, self.blocks(U, dtau), self.offdiag(U))

    def solve(self, U, verbose=False, **opts):
        o = dict(DEFAULTS, **opts)
        U = np.array(U, dtype=np.float64).reshape(self.N, 5)
        t0 = time.perf_counter()
        self.evals = 0
        R0 = self.residual(U)
        norm = self.stop_norm(R0)
        l20 = np.linalg.norm(R0)
        cfl = o["cfl0"]
        hist = []
        gm_total = 0
        it = 0
        while it < o["max_newton"] and norm > o["tol"]:
            if not np.isfinite(norm):
                raise FloatingPointError("non-finite residual")
            dtau = self.local_dt(U, cfl)
            if o["coupling"] == "dilu":
                M = self.dilu(U, dtau).apply
            else:
                if o["precond"] == "block":
                    Minv = np.linalg.inv(self.blocks(U, dtau))
                else:
                    Minv = np.eye(5)[None] * dtau[:, None, None]
                M = lambda v, Minv=Minv: np.einsum("nij,nj->ni", Minv,
                                                   v.reshape(self.N, 5)).reshape(-1)
            fd_num = o["fd_eps"] * (1.0 + np.linalg.norm(U))
            b = R0.reshape(-1)

            def A(z):
                eps = fd_num / max(np.linalg.norm(z), 1e-300)
                Rt = self.residual(U + eps * z.reshape(self.N, 5))
                return (z.reshape(self.N, 5) / dtau[:, None] - (Rt - R0) / eps).reshape(-1)

            x, its, rel = gmres(lambda v: A(M(v)), b, o["krylov_dims"],
                                o["max_restarts"], o["rtol"], o["reorth"])
            gm_total += its
            dU = M(x).reshape(self.N, 5)
            ratio = np.max(np.abs(dU[:, [0, 4]] / U[:, [0, 4]]))
            alpha = o["max_change"] / ratio if ratio > o["max_change"] else 1.0
            U = U + alpha * dU
            prev = norm
            R0 = self.residual(U)
            norm = self.stop_norm(R0)
            it += 1
            hist.append((prev, cfl, its, rel, alpha, time.perf_counter() - t0))
            if verbose:
                print("%4d  %.6e  cfl %.3e  gmres %3d  rel %.2e  alpha %.3f"
                      % (it, norm, cfl, its, rel, alpha), flush=True)
            if not norm < 3.0 * prev:
                cfl = max(cfl * o["cfl_cut"], 1.0)
            elif rel > o["cfl_stall"]:
                cfl = max(cfl * o["cfl_stall_cut"], 1.0)
            else:
                cfl = min(cfl * o["cfl_growth"], o["cfl_max"])
        stats = dict(status=0 if norm <= o["tol"] else 1, newton_steps=it,
                     gmres_iterations=gm_total, residual_evals=self.evals,
                     seconds=time.perf_counter() - t0, norm=norm,
                     l2_drop=np.linalg.norm(R0) / l20 if l20 > 0 else 0.0)
        return U, stats, np.array(hist)
