"""Batched cases (fvh_mesh_batch / fvh_batch_* of include/fv3d_hip.h):
many independent periodic Euler / Navier-Stokes cases of uneven size,
advanced together in one device state, one launch per kernel and stage
(one per (flux, viscous) group), optionally replayed from a HIP graph.

`c5_cases` draws the benchmark set: case sizes log-uniform in tetrahedra,
Mach numbers uniform, the flux alternating with the seed's parity and every
fourth case viscous.
"""
import ctypes
import math

import numpy as np

from ._lib import P_DOUBLE, P_INT
from . import solver as S

KAPPA = 1.4
FLUX_ID = {"AUSM": 0, "Rusanov": 1}


def c5_cases(n_cases=64, n_min=2e4, n_max=4e5, mach=(0.2, 2.0), seed0=0):
    """Case k (seed seed0 + k): N_k = exp(U[ln n_min, ln n_max]) tetrahedra ->
    n_hex = round((N_k / 6)^(1/3)) per side; Mach ~ U[mach]; Rusanov for even
    seeds, AUSM for odd; viscous when seed % 4 == 0."""
    out = []
    for k in range(n_cases):
        seed = seed0 + k
        rng = np.random.default_rng(seed)
        n_tets = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        out.append(dict(seed=seed, n_hex=max(3, int(round((n_tets / 6.0) ** (1.0 / 3.0)))),
                        mach=float(rng.uniform(*mach)),
                        flux="Rusanov" if seed % 2 == 0 else "AUSM",
                        viscous=seed % 4 == 0))
    return out


def case_state(x, mach, length=1.0):
    """Initial state of a case: rho = 1 + 0.2 sin(2 pi (x+y+z)/L),
    velocity Mach * c0 along (1,1,1)/sqrt(3) with c0 = sqrt(kappa), p = 1."""
    x = np.asarray(x).reshape(-1, 3)
    rho = 1.0 + 0.2 * np.sin(2 * math.pi * x.sum(1) / length)
    vel = mach * math.sqrt(KAPPA) / math.sqrt(3.0)
    U = np.zeros((len(x), 5))
    U[:, 0] = rho
    U[:, 1:4] = rho[:, None] * vel
    U[:, 4] = 1.0 / (KAPPA - 1.0) + 0.5 * rho * 3 * vel * vel
    return U


def physics(cfl=0.5):
    return dict(cfl_scale=cfl, dfl_scale=1.0, mu_mix=0.0, R_mix=1.0, Pr=0.718, kappa=KAPPA,
                flux="Rusanov", visc_dt_mode=1)


class Batch:
    """The cases, reordered so that equal (flux, viscous) groups are
    adjacent; `order[i]` is the input index of batch case i."""

    def __init__(self, cases, tile=8, mu_ns=1e-3, length=1.0, cfl=0.5):
        key = lambda i: (FLUX_ID[cases[i]["flux"]], int(cases[i]["viscous"]))
        self.order = sorted(range(len(cases)), key=key)
        self.cases = [cases[i] for i in self.order]
        self.length = length
        cfg = dict(mesh=("batch", dict(n_hex=[c["n_hex"] for c in self.cases], length=length,
                                       tile=tile)),
                   physics=physics(cfl), regions={"flow": ("flow", (1.0, 0.0, 0.0, 0.0, 1.0))},
                   numerics=dict(reconstruction="First-Order", limiter="None", scheme="RK3-3"))
        self.case = S.Case(**cfg)
        self.L = self.case.L
        x = S.mesh_arrays()["cell_x"].reshape(-1, 3)
        K = len(self.cases)
        offs = np.zeros(K + 1, dtype=np.int32)
        n_tets = [6 * c["n_hex"] ** 3 for c in self.cases]
        offs[1:] = np.cumsum(n_tets)
        self.offsets = offs
        U = np.zeros((offs[-1], 5))
        for k, c in enumerate(self.cases):
            U[offs[k]:offs[k + 1]] = case_state(x[offs[k]:offs[k + 1]], c["mach"], length)
        self.case.set_state(U)
        flux = (ctypes.c_int * K)(*[FLUX_ID[c["flux"]] for c in self.cases])
        vis = (ctypes.c_int * K)(*[int(c["viscous"]) for c in self.cases])
        mu = (ctypes.c_double * K)(*[mu_ns if c["viscous"] else 0.0 for c in self.cases])
        S.check(self.L.fvh_batch_setup(K, flux, vis, mu))
        got = np.zeros(K + 1, dtype=np.int32)
        n = ctypes.c_int(0)
        S.check(self.L.fvh_batch_info(ctypes.byref(n), got.ctypes.data_as(P_INT), K + 1))
        assert n.value == K and np.array_equal(got, offs)
        self.n_cells = int(offs[-1])

    def run(self, steps, graph=True):
        sec = ctypes.c_double(0)
        S.check(self.L.fvh_batch_run(steps, 1 if graph else 0, ctypes.byref(sec)))
        return sec.value

    def state(self, k):
        U = self.case.state()
        return U[self.offsets[k]:self.offsets[k + 1]]

    def scalars(self):
        K = len(self.cases)
        dt, t, res = np.zeros(K), np.zeros(K), np.zeros((K, 5))
        S.check(self.L.fvh_batch_case_scalars(dt.ctypes.data_as(P_DOUBLE),
                                              t.ctypes.data_as(P_DOUBLE),
                                              res.ctypes.data_as(P_DOUBLE)))
        return dt, t, res

    def close(self):
        self.case.close()


def autotune(cases, tiles=(0, 4, 8, 16), steps=10, warmup=2, graph_options=(False, True),
             **batch_kwargs):
    """Times the batch step for every cell-ordering tile (hexahedra per side
    of a locality tile; 0 = lexicographic) and with and without graph replay;
    returns (best dict, table of dicts with ms_per_step)."""
    table = []
    for tile in tiles:
        b = Batch(cases, tile=tile, **batch_kwargs)
        b.run(warmup, graph=False)
        for graph in graph_options:
            b.run(1, graph=graph)
            sec = b.run(steps, graph=graph)
            table.append(dict(tile=tile, graph=graph, ms_per_step=sec / steps * 1e3,
                              cells=b.n_cells))
        b.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    return best, table
