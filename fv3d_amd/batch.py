"""Batched cases (fvh_mesh_batch / fvh_batch_* of include/fv3d_hip.h):
many independent periodic Euler / Navier-Stokes cases of uneven size,
advanced together in one device state: per stage one persistent launch for
every case (a work queue of tiles, each with its case's flux and
viscosity; or, selectable, one launch per (flux, viscous) group), replayed
from a HIP graph.  `autotune` picks the workgroup size (tile size), the
persistent grid and the cell-ordering tile of the case meshes by timing,
and keeps the table in fv3d_amd/data/batch_tuning.json.

`c5_cases` draws the benchmark set: case sizes log-uniform in tetrahedra,
Mach numbers uniform, the flux alternating with the seed's parity and every
fourth case viscous.
"""
import ctypes
import json
import math
import os

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

    def __init__(self, cases, tile=8, mu_ns=1e-3, length=1.0, cfl=0.5, schedule=None):
        # viscous cases together (one gradient launch), then by flux
        key = lambda i: (int(cases[i]["viscous"]), FLUX_ID[cases[i]["flux"]])
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
        if schedule is not None:
            self.set_schedule(**schedule)

    def set_schedule(self, persistent=True, block=256, grid=0, sub_tiles=1, face=False):
        """persistent: the work-queue launches per class; face: the
        face-centred launches per group over case-aligned tiles; neither:
        the lane-per-slot launches per group."""
        mode = 2 if face else (1 if persistent else 0)
        S.check(self.L.fvh_batch_set_schedule(mode, block, grid, sub_tiles))

    def case_step_times(self, steps=5):
        """Per-case step time (ms) inside the batched launches."""
        ms = np.zeros(len(self.cases))
        S.check(self.L.fvh_batch_case_step_times(steps, ms.ctypes.data_as(P_DOUBLE)))
        return ms

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


TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                           "batch_tuning.json")


def signature(cases):
    """What the tuning depends on: the case sizes, fluxes and viscous flags."""
    return [[c["n_hex"], c["flux"], bool(c["viscous"])] for c in cases]


def _cus(L):
    cus = ctypes.c_int(0)
    name = ctypes.create_string_buffer(128)
    gbs = ctypes.c_double(0)
    S.check(L.fvh_device_info(name, 128, ctypes.byref(cus), ctypes.byref(gbs)))
    return max(1, cus.value)


def autotune(cases, tiles=(4, 8), blocks=(256, 512), grids=(2, 4, 8, -1), subs=(1, 4, 16),
             steps=10, warmup=2, progress=None, **batch_kwargs):
    """Times one graph-replayed batch step for every cell-ordering tile of
    the case meshes (hexahedra per side of a locality tile), persistent
    workgroup size (256 or 512 threads: runs of 64 or 128 tetrahedra),
    persistent grid (2, 4, 8 workgroups per CU, or -1: one workgroup per tile
    without the work queue) and runs per tile (1, 4, 16), and the two
    per-group schedules (lane per face slot; face-centred over case-aligned
    tiles); returns (best, table).
    `progress`, when given, is called with each table row as it is
    measured.  The explicit path has no cell colouring, so there is no
    colour-block size to tune."""
    table = []
    for tile in tiles:
        b = Batch(cases, tile=tile, **batch_kwargs)
        cus = _cus(b.L)
        b.run(warmup, graph=False)
        for block in blocks:
            for sub in subs:
                for g in grids:
                    grid = g * cus if g > 0 else -1
                    b.set_schedule(True, block, grid, sub)
                    b.run(1, graph=True)
                    sec = b.run(steps, graph=True)
                    table.append(dict(tile=tile, persistent=True, block=block, grid=grid,
                                      sub_tiles=sub, ms_per_step=sec / steps * 1e3))
                    if progress:
                        progress(table[-1])
        for face in (False, True):
            b.set_schedule(False, face=face)
            b.run(1, graph=True)
            sec = b.run(steps, graph=True)
            table.append(dict(tile=tile, persistent=False, face=face, block=0, grid=0,
                              sub_tiles=0, ms_per_step=sec / steps * 1e3))
            if progress:
                progress(table[-1])
        b.close()
    best = min(table, key=lambda r: r["ms_per_step"])
    return best, table


def load_tuning(cases):
    """The cached autotuner result for these cases, (best, table), or None."""
    try:
        with open(TUNING_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return (t["best"], t["table"]) if t.get("signature") == signature(cases) else None


def save_tuning(cases, best, table, device=""):
    os.makedirs(os.path.dirname(TUNING_FILE), exist_ok=True)
    with open(TUNING_FILE, "w") as f:
        json.dump({"signature": signature(cases), "device": device, "best": best,
                   "table": table}, f, indent=1)
        f.write("\n")
