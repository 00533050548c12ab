"""GPU parity of the batched mode (fvh_mesh_batch / fvh_batch_*): each case
of a batch of uneven, mixed-flux, Euler and Navier-Stokes periodic boxes,
advanced together (eagerly and by HIP-graph replay), equals the oracle run
of that case alone (its own mesh, physics and time steps) to 1e-12
relative per component; the per-case dt and residual norms agree too."""
import math

import numpy as np
import pytest

from fv3d_amd import batch as B
from fv3d_amd import solver as S
from fv3d_amd._lib import lib
from oracle import fvref
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

CASES = [dict(n_hex=4, mach=0.3, flux="Rusanov", viscous=False),
         dict(n_hex=3, mach=1.7, flux="AUSM", viscous=True),
         dict(n_hex=5, mach=0.8, flux="AUSM", viscous=False),
         dict(n_hex=4, mach=1.2, flux="Rusanov", viscous=True),
         dict(n_hex=3, mach=0.5, flux="Rusanov", viscous=False),
         # 81 tiles: the face schedule's workgroups take one full run of 64
         # tiles (8 adjacent tiles per XCD) and a partial one
         dict(n_hex=12, mach=0.6, flux="AUSM", viscous=False)]
TILE, MU, STEPS = 2, 2e-3, 6


def _oracle_runs():
    L = lib()
    L.init_solver(b"batch-oracle")
    out = []
    for c in CASES:
        S.check(L.fvh_mesh_box(c["n_hex"], c["n_hex"], c["n_hex"], 0.0, 0.0, 0.0, 1.0, 1.0,
                               1.0, 1, 7, TILE))
        mesh = S.mesh_arrays()
        phys = dict(B.physics(), mu_mix=MU if c["viscous"] else 0.0, flux=c["flux"])
        o = fvref.from_case(mesh, phys, {"flow": ("flow", (1.0, 0, 0, 0, 1.0))},
                            dict(reconstruction="First-Order", limiter="None"))
        o.set_state(B.case_state(mesh["cell_x"], c["mach"]).reshape(-1))
        n, log = o.timedisc("RK3-3", STEPS, True, 1e-10, 0.0, 1e300)
        assert n == STEPS
        out.append((o.state(), log))
    return out


SCHEDULES = {"group": dict(persistent=False),
             "face": dict(persistent=False, face=True),
             "persistent-128": dict(persistent=True, block=128, grid=0),
             "persistent-256-grid3": dict(persistent=True, block=256, grid=3),
             "persistent-512": dict(persistent=True, block=512, grid=0),
             "persistent-256-sub4": dict(persistent=True, block=256, grid=0, sub_tiles=4),
             "persistent-512-grid5-sub16": dict(persistent=True, block=512, grid=5,
                                                sub_tiles=16),
             "tiles-256": dict(persistent=True, block=256, grid=-1),
             "tiles-128-sub4": dict(persistent=True, block=128, grid=-1, sub_tiles=4)}


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("schedule", list(SCHEDULES))
def test_batch_matches_single_case_oracle(graph, schedule):
    """Every schedule (per-group launches, lane per face slot or
    face-centred over tiles of at most 128 cells of one case, the last tile
    of every case partial; persistent launches per class with tiles of one,
    4 or 16 runs of 32, 64 or 128 cells, some ending inside a run, and with
    3 or 5 workgroups looping over all tiles; one workgroup per tile) gives
    each case its own oracle run."""
    ref = _oracle_runs()
    b = B.Batch(CASES, tile=TILE, mu_ns=MU, schedule=SCHEDULES[schedule])
    b.run(STEPS, graph=graph)
    dt, t, res = b.scalars()
    worst = 0.0
    for k, i in enumerate(b.order):
        Uo, log = ref[i]
        e = rel_err(b.state(k), Uo)
        worst = max(worst, e)
        assert e <= 1e-12, (i, e)
        assert abs(t[k] - log[-1, 0]) <= 1e-13 * log[-1, 0]
        assert np.allclose(res[k], log[-1, 3:8], rtol=1e-10, atol=0)
    print("batch %s graph=%s worst rel=%g" % (schedule, graph, worst), flush=True)
    b.close()


@pytest.mark.parametrize("schedule", ["persistent-256-sub4", "face"])
def test_batch_case_step_times_inside_the_launch(schedule):
    """Per-case step times measured inside the shared launches (persistent
    or face schedule): one positive time per case, the larger cases taking
    longer; the steps they time are ordinary steps (the states still match
    the oracle)."""
    ref = _oracle_runs()
    b = B.Batch(CASES, tile=TILE, mu_ns=MU, schedule=SCHEDULES[schedule])
    b.run(STEPS - 2, graph=True)
    ms = b.case_step_times(2)
    assert ms.shape == (len(CASES),) and np.all(ms > 0) and np.all(np.isfinite(ms))
    sizes = np.array([c["n_hex"] ** 3 for c in b.cases])
    assert ms[np.argmax(sizes)] >= ms[np.argmin(sizes)]
    for k, i in enumerate(b.order):
        assert rel_err(b.state(k), ref[i][0]) <= 1e-12
    print("per-case step ms inside the batch", ms, flush=True)
    b.close()
