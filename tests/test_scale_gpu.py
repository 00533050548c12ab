"""The headline workload at its real size, on the GPU through the C-ABI:

- Taylor-Green vortex on the 96^3 x 6 mesh (5,308,416 tetrahedra, AUSM,
  Green-Gauss, Barth-Jespersen, viscous, RK3-3): on the periodic mesh every
  face flux leaves one cell and enters the other with the same bits, so the
  totals sum_c V_c U_c are conserved to round-off; two runs from the same
  state are bitwise equal (no atomics, fixed-order sums);
- the same workload on 16^3 x 6 cells for the full 200 steps against the
  oracle at 1e-12 relative per component.
"""
import math

import numpy as np
import pytest

from fv3d_amd import configs
from fv3d_amd import solver as S
from tests.helpers import make_pair, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu


def test_taylor_green_96_conservation_and_determinism():
    cfg = configs.taylor_green(n=96, steps=5, tile=4)
    case = S.Case(**cfg)
    V = S.mesh_arrays()["cell_volume"]
    U0 = case.state()
    rc, log = case.run()
    assert rc == 0 and len(log) == 5
    U1 = case.state()
    assert np.isfinite(U1).all()
    # exactly rounded totals (a running sum of 5.3M terms would itself err
    # by ~1e-13)
    tot0 = np.array([math.fsum(U0[:, j] * V) for j in range(5)])
    tot1 = np.array([math.fsum(U1[:, j] * V) for j in range(5)])
    # scales: the mass, the momentum magnitude, the energy
    mom = math.fsum((np.abs(U0[:, 1]) + np.abs(U0[:, 2]) + np.abs(U0[:, 3])) * V)
    scale = np.array([math.fsum(np.abs(U0[:, 0]) * V), mom, mom, mom,
                      math.fsum(np.abs(U0[:, 4]) * V)])
    drift = np.abs(tot1 - tot0) / scale
    print("96^3 x 6: drift of sum V U per variable", drift, "max |dU|",
          np.abs(U1 - U0).max(), flush=True)
    assert np.all(drift <= 1e-14)
    assert np.abs(U1 - U0).max() > 0.0
    case.set_state(U0)
    rc2, log2 = case.run()
    assert rc2 == 0
    assert n_bit_diff(case.state(), U1) == 0
    assert np.array_equal(log2, log)
    case.close()


def test_taylor_green_16_full_200_steps_vs_oracle():
    cfg = configs.taylor_green(n=16, steps=200, tile=4)
    case, o, _ = make_pair(cfg)
    rc, log_g = case.run()
    n, log_o = o.timedisc("RK3-3", 200, True, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 200 and len(log_g) == 200
    e = rel_err(case.state(), o.state())
    print("16^3 x 6, 200 steps: rel", e, "bitdiff", n_bit_diff(case.state(), o.state()),
          flush=True)
    assert e <= 1e-12
    assert np.array_equal(log_g[:, 1], log_o[:, 1])
    case.close()
