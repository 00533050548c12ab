"""CPU checks that pin the oracle's reference-semantics implicit step
(timestep_newton, src/Solver/src/timedisc/implicit.c:61-352) without the
reference binary:

- closed form: on a uniform periodic state R = 0 for every uniform
  perturbation, so the lumped Jacobian is exactly I/dt and the step leaves
  the state unchanged;
- consistency: implicit Euler and BDF2 on a smooth periodic density wave
  converge to the time-accurate solution (RK4-5 at a small step) as the
  step shrinks (a wrong sign, coefficient or history level in f breaks
  convergence);
- the BDF2 history: two single steps through the step API with a
  caller-kept history equal two steps of the driver.
"""
import math

import numpy as np
import pytest

from fv3d_amd import configs
from tests.test_oracle import _oracle_on, _prim_state


def _density_wave(o, mesh, amp=0.05):
    x = np.asarray(mesh["cell_x"]).reshape(-1, 3)
    rho = 1.0 + amp * np.sin(2 * math.pi * x.sum(1))
    return _prim_state(rho, 1.0, 1.0, 1.0, 1.0)


def test_uniform_state_closed_form():
    cfg = configs.density_wave_box(n=4, steps=1)
    o, mesh = _oracle_on(cfg)
    N = mesh["n_cells"]
    U = np.tile(_prim_state(1.2, 0.3, -0.1, 0.2, 0.9), (N, 1))
    o.set_state(U.reshape(-1))
    o.time_derivative()
    old = np.zeros((2, N * 5))
    dt = 1e-3
    k, jac = o.timestep_newton("Euler", 0, 0.0, dt, old, transient=False, want_jac=True)
    J = jac.reshape(N, 5, 5)
    # R of a uniform state cancels to rounding only (|R| ~ 1e-16 |F|/h)
    assert np.allclose(J, np.broadcast_to(np.eye(5) / dt, J.shape), rtol=0, atol=1e-9)
    assert k == 1
    assert np.allclose(o.state(), U, rtol=1e-14, atol=0)
    # transient mode waits for |f| to drop below its initial value, which
    # never happens from |f| = 0: the inner loop runs out (the reference
    # aborts there, implicit.c:340-351)
    k, _ = o.timestep_newton("Euler", 0, 0.0, dt, old, transient=True, max_iter_inner=3)
    assert k == -2


@pytest.mark.parametrize("scheme", ["Euler", "BDF2"])
def test_time_consistency(scheme):
    # Fixed step counts (no t_end clamp: the reference's BDF2 keeps constant
    # coefficients, so a shortened last step would break its consistency);
    # the time-accurate solution is RK4-5 at a small CFL to the same time.
    cfg = configs.density_wave_box(n=6, steps=1)
    o, mesh = _oracle_on(cfg)
    U0 = _density_wave(o, mesh)
    errs = []
    for cfl, steps in ((2.0, 8), (1.0, 16), (0.5, 32)):
        o.set_physics(cfl_scale=cfl, mu_mix=0.0, flux="Rusanov", visc_dt_mode=1)
        o.set_state(U0.reshape(-1))
        n, log = o.timedisc_implicit(scheme, steps, True, 1e-10, 0.0, 1e300)
        assert n == steps
        U = o.state()
        t_f = log[-1, 0]
        o.set_physics(cfl_scale=0.05, mu_mix=0.0, flux="Rusanov", visc_dt_mode=1)
        o.set_state(U0.reshape(-1))
        o.timedisc("RK4-5", 100000, True, 1e-10, 0.0, t_f)
        errs.append(np.abs(U - o.state()).max())
    # first order in time (one Newton step with the lumped Jacobian)
    assert errs[0] < 0.05
    assert errs[1] < 0.7 * errs[0] and errs[2] < 0.7 * errs[1]


def test_bdf2_history_single_steps_equal_driver():
    cfg = configs.density_wave_box(n=4, steps=1)
    o, mesh = _oracle_on(cfg)
    o.set_physics(cfl_scale=1.0, mu_mix=0.0, flux="Rusanov", visc_dt_mode=1)
    U0 = _density_wave(o, mesh)
    N = mesh["n_cells"]
    o.set_state(U0.reshape(-1))
    n, log = o.timedisc_implicit("BDF2", 3, True, 1e-10, 0.0, 1e300)
    U_driver = o.state()
    o.set_state(U0.reshape(-1))
    o.time_derivative()
    old = np.zeros((2, N * 5))
    t = 0.0
    for it in range(3):
        dt, _ = o.timestep()
        o.timestep_newton("BDF2", it, t, dt, old, transient=True)
        assert dt == log[it, 1]
        t += dt
    assert np.array_equal(o.state(), U_driver)
    # the history holds the last two step-start states
    assert not np.array_equal(old[0], old[1])
