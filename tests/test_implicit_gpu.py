"""GPU parity of the reference-semantics implicit mode (timedisc with
init_implicit: BDF1/BDF2, lumped finite-difference block Jacobian, the
linear solve by restarted GMRES or BiCGStab with the TimeDisc/Implicit
tolerance, iteration and Krylov-dimension parameters; src/Solver/src/
timedisc/implicit.c) against the oracle, through the C-ABI.

Bar: 1e-12 relative per component (as the explicit path) for the state
after N steps on the periodic density wave, dt and time to 1e-13, the
residual columns to 1e-11; on the walled channel a bar set by the
algorithm's own conditioning (see that test).  The finite-difference scale eps = 1e-4 |Y_k|_2 is a sum
over the cells, which the device adds in a different order: eps, and with
it the Jacobian, differ in the last bits.
"""
import numpy as np
import pytest

from fv3d_amd import configs
from tests.helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _print(tag, **kv):
    print(tag, " ".join("%s=%s" % (k, v) for k, v in kv.items()), flush=True)


def _implicit(cfg, scheme, steps, transient=True):
    cfg["numerics"] = dict(cfg["numerics"], timestep="Implicit", scheme=scheme,
                           max_iter=steps, transient=transient)
    return cfg


@pytest.mark.parametrize("scheme", ["Euler", "BDF2"])
def test_density_wave_implicit(scheme):
    cfg = _implicit(configs.density_wave_box(n=8, steps=6, flux="Rusanov"), scheme, 6)
    cfg["physics"]["cfl_scale"] = 2.0
    case, o, _ = make_pair(cfg)
    rc, log_g = case.run()
    n, log_o = o.timedisc_implicit(scheme, 6, True, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 6 and len(log_g) == 6
    e = rel_err(case.state(), o.state())
    _print("implicit " + scheme, rel=e)
    assert e <= TOL
    assert np.allclose(log_g[:, :2], log_o[:, :2], rtol=1e-13, atol=0)
    assert np.allclose(log_g[:, 3:8], log_o[:, 3:8], rtol=1e-11, atol=0)
    case.close()


def test_channel_walls_implicit_steady():
    """Slip walls, inflow/outflow, linear reconstruction.  The reference's
    lumped Jacobian differences R(Y + eps/2) - R(Y - eps/2) of whole-field
    shifts, whose terms nearly cancel, so the step amplifies rounding: the
    oracle itself moves by `sens` under a 1-ulp change of the initial state
    (about 6e-11 here; 2.5e-5 with the Barth-Jespersen limiter, whose
    switches the shifts cross).  The bar is 20x the oracle's own
    sensitivity, measured in the test, and at least 1e-12."""
    cfg = configs.bump_channel(nx=12, ny=4, nz=2, flux="AUSM", limiter="None")
    cfg = _implicit(cfg, "Euler", 3, transient=False)
    case, o, _ = make_pair(cfg)
    U0 = case.state()
    rc, log_g = case.run()
    n, log_o = o.timedisc_implicit("Euler", 3, False, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 3
    Uo = o.state()
    rng = np.random.default_rng(0)
    o.set_lsoe()  # the same initial guess of the first linear solve
    o.set_state((U0 * (1 + 2.0 ** -52 * rng.choice([-1, 0, 1], U0.shape))).reshape(-1))
    o.timedisc_implicit("Euler", 3, False, 1e-10, 0.0, 1e300)
    sens = rel_err(o.state(), Uo)
    e = rel_err(case.state(), Uo)
    _print("implicit channel", rel=e, oracle_ulp_sensitivity=sens)
    assert e <= max(TOL, 20 * sens)
    assert np.allclose(log_g[:, 3:8], log_o[:, 3:8], rtol=max(1e-11, 20 * sens), atol=0)
    case.close()


@pytest.mark.parametrize("solver,tol,dims,restarts", [
    ("GMRes", 1e-4, 5, 1), ("GMRes", 1e-12, 15, 2), ("BiCGStab", 1e-4, 15, 2)])
def test_density_wave_implicit_krylov_parameters(solver, tol, dims, restarts):
    """The linear solve honours tolerance_lsoe, max_krylov_dims and
    max_krylov_restarts: with a loose tolerance the step differs from the
    exact block solve by up to about the tolerance, and the device matches
    the oracle run with the same parameters.  The dots of the Krylov method
    are sums over all cells, which the device adds in a different order, so
    the bar is a few orders of magnitude below the solve's own tolerance."""
    cfg = _implicit(configs.density_wave_box(n=6, steps=4, flux="Rusanov"), "BDF2", 4)
    cfg["physics"]["cfl_scale"] = 2.0
    cfg["numerics"].update(solver=solver, tolerance_lsoe=tol, max_krylov_dims=dims,
                           max_krylov_restarts=restarts, max_iter_lsoe=100)
    case, o, _ = make_pair(cfg)
    U0 = case.state()
    o.set_lsoe(solver, tol=tol, max_iter=100, dims=dims, restarts=restarts)
    rc, log_g = case.run()
    n, log_o = o.timedisc_implicit("BDF2", 4, True, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 4
    e = rel_err(case.state(), o.state())
    o.set_lsoe("exact")
    o.set_state(U0.reshape(-1))
    o.timedisc_implicit("BDF2", 4, True, 1e-10, 0.0, 1e300)
    e_exact = rel_err(case.state(), o.state())
    _print("implicit krylov " + solver, tol=tol, rel_vs_oracle=e, rel_vs_exact_solve=e_exact)
    assert e <= max(1e-12, 1e-3 * tol)
    if tol > 1e-8:
        assert e_exact > 10 * e
    case.close()
