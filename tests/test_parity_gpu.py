"""GPU parity of the residual and explicit time-stepping path against the CPU
oracle, through the C-ABI (libfv3d_hip.so via ctypes).

Bar: 1e-12 relative per component, |a-b| <= 1e-12 *
max(|b|, scale_j) with scale_j = max_i |b_ij|, for dU/dt and for the state
after N steps.  With -ffp-contract=off on both sides most values are
bitwise equal; the count of differing bits is printed.
"""
import numpy as np
import pytest

from fv3d_amd import configs
from tests.helpers import make_pair, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _print(tag, **kv):
    print(tag, " ".join("%s=%s" % (k, v) for k, v in kv.items()), flush=True)


def test_density_wave_residual_bitwise():
    cfg = configs.density_wave_box(n=16, steps=100)
    case, o, _ = make_pair(cfg)
    Rg = case.time_derivative()
    Ro = o.time_derivative()
    e = rel_err(Rg, Ro)
    _print("density-wave dU/dt", rel=e, bitdiff=n_bit_diff(Rg, Ro), n=Rg.size)
    assert e <= TOL


def test_density_wave_100_steps():
    cfg = configs.density_wave_box(n=16, steps=100)
    case, o, _ = make_pair(cfg)
    rc, log_g = case.run()
    n, log_o = o.timedisc("RK3-3", 100, True, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 100 and len(log_g) == 100
    Ug = case.state()
    Uo = o.state()
    e = rel_err(Ug, Uo)
    _print("density-wave U(100)", rel=e, bitdiff=n_bit_diff(Ug, Uo))
    assert e <= TOL
    # dt and time are bitwise (min reductions are exact); residual norms are
    # sums in a different order
    assert np.array_equal(log_g[:, 1], log_o[:, 1])
    assert np.array_equal(log_g[:, 0], log_o[:, 0])
    assert np.allclose(log_g[:, 3:8], log_o[:, 3:8], rtol=1e-12, atol=0)
    # dU/dt of the last stage of step 100
    Rg = case.phi_dt()
    Ro = o.get(1).reshape(-1, 5)[: Rg.shape[0]]
    assert rel_err(Rg, Ro) <= TOL


@pytest.mark.parametrize("scheme", ["Euler", "RK4-5"])
def test_density_wave_other_schemes(scheme):
    cfg = configs.density_wave_box(n=8, steps=20, scheme=scheme, flux="AUSM")
    case, o, _ = make_pair(cfg)
    rc, log_g = case.run()
    n, log_o = o.timedisc(scheme, 20, True, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 20
    e = rel_err(case.state(), o.state())
    _print("density-wave", scheme=scheme, rel=e)
    assert e <= TOL


def test_taylor_green_gradients_and_residual():
    cfg = configs.taylor_green(n=6, steps=5, tile=2)
    case, o, _ = make_pair(cfg)
    Rg = case.time_derivative()
    Ro = o.time_derivative()
    Gg = case.field(0)
    Go = o.gradients()
    eg = rel_err(Gg, Go)
    er = rel_err(Rg, Ro)
    _print("tgv gradients/residual", grad_rel=eg, res_rel=er,
           res_bitdiff=n_bit_diff(Rg, Ro))
    assert eg <= TOL
    assert er <= TOL


@pytest.mark.parametrize("flux,mu,limiter", [("AUSM", True, "Barth-Jespersenn"),
                                              ("Rusanov", True, "Barth-Jespersenn"),
                                              ("Rusanov", False, "Barth-Jespersenn"),
                                              ("AUSM", False, "None")])
def test_taylor_green_steps(flux, mu, limiter):
    """Every instantiation of the face-centred residual kernel on face
    states (AUSM / Rusanov, viscous / inviscid) and the unlimited linear
    reconstruction, 10 RK3-3 steps on 6^3 x 6 tetrahedra (the last cell
    tile partial)."""
    cfg = configs.taylor_green(n=6, steps=10, tile=2)
    cfg["physics"]["flux"] = flux
    if not mu:
        cfg["physics"]["mu_mix"] = 0.0
        cfg["physics"]["visc_dt_mode"] = 1
    cfg["numerics"]["limiter"] = limiter
    case, o, _ = make_pair(cfg)
    rc, log_g = case.run()
    n, log_o = o.timedisc("RK3-3", 10, True, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 10
    e = rel_err(case.state(), o.state())
    _print("tgv U(10)", flux=flux, viscous=mu, limiter=limiter, rel=e,
           viscous_flags=log_g[:, 2].tolist())
    assert e <= TOL
    assert np.array_equal(log_g[:, 2], log_o[:, 2])


def test_taylor_green_full_runs_of_tiles():
    """The tile kernels hand each XCD 8 adjacent tiles of every run of 64
    workgroups (the last, partial run in launch order): 12^3 x 6 = 10,368
    cells in 4^3-hexahedron blocks are 81 tiles, one full run and a partial
    one.  5 RK3-3 steps (AUSM, viscous, Barth-Jespersen) against the oracle."""
    cfg = configs.taylor_green(n=12, steps=5, tile=4)
    case, o, _ = make_pair(cfg)
    rc, _ = case.run()
    n, _ = o.timedisc("RK3-3", 5, True, 1e-10, 0.0, 1e300)
    assert rc == 0 and n == 5
    e = rel_err(case.state(), o.state())
    _print("tgv 12^3 U(5)", rel=e, bitdiff=n_bit_diff(case.state(), o.state()))
    assert e <= TOL
    case.close()


def test_face_kernel_full_and_compact_frame_records(monkeypatch):
    """The face-centred kernel forms t1 and t2 from n when every face's
    stored tangents match that formula (the generated meshes), and loads
    them otherwise; FVH_FULL_FRAMES=1 forces the loads.  Both give the
    oracle's states after 5 RK3-3 steps, and the same bits."""
    states = {}
    for full in ("0", "1"):
        monkeypatch.setenv("FVH_FULL_FRAMES", full)
        cfg = configs.taylor_green(n=6, steps=5, tile=2)
        case, o, _ = make_pair(cfg)
        rc, _ = case.run()
        n, _ = o.timedisc("RK3-3", 5, True, 1e-10, 0.0, 1e300)
        assert rc == 0 and n == 5
        states[full] = case.state()
        e = rel_err(states[full], o.state())
        _print("tgv U(5)", full_frames=full, rel=e,
               bitdiff=n_bit_diff(states[full], o.state()))
        assert e <= TOL
        case.close()
    assert n_bit_diff(states["0"], states["1"]) == 0


@pytest.mark.parametrize("tets", [True, False])
def test_all_boundary_types(tets):
    cfg = configs.all_boundaries_box(tets=tets, steps=10)
    case, o, _ = make_pair(cfg)
    Rg = case.time_derivative()
    Ro = o.time_derivative()
    er = rel_err(Rg, Ro)
    ghosts_g = case.host_phi_total()
    ghosts_o = o.phi_total()
    eg = rel_err(ghosts_g, ghosts_o)
    rc, _ = case.run()
    n, _ = o.timedisc("RK3-3", 10, True, 1e-10, 0.0, 1e300)
    eu = rel_err(case.state(), o.state())
    _print("all-boundaries", tets=tets, res_rel=er, ghost_rel=eg, u_rel=eu)
    assert rc == 0 and n == 10
    assert er <= TOL and eg <= TOL and eu <= TOL


def test_bump_channel_small():
    cfg = configs.bump_channel(nx=16, ny=6, nz=3, steps=10)
    case, o, _ = make_pair(cfg)
    Rg = case.time_derivative()
    Ro = o.time_derivative()
    er = rel_err(Rg, Ro)
    rc, _ = case.run()
    n, _ = o.timedisc("RK3-3", 10, True, 1e-10, 0.0, 1e300)
    eu = rel_err(case.state(), o.state())
    _print("bump channel", res_rel=er, u_rel=eu)
    assert er <= TOL and eu <= TOL


def test_timestep_hook_matches_oracle():
    cfg = configs.taylor_green(n=5, steps=1, tile=0)
    case, o, _ = make_pair(cfg)
    dt_g = case.timestep()
    dt_o, _ = o.timestep()
    assert dt_g == dt_o


def test_runs_are_bitwise_deterministic():
    cfg = configs.taylor_green(n=8, steps=6, tile=4)
    case, _, _ = make_pair(cfg)
    U0 = case.state()
    case.run()
    U1 = case.state()
    case.set_state(U0)
    case.run()
    U2 = case.state()
    assert n_bit_diff(U1, U2) == 0


def test_nan_state_stops_with_status():
    cfg = configs.density_wave_box(n=6, steps=10)
    case, _, _ = make_pair(cfg)
    U = case.state()
    U[3, 0] = np.nan
    case.set_state(U)
    rc, log = case.run()
    assert rc == -5  # FVH_ERR_NAN
    assert len(log) == 0
