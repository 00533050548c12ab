"""GPU tests of the Solver API's hooks (src/Solver/include/solver/
solver_module.h:231-255) through the C-ABI.

- finite_volume_time_derivative leaves the reference's solver_data arrays
  (phi_total_left/right, grad_phi_total_*, flux) current on the host
  (src/Solver/src/finite_volume/finite_volume.c:39-52,81-124); they are
  compared with the oracle's.
- Hooks registered with set_calc_flux / set_update / set_update_gradients
  that are not the built-ins run on the host between the device passes
  (host-mirror path).  Each test registers a numpy restatement of the
  built-in it replaces and checks that the residual and a few RK steps
  equal the built-in device path.
- A time-step hook registered with set_calc_timestep is called by timedisc
  before every step (src/Solver/src/timedisc/timedisc.c:164); a hook that
  returns a constant is compared with the oracle driven by the same
  constant.
"""
import ctypes

import numpy as np
import pytest

from fv3d_amd import configs
from fv3d_amd import solver as S
from fv3d_amd._lib import VOID_T_FT, DOUBLE_T_FT
from tests.helpers import make_pair, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu
KAPPA = 1.4


def _data():
    return ctypes.POINTER(S.data_t).in_dll(S.lib(), "solver_data").contents


def _view(ptr, n, shape):
    return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape)


def _host_arrays(case):
    d = _data()
    N, F, B = case.n_cells, case.n_faces, case.n_boundaries
    return {"L": _view(d.phi_total_left, F * 10, (F, 10)),
            "R": _view(d.phi_total_right, F * 10, (F, 10)),
            "flux": _view(d.flux, F * 5, (F, 5)),
            "gx": _view(d.grad_phi_total_x, (N + B) * 10, (N + B, 10)),
            "gy": _view(d.grad_phi_total_y, (N + B) * 10, (N + B, 10)),
            "gz": _view(d.grad_phi_total_z, (N + B) * 10, (N + B, 10)),
            "phi": _view(d.phi_total, (N + B) * 10, (N + B, 10))}


@pytest.mark.parametrize("name", ["taylor_green", "channel"])
def test_solver_data_arrays_match_oracle(name):
    if name == "taylor_green":
        cfg = configs.taylor_green(n=4, steps=1, tile=0)
        tol = 0.0
    else:
        cfg = configs.bump_channel(nx=8, ny=3, nz=2, mu=1e-3)
        tol = 1e-12  # wall pressure through pow (ocml against glibc)
    case, o, mesh = make_pair(cfg)
    Rg = case.time_derivative()
    Ro = o.time_derivative()
    h = _host_arrays(case)
    N, F, B = case.n_cells, case.n_faces, case.n_boundaries
    ref = {"L": o.get(6).reshape(F, 10), "R": o.get(7).reshape(F, 10),
           "flux": o.get(5).reshape(F, 5), "gx": o.get(2).reshape(-1, 10),
           "gy": o.get(3).reshape(-1, 10), "gz": o.get(4).reshape(-1, 10)}
    for k, v in ref.items():
        e = rel_err(h[k], v)
        print(name, k, "rel", e, "bitdiff", n_bit_diff(h[k], v), flush=True)
        assert e <= max(tol, 0.0) if tol else n_bit_diff(h[k], v) == 0, k
    assert rel_err(Rg, Ro) <= 1e-12
    case.close()


def _rusanov_hook(case):
    """A host restatement of calc_ns_flux with the Rusanov solver, inviscid
    (flux.c:23-198, 320-336): rotate the momenta into the face frame,
    con_to_prim, Rusanov, rotate back; reads phi_total_left/right and writes
    flux."""
    mesh = S.mesh_arrays()
    F = mesh["n_faces"]
    n = mesh["face_n"].reshape(F, 3)
    t1 = mesh["face_t1"].reshape(F, 3)
    t2 = mesh["face_t2"].reshape(F, 3)
    h = _host_arrays(case)

    def rotate(phi):
        out = np.zeros((F, 10))
        m = phi[:, 1:4]
        out[:, 0] = phi[:, 0]
        for q, v in enumerate((n, t1, t2)):
            out[:, 1 + q] = ((0.0 + m[:, 0] * v[:, 0]) + m[:, 1] * v[:, 1]) + m[:, 2] * v[:, 2]
        out[:, 4] = phi[:, 4]
        rho = out[:, 0]
        out[:, 5] = out[:, 1] / rho
        out[:, 6] = out[:, 2] / rho
        out[:, 7] = out[:, 3] / rho
        ke = ((0.0 + out[:, 1] * out[:, 5]) + out[:, 2] * out[:, 6]) + out[:, 3] * out[:, 7]
        out[:, 8] = np.maximum(1e-10, (KAPPA - 1.0) * (out[:, 4] - 0.5 * ke))
        return out

    def euler(p):
        return np.stack([p[:, 1], p[:, 1] * p[:, 5] + p[:, 8], p[:, 1] * p[:, 6],
                         p[:, 1] * p[:, 7], (p[:, 4] + p[:, 8]) * p[:, 5]], axis=1)

    def hook(t):
        lr, rr = rotate(h["L"]), rotate(h["R"])
        cl = np.sqrt(KAPPA * lr[:, 8] / lr[:, 0])
        cr = np.sqrt(KAPPA * rr[:, 8] / rr[:, 0])
        lam = np.maximum(np.abs(lr[:, 5]) + cl, np.abs(rr[:, 5]) + cr)
        fc = 0.5 * (euler(lr) + euler(rr)) - 0.5 * lam[:, None] * (rr[:, :5] - lr[:, :5])
        out = h["flux"]
        out[:, 0] = fc[:, 0]
        for d in range(3):
            out[:, 1 + d] = (fc[:, 1] * n[:, d] + fc[:, 2] * t1[:, d]) + fc[:, 3] * t2[:, d]
        out[:, 4] = fc[:, 4]

    return VOID_T_FT(hook)


def test_custom_flux_hook_matches_builtin():
    cfg = configs.density_wave_box(n=8, steps=5)
    case, o, _ = make_pair(cfg)
    L = case.L
    U0 = case.state()
    R_builtin = case.time_derivative()
    rc, log_b = case.run()
    U_builtin = case.state()
    cb = _rusanov_hook(case)
    L.set_calc_flux(ctypes.cast(cb, ctypes.c_void_p))
    case.set_state(U0)
    R_host = case.time_derivative()
    e = rel_err(R_host, R_builtin)
    print("host Rusanov dU/dt rel", e, "bitdiff", n_bit_diff(R_host, R_builtin), flush=True)
    assert e <= 1e-12
    rc2, log_h = case.run()
    assert rc == rc2 == 0 and len(log_h) == len(log_b) == 5
    e = rel_err(case.state(), U_builtin)
    print("host Rusanov U(5) rel", e, flush=True)
    assert e <= 1e-12
    assert np.allclose(log_h[:, 1], log_b[:, 1], rtol=1e-13, atol=0)
    # the device-only paths refuse a host flux hook instead of ignoring it
    with pytest.raises(RuntimeError):
        case.jfnk(max_newton=1)
    case.close()
    del cb


def test_custom_update_hook_matches_builtin():
    """A host ns_update on a periodic mesh (no ghost rows): the primitives of
    the domain cells from their conservative variables (eos.c:81-92)."""
    cfg = configs.taylor_green(n=4, steps=3, tile=0)
    case, o, _ = make_pair(cfg)
    L = case.L
    N = case.n_cells
    R_mix = S.get_param("R_mix")
    U0 = case.state()
    R_builtin = case.time_derivative()
    rc, _ = case.run()
    U_builtin = case.state()

    def hook(t):
        phi = _host_arrays(case)["phi"][:N]
        rho = phi[:, 0]
        phi[:, 5] = phi[:, 1] / rho
        phi[:, 6] = phi[:, 2] / rho
        phi[:, 7] = phi[:, 3] / rho
        ke = ((0.0 + phi[:, 1] * phi[:, 5]) + phi[:, 2] * phi[:, 6]) + phi[:, 3] * phi[:, 7]
        p = np.maximum(1e-10, (KAPPA - 1.0) * (phi[:, 4] - 0.5 * ke))
        phi[:, 8] = p
        phi[:, 9] = p / (rho * R_mix)

    cb = VOID_T_FT(hook)
    L.set_update(ctypes.cast(cb, ctypes.c_void_p))
    case.set_state(U0)
    R_host = case.time_derivative()
    print("host update dU/dt bitdiff", n_bit_diff(R_host, R_builtin), flush=True)
    assert rel_err(R_host, R_builtin) <= 1e-13
    rc2, _ = case.run()
    assert rc == rc2 == 0
    assert rel_err(case.state(), U_builtin) <= 1e-13
    case.close()
    del cb


def test_custom_update_gradients_hook_matches_builtin():
    """A host ns_update_gradients on a box with outflow boundaries in x: the
    ghost gradient rows are copies of the interior cell's rows
    (boundary.c:355-366), which the viscous flux reads on boundary faces."""
    cfg = configs.density_wave_box(n=5, steps=3, flux="AUSM")
    cfg["mesh"][1]["periodic"] = 6
    cfg["physics"].update(mu_mix=2e-3, visc_dt_mode=0)
    cfg["numerics"].update(reconstruction="Linear", limiter="Barth-Jespersenn")
    cfg["regions"]["xmin"] = ("outflow", None)
    cfg["regions"]["xmax"] = ("outflow", None)
    case, o, mesh = make_pair(cfg)
    L = case.L
    N = case.n_cells
    bface = np.asarray(mesh["boundary_face"])
    fc = np.asarray(mesh["face_cells_file"]).reshape(-1, 2)
    inner = fc[bface, 0]
    U0 = case.state()
    R_builtin = case.time_derivative()
    rc, _ = case.run()
    U_builtin = case.state()

    def hook(t):
        h = _host_arrays(case)
        for k in ("gx", "gy", "gz"):
            h[k][N:] = h[k][inner]

    cb = VOID_T_FT(hook)
    L.set_update_gradients(ctypes.cast(cb, ctypes.c_void_p))
    case.set_state(U0)
    R_host = case.time_derivative()
    print("host update-gradients dU/dt bitdiff", n_bit_diff(R_host, R_builtin), flush=True)
    assert n_bit_diff(R_host, R_builtin) == 0
    rc2, _ = case.run()
    assert rc == rc2 == 0
    assert n_bit_diff(case.state(), U_builtin) == 0
    case.close()
    del cb


def test_custom_timestep_hook_honoured():
    """set_calc_timestep(hook): timedisc takes dt from the hook at every step
    (the last one clamped to t_end); the states equal the oracle's with the
    same constant step."""
    dt = 1.3e-3
    cfg = configs.density_wave_box(n=8, steps=6)
    cfg["numerics"]["t_end"] = 5.5 * dt
    case, o, _ = make_pair(cfg)
    calls = []

    def hook(t):
        calls.append(t)
        return dt

    cb = DOUBLE_T_FT(hook)
    case.L.set_calc_timestep(ctypes.cast(cb, ctypes.c_void_p))
    rc, log = case.run()
    assert rc == 0
    assert len(log) == 6 and len(calls) == 6
    assert np.allclose(log[:5, 1], dt, rtol=0, atol=0)
    assert log[5, 1] == pytest.approx(0.5 * dt, rel=1e-9)
    assert log[-1, 0] == pytest.approx(5.5 * dt, rel=1e-12)
    assert calls[0] == 0.0 and calls[1] == pytest.approx(dt, rel=1e-15)
    o.set_fixed_dt(dt)
    n, log_o = o.timedisc("RK3-3", 6, True, 1e-10, 0.0, 5.5 * dt)
    assert n == 6
    e = rel_err(case.state(), o.state())
    print("host dt hook U rel", e, "bitdiff", n_bit_diff(case.state(), o.state()), flush=True)
    assert e <= 1e-12
    assert np.array_equal(log[:, 1], log_o[:, 1])
    case.close()
    del cb
