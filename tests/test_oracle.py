"""Pins the CPU oracle (oracle/fvref.c) against what the mathematics fixes,
independently of the oracle's own formulas: closed forms, invariants,
inverse relations and special cases.  The reference ships no golden
vectors or known-answer tests for this path (it has no tests at all), so
these checks are what pins the oracle.  Each is chosen so that a plausible
restatement mistake (a dropped term, a wrong sign, index or factor) fails.
"""
import math

import numpy as np
import pytest

from fv3d_amd import configs
from fv3d_amd import solver as S
from fv3d_amd._lib import lib
from oracle import fvref

KAPPA = 1.4


def _oracle_on(cfg):
    """Mesh from the product generator (host code), oracle configured like
    the case; no device needed."""
    L = lib()
    L.init_solver(b"oracle-test")
    kind, p = cfg["mesh"]
    if kind == "box":
        n, o, ln = p["n"], p.get("origin", (0, 0, 0)), p.get("length", (1, 1, 1))
        S.check(L.fvh_mesh_box(n[0], n[1], n[2], o[0], o[1], o[2], ln[0], ln[1], ln[2],
                               1 if p.get("tets") else 0, p.get("periodic", 7), p.get("tile", 0)))
    else:
        n = p["n"]
        S.check(L.fvh_mesh_channel(n[0], n[1], n[2], p.get("jitter", 0.05), p.get("seed", 1)))
    mesh = S.mesh_arrays()
    o = fvref.from_case(mesh, cfg["physics"], cfg["regions"], cfg["numerics"])
    return o, mesh


def _prim_state(rho, u, v, w, p):
    rho = np.asarray(rho, dtype=float)
    U = np.zeros(rho.shape + (5,))
    U[..., 0] = rho
    U[..., 1] = rho * u
    U[..., 2] = rho * v
    U[..., 3] = rho * w
    U[..., 4] = p / (KAPPA - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return U


@pytest.fixture
def eos_oracle():
    cfg = configs.density_wave_box(n=3, steps=1)
    cfg["physics"]["R_mix"] = 287.0
    o, _ = _oracle_on(cfg)
    return o


# ---- EOS -------------------------------------------------------------------
def test_con_to_prim_textbook(eos_oracle):
    rho, u, v, w, p = 1.3, 0.4, -0.2, 0.7, 2.5
    phi = np.zeros(10)
    phi[:5] = _prim_state(rho, u, v, w, p)
    out = eos_oracle.con_to_prim(phi)
    assert np.allclose(out[5:8], [u, v, w], rtol=1e-15)
    assert out[8] == pytest.approx(p, rel=1e-14)
    assert out[9] == pytest.approx(p / (rho * 287.0), rel=1e-14)


def test_pressure_floor(eos_oracle):
    phi = np.zeros(10)
    phi[:5] = [1.0, 2.0, 0.0, 0.0, 1.0]  # kinetic energy 2 > total energy 1
    out = eos_oracle.con_to_prim(phi)
    assert out[8] == 1e-10


def test_prim_con_round_trip(eos_oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        phi = np.zeros(10)
        phi[0] = rng.uniform(0.1, 5)
        phi[5:8] = rng.uniform(-3, 3, 3)
        phi[8] = rng.uniform(0.1, 10)
        back = eos_oracle.con_to_prim(eos_oracle.prim_to_con(phi))
        assert np.allclose(back[5:9], phi[5:9], rtol=4e-15, atol=1e-15)


# ---- wall pressure (eos.c:55-75) ------------------------------------------
def test_riemann_p_expansion_inverse(eos_oracle):
    """u_n <= 0: p_w solves the isentropic relation
    u = 2c/(k-1) ((p_w/p)^((k-1)/(2k)) - 1)."""
    for u in (-0.5, -0.1, -1e-3, 0.0):
        rho, p = 1.2, 1.7
        c = math.sqrt(KAPPA * p / rho)
        phi = np.zeros(10)
        phi[0], phi[5], phi[8] = rho, u, p
        pw = eos_oracle.riemann_p(phi)
        u_back = 2 * c / (KAPPA - 1) * ((pw / p) ** ((KAPPA - 1) / (2 * KAPPA)) - 1)
        assert u_back == pytest.approx(u, abs=1e-13)
    # acoustic limit: dp/du = rho c at u = 0
    phi[5] = -1e-7
    assert (eos_oracle.riemann_p(phi) - p) / -1e-7 == pytest.approx(rho * c, rel=1e-5)


def test_riemann_p_compression_solves_shock_relation(eos_oracle):
    """u_n > 0: p_w solves the shock relation u = (p_w - p) sqrt(A/(p_w + B))
    with B = (k-1)/(k+1) p and the reference's A = 2/((k+1) u_n)."""
    for u in (1e-3, 0.2, 0.9, 3.0):
        rho, p = 1.1, 0.8
        phi = np.zeros(10)
        phi[0], phi[5], phi[8] = rho, u, p
        pw = eos_oracle.riemann_p(phi)
        A = 2.0 / ((KAPPA + 1) * u)
        B = (KAPPA - 1) / (KAPPA + 1) * p
        assert (pw - p) * math.sqrt(A / (pw + B)) == pytest.approx(u, rel=1e-12)


# ---- Riemann solvers -------------------------------------------------------
def _rot_state(rho, un, ut1, ut2, p):
    phi = np.zeros(10)
    phi[:5] = _prim_state(rho, un, ut1, ut2, p)
    phi[5:10] = [un, ut1, ut2, p, p / (rho * 287.0)]
    return phi


def _euler_flux(phi):
    rho, un, ut1, ut2, p = phi[0], phi[5], phi[6], phi[7], phi[8]
    E = phi[4]
    return np.array([rho * un, rho * un * un + p, rho * un * ut1, rho * un * ut2, (E + p) * un])


@pytest.mark.parametrize("flux", ["Rusanov", "AUSM"])
@pytest.mark.parametrize("un", [-2.5, -0.3, 0.0, 0.4, 1.8])
def test_flux_consistency(eos_oracle, flux, un):
    phi = _rot_state(1.1, un, 0.3, -0.2, 0.9)
    f = eos_oracle.riemann(flux, phi, phi)
    assert np.allclose(f, _euler_flux(phi), rtol=1e-13, atol=1e-13)


def test_ausm_supersonic_upwinding(eos_oracle):
    L = _rot_state(1.0, 3.0, 0.1, 0.2, 1.0)   # M_l > 1
    R = _rot_state(0.5, 2.5, -0.3, 0.0, 0.4)  # M_r > 1
    assert np.allclose(eos_oracle.riemann("AUSM", L, R), _euler_flux(L), rtol=1e-14)
    L2 = _rot_state(1.0, -3.0, 0.1, 0.2, 1.0)
    R2 = _rot_state(0.5, -2.5, -0.3, 0.0, 0.4)
    assert np.allclose(eos_oracle.riemann("AUSM", L2, R2), _euler_flux(R2), rtol=1e-14)


def test_rusanov_mirror_symmetry(eos_oracle):
    """Swapping sides and reversing the normal velocity mirrors the flux:
    mass, tangential momenta and energy change sign, normal momentum not."""
    L = _rot_state(1.2, 0.4, 0.2, -0.1, 1.1)
    R = _rot_state(0.9, -0.1, 0.3, 0.05, 0.8)
    Lm = _rot_state(0.9, 0.1, 0.3, 0.05, 0.8)
    Rm = _rot_state(1.2, -0.4, 0.2, -0.1, 1.1)
    f = eos_oracle.riemann("Rusanov", L, R)
    g = eos_oracle.riemann("Rusanov", Lm, Rm)
    assert np.allclose(g, [-f[0], f[1], -f[2], -f[3], -f[4]], rtol=1e-14, atol=1e-15)


def test_rusanov_dissipation(eos_oracle):
    """F = average of the two fluxes minus half the largest wave speed times
    the jump (checked on a pure density jump at rest)."""
    L = _rot_state(1.0, 0.0, 0.0, 0.0, 1.0)
    R = _rot_state(0.5, 0.0, 0.0, 0.0, 1.0)
    f = eos_oracle.riemann("Rusanov", L, R)
    lam = max(math.sqrt(KAPPA / 1.0), math.sqrt(KAPPA / 0.5))
    assert f[0] == pytest.approx(-0.5 * lam * (0.5 - 1.0), rel=1e-14)
    assert f[1] == pytest.approx(1.0, rel=1e-14)


# ---- mesh metrics ----------------------------------------------------------
def test_dist_vectors_through_periodic_image():
    cfg = configs.density_wave_box(n=4, steps=1)
    o, mesh = _oracle_on(cfg)
    F, N = mesh["n_faces"], mesh["n_cells"]
    d1 = o.get(8).reshape(F, 3)
    d2 = o.get(9).reshape(F, 3)
    h = 0.25
    # uniform hex: the face centre is half a cell from both centroids
    assert np.allclose(np.abs(d1).sum(1), 0.5 * h, atol=1e-15)
    assert np.allclose(np.abs(d2).sum(1), 0.5 * h, atol=1e-15)
    assert np.allclose(d1, -d2, atol=1e-15)
    # metrics agree bit for bit with the product's calc_mesh_metrics
    assert np.array_equal(d1.reshape(-1), mesh["dist_cell_1"])
    assert np.array_equal(d2.reshape(-1), mesh["dist_cell_2"])


# ---- gradients -------------------------------------------------------------
def test_green_gauss_exact_for_linear_fields():
    cfg = configs.density_wave_box(n=6, steps=1)
    cfg["mesh"][1]["periodic"] = 0
    cfg["regions"] = {k: ("outflow", None) for k in
                      ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    cfg["regions"]["flow"] = ("flow", (0, 0, 0, 1, 1))
    o, mesh = _oracle_on(cfg)
    xc = mesh["cell_x"].reshape(-1, 3)
    a = np.array([0.3, -0.7, 0.45])
    rho = 2.0 + xc @ a
    U = _prim_state(rho, 0.0, 0.0, 0.0, 1.0)
    o.set_state(U.reshape(-1))
    o.time_derivative()
    g = o.gradients()
    interior = np.all((xc > 1.0 / 6) & (xc < 5.0 / 6), axis=1)
    assert interior.sum() == 64
    assert np.allclose(g[interior][:, [0, 10, 20]], a, atol=1e-12)
    # constants have zero gradient everywhere (closed cells)
    assert np.allclose(g[:, [8, 18, 28]], 0.0, atol=1e-12)


# ---- residual invariants -------------------------------------------------
@pytest.mark.parametrize("tets", [False, True])
@pytest.mark.parametrize("recon", ["First-Order", "Linear"])
def test_free_stream_preservation(tets, recon):
    cfg = configs.taylor_green(n=4, steps=1, tile=0)
    cfg["mesh"][1]["tets"] = tets
    cfg["numerics"]["reconstruction"] = recon
    o, mesh = _oracle_on(cfg)
    U = _prim_state(np.full(mesh["n_cells"], 1.3), 0.4, -0.2, 0.1, 70.0)
    o.set_state(U.reshape(-1))
    R = o.time_derivative()
    assert np.abs(R).max() < 1e-9 * 70.0 / (2 * math.pi / 4)


@pytest.mark.parametrize("flux", ["Rusanov", "AUSM"])
def test_periodic_conservation(flux):
    cfg = configs.taylor_green(n=5, steps=1, tile=0)
    cfg["physics"]["flux"] = flux
    o, mesh = _oracle_on(cfg)
    rng = np.random.default_rng(11)
    N = mesh["n_cells"]
    U = _prim_state(1 + 0.1 * rng.random(N), 0.3 * rng.random(N), 0.2 * rng.random(N),
                    -0.1 * rng.random(N), 70 + rng.random(N))
    o.set_state(U.reshape(-1))
    R = o.time_derivative()
    V = mesh["cell_volume"]
    total = (R * V[:, None]).sum(0)
    scale = (np.abs(R) * V[:, None]).sum(0)
    assert np.all(np.abs(total) <= 1e-13 * scale)


def test_viscous_shear_closed_form():
    """u = y^2, rho = p = 1 on a uniform hex box, linear reconstruction, no
    limiter, Rusanov: the reconstructed states are continuous, so the only
    x-momentum residual on interior cells is the viscous one, 2*mu."""
    n = 8
    mu = 0.01
    cfg = configs.density_wave_box(n=n, steps=1)
    cfg["mesh"][1]["periodic"] = 0
    cfg["physics"].update(mu_mix=mu, R_mix=1.0)
    cfg["numerics"].update(reconstruction="Linear", limiter="None")
    cfg["regions"] = {k: ("outflow", None) for k in
                      ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    cfg["regions"]["flow"] = ("flow", (0, 0, 0, 1, 1))
    o, mesh = _oracle_on(cfg)
    xc = mesh["cell_x"].reshape(-1, 3)
    U = _prim_state(np.ones(len(xc)), xc[:, 1] ** 2, 0.0, 0.0, 1.0)
    o.set_state(U.reshape(-1))
    R = o.time_derivative()
    h = 1.0 / n
    inner = np.all((xc > 2 * h) & (xc < 1 - 2 * h), axis=1)
    assert inner.sum() == 64
    assert np.allclose(R[inner, 1], 2 * mu, rtol=1e-9)
    assert np.allclose(R[inner, 0], 0.0, atol=1e-12)
    # z-momentum: every z-face pair carries the same state (rho*v picks up
    # a pressure term because rho*E = p/(k-1) + y^4/2 is extrapolated
    # linearly, so it is not part of the closed form)
    assert np.allclose(R[inner, 3], 0.0, atol=1e-12)
    # and without viscosity the same state is steady in x-momentum
    o.set_physics(mu_mix=0.0, R_mix=1.0, flux="Rusanov", reconstruction="Linear",
                  limiter="None", visc_dt_mode=1)
    R0 = o.time_derivative()
    assert np.allclose(R0[inner, 1], 0.0, atol=1e-12)


def test_limited_faces_within_neighbourhood():
    cfg = configs.taylor_green(n=4, steps=1, tile=0)
    o, mesh = _oracle_on(cfg)
    rng = np.random.default_rng(5)
    N, F = mesh["n_cells"], mesh["n_faces"]
    U = _prim_state(1 + 0.3 * rng.random(N), rng.normal(size=N), rng.normal(size=N),
                    rng.normal(size=N), 70 + 5 * rng.random(N))
    o.set_state(U.reshape(-1))
    o.time_derivative()
    phi = o.phi_total()[:N]
    left = o.get(6).reshape(F, 10)
    fc = mesh["face_cells_file"].reshape(F, 2)
    cf = mesh["cell_faces"].reshape(N, -1)
    lo = phi.copy()
    hi = phi.copy()
    for c in range(N):
        for f in cf[c]:
            nb = fc[f, 1] if fc[f, 0] == c else fc[f, 0]
            lo[c] = np.minimum(lo[c], phi[nb])
            hi[c] = np.maximum(hi[c], phi[nb])
    c0 = fc[:, 0]
    tol = 1e-12 * np.abs(phi).max(0)
    assert np.all(left >= lo[c0] - tol) and np.all(left <= hi[c0] + tol)


# ---- time step -------------------------------------------------------------
def test_timestep_closed_form_uniform_hex():
    n = 8
    cfg = configs.density_wave_box(n=n, steps=1)
    o, mesh = _oracle_on(cfg)
    N = mesh["n_cells"]
    rho = 1.0 + 0.5 * np.random.default_rng(2).random(N)
    u, v, w, p = 0.3, -0.8, 0.1, 1.0
    o.set_state(_prim_state(rho, u, v, w, p).reshape(-1))
    dt, visc = o.timestep()
    h = 1.0 / n
    c = np.sqrt(KAPPA * p / rho)
    # dx = 2 h^2 (1,1,1), V = h^3  ->  V/ds1 = h / (2 sum_d (|u_d| + c))
    expect = 0.5 * (h / (2 * (abs(u) + abs(v) + abs(w) + 3 * c))).min()
    assert dt == pytest.approx(expect, rel=1e-14)
    assert visc == 0


def test_timestep_viscous_branch_reference_formula():
    """With mu > 0 the reference's viscous limit is dfl*mu*(kappa/Pr)*lv^2
    with lv = min (1/rho) dx.dx / V (navier_stokes.c:183-192)."""
    n = 4
    cfg = configs.density_wave_box(n=n, steps=1)
    cfg["physics"].update(mu_mix=1e-3, visc_dt_mode=0)
    o, mesh = _oracle_on(cfg)
    N = mesh["n_cells"]
    rho = np.full(N, 2.0)
    o.set_state(_prim_state(rho, 0.0, 0.0, 0.0, 1.0).reshape(-1))
    dt, visc = o.timestep()
    h = 1.0 / n
    lv = (1 / 2.0) * 3 * (2 * h * h) ** 2 / h ** 3
    lv_ref = lv * (1.0 * 1e-3 * KAPPA / 0.718 * lv)
    lc = 0.5 * h / (2 * 3 * math.sqrt(KAPPA / 2.0))
    assert dt == pytest.approx(min(lc, lv_ref), rel=1e-13)
    assert visc == int(lv_ref < lc)


# ---- LSERK -----------------------------------------------------------------
def test_lserk_stability_polynomials():
    L = fvref.lib()
    for z in (-0.1, -0.7, 0.3, -2.0):
        assert L.fvref_lserk_amplification(0, z) == pytest.approx(1 + z, rel=1e-15)
        assert L.fvref_lserk_amplification(1, z) == pytest.approx(
            1 + z + z * z / 2 + z ** 3 / 6, rel=1e-14)
    # RK4-5: fourth order, i.e. matches exp(z) up to z^4
    for z in (1e-2, -1e-2, 2e-2):
        err = abs(L.fvref_lserk_amplification(2, z) - math.exp(z))
        assert err < 2e-2 * abs(z) ** 5


def test_global_residual_definition():
    cfg = configs.taylor_green(n=3, steps=1, tile=0)
    o, mesh = _oracle_on(cfg)
    rng = np.random.default_rng(9)
    N = mesh["n_cells"]
    o.set_state(_prim_state(1 + 0.1 * rng.random(N), rng.random(N), 0.0, 0.0, 70.0).reshape(-1))
    R = o.time_derivative()
    dt = 0.37
    r = o.global_residual(dt)
    V = mesh["cell_volume"]
    expect = np.array([sum(abs(R[i, j]) * V[i] for i in range(N)) for j in range(5)])
    expect *= dt / V.sum()
    assert np.allclose(r, expect, rtol=1e-13)
