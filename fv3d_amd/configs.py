"""Case definitions of the workloads the path is measured and checked on.

Each function returns keyword arguments for :class:`fv3d_amd.solver.Case`
(mesh, physics, regions, ic, numerics).  Sizes are parameters so that the
tests can run the same set-up small.
"""
import math

R_AIR = 8.31446261815324 / 28.96e-3  # navier_stokes.c:16-24 default R_mix


def density_wave_box(n=16, steps=100, flux="Rusanov", scheme="RK3-3"):
    """3D Euler on [0,1]^3, n^3 periodic hexahedra, rho = 1 + 0.2 sin(2 pi
    (x+y+z)), u = v = w = 1, p = 1, kappa = 1.4, first order, explicit RK,
    CFL 0.5 (the 16^3 / 100-step case the CPU reference runs)."""
    R = R_AIR
    return dict(
        mesh=("box", dict(n=(n, n, n), origin=(0.0, 0.0, 0.0), length=(1.0, 1.0, 1.0),
                          tets=False, periodic=7, tile=0)),
        physics=dict(cfl_scale=0.5, dfl_scale=1.0, mu_mix=0.0, R_mix=R, Pr=0.718,
                     kappa=1.4, flux=flux, visc_dt_mode=1),
        regions={"flow": ("flow", (1.0, 1.0, 1.0, 1.0, 1.0 / R))},
        ic=(0, (1.0, 0.2, 1.0, 1.0)),
        numerics=dict(reconstruction="First-Order", limiter="None", scheme=scheme,
                      max_iter=steps, transient=True, t_end=1e300),
    )


def taylor_green(n=96, steps=200, tile=4, mach=0.1, reynolds=1600.0):
    """Navier-Stokes Taylor-Green vortex on [0, 2 pi]^3 periodic, n^3
    hexahedra each Kuhn-split into 6 tetrahedra, rho0 = U0 = 1, R = 1,
    p0 = 1/(kappa Ma^2), mu = 1/Re; AUSM, linear reconstruction with the
    Barth-Jespersen limiter, RK3-3, CFL 0.5."""
    kappa = 1.4
    p0 = 1.0 / (kappa * mach * mach)
    L = 2.0 * math.pi
    T0 = p0  # rho0 * R = 1
    return dict(
        mesh=("box", dict(n=(n, n, n), origin=(0.0, 0.0, 0.0), length=(L, L, L),
                          tets=True, periodic=7, tile=tile)),
        physics=dict(cfl_scale=0.5, dfl_scale=1.0, mu_mix=1.0 / reynolds, R_mix=1.0,
                     Pr=0.718, kappa=kappa, flux="AUSM", visc_dt_mode=0),
        regions={"flow": ("flow", (0.0, 0.0, 0.0, p0, T0))},
        ic=(1, (p0, 1.0, 1.0)),
        numerics=dict(reconstruction="Linear", limiter="Barth-Jespersenn",
                      scheme="RK3-3", max_iter=steps, transient=True, t_end=1e300),
    )


def bump_channel(nx=160, ny=40, nz=13, mach=0.5, jitter=0.05, seed=20261203,
                 mu=0.0, reconstruction="Linear", limiter="Barth-Jespersenn",
                 flux="AUSM", steps=100):
    """Subsonic channel over a Gaussian bump: inflow state at Mach `mach`
    (p = 1, T = 1, R = 1), zero-gradient outflow, slip walls at the bottom
    and top, symmetry in z."""
    kappa = 1.4
    u = mach * math.sqrt(kappa)
    inflow = (u, 0.0, 0.0, 1.0, 1.0)
    return dict(
        mesh=("channel", dict(n=(nx, ny, nz), jitter=jitter, seed=seed)),
        physics=dict(cfl_scale=0.5, dfl_scale=1.0, mu_mix=mu, R_mix=1.0, Pr=0.718,
                     kappa=kappa, flux=flux, visc_dt_mode=1),
        regions={"flow": ("flow", inflow), "inflow": ("inflow", inflow),
                 "outflow": ("outflow", None), "bottom": ("wall-slip", None),
                 "top": ("wall-slip", None), "zmin": ("symmetry", None),
                 "zmax": ("symmetry", None)},
        ic=None,
        numerics=dict(reconstruction=reconstruction, limiter=limiter, scheme="RK3-3",
                      max_iter=steps, transient=True, t_end=1e300),
    )


def bump_channel_steady(nx=160, ny=40, nz=13, mach=0.5, jitter=0.05, seed=20261203,
                        flux="AUSM", reconstruction="Linear"):
    """The bump channel set up for the steady Newton-Krylov solve: the outlet
    imposes the free-stream state (region type "state") instead of the
    zero-gradient copy, and the slopes are unlimited.  With the zero-gradient
    outlet next to a full-state inlet the steady Jacobian is near-singular
    (an eigenvalue of 2.7e-3 against a spectral radius of 6e2 on the 32x8x3
    mesh) and Newton stalls; the Barth-Jespersen limiter's min/max switches
    stall it below 1e-6 as well."""
    cfg = bump_channel(nx=nx, ny=ny, nz=nz, mach=mach, jitter=jitter, seed=seed,
                       reconstruction=reconstruction, limiter="None", flux=flux)
    cfg["regions"]["outflow"] = ("state", cfg["regions"]["inflow"][1])
    return cfg


def all_boundaries_box(n=(6, 5, 4), tets=True, mu=1e-3, flux="AUSM",
                       reconstruction="Linear", limiter="Barth-Jespersenn", steps=10):
    """Non-periodic box whose six sides carry six different boundary types
    (inflow, outflow, adiabatic wall, isothermal wall, slip wall, state), a
    perturbed initial state and viscosity, so every boundary branch, ghost
    gradient and limiter path is exercised."""
    R = 1.0
    flow = (0.3, 0.1, -0.05, 1.0, 1.0)
    return dict(
        mesh=("box", dict(n=n, origin=(0.0, 0.0, 0.0), length=(1.0, 0.8, 0.6),
                          tets=tets, periodic=0, tile=0)),
        physics=dict(cfl_scale=0.5, dfl_scale=1.0, mu_mix=mu, R_mix=R, Pr=0.718,
                     kappa=1.4, flux=flux, visc_dt_mode=1),
        regions={"flow": ("flow", flow),
                 "xmin": ("inflow", (0.35, 0.0, 0.0, 1.05, 1.0)),
                 "xmax": ("outflow", None),
                 "ymin": ("wall-adiabatic", None),
                 "ymax": ("wall-isothermal", (0, 0, 0, 0, 1.1)),
                 "zmin": ("wall-slip", None),
                 "zmax": ("state", (0.2, 0.0, 0.1, 0.95, 0.9))},
        ic=(0, (1.0, 0.15, 0.3, 1.0)),
        numerics=dict(reconstruction=reconstruction, limiter=limiter, scheme="RK3-3",
                      max_iter=steps, transient=True, t_end=1e300),
    )
