"""ctypes wrapper of the CPU oracle (oracle/libfvref.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker, never by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvref.so")
PD = ctypes.POINTER(ctypes.c_double)
PI = ctypes.POINTER(ctypes.c_int)
c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p

BC = {"flow": 0, "inflow": 1, "outflow": 2, "wall-adiabatic": 3,
      "wall-isothermal": 4, "wall-slip": 5, "symmetry": 6, "state": 7,
      "function": 8}

_L = None


def build():
    """Compile the oracle with the recipe in oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _L
    if _L is not None:
        return _L
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    L.fvref_create.restype = c_void_p
    L.fvref_create.argtypes = [c_int] * 5 + [PI, PI, PD, PD, PD, PI, PI, PD, PD, PD, PD,
                                             PD, PD, PD, PI, PI, PI]
    L.fvref_destroy.argtypes = [c_void_p]
    L.fvref_set_physics.argtypes = [c_void_p] + [c_double] * 6 + [c_int] * 4
    L.fvref_set_region.argtypes = [c_void_p, c_int, c_int, PD]
    L.fvref_set_state.argtypes = [c_void_p, PD]
    L.fvref_set_fixed_dt.argtypes = [c_void_p, c_double]
    L.fvref_set_lsoe.argtypes = [c_void_p, c_int, c_double, c_int, c_int, c_int]
    L.fvref_lsoe_iterations.restype = c_int
    L.fvref_lsoe_iterations.argtypes = [c_void_p]
    L.fvref_time_derivative.argtypes = [c_void_p, c_double]
    L.fvref_calc_timestep.restype = c_double
    L.fvref_calc_timestep.argtypes = [c_void_p, PI]
    L.fvref_global_residual.argtypes = [c_void_p, c_double, PD]
    L.fvref_timedisc_explicit.restype = c_int
    L.fvref_timedisc_explicit.argtypes = [c_void_p, c_int, c_int, c_int, c_double,
                                          c_double, c_double, PD]
    L.fvref_timestep_newton.restype = c_int
    L.fvref_timestep_newton.argtypes = [c_void_p, c_int, c_int, c_double, c_double,
                                        c_int, c_int, PD, PD]
    L.fvref_timedisc_implicit.restype = c_int
    L.fvref_timedisc_implicit.argtypes = [c_void_p, c_int, c_int, c_int, c_double,
                                          c_double, c_double, c_int, PD]
    L.fvref_get.argtypes = [c_void_p, c_int, PD]
    L.fvref_n_elements.restype = c_int
    L.fvref_n_elements.argtypes = [c_void_p]
    L.fvref_global_volume.restype = c_double
    L.fvref_global_volume.argtypes = [c_void_p]
    L.fvref_con_to_prim.argtypes = [c_void_p, PD]
    L.fvref_prim_to_con.argtypes = [c_void_p, PD]
    L.fvref_riemann_p.restype = c_double
    L.fvref_riemann_p.argtypes = [c_void_p, PD]
    L.fvref_riemann.argtypes = [c_void_p, c_int, PD, PD, PD]
    L.fvref_lserk_amplification.restype = c_double
    L.fvref_lserk_amplification.argtypes = [c_int, c_double]
    _L = L
    return L


def _pi(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(PI)


def _pd(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(PD)


class Oracle:
    """One oracle instance over a mesh given as mesh-file arrays."""

    def __init__(self, mesh):
        L = lib()
        self.L = L
        self.mesh = mesh
        keep = {}
        for k in ("cell_n_faces", "cell_faces", "face_cells_file", "face_boundary",
                  "boundary_id", "boundary_face", "region_is_boundary"):
            keep[k] = np.ascontiguousarray(mesh[k], dtype=np.int32)
        for k in ("cell_x", "cell_volume", "cell_dx", "face_x", "face_n", "face_t1",
                  "face_t2", "face_area", "face_lambda", "face_shift"):
            keep[k] = np.ascontiguousarray(mesh[k], dtype=np.float64)
        self._keep = keep
        self.N = mesh["n_cells"]
        self.F = mesh["n_faces"]
        self.B = mesh["n_boundaries"]
        self.h = L.fvref_create(
            self.N, mesh["max_cell_faces"], self.F, self.B, mesh["n_regions"],
            keep["cell_n_faces"].ctypes.data_as(PI), keep["cell_faces"].ctypes.data_as(PI),
            keep["cell_x"].ctypes.data_as(PD), keep["cell_volume"].ctypes.data_as(PD),
            keep["cell_dx"].ctypes.data_as(PD),
            keep["face_cells_file"].ctypes.data_as(PI), keep["face_boundary"].ctypes.data_as(PI),
            keep["face_x"].ctypes.data_as(PD), keep["face_n"].ctypes.data_as(PD),
            keep["face_t1"].ctypes.data_as(PD), keep["face_t2"].ctypes.data_as(PD),
            keep["face_area"].ctypes.data_as(PD), keep["face_lambda"].ctypes.data_as(PD),
            keep["face_shift"].ctypes.data_as(PD),
            keep["boundary_id"].ctypes.data_as(PI), keep["boundary_face"].ctypes.data_as(PI),
            keep["region_is_boundary"].ctypes.data_as(PI))

    def __del__(self):
        try:
            if self.h:
                self.L.fvref_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_physics(self, cfl_scale=1.0, dfl_scale=1.0, mu_mix=18.13e-6,
                    R_mix=8.31446261815324 / 28.96e-3, Pr=0.718, kappa=1.4,
                    flux="AUSM", reconstruction="First-Order", limiter="None",
                    visc_dt_mode=0):
        self.L.fvref_set_physics(self.h, cfl_scale, dfl_scale, mu_mix, R_mix, Pr, kappa,
                                 1 if flux == "Rusanov" else 0,
                                 1 if reconstruction == "Linear" else 0,
                                 0 if limiter == "Barth-Jespersenn" else 1, visc_dt_mode)

    def set_region(self, region, type_name, prim5):
        p = np.ascontiguousarray(prim5 if prim5 is not None else [0.0] * 5, dtype=np.float64)
        self.L.fvref_set_region(self.h, region, BC[type_name], p.ctypes.data_as(PD))

    def set_fixed_dt(self, dt):
        """timedisc with a time-step hook returning dt (0: calc_ns_timestep)."""
        self.L.fvref_set_fixed_dt(self.h, dt)

    def set_lsoe(self, solver="GMRes", tol=1e-12, max_iter=100, dims=15, restarts=2):
        """Linear solve of the implicit step: "GMRes", "BiCGStab" or "exact"."""
        self.L.fvref_set_lsoe(self.h, {"BiCGStab": 0, "GMRes": 1, "exact": -1}[solver], tol,
                              max_iter, dims, restarts)

    def lsoe_iterations(self):
        return self.L.fvref_lsoe_iterations(self.h)

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self.L.fvref_set_state(self.h, U.ctypes.data_as(PD))

    def get(self, what):
        sizes = {0: (self.N + self.B) * 10, 1: (self.N + self.B) * 5,
                 2: (self.N + self.B) * 10, 3: (self.N + self.B) * 10,
                 4: (self.N + self.B) * 10, 5: self.F * 5, 6: self.F * 10,
                 7: self.F * 10, 8: self.F * 3, 9: self.F * 3, 10: self.F * 2}
        out = np.zeros(sizes[what])
        self.L.fvref_get(self.h, what, out.ctypes.data_as(PD))
        return out

    def state(self):
        return self.get(0).reshape(-1, 10)[: self.N, :5].copy()

    def phi_total(self):
        return self.get(0).reshape(-1, 10)

    def time_derivative(self, t=0.0):
        self.L.fvref_time_derivative(self.h, t)
        return self.get(1).reshape(-1, 5)[: self.N].copy()

    def gradients(self):
        gx = self.get(2).reshape(-1, 10)[: self.N]
        gy = self.get(3).reshape(-1, 10)[: self.N]
        gz = self.get(4).reshape(-1, 10)[: self.N]
        return np.concatenate([gx, gy, gz], axis=1)

    def timestep(self):
        v = ctypes.c_int(0)
        dt = self.L.fvref_calc_timestep(self.h, ctypes.byref(v))
        return dt, v.value

    def global_residual(self, dt):
        r = np.zeros(5)
        self.L.fvref_global_residual(self.h, dt, r.ctypes.data_as(PD))
        return r

    def timedisc(self, scheme="RK3-3", max_iter=100, transient=True, abort_residual=1e-10,
                 t_start=0.0, t_end=1e300):
        sch = {"Euler": 0, "RK3-3": 1, "RK4-5": 2}[scheme]
        log = np.zeros((max_iter, 8))
        n = self.L.fvref_timedisc_explicit(self.h, sch, max_iter, 1 if transient else 0,
                                           abort_residual, t_start, t_end,
                                           log.ctypes.data_as(PD))
        return n, log[: max(n, 0)]

    def timedisc_implicit(self, scheme="Euler", max_iter=10, transient=True,
                          abort_residual=1e-10, t_start=0.0, t_end=1e300, max_iter_inner=100):
        """timedisc() with the implicit step (scheme "Euler" or "BDF2")."""
        sch = {"Euler": 0, "BDF2": 1}[scheme]
        log = np.zeros((max_iter, 8))
        n = self.L.fvref_timedisc_implicit(self.h, sch, max_iter, 1 if transient else 0,
                                           abort_residual, t_start, t_end, max_iter_inner,
                                           log.ctypes.data_as(PD))
        return n, log[: max(n, 0)]

    def timestep_newton(self, scheme, it, t, dt, old_levels, transient=True,
                        max_iter_inner=100, want_jac=False):
        """One implicit step on the current state; old_levels [2, N*5] is the
        BDF history (updated in place).  phi_dt must hold R of the state."""
        assert old_levels.dtype == np.float64 and old_levels.flags.c_contiguous
        jac = np.zeros(self.N * 25) if want_jac else None
        k = self.L.fvref_timestep_newton(self.h, {"Euler": 0, "BDF2": 1}[scheme], it, t, dt,
                                         1 if transient else 0, max_iter_inner,
                                         old_levels.ctypes.data_as(PD),
                                         jac.ctypes.data_as(PD) if want_jac else None)
        return k, jac

    # physics helpers
    def con_to_prim(self, phi10):
        a = np.array(phi10, dtype=np.float64)
        self.L.fvref_con_to_prim(self.h, a.ctypes.data_as(PD))
        return a

    def prim_to_con(self, phi10):
        a = np.array(phi10, dtype=np.float64)
        self.L.fvref_prim_to_con(self.h, a.ctypes.data_as(PD))
        return a

    def riemann_p(self, phi10):
        a = np.array(phi10, dtype=np.float64)
        return self.L.fvref_riemann_p(self.h, a.ctypes.data_as(PD))

    def riemann(self, flux, l10, r10):
        l = np.array(l10, dtype=np.float64)
        r = np.array(r10, dtype=np.float64)
        f = np.zeros(5)
        self.L.fvref_riemann(self.h, 1 if flux == "Rusanov" else 0, l.ctypes.data_as(PD),
                             r.ctypes.data_as(PD), f.ctypes.data_as(PD))
        return f


def from_case(case_mesh_arrays, physics, regions, numerics=None):
    """Oracle configured like a product Case (same mesh arrays, physics and
    region states)."""
    numerics = numerics or {}
    o = Oracle(case_mesh_arrays)
    o.set_physics(cfl_scale=physics.get("cfl_scale", 1.0), dfl_scale=physics.get("dfl_scale", 1.0),
                  mu_mix=physics.get("mu_mix", 18.13e-6),
                  R_mix=physics.get("R_mix", 8.31446261815324 / 28.96e-3),
                  Pr=physics.get("Pr", 0.718), kappa=physics.get("kappa", 1.4),
                  flux=physics.get("flux", "AUSM"),
                  reconstruction=numerics.get("reconstruction", "First-Order"),
                  limiter=numerics.get("limiter", "None"),
                  visc_dt_mode=physics.get("visc_dt_mode", 0))
    for i, nm in enumerate(case_mesh_arrays["region_names"]):
        t, prim = regions[nm]
        o.set_region(i, t, prim)
    return o
