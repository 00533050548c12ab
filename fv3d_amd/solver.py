"""Python host mirror of the Solver API (include/fv3d_solver.h).

ctypes views of the reference structs (src/Solver/include/solver/
solver_type.h:22-207) so that Python sees ``solver_mesh``/``solver_data``
exactly as a C application would, plus :class:`Case`, which runs the
reference's initialisation order (src/FV3D/src/fv3d.c:27-63 with the module
registration order fv3d -> mesh -> navier_stokes -> boundary -> flux -> fv
-> reconstruction -> limiter -> analyze -> output -> timedisc -> explicit /
implicit -> restart) against libfv3d_hip.so.
"""
import ctypes
import numpy as np

from ._lib import lib, P_DOUBLE, P_INT, JfnkOptions, JfnkStats

c_int, c_double = ctypes.c_int, ctypes.c_double
PD, PI = ctypes.POINTER(c_double), ctypes.POINTER(c_int)
c_ull = ctypes.c_ulonglong


class Boundaries_t(ctypes.Structure):
    _fields_ = [("n_global_boundaries", c_int), ("n_boundaries", c_int),
                ("max_boundary_vertices", c_int),
                ("id", PI), ("type", PI), ("n_vertices", PI), ("vertices", PI),
                ("face", PI), ("distance", PD), ("n", PD), ("t1", PD), ("t2", PD),
                ("stride", ctypes.POINTER(c_ull))]


class Cells_t(ctypes.Structure):
    _fields_ = [("n_global_cells", c_int), ("n_local_cells", c_int),
                ("n_domain_cells", c_int), ("max_cell_vertices", c_int),
                ("max_cell_faces", c_int),
                ("id", PI), ("type", PI), ("n_vertices", PI), ("vertices", PI),
                ("n_faces", PI), ("faces", PI), ("x", PD), ("volume", PD),
                ("dx", PD), ("stride", ctypes.POINTER(c_ull))]


class Faces_t(ctypes.Structure):
    _fields_ = [("n_global_faces", c_int), ("n_faces", c_int),
                ("max_face_vertices", c_int),
                ("type", PI), ("n_vertices", PI), ("vertices", PI), ("cells", PI),
                ("boundary", PI), ("area", PD), ("lambda_", PD), ("x", PD),
                ("n", PD), ("t1", PD), ("t2", PD),
                ("stride", ctypes.POINTER(c_ull)),
                ("dist_cell_1", PD), ("dist_cell_2", PD),
                ("n_internal_faces", c_int), ("internal_faces", PI),
                ("n_boundary_faces", c_int), ("boundary_faces", PI)]


class Regions_t(ctypes.Structure):
    _fields_ = [("n_regions", c_int), ("max_name_length", c_int),
                ("name", ctypes.POINTER(ctypes.c_char_p)), ("is_boundary", PI),
                ("flow_region", c_int), ("type", PI), ("function_id", PI),
                ("phi_total", PD)]


class Vertices_t(ctypes.Structure):
    _fields_ = [("n_vertices", c_int), ("x", PD)]


class Mesh_t(ctypes.Structure):
    _fields_ = [("dimension", c_int), ("is_partitioned", c_int),
                ("partition", ctypes.c_void_p),
                ("vertices", ctypes.POINTER(Vertices_t)),
                ("cells", ctypes.POINTER(Cells_t)),
                ("boundaries", ctypes.POINTER(Boundaries_t)),
                ("faces", ctypes.POINTER(Faces_t)),
                ("regions", ctypes.POINTER(Regions_t)),
                ("local_volume", c_double), ("global_volume", c_double)]


class data_t(ctypes.Structure):
    _fields_ = [("flux", PD), ("phi_total_left", PD), ("phi_total_right", PD),
                ("grad_phi_total_x", PD), ("grad_phi_total_y", PD),
                ("grad_phi_total_z", PD), ("phi_dt", PD), ("phi_total", PD)]


# enums (solver_type.h:212-262, navier_stokes_module.h:47-59)
FIRST_ORDER, LINEAR = 0, 1
BARTH_JESPERSENN, NO_LIMITER = 0, 1
EXPLICIT, IMPLICIT = 0, 1
EULER_EXPLICIT, RK33, RK45 = 0, 1, 2
EULER_IMPLICIT, BDF2 = 0, 1
BICGSTAB, GMRES = 0, 1
BOUNDARY_TYPES = ["flow", "inflow", "outflow", "wall-adiabatic",
                  "wall-isothermal", "wall-slip", "symmetry", "state", "function"]
SCHEMES = {"Euler": EULER_EXPLICIT, "RK3-3": RK33, "RK4-5": RK45}


def check(rc):
    if rc != 0:
        raise RuntimeError("fv3d_hip error %d: %s" % (rc, lib().fvh_last_error().decode()))
    return rc


def _arr(ptr, n, dtype):
    if not ptr or n == 0:
        return np.zeros(0, dtype=dtype)
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_int
    buf = ctypes.cast(ptr, ctypes.POINTER(ct * n)).contents
    return np.frombuffer(buf, dtype=dtype).copy()


def mesh_struct():
    return ctypes.POINTER(Mesh_t).in_dll(lib(), "solver_mesh").contents


def mesh_arrays():
    """Copy of solver_mesh as numpy arrays in the mesh-file layout (second
    face cell -1 on boundary faces), plus the periodic shift."""
    m = mesh_struct()
    c, f, b, r = m.cells.contents, m.faces.contents, m.boundaries.contents, m.regions.contents
    N, F, B, R = c.n_domain_cells, f.n_faces, b.n_boundaries, r.n_regions
    maxf = c.max_cell_faces
    out = dict(
        n_cells=N, n_faces=F, n_boundaries=B, n_regions=R, max_cell_faces=maxf,
        cell_n_faces=_arr(c.n_faces, N, np.int32),
        cell_faces=_arr(c.faces, N * maxf, np.int32),
        cell_x=_arr(c.x, 3 * N, np.float64),
        cell_volume=_arr(c.volume, N, np.float64),
        cell_dx=_arr(c.dx, 3 * N, np.float64),
        face_cells=_arr(f.cells, 2 * F, np.int32),
        face_boundary=_arr(f.boundary, F, np.int32),
        face_x=_arr(f.x, 3 * F, np.float64),
        face_n=_arr(f.n, 3 * F, np.float64),
        face_t1=_arr(f.t1, 3 * F, np.float64),
        face_t2=_arr(f.t2, 3 * F, np.float64),
        face_area=_arr(f.area, F, np.float64),
        face_lambda=_arr(f.lambda_, F, np.float64),
        dist_cell_1=_arr(f.dist_cell_1, 3 * F, np.float64),
        dist_cell_2=_arr(f.dist_cell_2, 3 * F, np.float64),
        boundary_id=_arr(b.id, B, np.int32),
        boundary_face=_arr(b.face, B, np.int32),
        region_is_boundary=_arr(r.is_boundary, R, np.int32),
        region_names=[r.name[i].decode() for i in range(R)],
        global_volume=m.global_volume,
    )
    fc = out["face_cells"].reshape(F, 2).copy()
    fc[out["face_boundary"] >= 0, 1] = -1
    out["face_cells_file"] = fc.reshape(-1)
    shift = np.zeros(3 * F)
    check(lib().fvh_get_face_shift(shift.ctypes.data_as(P_DOUBLE)))
    out["face_shift"] = shift
    return out


def set_param(name, value):
    ctypes.c_double.in_dll(lib(), name).value = value


def get_param(name):
    return ctypes.c_double.in_dll(lib(), name).value


_JFNK_NAMES = {"precision": {"fp64": 0, "fp32": 1},
               "precond": {"scalar": 0, "block": 1, "learned": 2},
               "coupling": {"none": 0, "dilu": 1}}


class Case:
    """One solver set-up in the reference's initialisation order.

    mesh: ("box", dict(n=(nx,ny,nz), origin=..., length=..., tets=bool,
           periodic=mask, tile=int)) or ("channel", dict(n=..., jitter=,
           seed=)) or ("file", path) or ("gmsh", dict(path=, scale=(sx,sy,sz)))
    physics: cfl_scale, dfl_scale, mu_mix, R_mix, Pr, kappa, flux
           ("AUSM"/"Rusanov"), visc_dt_mode
    regions: {name: (type_string, (u, v, w, p, T))}
    ic: None (reference exact function: uniform flow state) or
        (kind, params) for fvh_set_ic
    numerics: reconstruction ("First-Order"/"Linear"), limiter
           ("Barth-Jespersenn"/"None"), scheme ("RK3-3", ...), max_iter,
           transient, abort_residual, t_start, t_end, i_output_data
    """

    _active = None

    def __init__(self, mesh, physics=None, regions=None, ic=None, numerics=None,
                 title="fv3d", verbose=False):
        L = lib()
        if Case._active is not None:
            Case._active.close()
        physics = dict(physics or {})
        numerics = dict(numerics or {})
        self.L = L
        L.fvh_set_verbose(1 if verbose else 0)
        L.init_solver(title.encode())
        kind = mesh[0]
        if kind == "box":
            p = mesh[1]
            n = p["n"]
            o = p.get("origin", (0.0, 0.0, 0.0))
            ln = p.get("length", (1.0, 1.0, 1.0))
            check(L.fvh_mesh_box(n[0], n[1], n[2], o[0], o[1], o[2], ln[0], ln[1], ln[2],
                                 1 if p.get("tets") else 0, p.get("periodic", 7),
                                 p.get("tile", 0)))
        elif kind == "channel":
            p = mesh[1]
            n = p["n"]
            check(L.fvh_mesh_channel(n[0], n[1], n[2], p.get("jitter", 0.05),
                                     p.get("seed", 20261203)))
        elif kind == "batch":
            p = mesh[1]
            nh = (ctypes.c_int * len(p["n_hex"]))(*p["n_hex"])
            check(L.fvh_mesh_batch(len(p["n_hex"]), nh, p.get("length", 1.0), p.get("tile", 0)))
        elif kind == "file":
            L.init_mesh(mesh[1].encode())
        elif kind == "gmsh":
            p = mesh[1]
            check(L.fvh_mesh_gmsh(p["path"].encode(), *p.get("scale", (1.0, 1.0, 1.0))))
        else:
            raise ValueError(kind)

        for k in ("cfl_scale", "dfl_scale", "mu_mix", "R_mix", "Pr", "kappa"):
            if k in physics:
                set_param(k, physics[k])
        L.navier_stokes_initialize()
        check(L.fvh_set_visc_dt_mode(physics.get("visc_dt_mode", 0)))
        m = mesh_struct()
        names = [m.regions.contents.name[i].decode() for i in range(m.regions.contents.n_regions)]
        regions = dict(regions or {})
        for i, nm in enumerate(names):
            if nm not in regions:
                raise ValueError("no boundary condition for region %r" % nm)
            t, prim = regions[nm]
            prim5 = (ctypes.c_double * 5)(*(prim if prim is not None else (0, 0, 0, 0, 0)))
            L.boundary_set_region(i, BOUNDARY_TYPES.index(t), prim5)
        L.flux_initialize_scheme(physics.get("flux", "AUSM").encode())
        if ic is not None:
            pk, params = ic
            arr = (ctypes.c_double * max(1, len(params)))(*params)
            check(L.fvh_set_ic(pk, arr, len(params)))
            L.set_exact_function(ctypes.cast(L.fvh_exact_function_builtin, ctypes.c_void_p))
        L.init_finite_volume()
        L.init_reconstruction(LINEAR if numerics.get("reconstruction", "First-Order") == "Linear"
                              else FIRST_ORDER)
        L.init_limiter(BARTH_JESPERSENN if numerics.get("limiter", "None") == "Barth-Jespersenn"
                       else NO_LIMITER)
        L.init_analyze()
        L.init_output(numerics.get("i_output_data", 0))
        L.init_timedisc(EXPLICIT if numerics.get("timestep", "Explicit") == "Explicit" else IMPLICIT,
                        numerics.get("max_iter", 100), 1 if numerics.get("transient", True) else 0,
                        numerics.get("abort_residual", 1e-10), numerics.get("t_start", 0.0),
                        numerics.get("t_end", 1e300))
        if numerics.get("timestep", "Explicit") == "Explicit":
            L.init_explicit(SCHEMES[numerics.get("scheme", "RK3-3")])
        else:
            # TimeDisc/Implicit keys (src/FV3D/src/timedisc/implicit.c:36-56): scheme,
            # solver, max_iter_inner, tolerance_lsoe, max_iter_lsoe, krylov dims/restarts
            L.init_implicit({"Euler": EULER_IMPLICIT, "BDF2": BDF2}[numerics.get("scheme", "Euler")],
                            {"BiCGStab": 0, "GMRes": 1}[numerics.get("solver", "GMRes")],
                            numerics.get("max_iter_inner", 100),
                            numerics.get("tolerance_lsoe", 1e-12),
                            numerics.get("max_iter_lsoe", 100),
                            numerics.get("max_krylov_dims", 15),
                            numerics.get("max_krylov_restarts", 2))
        L.init_restart(0, b"")
        self.n_cells = m.cells.contents.n_domain_cells
        self.n_faces = m.faces.contents.n_faces
        self.n_boundaries = m.boundaries.contents.n_boundaries
        Case._active = self
        self.initial_state = self.state()

    # -- state ---------------------------------------------------------------
    def state(self):
        U = np.zeros(self.n_cells * 5)
        check(self.L.fvh_get_state(U.ctypes.data_as(P_DOUBLE)))
        return U.reshape(self.n_cells, 5)

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        check(self.L.fvh_set_state(U.ctypes.data_as(P_DOUBLE)))

    def phi_dt(self):
        R = np.zeros(self.n_cells * 5)
        check(self.L.fvh_get_phi_dt(R.ctypes.data_as(P_DOUBLE)))
        return R.reshape(self.n_cells, 5)

    def field(self, what):
        rows = 30 if what == 0 else 20
        out = np.zeros(self.n_cells * rows)
        check(self.L.fvh_get_field(what, out.ctypes.data_as(P_DOUBLE)))
        return out.reshape(self.n_cells, rows)

    def time_derivative(self, t=0.0):
        self.L.finite_volume_time_derivative(t)
        return self.phi_dt()

    def timestep(self):
        return self.L.calc_ns_timestep(0.0)

    def run(self):
        """timedisc(); returns (status, per-step log rows)."""
        rc = self.L.fvh_timedisc()
        it = ctypes.c_int(0)
        self.L.fvh_last_status(ctypes.byref(it))
        rows = np.zeros((max(1, it.value), 8))
        n = self.L.fvh_get_log(rows.ctypes.data_as(P_DOUBLE), it.value)
        return rc, rows[:n]

    def bench(self, steps, warmup):
        sec = ctypes.c_double(0)
        kus = (ctypes.c_double * 3)()
        check(self.L.fvh_bench_explicit(steps, warmup, ctypes.byref(sec), kus))
        return sec.value, list(kus)

    def jfnk(self, **options):
        """Steady Newton-Krylov solve of the current state (fvh_jfnk_solve).
        options override fvh_jfnk_default_options fields; precision,
        precond and coupling also take the names "fp64"/"fp32",
        "scalar"/"block"/"learned" and "none"/"dilu".  Returns (stats dict,
        history rows)."""
        o = JfnkOptions()
        check(self.L.fvh_jfnk_default_options(ctypes.byref(o)))
        names = _JFNK_NAMES
        for k, v in options.items():
            if not hasattr(o, k):
                raise KeyError("unknown JFNK option %r" % k)
            if k in names and isinstance(v, str):
                v = names[k][v]
            setattr(o, k, v)
        st = JfnkStats()
        check(self.L.fvh_jfnk_solve(ctypes.byref(o), ctypes.byref(st)))
        n = self.L.fvh_jfnk_history(None, 0)
        rows = np.zeros((n, 6))
        if n:
            self.L.fvh_jfnk_history(rows.ctypes.data_as(P_DOUBLE), n)
        stats = {f: getattr(st, f) for f, _ in JfnkStats._fields_}
        return stats, rows

    def jfnk_blocks(self, cfl, precond, dtau, minv):
        """Pseudo time steps and inverse preconditioner blocks of one Newton
        step at the current state (fvh_jfnk_blocks) into dtau [N] and
        minv [N,5,5] (float64, C-contiguous)."""
        kinds = {"scalar": 0, "block": 1}
        assert dtau.dtype == np.float64 and dtau.flags.c_contiguous
        assert minv.dtype == np.float64 and minv.flags.c_contiguous
        assert dtau.size == self.n_cells and minv.size == 25 * self.n_cells
        check(self.L.fvh_jfnk_blocks(cfl, kinds.get(precond, precond),
                                     dtau.ctypes.data_as(P_DOUBLE),
                                     minv.ctypes.data_as(P_DOUBLE)))

    def jfnk_precond_apply(self, cfl, v, precond="block", coupling="none"):
        """z = M^-1 v (fvh_jfnk_precond_apply) for the preconditioner a
        Newton step at the current state and pseudo CFL builds; v [N,5]."""
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        assert v.size == 5 * self.n_cells
        z = np.zeros_like(v)
        check(self.L.fvh_jfnk_precond_apply(
            cfl, _JFNK_NAMES["precond"].get(precond, precond),
            _JFNK_NAMES["coupling"].get(coupling, coupling),
            v.ctypes.data_as(P_DOUBLE), z.ctypes.data_as(P_DOUBLE)))
        return z.reshape(self.n_cells, 5)

    def host_phi_total(self):
        check(self.L.fvh_sync_host())
        d = ctypes.POINTER(data_t).in_dll(self.L, "solver_data").contents
        return _arr(d.phi_total, (self.n_cells + self.n_boundaries) * 10,
                    np.float64).reshape(-1, 10)

    def close(self):
        L = self.L
        L.free_restart()
        L.free_explicit()
        L.free_timedisc()
        L.free_output()
        L.free_analyze()
        L.free_limiter()
        L.free_reconstruction()
        L.free_finite_volume()
        L.free_equation()
        L.free_mesh()
        L.free_solver()
        if Case._active is self:
            Case._active = None
