"""Loader of libfv3d_hip.so with the C-ABI signatures of include/*.h."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_int, c_double, c_char_p, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p
P_DOUBLE = ctypes.POINTER(ctypes.c_double)
P_INT = ctypes.POINTER(ctypes.c_int)
P_FLOAT = ctypes.POINTER(ctypes.c_float)



class JfnkOptions(ctypes.Structure):
    """fvh_jfnk_options_t (include/fv3d_hip.h)."""
    _fields_ = [("max_newton", c_int), ("tol", c_double), ("cfl0", c_double),
                ("cfl_growth", c_double), ("cfl_cut", c_double), ("cfl_stall", c_double),
                ("cfl_stall_cut", c_double), ("cfl_max", c_double),
                ("krylov_dims", c_int), ("max_restarts", c_int), ("rtol", c_double),
                ("fd_eps", c_double), ("fd_eps32", c_double), ("max_change", c_double),
                ("precision", c_int),
                ("precond", c_int), ("verbose", c_int), ("learned_check", c_int),
                ("learned_tol", c_double), ("coupling", c_int),
                ("reorth", c_double)]


class JfnkStats(ctypes.Structure):
    """fvh_jfnk_stats_t (include/fv3d_hip.h)."""
    _fields_ = [("status", c_int), ("newton_steps", c_int), ("gmres_iterations", c_int),
                ("residual_evals", c_int), ("seconds", c_double), ("norm0", c_double),
                ("norm", c_double), ("l2_drop", c_double), ("precond_fallbacks", c_int),
                ("learned_error", c_double), ("precond_seconds", c_double)]


EXACT_FT = ctypes.CFUNCTYPE(None, c_int, c_double, P_DOUBLE, P_DOUBLE)
VOID_T_FT = ctypes.CFUNCTYPE(None, c_double)
DOUBLE_T_FT = ctypes.CFUNCTYPE(c_double, c_double)

# (name, restype, argtypes) of every function declared in include/*.h
SIGNATURES = [
    # fv3d_solver.h — Solver API (reference names)
    ("add_sol_variable", c_int, [c_char_p]),
    ("add_dep_variable", c_int, [c_char_p]),
    ("calc_global_residual", None, [c_double]),
    ("create_file_header", None, [c_char_p]),
    ("finite_volume_time_derivative", None, [c_double]),
    ("free_analyze", None, []),
    ("free_equation", None, []),
    ("free_explicit", None, []),
    ("free_finite_volume", None, []),
    ("free_implicit", None, []),
    ("free_limiter", None, []),
    ("free_mesh", None, []),
    ("free_reconstruction", None, []),
    ("free_output", None, []),
    ("free_restart", None, []),
    ("free_solver", None, []),
    ("free_timedisc", None, []),
    ("get_simulation_title", c_char_p, []),
    ("init_analyze", None, []),
    ("init_equation", None, []),
    ("init_explicit", None, [c_int]),
    ("init_finite_volume", None, []),
    ("init_implicit", None, [c_int, c_int, c_int, c_double, c_int, c_int, c_int]),
    ("init_limiter", None, [c_int]),
    ("init_mesh", None, [c_char_p]),
    ("init_reconstruction", None, [c_int]),
    ("init_output", None, [c_int]),
    ("init_restart", None, [c_int, c_char_p]),
    ("init_solver", None, [c_char_p]),
    ("init_timedisc", None, [c_int, c_int, c_int, c_double, c_double, c_double]),
    ("is_explicit", c_int, []),
    ("print_mesh_info", None, []),
    ("print_variables", None, []),
    ("read_restart_data", None, [c_char_p]),
    ("set_calc_flux", None, [c_void_p]),
    ("set_calc_timestep", None, [c_void_p]),
    ("set_exact_function", None, [c_void_p]),
    ("set_update", None, [c_void_p]),
    ("set_update_gradients", None, [c_void_p]),
    ("set_viscous_dt", None, [c_int]),
    ("timedisc", None, []),
    ("write_output", None, [c_int, c_double]),
    ("navier_stokes_initialize", None, []),
    ("boundary_set_region", None, [c_int, c_int, P_DOUBLE]),
    ("flux_initialize_scheme", None, [c_char_p]),
    ("calc_ns_flux", None, [c_double]),
    ("calc_ns_timestep", c_double, [c_double]),
    ("calc_ns_exact_function", None, [c_int, c_double, P_DOUBLE, P_DOUBLE]),
    ("ns_update", None, [c_double]),
    ("ns_update_gradients", None, [c_double]),
    # fv3d_hip.h — extension API
    ("fvh_last_error", c_char_p, []),
    ("fvh_mesh_box", c_int, [c_int, c_int, c_int, c_double, c_double, c_double,
                             c_double, c_double, c_double, c_int, c_int, c_int]),
    ("fvh_mesh_channel", c_int, [c_int, c_int, c_int, c_double, ctypes.c_ulonglong]),
    ("fvh_mesh_write", c_int, [c_char_p]),
    ("fvh_get_face_shift", c_int, [P_DOUBLE]),
    ("fvh_set_visc_dt_mode", c_int, [c_int]),
    ("fvh_set_ic", c_int, [c_int, P_DOUBLE, c_int]),
    ("fvh_time_derivative_noalloc", c_int, [c_void_p, c_void_p, c_void_p]),
    ("fvh_exact_function_builtin", None, [c_int, c_double, P_DOUBLE, P_DOUBLE]),
    ("fvh_get_state", c_int, [P_DOUBLE]),
    ("fvh_set_state", c_int, [P_DOUBLE]),
    ("fvh_get_phi_dt", c_int, [P_DOUBLE]),
    ("fvh_get_field", c_int, [c_int, P_DOUBLE]),
    ("fvh_sync_host", c_int, []),
    ("fvh_set_verbose", c_int, [c_int]),
    ("fvh_set_sync_interval", c_int, [c_int]),
    ("fvh_timedisc", c_int, []),
    ("fvh_get_log", c_int, [P_DOUBLE, c_int]),
    ("fvh_last_status", c_int, [P_INT]),
    ("fvh_bench_explicit", c_int, [c_int, c_int, P_DOUBLE, P_DOUBLE]),
    ("fvh_device_info", c_int, [c_char_p, c_int, P_INT, P_DOUBLE]),
    ("fvh_jfnk_default_options", c_int, [ctypes.POINTER(JfnkOptions)]),
    ("fvh_jfnk_solve", c_int, [ctypes.POINTER(JfnkOptions), ctypes.POINTER(JfnkStats)]),
    ("fvh_jfnk_history", c_int, [P_DOUBLE, c_int]),
    ("fvh_jfnk_blocks", c_int, [c_double, c_int, P_DOUBLE, P_DOUBLE]),
    ("fvh_jfnk_precond_apply", c_int, [c_double, c_int, c_int, P_DOUBLE, P_DOUBLE]),
    ("fvh_hdf5_available", c_int, []),
    ("fvh_set_output_format", c_int, [c_int]),
    ("fvh_mesh_gmsh", c_int, [c_char_p, c_double, c_double, c_double]),
    ("fvh_mesh_batch", c_int, [c_int, P_INT, c_double, c_int]),
    ("fvh_batch_setup", c_int, [c_int, P_INT, P_INT, P_DOUBLE]),
    ("fvh_batch_run", c_int, [c_int, c_int, P_DOUBLE]),
    ("fvh_batch_info", c_int, [P_INT, P_INT, c_int]),
    ("fvh_batch_case_scalars", c_int, [P_DOUBLE, P_DOUBLE, P_DOUBLE]),
    ("fvh_batch_set_schedule", c_int, [c_int, c_int, c_int, c_int]),
    ("fvh_batch_case_step_times", c_int, [c_int, P_DOUBLE]),
    ("fvh_nn_param_count", c_int, []),
    ("fvh_nn_init", c_int, [ctypes.c_ulonglong]),
    ("fvh_nn_get_params", c_int, [P_FLOAT]),
    ("fvh_nn_set_params", c_int, [P_FLOAT]),
    ("fvh_nn_features", c_int, [c_double, P_FLOAT]),
    ("fvh_nn_collect", c_int, [c_double, ctypes.POINTER(ctypes.c_longlong)]),
    ("fvh_nn_clear", c_int, []),
    ("fvh_nn_train", c_int, [c_int, c_int, c_double, ctypes.c_ulonglong, P_DOUBLE]),
    ("fvh_nn_forward", c_int, [P_FLOAT, c_int, P_FLOAT]),
    ("fvh_nn_loss_grad", c_int, [P_FLOAT, P_FLOAT, c_int, P_FLOAT, P_DOUBLE]),
    ("fvh_nn_validate", c_int, [c_double, P_DOUBLE]),
]

# exported data symbols (globals) of fv3d_solver.h
DATA_SYMBOLS = [
    "solver_mesh", "solver_variables", "solver_data",
    "cfl_scale", "dfl_scale", "mu_mix", "R_mix", "Pr", "kappa",
    "kappa_m1", "kappa_p1", "s_kappa", "s_kappa_m1", "s_kappa_p1", "cv", "cp",
    "kappa_pr", "lambda",
    "ic_rho", "ic_rho_u", "ic_rho_v", "ic_rho_w", "ic_rho_e",
    "ip_u", "ip_v", "ip_w", "ip_p", "ip_T",
]


def lib_path():
    """The in-tree build, or an alternate build named by FV3D_HIP_LIB."""
    return os.environ.get("FV3D_HIP_LIB") or os.path.join(_HERE, "libfv3d_hip.so")


def lib():
    """Load libfv3d_hip.so (once).  Raises if the library was not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(
            "libfv3d_hip.so is not built (%s); run python -c "
            "'import __graft_entry__; __graft_entry__.build()'" % path)
    L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    for name, res, args in SIGNATURES:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L
