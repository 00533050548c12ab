"""Run a case from the reference's JSON parameter file.

The keys are those of the reference application's parameter file (the
basec parameter tree of src/FV3D/src/*: General/title, Mesh/mesh_file,
Equation/Navier-Stokes/{cfl_scale, dfl_scale, mu_mix, R_mix, Pr, kappa,
Flux/flux_scheme, Boundary/<region>}, FV/Reconstruction/reconstruction,
FV/Limiter/limiter, Output/i_output_data, Restart/{use_restart,
restart_file}, TimeDisc/{timestep, max_iter, transient, abort_residual,
t_start, t_end, Explicit/scheme, Implicit/{scheme, max_iter, solver,
tolerance_lsoe, max_iter_lsoe, max_krylov_dims, max_krylov_restarts}}), as
in misc/gmsh/naca0012/naca0012_simulation.json.  Values are read from the
"value" entries; boundary regions hold type and u, v, w, p, T.  The mesh
file is HDF5 (.h5, the reference layout), the raw container, or a Gmsh 4.1
ASCII file (.msh, preprocessed on load with unit scale).

    python -m fv3d_amd.driver case.json
"""
import argparse
import json
import os
import sys

from . import solver as S


def _v(node, *path, default=None):
    for k in path:
        if not isinstance(node, dict) or k not in node:
            return default
        node = node[k]
    if isinstance(node, dict) and "value" in node:
        return node["value"]
    return node


def config_from_json(params, base_dir="."):
    """Case keyword arguments (fv3d_amd.solver.Case) from the parameter tree."""
    ns = _v(params, "Equation", "Navier-Stokes", default={}) or {}
    physics = {}
    for k in ("cfl_scale", "dfl_scale", "mu_mix", "R_mix", "Pr", "kappa"):
        val = _v(ns, k)
        if val is not None:
            physics[k] = float(val)
    physics["flux"] = _v(ns, "Flux", "flux_scheme", default="AUSM")
    regions = {}
    for name, bc in (_v(ns, "Boundary", default={}) or {}).items():
        t = bc.get("type", "flow")
        prim = None
        if any(k in bc for k in ("u", "v", "w", "p", "T")):
            prim = tuple(float(bc.get(k, 0.0)) for k in ("u", "v", "w", "p", "T"))
        regions[name] = (t, prim)
    timestep = _v(params, "TimeDisc", "timestep", default="Explicit")
    numerics = dict(
        reconstruction=_v(params, "FV", "Reconstruction", "reconstruction", default="First-Order"),
        limiter=_v(params, "FV", "Limiter", "limiter", default="None"),
        i_output_data=int(_v(params, "Output", "i_output_data", default=0)),
        timestep=timestep,
        max_iter=int(_v(params, "TimeDisc", "max_iter", default=100)),
        transient=bool(_v(params, "TimeDisc", "transient", default=True)),
        abort_residual=float(_v(params, "TimeDisc", "abort_residual", default=1e-10)),
        t_start=float(_v(params, "TimeDisc", "t_start", default=0.0)),
        t_end=float(_v(params, "TimeDisc", "t_end", default=1e300)),
    )
    if timestep == "Explicit":
        numerics["scheme"] = _v(params, "TimeDisc", "Explicit", "scheme", default="RK3-3")
    else:
        imp = _v(params, "TimeDisc", "Implicit", default={}) or {}
        scheme = _v(imp, "scheme", default="Euler")
        numerics.update(
            scheme="BDF2" if scheme in ("BDF-2", "BDF2") else "Euler",
            solver=_v(imp, "solver", default="GMRes"),
            max_iter_inner=int(_v(imp, "max_iter", default=100)),
            tolerance_lsoe=float(_v(imp, "tolerance_lsoe", default=1e-12)),
            max_iter_lsoe=int(_v(imp, "max_iter_lsoe", default=100)),
            max_krylov_dims=int(_v(imp, "max_krylov_dims", default=15)),
            max_krylov_restarts=int(_v(imp, "max_krylov_restarts", default=2)))
    mesh_file = _v(params, "Mesh", "mesh_file")
    if mesh_file is None:
        raise KeyError("Mesh/mesh_file is required")
    if not os.path.isabs(mesh_file):
        mesh_file = os.path.join(base_dir, mesh_file)
    restart = None
    if _v(params, "Restart", "use_restart", default=False):
        restart = _v(params, "Restart", "restart_file")
    mesh = ("gmsh", dict(path=mesh_file)) if mesh_file.endswith(".msh") else ("file", mesh_file)
    return dict(mesh=mesh, physics=physics, regions=regions, numerics=numerics,
                title=_v(params, "General", "title", default="untitled")), restart


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run an fv3d case from its JSON parameter file.")
    ap.add_argument("parameter_file")
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    with open(a.parameter_file) as f:
        params = json.load(f)
    cfg, restart = config_from_json(params, os.path.dirname(os.path.abspath(a.parameter_file)))
    case = S.Case(**cfg, verbose=not a.quiet)
    if restart:
        case.L.init_restart(1, restart.encode())
    rc, log = case.run()
    case.close()
    return 0 if rc == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
