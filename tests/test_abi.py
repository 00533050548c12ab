"""The C-ABI boundary: libfv3d_hip.so loads and exports every function and
data symbol declared in include/*.h (no compute calls; CPU only)."""
import ctypes
import os
import re

from fv3d_amd import _lib
from fv3d_amd._lib import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    funcs, data = set(), set()
    for h in ("fv3d_solver.h", "fv3d_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text)
        text = re.sub(r"typedef\s+(struct|enum)[^}]*}[^;]*;", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{]*)\)\s*;", text):
            name = m.group(1)
            if name not in ("if", "while", "sizeof"):
                funcs.add(name)
        for m in re.finditer(r"extern\s+[A-Za-z_][A-Za-z0-9_ \*]*?\b([A-Za-z_][A-Za-z0-9_,\s]*);", text):
            for nm in m.group(1).split(","):
                nm = nm.strip()
                if nm and nm != '"C"':
                    data.add(nm)
    return funcs, data


def test_header_parse_found_the_api():
    funcs, data = _declared()
    assert "finite_volume_time_derivative" in funcs and "timedisc" in funcs
    assert "fvh_mesh_box" in funcs
    assert "solver_mesh" in data and "ip_T" in data


def test_library_exports_every_declared_symbol():
    L = lib()
    funcs, data = _declared()
    missing = [f for f in sorted(funcs | data) if not hasattr(L, f)]
    assert missing == []


def test_python_bindings_cover_the_headers():
    funcs, data = _declared()
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert funcs <= bound, sorted(funcs - bound)
    assert data <= set(_lib.DATA_SYMBOLS), sorted(data - set(_lib.DATA_SYMBOLS))


def test_reference_defaults_of_parameter_globals():
    L = lib()
    # navier_stokes.c:16-24
    assert ctypes.c_double.in_dll(L, "kappa").value == 1.4
    assert ctypes.c_double.in_dll(L, "Pr").value == 0.718
    assert ctypes.c_double.in_dll(L, "R_mix").value == 8.31446261815324 / 28.96e-3


def test_extension_errors_are_status_codes():
    L = lib()
    assert L.fvh_set_visc_dt_mode(7) == -1
    assert L.fvh_set_sync_interval(0) == -1
    assert L.fvh_mesh_box(2, 2, 2, 0, 0, 0, 1, 1, 1, 0, 7, 0) != 0  # periodic needs n >= 3
    assert b"periodic" in L.fvh_last_error()
