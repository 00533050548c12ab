"""State files of write_output (src/Solver/src/output/output.c:20-191) and
restart from them (src/Solver/src/restart/restart.c:50-178), through the
C-ABI.

- Layout: the attributes n_tot_variables, n_sol_variables, iter, t (and
  n_stages in implicit mode) and the datasets tot_variables, sol_variables,
  phi_total [N x 10], phi_dt [N x 5] (and phi_old:<i> [N x 5]), checked with
  the independent h5dump tool.
- Contents: phi_total is the reference's array at output time, whose
  primitive columns come from the last RK stage's input (the update hook
  ran on that state); the oracle keeps the same array, and the file's rows
  equal it.
- Restart: an explicit run restarted from its step-3 file repeats the
  uninterrupted run bit for bit; a BDF2 run restarts with its history
  (phi_old) and matches to the Krylov solve's tolerance (its first solve after
  the restart starts from zero instead of the previous step's solution, in
  the reference as here).

libhdf5 1.10 ships with this image (/opt/conda); the HDF5 cases fail rather
than skip when that library is present but does not load.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from fv3d_amd import configs
from fv3d_amd import solver as S
from fv3d_amd._lib import lib
from tests.helpers import make_pair, n_bit_diff, rel_err

pytestmark = pytest.mark.gpu
CONDA_HDF5 = "/opt/conda/lib/libhdf5.so"


def _hdf5():
    L = lib()
    ok = bool(L.fvh_hdf5_available())
    if os.path.exists(CONDA_HDF5):
        assert ok, "libhdf5 is installed at %s but does not load" % CONDA_HDF5
    dump = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                       if os.path.exists("/opt/conda/bin/h5dump") else None)
    return ok, dump


def _h5_dataset(dump, path, name, shape):
    out = path + "." + name.replace(":", "_") + ".bin"
    subprocess.run([dump, "-d", "/" + name, "-b", "LE", "-o", out, path], check=True,
                   capture_output=True)
    return np.fromfile(out, dtype="<f8").reshape(shape)


def _channel_cfg(steps, scheme="RK3-3", timestep="Explicit", i_output=3):
    cfg = configs.bump_channel(nx=8, ny=3, nz=2, mu=1e-3)
    cfg["numerics"].update(max_iter=steps, i_output_data=i_output, timestep=timestep,
                           scheme=scheme)
    return cfg


@pytest.mark.parametrize("fmt", ["h5", "raw"])
def test_explicit_state_file_and_bitwise_restart(tmp_path, fmt):
    ok, dump = _hdf5()
    if fmt == "h5" and not ok:
        pytest.skip("libhdf5 is not installed in this image")
    L = lib()
    cwd = os.getcwd()
    os.chdir(str(tmp_path))
    try:
        L.fvh_set_output_format(2 if fmt == "h5" else 1)
        cfg = _channel_cfg(6)
        case, o, mesh = make_pair(cfg)
        N = case.n_cells
        rc, log = case.run()
        U6 = case.state()
        assert rc == 0 and len(log) == 6
        ext = ".h5" if fmt == "h5" else ".fvs"
        f3 = os.path.abspath("fv3d.%09d%s" % (3, ext))
        assert os.path.exists(f3) and os.path.exists("fv3d.%09d%s" % (6, ext))
        case.close()
        # the oracle's phi_total after 3 steps: conservative columns of the
        # new state, primitives of the third step's last stage input
        n, _ = o.timedisc("RK3-3", 3, True, 1e-10, 0.0, 1e300)
        assert n == 3
        phi_o = o.phi_total()[:N]
        dt_o = o.get(1).reshape(-1, 5)[:N]
        if fmt == "h5" and dump:
            head = subprocess.run([dump, "-H", f3], capture_output=True, text=True,
                                  check=True).stdout
            for a in ("n_tot_variables", "n_sol_variables", "iter", "t"):
                assert 'ATTRIBUTE "%s"' % a in head, a
            for d in ("tot_variables", "sol_variables", "phi_total", "phi_dt"):
                assert 'DATASET "%s"' % d in head, d
            assert "( %d, 10 )" % N in head and "( %d, 5 )" % N in head
            assert "H5T_STRING" in head
            assert "n_stages" not in head
            phi_f = _h5_dataset(dump, f3, "phi_total", (N, 10))
            dt_f = _h5_dataset(dump, f3, "phi_dt", (N, 5))
            print("file phi_total vs oracle rel", rel_err(phi_f, phi_o), "bitdiff",
                  n_bit_diff(phi_f, phi_o), flush=True)
            assert rel_err(phi_f, phi_o) <= 1e-12
            assert rel_err(dt_f, dt_o) <= 1e-12
        # restart from step 3 and run to 6: the same bits as the
        # uninterrupted run
        cfg2 = _channel_cfg(3, i_output=0)
        case = S.Case(**cfg2)
        L.init_restart(1, f3.encode())
        rc2, log2 = case.run()
        assert rc2 == 0 and len(log2) == 3
        assert log2[0, 0] == pytest.approx(log[3, 0], rel=0, abs=0)
        assert np.array_equal(log2[:, 1], log[3:, 1])
        assert n_bit_diff(case.state(), U6) == 0
        case.close()
    finally:
        L.fvh_set_output_format(0)
        os.chdir(cwd)


def test_implicit_state_file_history_and_restart(tmp_path):
    ok, dump = _hdf5()
    L = lib()
    cwd = os.getcwd()
    os.chdir(str(tmp_path))
    try:
        L.fvh_set_output_format(2 if ok else 1)
        ext = ".h5" if ok else ".fvs"
        cfg = configs.density_wave_box(n=6, steps=6, flux="Rusanov")
        cfg["physics"]["cfl_scale"] = 2.0
        cfg["numerics"].update(timestep="Implicit", scheme="BDF2", max_iter=6,
                               i_output_data=3)
        case = S.Case(**cfg)
        rc, log = case.run()
        U6 = case.state()
        assert rc == 0
        f3 = os.path.abspath("fv3d.%09d%s" % (3, ext))
        case.close()
        if ok and dump:
            head = subprocess.run([dump, "-H", f3], capture_output=True, text=True,
                                  check=True).stdout
            for d in ("phi_old:0", "phi_old:1"):
                assert 'DATASET "%s"' % d in head, d
            assert 'ATTRIBUTE "n_stages"' in head
        cfg2 = dict(cfg)
        cfg2["numerics"] = dict(cfg["numerics"], max_iter=3, i_output_data=0)
        case = S.Case(**cfg2)
        L.init_restart(1, f3.encode())
        rc2, log2 = case.run()
        assert rc2 == 0 and len(log2) == 3
        e = rel_err(case.state(), U6)
        print("implicit restart rel", e, flush=True)
        assert e <= 1e-10
        case.close()
    finally:
        L.fvh_set_output_format(0)
        os.chdir(cwd)
