"""GPU test of the allocation-free residual fvh_time_derivative_noalloc
(include/fv3d_hip.h) through the C-ABI: its dU/dt equals
finite_volume_time_derivative's bit for bit, on two calls in a row (the
second on a smaller mesh, so the shared scratch array holds the first call's
larger contents), with boundaries and without, and a mesh too large for the
scratch array is rejected with a message.  The caller's device memory comes
from torch."""
import ctypes

import numpy as np
import pytest

from fv3d_amd import configs
from fv3d_amd import solver as S
from tests.helpers import n_bit_diff

pytestmark = pytest.mark.gpu


def _noalloc(case):
    import torch
    U = torch.tensor(case.state(), dtype=torch.float64, device="cuda:0")
    R = torch.full(U.shape, float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = case.L.fvh_time_derivative_noalloc(ctypes.c_void_p(U.data_ptr()),
                                            ctypes.c_void_p(R.data_ptr()), None)
    assert rc == 0, case.L.fvh_last_error().decode()
    torch.cuda.synchronize()
    return R.cpu().numpy()


@pytest.mark.parametrize("cfgs", [
    (configs.taylor_green(n=8, steps=1, tile=2), configs.taylor_green(n=4, steps=1, tile=0)),
    (configs.bump_channel(nx=12, ny=4, nz=3, mu=1e-3), configs.density_wave_box(n=6, steps=1)),
], ids=["periodic", "walls-then-first-order"])
def test_noalloc_equals_time_derivative_two_calls(cfgs):
    for i, cfg in enumerate(cfgs):
        case = S.Case(**cfg)
        R_free = _noalloc(case)
        R_ref = case.time_derivative()
        nb = n_bit_diff(R_free, R_ref)
        print("call", i, "cells", case.n_cells, "bitdiff", nb, flush=True)
        assert np.isfinite(R_free).all()
        assert nb == 0
        case.close()


def test_noalloc_rejects_a_mesh_beyond_its_scratch():
    import torch
    case = S.Case(**configs.taylor_green(n=34, steps=1, tile=0))  # 235,824 tetrahedra
    # valid device arrays of the right size, so that nothing but the size
    # check stands between the call and its kernels
    U = torch.tensor(case.state(), dtype=torch.float64, device="cuda:0")
    R = torch.empty_like(U)
    torch.cuda.synchronize()
    rc = case.L.fvh_time_derivative_noalloc(ctypes.c_void_p(U.data_ptr()),
                                            ctypes.c_void_p(R.data_ptr()), None)
    torch.cuda.synchronize()
    assert rc != 0
    assert "scratch" in case.L.fvh_last_error().decode()
    case.close()
