"""Python side of the learned block preconditioner (fvh_nn_* of
include/fv3d_hip.h): parameters, features, training-set collection along
the solver's own Newton path, GPU training and validation.

The network and its kernels live in libfv3d_hip.so; this module only moves
arrays and drives the steps.  Pretrained parameters ship as
fv3d_amd/data/precond_mlp.npy (see `load_pretrained`).
"""
import ctypes
import os

import numpy as np

from ._lib import lib, P_DOUBLE, P_FLOAT
from .solver import check

K_IN, K_OUT = 32, 32
PRETRAINED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "precond_mlp.npy")


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


def param_count():
    return lib().fvh_nn_param_count()


def init(seed):
    check(lib().fvh_nn_init(seed))


def set_params(theta):
    th = _f32(theta, (param_count(),))
    check(lib().fvh_nn_set_params(th.ctypes.data_as(P_FLOAT)))


def get_params():
    th = np.zeros(param_count(), dtype=np.float32)
    check(lib().fvh_nn_get_params(th.ctypes.data_as(P_FLOAT)))
    return th


def load_pretrained(path=PRETRAINED):
    set_params(np.load(path, allow_pickle=False))


def features(case, cfl):
    F = np.zeros((case.n_cells, K_IN), dtype=np.float32)
    check(lib().fvh_nn_features(cfl, F.ctypes.data_as(P_FLOAT)))
    return F


def forward(X):
    X = _f32(X)
    n = X.shape[0]
    Y = np.zeros((n, K_OUT), dtype=np.float32)
    check(lib().fvh_nn_forward(X.ctypes.data_as(P_FLOAT), n, Y.ctypes.data_as(P_FLOAT)))
    return Y


def loss_grad(X, T):
    X, T = _f32(X), _f32(T)
    g = np.zeros(param_count(), dtype=np.float32)
    loss = ctypes.c_double(0)
    check(lib().fvh_nn_loss_grad(X.ctypes.data_as(P_FLOAT), T.ctypes.data_as(P_FLOAT),
                                 X.shape[0], g.ctypes.data_as(P_FLOAT), ctypes.byref(loss)))
    return loss.value, g


def clear():
    check(lib().fvh_nn_clear())


def collect(cfl):
    n = ctypes.c_longlong(0)
    check(lib().fvh_nn_collect(cfl, ctypes.byref(n)))
    return n.value


def collect_along_solve(case, cfls, newton_steps, **jfnk_options):
    """Adds the features and block-Jacobi targets of the states the JFNK
    solve passes through after each of `newton_steps` steps (block-Jacobi
    preconditioner), each at every CFL of `cfls`.  Restores the state."""
    U0 = case.state()
    n = 0
    for k in sorted(newton_steps):
        case.set_state(U0)
        if k > 0:
            case.jfnk(max_newton=k, tol=0.0, precond="block", **jfnk_options)
        for cfl in cfls:
            n = collect(cfl)
    case.set_state(U0)
    return n


def train(epochs, batch, lr, seed=0):
    loss = np.zeros(epochs)
    check(lib().fvh_nn_train(epochs, batch, lr, seed, loss.ctypes.data_as(P_DOUBLE)))
    return loss


def validate(case, cfl):
    """(max, mean) per-cell relative Frobenius error of the network's blocks
    against the block-Jacobi blocks at the case's current state."""
    err = np.zeros(2)
    check(lib().fvh_nn_validate(cfl, err.ctypes.data_as(P_DOUBLE)))
    return float(err[0]), float(err[1])
