"""CPU checks that pin the oracle's restatement of the learned
preconditioner's network and features (no GPU):

- the analytic FP64 gradient of the loss equals central finite differences
  (a transposed weight, a missing tanh' factor or a wrong loss scale fails);
- bfloat16 rounding: known answers, including ties to even;
- features of a uniform state on a periodic hexahedral box: zero gradients,
  face moments sum_f A n n^T / h^2 = 2 I exactly (closed form for a cube),
  Mach components u/c;
- targets of I/dtau blocks (no Jacobian) are zero.
"""
import math

import numpy as np

from fv3d_amd import configs
from oracle import mlp
from tests.test_oracle import _oracle_on, _prim_state


def _params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    th = rng.standard_normal(mlp.N_PARAMS) * scale
    p = mlp.unpack(th)
    p["W1"][mlp.K_FEAT:, :] = 0.0
    p["W3"][:, mlp.K_BLK:] = 0.0
    p["b3"][mlp.K_BLK:] = 0.0
    return mlp.pack(p)


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    th = _params(2)
    X = rng.standard_normal((7, mlp.K_IN)) * 0.5
    X[:, mlp.K_FEAT:] = 0.0
    T = rng.standard_normal((7, mlp.K_OUT)) * 0.3
    loss, g = mlp.loss_grad(th, X, T, bf16_ops=False)
    idx = list(rng.choice(mlp.N_PARAMS, 60, replace=False))
    idx += [0, 2048 + 3, 2112 + 5, 6208 + 1, 6272 + 24, 8320 + 24]
    for i in idx:
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mlp.loss_grad(tp, X, T, False)[0] - mlp.loss_grad(tm, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_known_answers():
    vals = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -7, -2.5,
                     3.0e38, 1 + 2.0 ** -8 + 2.0 ** -20])
    want = np.array([1.0, 1.0, 1 + 2.0 ** -6, 1 + 2.0 ** -7, -2.5, 2.9941e38,
                     1 + 2.0 ** -7])
    got = mlp.bf16(vals)
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2.0 ** -8
    assert got[6] == want[6]


def test_uniform_box_features_closed_form():
    cfg = configs.density_wave_box(n=4, steps=1)
    o, mesh = _oracle_on(cfg)
    N = mesh["n_cells"]
    rho, u, v, w, p = 1.3, 0.4, -0.2, 0.1, 0.8
    U = np.tile(_prim_state(rho, u, v, w, p), (N, 1))
    o.set_state(U.reshape(-1))
    o.time_derivative()
    X = mlp.features(mesh, o.phi_total(), 50.0, 1.0, 1.0, 1.4)
    c = math.sqrt(1.4 * p / rho)
    assert np.allclose(X[:, 0], rho - 1.0)
    assert np.allclose(X[:, 1:4], np.array([u, v, w]) / c)
    assert np.allclose(X[:, 4], p - 1.0)
    assert np.abs(X[:, 5:20]).max() < 1e-12
    assert np.allclose(X[:, 20:23], 2.0, rtol=1e-12)
    assert np.abs(X[:, 23:26]).max() < 1e-12
    assert np.all(X[:, 26] == 0.0)
    assert np.allclose(X[:, 27], math.sqrt(u * u + v * v + w * w) / c)
    assert np.all(X[:, 28:] == 0.0)


def test_scalar_blocks_give_zero_targets():
    dtau = np.array([0.1, 2.0, 3e-3])
    Minv = np.eye(5)[None] * dtau[:, None, None]
    T = mlp.targets(Minv.reshape(-1), dtau, 30.0, 1.2, 0.7, 1.4)
    assert np.abs(T).max() < 1e-15


def test_targets_and_blocks_are_inverse():
    rng = np.random.default_rng(8)
    n, cfl = 6, 75.0
    dtau = rng.uniform(0.1, 2.0, n)
    K = rng.standard_normal((n, 5, 5)) * 0.3 + np.eye(5)[None]
    Minv = dtau[:, None, None] * np.linalg.inv(np.eye(5)[None] + cfl * K)
    T = mlp.targets(Minv.reshape(-1), dtau, cfl, 1.3, 0.9, 1.4)
    back = mlp.blocks(T, dtau, cfl, 1.3, 0.9, 1.4)
    assert np.allclose(back, Minv, rtol=1e-10, atol=1e-14)
