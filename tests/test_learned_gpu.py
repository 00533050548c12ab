"""GPU tests of the learned block preconditioner (fvh_nn_*), through the
C-ABI, against the CPU oracle (oracle/mlp.py):

- inference: the BF16-MFMA forward pass equals the oracle's network with
  the same bfloat16 operand roundings (FP32 accumulation order is the only
  difference): |dy| <= 2e-3 max|y|;
- training: the loss and its gradient from the forward + backward kernels
  equal the oracle's (|dg| <= 1e-2 max|g|, loss to 1e-4);
- features of a non-uniform channel state equal the oracle's features
  computed from the oracle's own phi_total (FP32 storage: 1e-5);
- trained on the solver's own block-Jacobi blocks, the network's blocks
  precondition GMRES: the learned JFNK converges to the block-Jacobi JFNK
  solution, needs far fewer Krylov steps than the scalar preconditioner
  and at most 1.5x those of block Jacobi; a network that cannot predict
  the blocks (random weights) is rejected and the solve falls back;
- what the trained network gains: with the shipped network, the coupled
  preconditioner (multicolour block ILU(0)) built on the predicted blocks
  reaches the 1e-8 stop norm in fewer Newton steps and fewer GMRES
  iterations than the same preconditioner on the finite-difference blocks,
  on a mid-size channel that is not one of its training meshes
  (80x20x8 hexahedra x 6; measured 33 against 37 Newton steps, 0.45 against
  0.52 s), and the two solutions agree.
"""
import numpy as np
import pytest

from fv3d_amd import configs
from fv3d_amd import learned
from oracle import mlp
from tests.helpers import make_pair

pytestmark = pytest.mark.gpu


def _print(tag, **kv):
    print(tag, " ".join("%s=%s" % (k, v) for k, v in kv.items()), flush=True)


def _random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = mlp.unpack(rng.standard_normal(mlp.N_PARAMS) * scale)
    p["W1"][mlp.K_FEAT:, :] = 0.0
    p["W3"][:, mlp.K_BLK:] = 0.0
    p["b3"][mlp.K_BLK:] = 0.0
    return mlp.pack(p).astype(np.float32)


@pytest.fixture(scope="module")
def channel():
    cfg = configs.bump_channel_steady(nx=16, ny=4, nz=3)
    case, o, mesh = make_pair(cfg)
    yield cfg, case, o, mesh
    case.close()


def test_forward_matches_oracle(channel):
    _, case, _, _ = channel
    th = _random_params(3)
    rng = np.random.default_rng(4)
    X = (rng.standard_normal((101, mlp.K_IN)) * 0.7).astype(np.float32)
    X[:, mlp.K_FEAT:] = 0.0
    learned.set_params(th)
    Y = learned.forward(X)
    Yo, _ = mlp.forward(th.astype(np.float64), X.astype(np.float64), bf16_ops=True)
    e = float(np.abs(Y - Yo).max() / np.abs(Yo).max())
    _print("mlp forward", rel=e)
    assert e <= 2e-3


def test_gradient_matches_oracle(channel):
    _, case, _, _ = channel
    th = _random_params(5)
    rng = np.random.default_rng(6)
    n = 333
    X = (rng.standard_normal((n, mlp.K_IN)) * 0.7).astype(np.float32)
    X[:, mlp.K_FEAT:] = 0.0
    T = (rng.standard_normal((n, mlp.K_OUT)) * 0.3).astype(np.float32)
    T[:, mlp.K_BLK:] = 0.0
    learned.set_params(th)
    loss, g = learned.loss_grad(X, T)
    lo, go = mlp.loss_grad(th.astype(np.float64), X.astype(np.float64), T.astype(np.float64),
                           bf16_ops=True)
    e = float(np.abs(g - go).max() / np.abs(go).max())
    _print("mlp gradient", rel=e, loss=loss, loss_oracle=lo)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert e <= 1e-2


def test_features_match_oracle(channel):
    cfg, case, o, mesh = channel
    rng = np.random.default_rng(7)
    U = case.state()
    U[:, 0] *= 1.0 + 0.05 * rng.standard_normal(len(U))
    U[:, 1:4] += 0.05 * rng.standard_normal((len(U), 3))
    case.set_state(U)
    o.set_state(U.reshape(-1))
    o.time_derivative()
    F = learned.features(case, 30.0)
    rho_r, p_r = 1.0, 1.0
    Fo = mlp.features(mesh, o.phi_total(), 30.0, rho_r, p_r, cfg["physics"]["kappa"])
    e = float(np.abs(F - Fo).max() / np.abs(Fo).max())
    _print("features", rel=e)
    assert e <= 1e-5


def test_trained_network_preconditions(channel):
    cfg, case, _, _ = channel
    case.set_state(case.initial_state)
    n = learned.collect_along_solve(case, cfls=(10.0,), newton_steps=range(0, 40, 3))
    learned.init(11)
    err0 = learned.validate(case, 100.0)
    losses = learned.train(epochs=80, batch=1024, lr=3e-3, seed=1)
    err1 = learned.validate(case, 100.0)
    _print("training", samples=n, loss_first=losses[0], loss_last=losses[-1],
           err_before=err0, err_after=err1)
    assert losses[-1] < 0.2 * losses[0]
    assert err1[1] < 0.5 * err0[1]
    case.set_state(case.initial_state)
    stb, _ = case.jfnk(tol=1e-8, precond="block")
    Ub = case.state()
    case.set_state(case.initial_state)
    stl, _ = case.jfnk(tol=1e-8, precond="learned", learned_tol=0.5)
    Ul = case.state()
    case.set_state(case.initial_state)
    sts, _ = case.jfnk(tol=1e-8, precond="scalar", max_newton=400)
    _print("learned vs block vs scalar", gmres_learned=stl["gmres_iterations"],
           gmres_block=stb["gmres_iterations"], gmres_scalar=sts["gmres_iterations"],
           fallback=stl["precond_fallbacks"], learned_err=stl["learned_error"],
           setup_learned=stl["precond_seconds"], setup_block=stb["precond_seconds"],
           s_learned=stl["seconds"], s_block=stb["seconds"])
    assert stl["status"] == 0 and stl["precond_fallbacks"] == 0
    assert np.linalg.norm(Ul - Ub) / np.linalg.norm(Ub) <= 1e-6
    assert stl["gmres_iterations"] <= 1.5 * stb["gmres_iterations"]
    assert stl["gmres_iterations"] < sts["gmres_iterations"]
    # an untrained network is rejected: the solve falls back to block Jacobi
    learned.init(12)
    case.set_state(case.initial_state)
    str_, _ = case.jfnk(tol=1e-8, precond="learned", learned_tol=0.5)
    assert str_["precond_fallbacks"] == 1 and str_["status"] == 0


def test_learned_blocks_gain_in_the_coupled_preconditioner():
    cfg = configs.bump_channel_steady(nx=80, ny=20, nz=8)
    case, _, _ = make_pair(cfg)
    learned.load_pretrained()
    U0 = case.state()
    coupled = dict(coupling="dilu", krylov_dims=40, max_restarts=2, precision="fp32",
                   tol=1e-8, max_newton=400)
    stf, _ = case.jfnk(precond="block", **coupled)
    Uf = case.state()
    case.set_state(U0)
    stl, _ = case.jfnk(precond="learned", **coupled)
    Ul = case.state()
    _print("learned vs FD blocks in DILU", newton_learned=stl["newton_steps"],
           newton_fd=stf["newton_steps"], gmres_learned=stl["gmres_iterations"],
           gmres_fd=stf["gmres_iterations"], s_learned=stl["seconds"], s_fd=stf["seconds"],
           fallback=stl["precond_fallbacks"], learned_err=stl["learned_error"])
    assert stf["status"] == 0 and stl["status"] == 0 and stl["precond_fallbacks"] == 0
    assert stl["newton_steps"] < stf["newton_steps"]
    assert stl["gmres_iterations"] < stf["gmres_iterations"]
    assert np.linalg.norm(Ul - Uf) / np.linalg.norm(Uf) <= 1e-7
    case.close()
