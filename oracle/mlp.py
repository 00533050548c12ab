"""CPU restatement (numpy, FP64 accumulation) of the learned block
preconditioner's network, features and targets.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the device's
BF16-MFMA kernels, never by the product.

The reference has no learned component; the network is this build's
(28 -> 64 -> 64 -> 25, tanh).  The device rounds every MFMA operand to
bfloat16 (round to nearest even) and accumulates in FP32; `bf16=True`
emulates exactly those roundings, `bf16=False` gives the plain FP64 network
whose analytic gradient the tests check against finite differences.
Parameter layout: W1 [32][64], b1 [64], W2 [64][64], b2 [64], W3 [64][32],
b3 [32], row-major [in][out], one vector.
"""
import math

import numpy as np

K_IN, K_HID, K_OUT, K_BLK, K_FEAT = 32, 64, 32, 25, 28
SHAPES = [("W1", (K_IN, K_HID)), ("b1", (K_HID,)), ("W2", (K_HID, K_HID)), ("b2", (K_HID,)),
          ("W3", (K_HID, K_OUT)), ("b3", (K_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, off = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[off:off + n], dtype=np.float64).reshape(shp)
        off += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def _tanh(z, on):
    return np.tanh(z.astype(np.float32)).astype(np.float64) if on else np.tanh(z)


def forward(theta, X, bf16_ops=True):
    p = unpack(theta)
    X = np.asarray(X, dtype=np.float64)
    A1 = _tanh(_r(X, bf16_ops) @ _r(p["W1"], bf16_ops) + p["b1"], bf16_ops)
    A2 = _tanh(_r(A1, bf16_ops) @ _r(p["W2"], bf16_ops) + p["b2"], bf16_ops)
    Y = _r(A2, bf16_ops) @ _r(p["W3"], bf16_ops) + p["b3"]
    return Y, (X, A1, A2)


def loss_grad(theta, X, T, bf16_ops=True):
    """Loss sum_{c, j<25} (y - t)^2 / (n 25) and its gradient (vector)."""
    p = unpack(theta)
    n = X.shape[0]
    Y, (X, A1, A2) = forward(theta, X, bf16_ops)
    E = np.zeros_like(Y)
    E[:, :K_BLK] = Y[:, :K_BLK] - np.asarray(T, dtype=np.float64)[:, :K_BLK]
    loss = float((E ** 2).sum() / (n * K_BLK))
    D3 = 2.0 * E / (n * K_BLK)
    r = lambda a: _r(a, bf16_ops)
    g = {}
    g["W3"] = r(A2).T @ r(D3)
    g["b3"] = D3.sum(0)
    D2 = (r(D3) @ r(p["W3"]).T) * (1.0 - A2 * A2)
    g["W2"] = r(A1).T @ r(D2)
    g["b2"] = D2.sum(0)
    D1 = (r(D2) @ r(p["W2"]).T) * (1.0 - A1 * A1)
    g["W1"] = r(X).T @ r(D1)
    g["b1"] = D1.sum(0)
    return loss, pack(g)


def scales(rho, p, kappa):
    a = math.sqrt(kappa * p / rho)
    return np.array([1.0 / rho, 1.0 / (rho * a), 1.0 / (rho * a), 1.0 / (rho * a),
                     1.0 / (rho * a * a)])


def targets(Minv, dtau, cfl, rho, p, kappa):
    """S K S^-1 with K = (dtau Minv^-1 - I) / cfl, [N, 32] (25 used)."""
    s = scales(rho, p, kappa)
    M = np.linalg.inv(np.asarray(Minv).reshape(-1, 5, 5))
    K = (np.asarray(dtau)[:, None, None] * M - np.eye(5)[None]) / cfl
    T = s[None, :, None] * K / s[None, None, :]
    out = np.zeros((T.shape[0], K_OUT))
    out[:, :K_BLK] = T.reshape(-1, 25)
    return out


def blocks(out, dtau, cfl, rho, p, kappa):
    """Minv = dtau (I + cfl S^-1 out S)^-1, [N, 5, 5]."""
    s = scales(rho, p, kappa)
    K = np.asarray(out)[:, :K_BLK].reshape(-1, 5, 5) * s[None, None, :] / s[None, :, None]
    return np.asarray(dtau)[:, None, None] * np.linalg.inv(np.eye(5)[None] + cfl * K)


def features(mesh, phi_total, cfl, rho_r, p_r, kappa):
    """Features [N, 32] from the oracle's phi_total [(N+B), 10] (primitive
    columns refreshed, ghost rows filled) on the mesh arrays."""
    N, F = mesh["n_cells"], mesh["n_faces"]
    a_r = math.sqrt(kappa * p_r / rho_r)
    fc = np.asarray(mesh["face_cells"]).reshape(F, 2)
    n = np.asarray(mesh["face_n"]).reshape(F, 3)
    A = np.asarray(mesh["face_area"])
    V = np.asarray(mesh["cell_volume"])
    phi = np.asarray(phi_total).reshape(-1, 10)
    q = np.stack([phi[:, 0] / rho_r, phi[:, 5] / a_r, phi[:, 6] / a_r, phi[:, 7] / a_r,
                  phi[:, 8] / p_r], axis=1)
    qf = 0.5 * (q[fc[:, 0]] + q[fc[:, 1]])
    g = np.zeros((N, 5, 3))
    mom = np.zeros((N, 6))
    nbnd = np.zeros(N)
    outer = np.stack([n[:, 0] ** 2, n[:, 1] ** 2, n[:, 2] ** 2, n[:, 0] * n[:, 1],
                      n[:, 0] * n[:, 2], n[:, 1] * n[:, 2]], axis=1) * A[:, None]
    contrib = qf[:, :, None] * (n * A[:, None])[:, None, :]
    np.add.at(g, fc[:, 0], contrib)
    inner = fc[:, 1] < N
    np.add.at(g, fc[inner, 1], -contrib[inner])
    np.add.at(mom, fc[:, 0], outer)
    np.add.at(mom, fc[inner, 1], outer[inner])
    np.add.at(nbnd, fc[~inner, 0], 1.0)
    h = np.cbrt(V)
    c = np.sqrt(kappa * phi[:N, 8] / phi[:N, 0])
    X = np.zeros((N, K_IN))
    X[:, 0] = q[:N, 0] - 1.0
    X[:, 1:4] = phi[:N, 5:8] / c[:, None]
    X[:, 4] = q[:N, 4] - 1.0
    X[:, 5:20] = (g * (h / V)[:, None, None]).reshape(N, 15)
    X[:, 20:26] = mom / (h * h)[:, None]
    X[:, 26] = nbnd * 0.25
    X[:, 27] = np.sqrt((phi[:N, 5:8] ** 2).sum(1)) / c
    return X
