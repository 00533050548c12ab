"""Train the learned block preconditioner on the solver's own data.

Training set: bump-channel solves (the JFNK benchmark's geometry at several
resolutions, the benchmark's own among them) at Mach numbers drawn uniformly
from [0.3, 0.7] with jitter seeds 1..n_solves.  Each
solve starts from the free stream and runs block-Jacobi JFNK in short
bursts; after each burst the features and the block-Jacobi targets of the
current state are appended to the device training set (fvh_nn_collect).
The network is then trained on the GPU (BF16-MFMA forward/backward, Adam)
and checked on held-out seeds.  The parameters are saved as a float32 .npy
(fv3d_amd/data/precond_mlp.npy by default), which learned.load_pretrained()
reads.

    python -m fv3d_amd.train_precond --solves 200 --out fv3d_amd/data/precond_mlp.npy
"""
import argparse
import math
import os
import sys
import time

import numpy as np

from . import configs
from . import learned
from . import solver as S


def channel_case(n, seed, mach):
    cfg = configs.bump_channel_steady(nx=n[0], ny=n[1], nz=n[2], mach=mach, seed=seed)
    return S.Case(**cfg)


def collect_solve(n, seed, mach, bursts, burst_steps):
    case = channel_case(n, seed, mach)
    total = learned.collect(10.0)
    for _ in range(bursts):
        st, _ = case.jfnk(max_newton=burst_steps, tol=1e-8, precond="block")
        total = learned.collect(10.0)
        if st["status"] == 0:
            break
    case.close()
    return total


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--solves", type=int, default=200)
    ap.add_argument("--mesh", default="16,4,3;32,8,3;64,16,6;160,40,13",
                    help="channel meshes (hexahedra nx,ny,nz; ';'-separated), used in turn "
                         "by the solves in the proportions of --mesh-weights")
    ap.add_argument("--mesh-weights", default="10,6,3,1")
    ap.add_argument("--bursts", type=int, default=6)
    ap.add_argument("--burst-steps", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--holdout", type=int, default=4)
    ap.add_argument("--out", default=learned.PRETRAINED)
    a = ap.parse_args(argv)
    meshes = [tuple(int(x) for x in m.split(",")) for m in a.mesh.split(";")]
    weights = [int(w) for w in a.mesh_weights.split(",")][:len(meshes)]
    cycle = [m for m, w in zip(meshes, weights) for _ in range(w)]
    n = meshes[0]
    rng = np.random.default_rng(a.seed)
    machs = rng.uniform(0.3, 0.7, a.solves + a.holdout)
    t0 = time.time()
    # a case must exist before the first collect (it owns the device state)
    first = channel_case(n, 1, float(machs[0]))
    learned.init(a.seed)
    learned.clear()
    first.close()
    total = 0
    for i in range(a.solves):
        total = collect_solve(cycle[i % len(cycle)], i + 1, float(machs[i]), a.bursts,
                              a.burst_steps)
        if (i + 1) % 20 == 0 or i + 1 == a.solves:
            print("solves %d  samples %d  %.1f s" % (i + 1, total, time.time() - t0), flush=True)
    t1 = time.time()
    case = channel_case(n, 1, 0.5)
    losses = learned.train(a.epochs, a.batch, a.lr, a.seed)
    print("training: %d epochs, loss %.3e -> %.3e, %.1f s" % (a.epochs, losses[0], losses[-1],
                                                             time.time() - t1), flush=True)
    case.close()
    # held-out check: mean/max relative Frobenius error of the inverse blocks
    errs = []
    for j in range(a.holdout):
        seed = 10000 + j
        case = channel_case(meshes[j % len(meshes)], seed, float(machs[a.solves + j]))
        case.jfnk(max_newton=10, tol=1e-8, precond="block")
        for cfl in (10.0, 1e3):
            errs.append(learned.validate(case, cfl))
        case.close()
    errs = np.array(errs)
    print("held-out block error: mean %.3f  max %.3f" % (errs[:, 1].mean(), errs[:, 0].max()),
          flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, learned.get_params())
    print("saved", a.out, flush=True)


if __name__ == "__main__":
    sys.exit(main())
