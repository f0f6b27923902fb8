"""Writes tests/golden/adjoint_fd_128.json: FP64 central finite differences
of L = |x_T - target|^2 on the differentiable-rollout workload (128x128 cloth,
120 steps, dt 1/60, 40 CG iterations, per-spring ks log-uniform in [20,200])
along 32 random directions in (ks, v0), with eps = 1e-3 (the workload's
stated step; its O(eps^2) truncation reaches ~1e-2 relative on the smallest
projections) and eps = 1e-5 (truncation below 1e-5), ks perturbed relatively
and v0 absolutely, computed with the FP64 oracle on a CPU process pool.

    python tests/golden/make_adjoint_fd.py
"""
import json
import os
import sys
from multiprocessing import get_context

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

N, STEPS, SEED, KS_SEED, DIR_SEED, NDIR, EPSS, EVERY = 128, 120, 0, 7, 11, 32, (1e-3, 1e-5), 16


def setup():
    from oracle import Oracle, make_params
    from opencloth_amd import scenes
    from tests.test_gpu_adjoint import random_ks

    sc = scenes.diff_scene(n=N, steps=STEPS, seed=SEED)
    ks = random_ks(sc, KS_SEED).astype(np.float64)
    s = Oracle(64).system(sc.n_vertices, sc.springs, sc.pinned)
    return sc, ks, s, make_params(sc.params)


def directions(sc, ks):
    rng = np.random.default_rng(DIR_SEED)
    out = []
    for _ in range(NDIR):
        dk = rng.standard_normal(len(ks)) * ks
        dv = rng.standard_normal(sc.X.shape)
        out.append((dk, dv))
    return out


def loss_at(args):
    k, sign, EPS = args
    sc, ks, s, prm = setup()
    dk, dv = directions(sc, ks)[k]
    s.set_ks(ks + sign * EPS * dk)
    X, _, _, _, _, st = s.rollout(sc.X, sign * EPS * dv, prm, STEPS)
    assert st == 0
    return float(((X - sc.extra["target"].astype(np.float64)) ** 2).sum())


def main():
    sc, ks, s, prm = setup()
    s.set_ks(ks)
    XT, _, states = s.rollout_forward(sc.X, np.zeros_like(sc.X), prm, STEPS)
    gks, _, gv0 = s.rollout_grad(sc.X, np.zeros_like(sc.X), prm, STEPS,
                                 2 * (XT - sc.extra["target"]), states=states)
    dirs = directions(sc, ks)
    fd = {}
    with get_context("spawn").Pool(min(8, os.cpu_count() or 1)) as pool:
        for eps in EPSS:
            vals = pool.map(loss_at, [(k, sg, eps) for k in range(NDIR) for sg in (1, -1)])
            fd[repr(eps)] = [(vals[2 * k] - vals[2 * k + 1]) / (2 * eps) for k in range(NDIR)]
    adj = [float(gks @ dk + (gv0 * dv).sum()) for dk, dv in dirs]
    out = {"n": N, "steps": STEPS, "seed": SEED, "ks_seed": KS_SEED, "dir_seed": DIR_SEED,
           "checkpoint_every": EVERY, "loss": float(((XT - sc.extra["target"]) ** 2).sum()),
           "fd": fd, "adjoint_fp64": adj}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_fd_128.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    for eps, v in fd.items():
        rel = max(abs(a - b) / abs(b) for a, b in zip(adj, v))
        print(f"eps {eps}: worst FP64 adjoint vs FD relative difference {rel:.3e}")
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
