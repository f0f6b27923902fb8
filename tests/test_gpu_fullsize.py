"""Parity at the benchmarked sizes.

- The 512x512 bench workload (top edge pinned, 30 fixed CG iterations,
  dt 1/120): 100 GPU steps bit-exact against the FP32 oracle with the
  kernels' reduction order, and within 1e-5 (norm-wise relative) of the FP64
  oracle.
- The 4096-cloth batched workload (sides U[16,128], 25M vertices): after
  its configured 200 steps, eight sampled cloths (the largest, the smallest
  and six random ones) bit-exact against the FP32 oracle with the batched
  kernel's order, and every cloth finite with status 0 in every step
  (sticky status).

The oracle runs its OpenMP loops on the host's cores here; its results do
not depend on the thread count."""
import os

import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params, rel_pos_err, say

pytestmark = pytest.mark.gpu


def host_threads():
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() else min(n, 16)


@pytest.fixture(scope="module")
def oracles():
    from oracle import Oracle

    o32, o64 = Oracle(32), Oracle(64)
    o32.set_threads(host_threads())
    o64.set_threads(host_threads())
    yield o32, o64
    o32.set_threads(1)
    o64.set_threads(1)


def test_bench_workload_100_steps(oracles):
    from oracle import make_params

    from opencloth_amd.cloth import ImplicitCloth

    o32, o64 = oracles
    sc = scenes.large_scene()
    say("512x512 bench workload x100, oracle threads", host_threads())
    cl = ImplicitCloth.from_scene(sc)
    cl.step(100)
    Xg, Vg = cl.get_state()
    it, st = cl.last_status()
    s32 = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    X, V, _, _, iters, st_r = s32.rollout(sc.X, sc.V, oracle_params(sc.params, sc.nx, sc.ny), 100)
    assert (it, st) == (int(iters[-1]), st_r) == (30, 0)
    assert np.array_equal(Xg.view(np.uint32), X.view(np.uint32))
    assert np.array_equal(Vg.view(np.uint32), V.view(np.uint32))
    s64 = o64.system(sc.n_vertices, sc.springs, sc.pinned)
    X64 = s64.rollout(sc.X, sc.V, make_params(sc.params, nx=sc.nx, ny=sc.ny, dot_mode=0), 100)[0]
    err = rel_pos_err(Xg, X64)
    say("vs FP64 after 100 steps:", err)
    assert err <= 1e-5


def test_batch_workload_sampled_cloths(oracles):
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    o32, _ = oracles
    sc = scenes.batch_scene(n_cloths=4096)
    say("4096-cloth batch,", sc.X.shape[0], "vertices")
    b = ClothBatch.from_scene(sc)
    b.set_schedule(1024, 1)
    T = b.threads_per_cloth()
    nsteps = sc.steps  # the configured 200 steps
    b.step(nsteps)
    X, V = b.get_state()
    it, st = b.status()
    assert np.isfinite(X).all() and np.isfinite(V).all()
    assert not st.any()
    assert not b.sticky_status().any()  # no step of any cloth broke down or failed to converge
    n = sc.nx.astype(np.int64) * sc.ny
    rng = np.random.default_rng(1)
    sample = sorted({int(np.argmax(n)), int(np.argmin(n)), *rng.choice(sc.n_cloths, 6, replace=False).tolist()})
    for c in sample:
        cs = sc.cloth(c)
        say(f"cloth {c} ({cs.nx}x{cs.ny}): FP32 oracle, {nsteps} steps")
        s = o32.system(cs.n_vertices, cs.springs, cs.pinned)
        prm = make_params(cs.params, dot_mode=2, tx=int(T[c]))
        Xr, _, _, _, iters, st_r = s.rollout(cs.X, cs.V, prm, nsteps)
        Xg = X[sc.voff[c]:sc.voff[c] + cs.n_vertices]
        assert it[c] == iters[-1] and st[c] == st_r, c
        assert np.array_equal(Xg.view(np.uint32), Xr.view(np.uint32)), (c, cs.nx, cs.ny)
    b.close()
