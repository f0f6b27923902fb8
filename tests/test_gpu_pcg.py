"""GPU parity of the Jacobi-preconditioned CG (OC_FLAG_JACOBI, Baraff and
Witkin's modified PCG, the intent of OpenCloth_Implicit/.../main.cpp:308-330)
through the C ABI: single-cloth steps (per-iteration launches and the
persistent launch), the single-cloth adjoint, the batched step and the
batched adjoint, each bit-exact against the FP32 oracle with the same
preconditioner and reduction order; and the iteration counts of the batched
workload with and without the preconditioner (reported)."""
import dataclasses

import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def jac(p):
    return dataclasses.replace(p, cg_jacobi=True)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("scene,persistent", [("plain", False), ("plain", True), ("large", False),
                                              ("large", True)])
def test_step_bitwise(o32, scene, persistent):
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene() if scene == "plain" else scenes.large_scene()
    p = dataclasses.replace(jac(sc.params), cg_persistent=persistent)
    say("jacobi step", scene, "persistent", persistent)
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, sc.V)
    nsteps = 3
    cl.step(nsteps)
    Xg, Vg = cl.get_state()
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    X, V, _, _, iters, st = s.rollout(sc.X, sc.V, oracle_params(p, sc.nx, sc.ny), nsteps)
    assert cl.last_status() == (int(np.ravel(iters)[-1]), int(np.ravel(st)[-1]))
    assert np.array_equal(bits(Xg), bits(X)) and np.array_equal(bits(Vg), bits(V))


def test_adjoint_bitwise(o32):
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.diff_scene(n=24, steps=10, seed=1)
    sc.params = jac(sc.params)
    rng = np.random.default_rng(2)
    ks = (20.0 * np.exp(np.log(10.0) * rng.random(len(sc.springs)))).astype(np.float32)
    say("jacobi adjoint 24x24 x10")
    cl = ImplicitCloth.from_scene(sc)
    cl.set_spring_stiffness(ks)
    cl.rollout_forward(10, 4)
    XT = cl.X
    gx = (2.0 * (XT - sc.extra["target"])).astype(np.float32)
    gks, gx0, gv0 = cl.rollout_backward(gx)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    s.set_ks(ks.astype(np.float64))
    prm = oracle_params(sc.params, sc.nx, sc.ny)
    XTr, _, states = s.rollout_forward(sc.X, sc.V, prm, 10)
    assert np.array_equal(bits(XT), bits(XTr))
    rks, rx0, rv0 = s.rollout_grad(sc.X, sc.V, prm, 10, gx, states=states)
    assert np.array_equal(bits(gks), bits(rks)) and np.array_equal(bits(gv0), bits(rv0))
    assert np.array_equal(bits(gx0), bits(rx0))


@pytest.mark.parametrize("sched", [(256, 4), (1024, 1000)])
def test_batch_bitwise(o32, sched):
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sc = scenes.batch_scene(n_cloths=10, sides=(16, 40), seed=3, steps=3)
    sc.params = jac(sc.params)
    say("jacobi batch", sched)
    b = ClothBatch.from_scene(sc)
    b.set_schedule(*sched)
    T = b.threads_per_cloth()
    b.step(3)
    X, _ = b.get_state()
    it, st = b.status()
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        s = o32.system(cs.n_vertices, cs.springs, cs.pinned)
        Xr, _, _, _, iters, str_ = s.rollout(cs.X, cs.V, make_params(cs.params, dot_mode=2, tx=int(T[c])), 3)
        assert it[c] == iters[-1] and st[c] == str_, c
        v0 = sc.voff[c]
        assert np.array_equal(bits(X[v0:v0 + cs.n_vertices]), bits(Xr)), c


def test_batch_adjoint_bitwise(o32):
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sc = scenes.batch_diff_scene(n_cloths=5, sides=(16, 30), steps=6, seed=2)
    sc.params = jac(sc.params)
    say("jacobi batch adjoint")
    b = ClothBatch.from_scene(sc)
    T = b.threads_per_cloth()
    b.rollout_forward(6, 3)
    XT, _ = b.get_state()
    gx = (2.0 * (XT - sc.extra["target"])).astype(np.float32)
    gks, gx0, gv0 = b.rollout_backward(gx, n_springs=len(sc.springs))
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        v0, s0, n = int(sc.voff[c]), int(sc.soff[c]), cs.n_vertices
        s = o32.system(n, cs.springs, cs.pinned)
        prm = make_params(cs.params, dot_mode=2, tx=int(T[c]))
        _, _, states = s.rollout_forward(cs.X, cs.V, prm, 6)
        rks, rx0, rv0 = s.rollout_grad(cs.X, cs.V, prm, 6, gx[v0:v0 + n], states=states)
        assert np.array_equal(bits(gks[s0:s0 + len(cs.springs)]), bits(rks)), c
        assert np.array_equal(bits(gv0[v0:v0 + n]), bits(rv0)), c
        assert np.array_equal(bits(gx0[v0:v0 + n]), bits(rx0)), c


def test_batch_iterations_with_and_without_jacobi():
    """The batched workload's cloths (relative tolerance 1e-4, at most 40
    iterations) on 64 cloths: iterations per cloth with and without the
    preconditioner, reported; the preconditioned solve converges at least
    as often."""
    from opencloth_amd.batch import ClothBatch

    sc = scenes.batch_scene(n_cloths=64, seed=7, steps=2)
    res = {}
    for j in (False, True):
        p = dataclasses.replace(sc.params, cg_jacobi=j)
        b = ClothBatch(sc.nx, sc.ny, sc.X, sc.springs, sc.n_springs, sc.pinned, p, sc.V)
        b.step(2)
        it, st = b.status()
        res[j] = (it.mean(), int((st == 0).sum()))
    say("mean iterations plain %.2f (%d converged), Jacobi %.2f (%d converged)"
        % (res[False][0], res[False][1], res[True][0], res[True][1]))
    assert res[True][1] >= res[False][1]
