"""The single-reduction CG on the GPU (OC_FLAG_CG_SINGLE_REDUCTION: one launch
per iteration, the Chronopoulos-Gear recurrence), bit-exact against the FP32
oracle's restatement of it (cg_core_cg1) with the kernels' reduction order:
the CG solution (cold and warm start, plain and Jacobi), the 21x21
reference rollout with its stopping test, the 512x512 workload (fixed 30
iterations), the fused multi-step path, a rollout adjoint, and the
breakdown flag; and within 1e-4 of the standard recurrence on the 21x21
rollout (300 steps)."""
import dataclasses

import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params, rel_pos_err, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("jacobi", [False, True])
def test_cg1_rollout_21x21(o32, jacobi):
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene()
    p = dataclasses.replace(sc.params, cg_single_reduction=True, cg_jacobi=jacobi)
    say("cg1 21x21 x300 jacobi", jacobi)
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, sc.V)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(p, sc.nx, sc.ny)
    X, V, DV = sc.X.copy(), sc.V.copy(), np.zeros_like(sc.X)
    _, _, _, h0, _, _ = s.rollout(sc.X, sc.V, oracle_params(dataclasses.replace(sc.params, cg_jacobi=jacobi),
                                                           sc.nx, sc.ny), 300, history=True)
    worst = 0.0
    for k in range(300):
        it, st = s.step(X, V, DV, prm)
        cl.step(1)
        assert cl.last_status() == (it, st), k
        if k % 50 == 49:
            Xg, Vg = cl.get_state()
            assert np.array_equal(bits(Xg), bits(X)) and np.array_equal(bits(Vg), bits(V)), k
        worst = max(worst, rel_pos_err(X, h0[k]))
    say("single-reduction vs standard recurrence: %.3e" % worst)
    assert worst <= 1e-4


@pytest.mark.parametrize("warm", [False, True])
def test_cg1_solution_bitwise(o32, warm):
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene()
    rng = np.random.default_rng(4)
    X = scenes.perturb(sc.X, 0.05, 3)
    V = (0.3 * rng.standard_normal(X.shape)).astype(np.float32)
    V[sc.pinned.astype(bool)] = 0
    p = scenes.StepParams(dt=1 / 60, cg_max_iters=25, cg_tol=1e-6, cg_tol_mode=scenes.TOL_ABSOLUTE,
                          warm_start=warm, cg_single_reduction=True)
    cl = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, p, V)
    dv0 = (1e-2 * rng.standard_normal(X.shape)).astype(np.float32)
    dv0[sc.pinned.astype(bool)] = 0
    cl.set_dv(dv0)
    cl.assemble()
    it, st = cl.solve_conjugate_gradient()
    dv = cl.get_dv()
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(p, sc.nx, sc.ny)
    B, Cp = s.assemble(X, V, prm)
    dvr, itr, str_, _ = s.cg(Cp, prm, B, dv0 if warm else None)
    assert (it, st) == (itr, str_)
    assert np.array_equal(bits(dv), bits(dvr))


def test_cg1_large_fused_steps(o32):
    """512x512, 30 fixed iterations, three steps in one call (the fused
    path): bit-exact against the oracle."""
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.large_scene()
    p = dataclasses.replace(sc.params, cg_single_reduction=True)
    say("cg1 512x512 x3")
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, sc.V)
    cl.step(3)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    X, V, _, _, it, st = s.rollout(sc.X, sc.V, oracle_params(p, sc.nx, sc.ny), 3)
    Xg, Vg = cl.get_state()
    assert cl.last_status() == (int(np.ravel(it)[-1]), int(np.ravel(st)[-1]))
    assert np.array_equal(bits(Xg), bits(X)) and np.array_equal(bits(Vg), bits(V))


def test_cg1_adjoint_bitwise(o32):
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.diff_scene(n=24, steps=8, seed=2)
    p = dataclasses.replace(sc.params, cg_single_reduction=True)
    V0 = (0.2 * np.random.default_rng(5).standard_normal(sc.X.shape)).astype(np.float32)
    V0[sc.pinned.astype(bool)] = 0
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, V0)
    cl.rollout_forward(8, 3)
    XT = cl.X
    gx = (2.0 * (XT - sc.extra["target"])).astype(np.float32)
    gks, gx0, gv0 = cl.rollout_backward(gx)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(p, sc.nx, sc.ny)
    XTr, _, states = s.rollout_forward(sc.X, V0, prm, 8)
    assert np.array_equal(bits(XT), bits(XTr))
    rks, rx0, rv0 = s.rollout_grad(sc.X, V0, prm, 8, gx, states=states)
    assert np.array_equal(bits(gks), bits(rks))
    assert np.array_equal(bits(gv0), bits(rv0)) and np.array_equal(bits(gx0), bits(rx0))


def test_cg1_breakdown_flag(o32):
    """The compressed stiff cloth of the edge tests: the single-reduction
    recurrence meets a negative denominator at the oracle's iteration."""
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene()
    X = (sc.X * 0.4).astype(np.float32)
    springs = sc.springs.copy()
    springs["Ks"] = 5000.0
    p = dataclasses.replace(sc.params, dt=0.1, cg_max_iters=40, cg_tol=1e-9, cg_single_reduction=True)
    cl = ImplicitCloth(sc.nx, sc.ny, X, springs, sc.pinned, p, np.zeros_like(X))
    cl.step(1)
    s = o32.system(sc.n_vertices, springs, sc.pinned)
    Xr, Vr, DV = X.copy(), np.zeros_like(X), np.zeros_like(X)
    it, st = s.step(Xr, Vr, DV, oracle_params(p, sc.nx, sc.ny))
    assert st & 2 and cl.last_status() == (it, st)
    Xg, _ = cl.get_state()
    assert np.array_equal(bits(Xg), bits(Xr))
