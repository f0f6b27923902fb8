"""Velocity damping on the GPU (BW:706-708's DEFAULT_DAMPING*V with its
Jacobian c*I in A), bit-exact against the FP32 oracle: the assembled system
and damped rollouts with both assembly kernels (the stencil form and the
per-end form of a shuffled spring list) and with the Jacobi preconditioner,
the single-cloth and batched adjoints, and the Baraff-Witkin demo's full
StepPhysics (damping, Jacobi-preconditioned CG to a relative tolerance,
ground clamp, collider, Provot slack 1.01)."""
import dataclasses

import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params, say

pytestmark = pytest.mark.gpu
C_DAMP = -0.125


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def damped_scene(shuffle=False, jacobi=False):
    sc = scenes.plain_scene()
    sp = sc.springs
    if shuffle:
        sp = np.ascontiguousarray(sp[np.random.default_rng(8).permutation(len(sp))])
    rng = np.random.default_rng(9)
    X = scenes.perturb(sc.X, 0.02, 4)
    V = (0.3 * rng.standard_normal(sc.X.shape)).astype(np.float32)
    V[sc.pinned.astype(bool)] = 0
    p = dataclasses.replace(sc.params, damping=C_DAMP, cg_tol=1e-6, cg_jacobi=jacobi)
    return sc, sp, X, V, p


@pytest.mark.parametrize("shuffle,jacobi", [(False, False), (True, False), (False, True)])
def test_damped_system_and_rollout_bitwise(o32, shuffle, jacobi):
    from opencloth_amd.cloth import ImplicitCloth

    sc, sp, X, V, p = damped_scene(shuffle, jacobi)
    say("damped", "shuffled" if shuffle else "stencil", "jacobi" if jacobi else "cg")
    cl = ImplicitCloth(sc.nx, sc.ny, X, sp, sc.pinned, p, V)
    s = o32.system(sc.n_vertices, sp, sc.pinned)
    prm = oracle_params(p, sc.nx, sc.ny)
    cl.assemble(p.dt)
    Cg, Bg = cl.system()
    Br, Cr = s.assemble(X, V, prm)
    a, b, _ = s.pairs()
    assert np.array_equal(bits(Bg), bits(Br))
    assert np.array_equal(bits(Cg), bits(Cr))
    F = cl.compute_forces()
    assert np.array_equal(bits(F), bits(s.forces_gather(X, V, p.mass, p.gravity, C_DAMP)[0]))
    cl.set_state(X, V)
    cl.step(20)
    Xr, Vr, _, _, it, st = s.rollout(X, V, prm, 20)
    Xg, Vg = cl.get_state()
    assert cl.last_status() == (int(np.ravel(it)[-1]), int(np.ravel(st)[-1]))
    assert np.array_equal(bits(Xg), bits(Xr)) and np.array_equal(bits(Vg), bits(Vr))
    # the damping acts: the undamped rollout differs
    X0r = s.rollout(X, V, oracle_params(dataclasses.replace(p, damping=0.0), sc.nx, sc.ny), 20)[0]
    assert not np.array_equal(X0r, Xr)


def test_bw_demo_step_physics(o32):
    """The Baraff-Witkin demo (scenes.bw_demo_scene): 1200 StepPhysics steps
    (the cloth drapes over the collider after about 800) compared bit for bit
    every 200."""
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.bw_demo_scene()
    ell, slack = sc.extra["ellipsoid"], sc.extra["provot_slack"]
    say("BW demo StepPhysics x1200")
    cl = ImplicitCloth.from_scene(sc)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(sc.params, sc.nx, sc.ny)
    X, V, DV = sc.X.copy(), sc.V.copy(), np.zeros_like(sc.X)
    for k in range(1200):
        s.step_physics(X, V, DV, prm, ell, slack)
        if k % 200 == 199:
            cl.step_physics(200, ell, slack)
            Xg, Vg = cl.get_state()
            assert np.array_equal(bits(Xg), bits(X)), k
            assert np.array_equal(bits(Vg), bits(V)), k
    assert cl.sticky_status() & 6 == 0
    assert X[:, 1].min() < 1.0  # fell past the collider's top


def test_damped_adjoint_bitwise(o32):
    """Single-cloth rollout adjoint with damping: gradients bit-exact against
    the FP32 oracle's adjoint in the kernels' reduction order."""
    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.diff_scene(n=24, steps=10, seed=1)
    p = dataclasses.replace(sc.params, damping=-0.4)
    V0 = (0.2 * np.random.default_rng(3).standard_normal(sc.X.shape)).astype(np.float32)
    V0[sc.pinned.astype(bool)] = 0
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, V0)
    cl.rollout_forward(10, 4)
    XT = cl.X
    gx = (2.0 * (XT - sc.extra["target"])).astype(np.float32)
    gks, gx0, gv0 = cl.rollout_backward(gx)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(p, sc.nx, sc.ny)
    XTr, _, states = s.rollout_forward(sc.X, V0, prm, 10)
    assert np.array_equal(bits(XT), bits(XTr))
    rks, rx0, rv0 = s.rollout_grad(sc.X, V0, prm, 10, gx, states=states)
    assert np.array_equal(bits(gks), bits(rks))
    assert np.array_equal(bits(gv0), bits(rv0)) and np.array_equal(bits(gx0), bits(rx0))


def test_damped_batch_and_batch_adjoint_bitwise(o32):
    """A batch of damped cloths: forward steps and the batched adjoint, each
    cloth bit-exact against the FP32 oracle in the batched order."""
    from oracle import make_params

    from opencloth_amd.batch import ClothBatch

    sc = scenes.batch_diff_scene(n_cloths=5, sides=(16, 30), steps=6, seed=4)
    sc.params = dataclasses.replace(sc.params, damping=-0.3)
    b = ClothBatch.from_scene(sc)
    b.set_schedule(256, 4)
    T = b.threads_per_cloth()
    b.rollout_forward(6, 3)
    XT, _ = b.get_state()
    gx = (2.0 * (XT - sc.extra["target"])).astype(np.float32)
    gks, gx0, gv0 = b.rollout_backward(gx, n_springs=len(sc.springs))
    for c in range(sc.n_cloths):
        cs = sc.cloth(c)
        v0, s0, n = int(sc.voff[c]), int(sc.soff[c]), cs.n_vertices
        s = o32.system(n, cs.springs, cs.pinned)
        prm = make_params(sc.params, dot_mode=2, tx=int(T[c]))
        XTr, _, states = s.rollout_forward(cs.X, cs.V, prm, 6)
        assert np.array_equal(bits(XT[v0:v0 + n]), bits(XTr)), c
        rks, rx0, rv0 = s.rollout_grad(cs.X, cs.V, prm, 6, gx[v0:v0 + n], states=states)
        assert np.array_equal(bits(gks[s0:s0 + len(cs.springs)]), bits(rks)), c
        assert np.array_equal(bits(gv0[v0:v0 + n]), bits(rv0)), c
        assert np.array_equal(bits(gx0[v0:v0 + n]), bits(rx0)), c
