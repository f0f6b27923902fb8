"""Velocity damping in the oracle (BW:706-708 adds DEFAULT_DAMPING*V[i] to
every force; OC params `damping`), pinned without restating its formulas:
FP64 central differences of the damped force function give the system
matrix A = M - h J_v - h^2 J_x (J_v now holds c*I) and the right-hand side
b = h (F + h J_x V); the gather and scatter forces agree bit for bit; the
adjoint with damping matches finite differences of FP64 rollouts."""
import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params
from tests.test_oracle_physics import G, dense_A, fd_jacobian, small_scene

C_DAMP = -0.125


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


@pytest.fixture(scope="module")
def o64():
    from oracle import Oracle

    return Oracle(64)


def test_damped_gather_equals_scatter(o32):
    X, V, springs, pinned = small_scene(9, seed=11)
    s = o32.system(len(X), springs, pinned)
    Fs = s.forces_scatter(X, V, 1.0, G, C_DAMP)
    Fg, _ = s.forces_gather(X, V, 1.0, G, C_DAMP)
    assert np.array_equal(Fs.view(np.uint32), Fg.view(np.uint32))
    F0 = s.forces_scatter(X, V, 1.0, G)
    assert not np.array_equal(Fs, F0)
    assert np.array_equal(Fs[pinned == 1], F0[pinned == 1])  # pins get no force


def test_damping_force_is_c_times_v(o64):
    """Without springs the force is m g + c v on free vertices, 0 on pins."""
    X, V, _, pinned = small_scene(5, seed=12)
    s = o64.system(len(X), np.zeros(0, scenes.SPRING_DTYPE), pinned)
    F = s.forces_scatter(X, V, 2.0, G, C_DAMP)
    free = pinned == 0
    want = 2.0 * np.asarray(G) + C_DAMP * V.astype(np.float64)
    assert np.abs(F[free] - want[free]).max() < 1e-15
    assert not F[~free].any()


def test_damped_system_matrix_and_rhs(o64):
    """A = M - h J_v - h^2 J_x and b = h (F + h J_x V), J_v and J_x central
    differences of the damped forces; the matvec and the Jacobi diagonal use
    the same A."""
    X, V, springs, _ = small_scene(6, seed=13, pins=False)
    N = len(X)
    p = scenes.StepParams(dt=1 / 30, mass=1.3, damping=C_DAMP)
    prm = oracle_params(p, 6, 6, gpu_order=False)
    s = o64.system(N, springs, None)
    B, Cp, diag = s.assemble(X, V, prm, with_diag=True)
    a, b, _ = s.pairs()
    h = float(prm.h)
    md = p.mass - h * C_DAMP
    A = dense_A(N, a, b, Cp, md)
    Z = np.zeros((N, 3))
    Jx = fd_jacobian(lambda x: s.forces_scatter(x, Z, p.mass, G, C_DAMP), X, 1e-6)
    Jv = fd_jacobian(lambda v: s.forces_scatter(X, v, p.mass, G, C_DAMP), V, 1e-3)
    assert np.abs(np.diag(Jv) - np.diag(fd_jacobian(lambda v: s.forces_scatter(X, v, p.mass, G), V, 1e-3))
                  - C_DAMP).max() < 1e-9
    Aref = p.mass * np.eye(3 * N) - h * Jv - h * h * Jx
    assert np.abs(A - Aref).max() <= 1e-8 * np.abs(Aref).max()
    F = s.forces_scatter(X, V, p.mass, G, C_DAMP)
    bref = h * (F.ravel() + h * Jx @ np.asarray(V, np.float64).ravel())
    assert np.abs(B.ravel() - bref).max() <= 1e-8 * np.abs(bref).max()
    assert np.abs(np.asarray(diag).ravel() - np.diag(A)).max() <= 1e-12 * np.abs(np.diag(A)).max()
    d = np.random.default_rng(1).standard_normal((N, 3))
    q = s.matvec(Cp, md, d)
    assert np.abs(q.ravel() - A @ d.ravel()).max() < 1e-12 * np.abs(A @ d.ravel()).max()


def test_damped_step_solves_dense_system(o64):
    """One damped step's dv equals the dense FP64 solve of the filtered
    system (CG to 1e-14 relative)."""
    X, V, springs, pinned = small_scene(6, seed=14)
    N = len(X)
    p = scenes.StepParams(dt=1 / 30, damping=C_DAMP, cg_max_iters=500, cg_tol=1e-14,
                          cg_tol_mode=scenes.TOL_RELATIVE, warm_start=False)
    prm = oracle_params(p, 6, 6, gpu_order=False)
    s = o64.system(N, springs, pinned)
    B, Cp = s.assemble(X, V, prm)
    a, b, _ = s.pairs()
    A = dense_A(N, a, b, Cp, p.mass - float(prm.h) * C_DAMP)
    free = np.repeat(pinned == 0, 3)
    dv_ref = np.zeros(3 * N)
    dv_ref[free] = np.linalg.solve(A[np.ix_(free, free)], B.ravel()[free])
    Xs, Vs, DV = X.astype(np.float64).copy(), V.astype(np.float64).copy(), np.zeros((N, 3))
    s.step(Xs, Vs, DV, prm)
    assert np.abs(DV.ravel() - dv_ref).max() <= 1e-10 * np.abs(dv_ref).max()


@pytest.mark.parametrize("pins", [True, False])
def test_damped_adjoint_matches_finite_differences(o64, pins):
    from tests.test_oracle_adjoint import loss, setup

    X0, V0, springs, pinned, target, prm, steps = setup(pins=pins)
    prm.damping = -0.5
    s = o64.system(len(X0), springs, pinned)
    ks = springs["Ks"].astype(np.float64)
    s.set_ks(ks)
    XT, _, states = s.rollout_forward(X0, V0, prm, steps)
    gks, gx0, gv0 = s.rollout_grad(X0, V0, prm, steps, 2 * (XT - target), states=states)
    rng = np.random.default_rng(43)
    for trial in range(3):
        dk = rng.standard_normal(len(ks)) * ks
        dv = rng.standard_normal(V0.shape)
        eps = 1e-4
        Lp = loss(s, X0, V0 + eps * dv, prm, steps, target, ks + eps * dk)
        Lm = loss(s, X0, V0 - eps * dv, prm, steps, target, ks - eps * dk)
        fd = (Lp - Lm) / (2 * eps)
        an = float(gks @ dk + (gv0 * dv).sum())
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-8), (trial, fd, an)
    # the damping changes the gradient (the term is not dropped)
    prm.damping = 0.0
    XT0, _, st0 = s.rollout_forward(X0, V0, prm, steps)
    g0 = s.rollout_grad(X0, V0, prm, steps, 2 * (XT0 - target), states=st0)[2]
    assert np.abs(g0 - gv0).max() > 1e-3 * np.abs(gv0).max()
