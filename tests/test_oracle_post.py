"""Pins the oracle's post-step passes (EllipsoidCollision IE:727-751,
ApplyProvotDynamicInverse IE:702-726) and the position-Verlet step
(VL:428-484) with closed forms and invariants."""
import numpy as np
import pytest

from opencloth_amd import scenes


@pytest.fixture(scope="module")
def o64():
    from oracle import Oracle

    return Oracle(64)


def two_point_system(o, pinned=None, rest=1.0):
    sp = np.zeros(1, scenes.SPRING_DTYPE)
    sp["p1"], sp["p2"], sp["rest_length"], sp["Ks"], sp["Kd"] = 0, 1, rest, 10.0, -0.1
    return o.system(2, sp, pinned)


def test_ellipsoid_pushes_to_sphere_surface(o64):
    """With a pure translation the collider is the unit sphere around c: an
    inside point moves radially to distance 1, its velocity is zeroed; an
    outside point is untouched."""
    c = np.array([0.5, -1.0, 2.0])
    M = np.eye(4); M[:3, 3] = c
    ell = {"M": M.T.astype(np.float32).ravel(), "Minv": np.linalg.inv(M).T.astype(np.float32).ravel(),
           "center": np.zeros(3, np.float32), "radius": 1.0}
    s = o64.system(3, np.zeros(0, scenes.SPRING_DTYPE), None)
    X = np.array([c + [0.3, 0.0, 0.0], c + [0.0, -0.2, 0.4], c + [1.5, 0.0, 0.0]], np.float64)
    V = np.ones((3, 3))
    Xo = X.copy()
    s.ellipsoid(X, V, ell, 0)
    for i in range(2):
        d = Xo[i] - c
        assert np.allclose(X[i], c + d / np.linalg.norm(d), atol=1e-12)
        assert np.all(V[i] == 0)
    assert np.array_equal(X[2], Xo[2]) and np.all(V[2] == 1)
    XL = np.zeros((3, 3))
    s.ellipsoid(Xo.copy(), XL, ell, 1)  # Verlet mode: previous position = new position
    assert np.allclose(XL[0], c + [1.0, 0, 0], atol=1e-12)


def test_ellipsoid_axis_scaling(o64):
    """scale(1,1,s): the reference maps the object-space push d0 back with
    row_a/|row_a|^2 of the transform (IE:738-746), i.e. d0_z / s, not s*d0_z;
    a point (0,0,z) inside therefore lands at z + (1 - z/s)/s."""
    sz = 0.5
    M = np.diag([1.0, 1.0, sz, 1.0])
    ell = {"M": M.T.astype(np.float32).ravel(), "Minv": np.linalg.inv(M).T.astype(np.float32).ravel(),
           "center": np.zeros(3, np.float32), "radius": 1.0}
    s = o64.system(1, np.zeros(0, scenes.SPRING_DTYPE), None)
    z = 0.2
    X = np.array([[0.0, 0.0, z]]); V = np.zeros((1, 3))
    s.ellipsoid(X, V, ell, 0)
    assert np.allclose(X[0], [0, 0, z + (1 - z / sz) / sz], atol=1e-12)


@pytest.mark.parametrize("slack", [1.0, 1.01])
def test_provot_closed_form(o64, slack):
    s = two_point_system(o64)
    X = np.array([[0.0, 0.0, 0.0], [1.3, 0.0, 0.0]])
    V = np.zeros((2, 3))
    s.provot(X, V, slack)
    ex = (1.3 - slack) / 2
    assert np.allclose(V[0], [ex, 0, 0]) and np.allclose(V[1], [-ex, 0, 0])
    assert np.allclose(V.sum(0), 0)  # momentum
    V = np.zeros((2, 3))
    s.provot(np.array([[0.0, 0, 0], [0.9, 0, 0]]), V, slack)  # below rest: nothing
    assert np.all(V == 0)


def test_provot_pins(o64):
    X = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    s = two_point_system(o64, np.array([1, 0], np.uint8))
    V = np.zeros((2, 3)); s.provot(X, V, 1.0)
    assert np.all(V[0] == 0) and np.allclose(V[1], [-0.5, 0, 0])
    s = two_point_system(o64, np.array([0, 1], np.uint8))
    V = np.zeros((2, 3)); s.provot(X, V, 1.0)
    assert np.allclose(V[0], [0.5, 0, 0]) and np.all(V[1] == 0)


def test_verlet_free_fall_closed_form(o64):
    """No springs, no damping: x_n = x_0 + g dt^2 n(n+1)/2 from rest."""
    s = o64.system(4, np.zeros(0, scenes.SPRING_DTYPE), None)
    X0 = np.arange(12, dtype=np.float64).reshape(4, 3)
    dt, n = 0.01, 25
    X, XL = s.verlet_rollout(X0, X0, dt, 2.0, (0.0, -9.81, 0.0), 0.0, 0, n)
    expect = X0.copy(); expect[:, 1] += -9.81 * dt * dt * n * (n + 1) / 2
    assert np.allclose(X, expect, atol=1e-12)


def test_verlet_forces_are_spring_forces_at_verlet_velocity(o64):
    n = 7
    X0 = scenes.grid_positions(n, n)
    springs = scenes.grid_springs(n, n, X0)
    rng = np.random.default_rng(0)
    X = X0 + 0.05 * rng.standard_normal(X0.shape)
    XL = X - 0.01 * rng.standard_normal(X0.shape)
    pinned = np.zeros(n * n, np.uint8); pinned[0] = 1
    s = o64.system(n * n, springs, pinned)
    dt, m, kd = 1 / 60, 2.0, -0.0125
    F = s.verlet_forces(X, XL, dt, m, (0, -9.81, 0), kd)
    V = (X - XL) / dt
    Fs = s.forces_scatter(X, V, m, (0, -9.81, 0))
    expect = Fs + kd * V
    expect[0] = kd * V[0]  # pinned: damping only (VL:455-458, VL:479-482)
    assert np.allclose(F, expect, rtol=1e-12, atol=1e-12)
