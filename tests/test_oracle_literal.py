"""The oracle's restatement of the reference's own step, bug for bug
(orc_ie_literal_step: IE's IntegrateImplicit puts spring i's Jacobian on
vertex i).  It is kept for the record and the CPU timing; these checks pin
it independently of its code (CPU only)."""
import numpy as np
import pytest

from opencloth_amd import scenes
from oracle import Oracle


def test_free_fall_with_inert_springs():
    """Ks = Kd = 0: A = M, b = M V + h m g, so after n steps every free vertex
    has fallen h^2 g n(n+1)/2 and pinned ones stay (closed form)."""
    o = Oracle(64)
    sc = scenes.plain_scene()
    sp = sc.springs.copy()
    sp["Ks"] = 0.0
    sp["Kd"] = 0.0
    s = o.system(sc.n_vertices, sp, sc.pinned)
    X = sc.X.astype(np.float64) + 10.0  # above the ground clamp
    V = np.zeros_like(X); Vn = np.zeros_like(X)
    h, g, n = 1.0 / 60.0, -9.81, 20
    for _ in range(n):
        s.ie_literal_step(X, V, Vn, 1.0, (0.0, g, 0.0), h, i_max=10, eps=1e-12)
    free = sc.pinned == 0
    drop = h * h * g * n * (n + 1) / 2
    assert np.allclose(X[free, 1] - (sc.X[free, 1] + 10.0), drop, rtol=1e-10, atol=0)
    assert np.array_equal(X[~free], sc.X[~free].astype(np.float64) + 10.0)
    assert np.allclose(V[free, 1], n * h * g, rtol=1e-10)


def spectral_K(xa, xb, L0, ks):
    """df1/dx1 of one spring from its spectrum (IE:671): -ks along the spring,
    -ks(1 - L0/l) across."""
    d = xa - xb
    l = np.linalg.norm(d)
    e = d / l
    P = np.outer(e, e)
    return -ks * P - ks * (1.0 - L0 / l) * (np.eye(3) - P)


def test_blocks_use_each_springs_jacobian():
    """Three vertices, three stretched springs, at rest: one literal step
    equals a dense float64 solve of the block-diagonal system whose block i
    is m I - h^2 K(spring i), and differs from the one whose block i sums
    the springs at vertex i (so the test tells the two apart)."""
    o = Oracle(64)
    X0 = np.array([[0.0, 1.0, 0.0], [1.3, 1.1, 0.2], [0.4, 1.9, -0.3]])
    springs = np.zeros(3, scenes.SPRING_DTYPE)
    springs["p1"] = [0, 1, 0]
    springs["p2"] = [1, 2, 2]
    springs["rest_length"] = [1.0, 0.8, 0.9]
    springs["Ks"] = [40.0, 25.0, 60.0]
    springs["Kd"] = [-0.25, -0.25, -0.25]
    m, h, g = 0.5, 0.05, np.array([0.0, -9.81, 0.0])
    s = o.system(3, springs, np.zeros(3, np.uint8))
    X = X0.copy(); V = np.zeros_like(X); Vn = np.zeros_like(X)
    s.ie_literal_step(X, V, Vn, m, g, h, i_max=60, eps=0.0)

    L0 = springs["rest_length"].astype(np.float64)
    ks = springs["Ks"].astype(np.float64)
    F = np.tile(m * g, (3, 1))
    Ks = []
    for k in range(3):
        a, b = springs["p1"][k], springs["p2"][k]
        d = X0[a] - X0[b]
        f = -ks[k] * (np.linalg.norm(d) - L0[k]) * d / np.linalg.norm(d)
        F[a] += f
        F[b] -= f
        Ks.append(spectral_K(X0[a], X0[b], L0[k], ks[k]))
    bvec = h * F  # V = 0
    want = np.stack([np.linalg.solve(m * np.eye(3) - h * h * Ks[i], bvec[i]) for i in range(3)])
    assert np.allclose(Vn, want, rtol=1e-9, atol=1e-12)
    assert np.allclose(X, X0 + h * want, rtol=1e-12, atol=1e-12)

    # the vertex-assembled blocks (the physically intended A_ii) differ
    Kv = [np.zeros((3, 3)) for _ in range(3)]
    for k in range(3):
        Kv[springs["p1"][k]] += Ks[k]
        Kv[springs["p2"][k]] += Ks[k]
    other = np.stack([np.linalg.solve(m * np.eye(3) - h * h * Kv[i], bvec[i]) for i in range(3)])
    assert np.abs(other - want).max() > 1e-3


def test_needs_as_many_springs_as_vertices():
    o = Oracle(32)
    springs = np.zeros(1, scenes.SPRING_DTYPE)
    springs["p1"], springs["p2"], springs["rest_length"], springs["Ks"] = 0, 1, 1.0, 1.0
    s = o.system(3, springs, None)
    X = np.zeros((3, 3), np.float32); X[1, 0] = 1.0
    with pytest.raises(ValueError):
        s.ie_literal_step(X, np.zeros_like(X), np.zeros_like(X), 1.0, (0, -9.81, 0), 0.01)


def test_fp32_demo_runs_and_stays_finite():
    """The demo scene in FP32 for 300 literal steps (10 CG iterations, EPS
    1e-3, as IE:75-77): finite, pins fixed, cloth below its start."""
    o = Oracle(32)
    sc = scenes.demo_scene()
    s = o.system(sc.n_vertices, sc.springs, sc.pinned)
    X = sc.X.copy(); V = np.zeros_like(X); Vn = np.zeros_like(X)
    for _ in range(300):
        s.ie_literal_step(X, V, Vn, sc.params.mass, sc.params.gravity, sc.params.dt)
    assert np.isfinite(X).all()
    assert np.array_equal(X[sc.pinned == 1], sc.X[sc.pinned == 1])
    assert X[:, 1].mean() < sc.X[:, 1].mean()
