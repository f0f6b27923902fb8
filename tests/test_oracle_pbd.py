"""Pins the oracle's position-based-dynamics step (oracle/pbd.c, after
OpenCloth_PositionBasedDynamics/.../main.cpp, PBD) with closed forms and
invariants: one distance projection (PBD:680-701), one triangle bending
projection (PBD:703-731), the centre-of-mass damping (PBD:612-668) against
an independent rigid-body derivation, momentum conservation, colour
validity, and the difference between the reference's sequential order and
the kernels' colour order on the demo scene."""
import numpy as np
import pytest

from opencloth_amd import scenes


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


@pytest.fixture(scope="module")
def o64():
    from oracle import Oracle

    return Oracle(64)


def params(**over):
    from oracle import make_pbd_params

    return make_pbd_params(scenes.PBDParams(), **over)


def dist_list(pairs, k=1.0):
    d = np.zeros(len(pairs), scenes.PBD_DISTANCE_DTYPE)
    if len(pairs):
        d["p1"], d["p2"] = np.asarray(pairs).T
    d["k"] = k
    return d


def bend_list(triples, k=1.0):
    b = np.zeros(len(triples), scenes.PBD_BENDING_DTYPE)
    if len(triples):
        b["p1"], b["p2"], b["p3"] = np.asarray(triples).T
    b["k"] = k
    return b


# ---- scene ---------------------------------------------------------------
def test_demo_constraint_lists():
    """PBD:399-433 on the 21x21 demo: 420 horizontal, 420 vertical, 800
    shear distance constraints; 399 + 399 bending triples along columns and
    rows; every constraint joins neighbours at the grid spacing."""
    sc = scenes.pbd_demo_scene()
    d, b = sc.distance, sc.bending
    assert len(d) == 420 + 420 + 800 and len(b) == 399 + 399
    assert (d["p1"][:3] == [0, 1, 2]).all() and (d["p2"][:3] == [1, 2, 3]).all()
    assert d["p1"][420] == 0 and d["p2"][420] == 21
    assert (d["p1"][840], d["p2"][840], d["p1"][841], d["p2"][841]) == (0, 22, 21, 1)
    assert tuple(b[0])[:3] == (0, 21, 42) and tuple(b[399])[:3] == (0, 1, 2)
    h = 4.0 / 20
    L = np.linalg.norm(sc.X[d["p1"]] - sc.X[d["p2"]], axis=1)
    assert np.allclose(L[:840], h, rtol=1e-6) and np.allclose(L[840:], h * np.sqrt(2), rtol=1e-6)
    mid = sc.X[b["p2"]]
    assert np.allclose(mid, 0.5 * (sc.X[b["p1"]] + sc.X[b["p3"]]), atol=1e-6)
    assert set(np.nonzero(sc.pinned)[0]) == {0, 20}


def test_k_prime(o32):
    """k' = 1 - (1 - k)^(1/ns) (PBD:153): ns = 2, k = 0.25 gives 1 - sqrt(0.75);
    ns = 1 gives k itself; k = 1 gives 1."""
    X = np.zeros((3, 3), np.float32); X[1, 0] = 1; X[2, 0] = 2
    for ns, k, want in [(2, 0.25, 1 - np.sqrt(0.75)), (1, 0.25, 0.25), (2, 1.0, 1.0), (4, 0.5, 1 - 0.5 ** 0.25)]:
        s = o32.pbd(X, None, 1.0, dist_list([(0, 1)], k), bend_list([(0, 1, 2)], k), ns)
        _, dkp, _, _, bkp = s.rest()
        assert abs(float(dkp[0]) - want) <= 2e-7 and abs(float(bkp[0]) - want) <= 2e-7


# ---- projections -----------------------------------------------------------
@pytest.mark.parametrize("stretch", [2.5, 0.2])
def test_distance_projection_closed_form(o64, stretch):
    """k' = 1, equal inverse masses: one projection restores the rest length
    and keeps the midpoint; with p1 pinned, p1 stays and p2 moves alone."""
    X0 = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    T = np.array([[0.1, 0.2, -0.3], [0.1 + stretch * 0.6, 0.2 + stretch * 0.8, -0.3]])
    s = o64.pbd(X0, None, 1.0, dist_list([(0, 1)]), bend_list([]), 1)
    out = s.project(T, params(solver_iterations=1, ground=0), 0)
    assert abs(np.linalg.norm(out[0] - out[1]) - 1.0) < 1e-12
    assert np.allclose(out.mean(0), T.mean(0), atol=1e-12)
    s = o64.pbd(X0, np.array([1, 0], np.uint8), 1.0, dist_list([(0, 1)]), bend_list([]), 1)
    out = s.project(T, params(solver_iterations=1, ground=0), 0)
    assert np.array_equal(out[0], T[0])
    assert np.allclose(out[1], T[0] + (T[1] - T[0]) / np.linalg.norm(T[1] - T[0]), atol=1e-12)


def test_distance_projection_partial_stiffness(o64):
    """k' < 1 moves the pair by k' of the full correction."""
    X0 = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    T = np.array([[0.0, 0, 0], [3.0, 0, 0]])
    s = o64.pbd(X0, None, 1.0, dist_list([(0, 1)], 0.25), bend_list([]), 1)
    out = s.project(T, params(solver_iterations=1, ground=0), 0)
    # full correction moves each end by 1 (length 3 -> 1); k' = 0.25 moves 0.25
    assert np.allclose(out, [[0.25, 0, 0], [2.75, 0, 0]], atol=1e-12)


def test_bending_projection_closed_form(o64):
    """k' = 1, equal inverse masses: the centroid of the three particles is
    kept and p3 ends at distance rest + global_k from the 0.3333-weighted
    centre (global_k = 0.01 * global_damping, PBD:712)."""
    X0 = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)
    s = o64.pbd(X0, None, 1.0, dist_list([]), bend_list([(0, 1, 2)]), 1)
    rest = s.rest()[2][0]
    T = np.array([[0.0, 0, 0], [1.0, 0.6, 0.1], [2.1, -0.2, 0.3]])
    prm = params(solver_iterations=1, ground=0, global_damping=0.98)
    out = s.project(T, prm, 0)
    ctr = float(np.float32(0.3333)) * out.sum(0)  # the float constant of PBD:713
    global_k = float(np.float32(0.98)) * float(np.float32(0.01))  # float constants of PBD:712
    assert abs(np.linalg.norm(out[2] - ctr) - (rest + global_k)) < 1e-9
    assert np.allclose(out.sum(0), T.sum(0), atol=1e-12)
    # the first two particles move together, the third against them, 2:2:4
    d = out - T
    assert np.allclose(d[0], d[1], atol=1e-12) and np.allclose(d[2], -2 * d[0], atol=1e-12)


def test_ground_clamp(o64):
    X0 = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    s = o64.pbd(X0, None, 1.0, dist_list([]), bend_list([]), 2)
    T = np.array([[0.0, -0.5, 0], [1.0, 0.25, 0]])
    out = s.project(T, params(solver_iterations=2, ground=1), 0)
    assert out[0, 1] == 0 and out[1, 1] == 0.25
    out = s.project(T, params(solver_iterations=2, ground=0), 0)
    assert np.array_equal(out, T)


# ---- prediction and damping -------------------------------------------------
def cloud(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, 3)), rng.standard_normal((n, 3))


@pytest.mark.parametrize("order", [0, 1])
def test_damping_of_a_rigid_motion(o64, order):
    """kDamp = 1, no gravity, no global damping, V = v0 + w0 x R: then
    L = I_true w0 and the reference's I = identity + I_true, so the damped
    velocity is v0 + w x R with w = (1 + I_true)^-1 I_true w0 (derived here
    in float64 from the definitions, not from the oracle's code)."""
    n, m = 300, float(np.float32(0.37))  # the system stores the mass as a float
    X, _ = cloud(n, 1)
    v0, w0 = np.array([0.3, -0.2, 0.5]), np.array([0.7, 0.1, -0.4])
    s = o64.pbd(X.astype(np.float32), None, m, dist_list([]), bend_list([]), 1)
    X32 = X.astype(np.float32).astype(np.float64)  # the system's particle positions
    R = X32 - X32.mean(0)
    V = v0 + np.cross(w0, R)
    Itrue = m * (np.sum(R * R) * np.eye(3) - R.T @ R)
    w = np.linalg.solve(np.eye(3) + Itrue, Itrue @ w0)
    want = v0 + np.cross(w, R)
    Vout, T, wout = s.predict(X32, V, params(mass=m, gx=0, gy=0, gz=0, global_damping=1.0, k_damp=1.0), order)
    assert np.allclose(wout, w, rtol=1e-9, atol=1e-12)
    assert np.allclose(Vout, want, rtol=1e-9, atol=1e-12)
    dt = float(np.float32(1 / 60))
    assert np.allclose(T, X32 + Vout * dt, atol=1e-12)


@pytest.mark.parametrize("order", [0, 1])
def test_damping_conserves_momentum(o32, order):
    """Centre-of-mass damping changes velocities by Vcm + w x R - V, whose
    mass-weighted sum is zero: total momentum is kept (FP32, to rounding)."""
    X, V = cloud(1000, 2)
    s = o32.pbd(X.astype(np.float32), None, 0.01, dist_list([]), bend_list([]), 1)
    Vout, _, _ = s.predict(X, V, params(mass=0.01, gx=0, gy=0, gz=0, global_damping=1.0, k_damp=0.6), order)
    p_in = V.astype(np.float32).astype(np.float64).sum(0)
    assert np.abs(Vout.astype(np.float64).sum(0) - p_in).max() < 2e-4 * np.abs(V).sum(0).max() / 100
    assert np.abs(Vout - V).max() > 0.1  # the damping did act


def test_prediction_gravity_and_pins(o64):
    """kDamp = 0: V' = gd V + g dt W for free particles, V' = gd V for pinned
    ones (no gravity: F = 0 where W = 0, PBD:601-611), T = X + V' dt, pinned
    particles predicted in place."""
    X, V = cloud(50, 3)
    pins = np.zeros(50, np.uint8); pins[[0, 7]] = 1
    m = 0.5
    s = o64.pbd(X.astype(np.float32), pins, m, dist_list([]), bend_list([]), 1)
    X = X.astype(np.float32).astype(np.float64)
    p = params(mass=m, global_damping=0.9, k_damp=0.0, gx=0.1, gy=-2.0, gz=0.3, dt=0.02)
    Vout, T, _ = s.predict(X, V, p, 0)
    g = np.array([0.1, -2.0, 0.3], np.float32).astype(np.float64)
    dt = float(np.float32(0.02)); gd = float(np.float32(0.9))
    want = gd * V + (1 - pins[:, None]) * g * dt / float(np.float32(m))
    assert np.allclose(Vout, want, rtol=1e-12, atol=1e-12)
    assert np.array_equal(T[pins == 1], X[pins == 1])
    assert np.allclose(T[pins == 0], X[pins == 0] + Vout[pins == 0] * dt, atol=1e-12)


# ---- colours and orders ------------------------------------------------------
def test_colours_are_valid(o32):
    """Each colour's constraints share no particle; every constraint appears
    once; list order is kept inside a colour; the demo needs 8 distance and
    6 bending colours."""
    sc = scenes.pbd_demo_scene()
    s = o32.pbd(sc.X, sc.pinned, sc.params.mass, sc.distance, sc.bending, 2)
    do, dp, bo, bp = s.colours()
    assert len(dp) - 1 == 8 and len(bp) - 1 == 6
    for order, ptr, lst, ar in [(do, dp, sc.distance, 2), (bo, bp, sc.bending, 3)]:
        assert np.array_equal(np.sort(order), np.arange(len(lst)))
        cols = ["p1", "p2", "p3"][:ar]
        for c in range(len(ptr) - 1):
            ids = order[ptr[c]:ptr[c + 1]]
            assert np.all(np.diff(ids) > 0)
            parts = np.concatenate([lst[k][ids] for k in cols])
            assert len(np.unique(parts)) == len(parts)


def test_orders_agree_on_disjoint_constraints(o32):
    """When every constraint of a family touches different particles there is
    one colour holding the list order, so the two projection orders give the
    same bits."""
    X = np.random.default_rng(4).standard_normal((12, 3)).astype(np.float32)
    s = o32.pbd(X, None, 1.0, dist_list([(0, 1), (2, 3), (4, 5)], 0.7),
                bend_list([(6, 7, 8), (9, 10, 11)], 0.6), 3)
    T = X * 1.3 + 0.1
    a = s.project(T, params(solver_iterations=3, ground=0), 0)
    b = s.project(T, params(solver_iterations=3, ground=0), 1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def demo_runs(o, order, steps):
    from oracle import make_pbd_params

    sc = scenes.pbd_demo_scene()
    s = o.pbd(sc.X, sc.pinned, sc.params.mass, sc.distance, sc.bending, sc.params.solver_iterations)
    return sc, s.step(sc.X, sc.V, make_pbd_params(sc.params), order, steps, sc.extra["ellipsoid"])


def test_demo_order_difference_is_small(o32):
    """The reference's sequential Gauss-Seidel order against the colour order
    on the demo (600 steps, cloth draped on the ellipsoid): the trajectories
    differ by 2.2e-3 relative (norm-wise); the bar is 5e-3."""
    from tests.helpers import rel_pos_err

    sc, (X0, _) = demo_runs(o32, 0, 600)
    _, (X1, _) = demo_runs(o32, 1, 600)
    err = rel_pos_err(X1, X0)
    print("literal vs colour order, 600 steps: %.3e" % err)
    assert 0 < err < 5e-3
    assert np.array_equal(X1[[0, 20]], sc.X[[0, 20]])  # pins stay
    assert X1[:, 1].min() >= 0 and X1[:, 1].max() <= 5.0
    assert X1[:, 1].min() < 2.0  # the cloth fell and draped


def test_demo_fp32_against_fp64(o32, o64):
    """Colour order, 600 demo steps: FP32 restatement against the FP64 one."""
    from tests.helpers import rel_pos_err

    _, (X32, _) = demo_runs(o32, 1, 600)
    _, (X64, _) = demo_runs(o64, 1, 600)
    err = rel_pos_err(X32, X64)
    print("fp32 vs fp64, 600 steps: %.3e" % err)
    assert err < 1e-3
