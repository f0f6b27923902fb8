"""The C ABI's other ways of driving the implicit step, against the plain
step and the oracle (bit for bit):

- the reference's split (assemble, SolveConjugateGradient, update) and
  IntegrateImplicit equal oc_cloth_step;
- parameters changed between steps (iteration cap, tolerance, dt) follow
  the oracle run with the same schedule;
- work ordered on a caller's HIP stream gives the same bits;
- state and CG start round-trip through set/get;
- the profiling and timing entry points leave x and v alone and the next
  step unchanged."""
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


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def scene():
    sc = scenes.plain_scene(n=40)
    X = scenes.perturb(sc.X, 0.02, 21)
    return sc, X


def test_split_and_integrate_equal_step():
    from opencloth_amd.cloth import ImplicitCloth

    sc, X = scene()
    a = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    b = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    c = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    for _ in range(4):
        a.step(1)
        b.assemble()
        it_b, st_b = b.solve_conjugate_gradient()
        b.update()
        c.integrate_implicit()
        assert (it_b, st_b) == a.last_status() == c.last_status()
    Xa, Va = a.get_state()
    for o in (b, c):
        Xo, Vo = o.get_state()
        assert np.array_equal(bits(Xo), bits(Xa)) and np.array_equal(bits(Vo), bits(Va))


@pytest.mark.parametrize("warm,tol_mode", [(True, 0), (False, 1), (True, 2)])
def test_multi_step_call_equals_single_steps(o32, warm, tol_mode):
    """With a warm CG start oc_cloth_step(n) applies each step's update
    inside the next step's assembly (a cold start keeps an update launch per
    step); either way n-step calls equal n calls of oc_cloth_step(1) and the
    oracle bit for bit, with the ground clamp acting (the cloth starts just
    above the ground and falls onto it) and each stopping rule."""
    from opencloth_amd.cloth import ImplicitCloth

    sc, X = scene()
    X = X.copy()
    X[:, 1] = 0.002 * np.arange(len(X)) % 0.05  # heights 0 .. 0.05
    prm = dataclasses.replace(sc.params, ground_clamp=True, warm_start=warm, cg_tol_mode=tol_mode,
                              cg_tol=1e-4, cg_max_iters=12)
    pins = np.zeros_like(sc.pinned)
    a = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, pins, prm)
    b = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, pins, prm)
    n = 12
    a.step(n)
    for _ in range(n):
        b.step(1)
    Xa, Va = a.get_state()
    Xb, Vb = b.get_state()
    assert np.array_equal(bits(Xa), bits(Xb)) and np.array_equal(bits(Va), bits(Vb))
    assert a.last_status() == b.last_status()
    assert (Xa[:, 1] == 0.0).sum() > 10  # clamped vertices
    s = o32.system(sc.n_vertices, sc.springs, pins)
    Xr, Vr, DV = X.copy(), np.zeros_like(X), np.zeros_like(X)
    for _ in range(n):
        it_r, st_r = s.step(Xr, Vr, DV, oracle_params(prm, sc.nx, sc.ny))
    assert a.last_status() == (it_r, st_r)
    assert np.array_equal(bits(Xa), bits(Xr)) and np.array_equal(bits(Va), bits(Vr))
    a.step(3)  # the swapped state buffers keep working across calls
    b.step(1); b.step(2)
    assert np.array_equal(bits(a.get_state()[0]), bits(b.get_state()[0]))


def test_params_changed_between_steps(o32):
    from opencloth_amd.cloth import ImplicitCloth

    sc, X = scene()
    schedule = [sc.params,
                dataclasses.replace(sc.params, cg_max_iters=3),
                dataclasses.replace(sc.params, dt=1 / 120, cg_tol=1e-6, cg_max_iters=25),
                dataclasses.replace(sc.params, cg_tol_mode=scenes.TOL_RELATIVE, cg_tol=1e-3),
                dataclasses.replace(sc.params, cg_jacobi=True, cg_tol=1e-7, cg_max_iters=15)]
    cl = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    s = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    Xr, Vr, DV = X.copy(), np.zeros_like(X), np.zeros_like(X)
    for p in schedule:
        say("params", p)
        cl.set_params(p)
        cl.step(2)
        for _ in range(2):
            it_r, st_r = s.step(Xr, Vr, DV, oracle_params(p, sc.nx, sc.ny))
        assert cl.last_status() == (it_r, st_r)
    Xg, Vg = cl.get_state()
    assert np.array_equal(bits(Xg), bits(Xr)) and np.array_equal(bits(Vg), bits(Vr))


def test_caller_stream():
    from opencloth_amd.cloth import ImplicitCloth
    from opencloth_amd.hiptime import Stream

    sc, X = scene()
    ref = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    ref.step(5)
    cl = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    stream = Stream()
    cl.set_stream(stream.handle)
    cl.step(5)
    cl.synchronize()
    Xs, Vs = cl.get_state()
    cl.set_stream(None)
    Xr, Vr = ref.get_state()
    assert np.array_equal(bits(Xs), bits(Xr)) and np.array_equal(bits(Vs), bits(Vr))


def test_state_and_dv_round_trip():
    from opencloth_amd.cloth import ImplicitCloth

    sc, X = scene()
    cl = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    rng = np.random.default_rng(3)
    X2 = (X + 0.01 * rng.standard_normal(X.shape)).astype(np.float32)
    V2 = (0.1 * rng.standard_normal(X.shape)).astype(np.float32)
    D2 = (0.01 * rng.standard_normal(X.shape)).astype(np.float32)
    cl.set_state(X2, V2)
    cl.set_dv(D2)
    Xg, Vg = cl.get_state()
    assert np.array_equal(Xg, X2) and np.array_equal(Vg, V2)
    assert np.array_equal(cl.get_dv(), D2)
    cl.set_state(V=np.zeros_like(V2))  # one of the two only
    Xg, Vg = cl.get_state()
    assert np.array_equal(Xg, X2) and not Vg.any()


def test_profiling_entry_points_leave_the_step_alone():
    from opencloth_amd.cloth import ImplicitCloth

    sc, X = scene()
    a = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    b = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, sc.params)
    a.step(2)
    b.step(2)
    Xa, Va = a.get_state()
    us = a.time_cg_spmv(20)
    assert us > 0
    Xt, Vt = a.get_state()
    assert np.array_equal(bits(Xt), bits(Xa)) and np.array_equal(bits(Vt), bits(Va))
    prof = a.profile_steps(3)
    assert prof["cg_spmv"][1] > 0 and prof["assemble"][1] == 3 and all(ms >= 0 for ms, _ in prof.values())
    b.step(3)
    # time_cg_spmv leaves dv as its iteration left it; reset both starts
    a.set_state(*b.get_state())
    a.set_dv(b.get_dv())
    a.step(1)
    b.step(1)
    assert np.array_equal(bits(a.get_state()[0]), bits(b.get_state()[0]))
