"""GPU parity of the position-based-dynamics step (oc_pbd, after
OpenCloth_PositionBasedDynamics/.../main.cpp:601-925) against the FP32
oracle in the kernels' colour order and reduction tree: bit-exact for one
step, for the demo's 600 steps with the ellipsoid and on a perturbed 64x64
cloth; the reference's sequential order is compared at the documented
tolerance."""
import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import rel_pos_err, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def oracle_system(o, sc):
    return o.pbd(sc.X, sc.pinned, sc.params.mass, sc.distance, sc.bending, sc.params.solver_iterations)


def test_colours_match_oracle(o32):
    from opencloth_amd.cloth import PBDCloth

    sc = scenes.pbd_demo_scene()
    cl = PBDCloth.from_scene(sc)
    _, dp, _, bp = oracle_system(o32, sc).colours()
    assert cl.num_colors() == (len(dp) - 1, len(bp) - 1)


def test_demo_600_steps_bitwise(o32):
    """The PBD demo with its collider, checked bit for bit every 100 steps."""
    from oracle import make_pbd_params
    from opencloth_amd.cloth import PBDCloth

    sc = scenes.pbd_demo_scene()
    ell = sc.extra["ellipsoid"]
    say("pbd demo x600")
    cl = PBDCloth.from_scene(sc)
    s = oracle_system(o32, sc)
    prm = make_pbd_params(sc.params)
    X, V = sc.X.copy(), sc.V.copy()
    cl.step(1, ell)
    X, V = s.step(X, V, prm, 1, 1, ell)
    Xg, Vg = cl.get_state()
    assert np.array_equal(bits(Xg), bits(X)) and np.array_equal(bits(Vg), bits(V))
    for k in range(1, 600, 100):
        n = min(100, 600 - k)
        cl.step(n, ell)
        X, V = s.step(X, V, prm, 1, n, ell)
        Xg, Vg = cl.get_state()
        assert np.array_equal(bits(Xg), bits(X)), k
        assert np.array_equal(bits(Vg), bits(V)), k
    assert X[:, 1].min() < 2.0  # draped on the collider
    # the reference's sequential order, same 600 steps
    Xs, _ = s.step(sc.X, sc.V, prm, 0, 600, ell)
    err = rel_pos_err(Xg, Xs)
    say("colour order vs sequential order: %.3e" % err)
    assert err < 5e-3


def test_perturbed_64x64_bitwise(o32):
    """64x64 cloth (17 reduction blocks), random positions and velocities,
    top edge pinned, no collider: 20 steps bit-exact."""
    from oracle import make_pbd_params
    from opencloth_amd.cloth import PBDCloth

    sc = scenes.pbd_large_scene(64, seed=5)
    rng = np.random.default_rng(6)
    V = (0.05 * rng.standard_normal(sc.X.shape)).astype(np.float32)
    V[sc.pinned.astype(bool)] = 0
    say("pbd 64x64 x20")
    cl = PBDCloth(sc.X, sc.distance, sc.bending, sc.pinned, sc.params, V)
    s = oracle_system(o32, sc)
    cl.step(20)
    X, Vr = s.step(sc.X, V, make_pbd_params(sc.params), 1, 20)
    Xg, Vg = cl.get_state()
    assert np.array_equal(bits(Xg), bits(X)) and np.array_equal(bits(Vg), bits(Vr))
    assert np.array_equal(Xg[:64], sc.X[:64])  # pins stay


def test_partial_stiffness_and_more_iterations(o32):
    """Other constants: 5 solver iterations, k 0.9 / 0.1, stronger damping,
    no ground clamp: 30 steps bit-exact."""
    import dataclasses

    from oracle import make_pbd_params
    from opencloth_amd.cloth import PBDCloth

    sc = scenes.pbd_demo_scene(33)
    sc.distance["k"] = 0.9
    sc.bending["k"] = 0.1
    sc.params = dataclasses.replace(sc.params, solver_iterations=5, k_damp=0.05, global_damping=0.95,
                                    ground=False)
    say("pbd 33x33 ns5 x30")
    cl = PBDCloth.from_scene(sc)
    s = oracle_system(o32, sc)
    cl.step(30)
    X, V = s.step(sc.X, sc.V, make_pbd_params(sc.params), 1, 30)
    Xg, Vg = cl.get_state()
    assert np.array_equal(bits(Xg), bits(X)) and np.array_equal(bits(Vg), bits(V))


def levels_of(n_particles, constraints):
    """Level of each constraint in list order: one more than the highest level
    among the earlier constraints sharing a particle (a sequential scan)."""
    last = np.full(n_particles, -1)
    lev = np.zeros(len(constraints), np.int64)
    for c, ps in enumerate(constraints):
        lev[c] = max(last[p] for p in ps) + 1
        for p in ps:
            last[p] = lev[c]
    return lev


def test_list_order_levels_match(o32):
    """OC_FLAG_LIST_ORDER: the level counts the library reports equal a
    sequential scan of the lists; every level's constraints share no
    particle and each constraint comes after every earlier one it touches."""
    import dataclasses

    from opencloth_amd.cloth import PBDCloth

    sc = scenes.pbd_demo_scene()
    sc.params = dataclasses.replace(sc.params, list_order=True)
    cl = PBDCloth.from_scene(sc)
    dl = levels_of(sc.n_vertices, [(c["p1"], c["p2"]) for c in sc.distance])
    bl = levels_of(sc.n_vertices, [(c["p1"], c["p2"], c["p3"]) for c in sc.bending])
    assert cl.num_colors() == (int(dl.max()) + 1, int(bl.max()) + 1)
    for L in range(int(dl.max()) + 1):
        ps = np.concatenate([[c["p1"], c["p2"]] for c in sc.distance[dl == L]])
        assert len(np.unique(ps)) == len(ps)
    cl.close()


@pytest.mark.parametrize("steps", [1, 600])
def test_list_order_equals_sequential_oracle(o32, steps):
    """The reference's sequential order (list order and sequential sums, the
    oracle's order 0) on the PBD demo with its collider: bit-exact after one
    step and over 600 steps (checked every 100)."""
    import dataclasses

    from oracle import make_pbd_params
    from opencloth_amd.cloth import PBDCloth

    sc = scenes.pbd_demo_scene()
    sc.params = dataclasses.replace(sc.params, list_order=True)
    ell = sc.extra["ellipsoid"]
    say("pbd list order x", steps)
    cl = PBDCloth.from_scene(sc)
    s = oracle_system(o32, sc)
    prm = make_pbd_params(sc.params)
    X, V = sc.X.copy(), sc.V.copy()
    done = 0
    while done < steps:
        n = min(100, steps - done)
        cl.step(n, ell)
        X, V = s.step(X, V, prm, 0, n, ell)
        done += n
        Xg, Vg = cl.get_state()
        assert np.array_equal(bits(Xg), bits(X)), done
        assert np.array_equal(bits(Vg), bits(V)), done
    if steps == 600:
        assert X[:, 1].min() < 2.0  # draped on the collider
    cl.close()


def test_list_order_perturbed_64x64(o32):
    """A perturbed 64x64 cloth with random velocities, 5 solver iterations:
    10 steps in list order bit-exact against the oracle's order 0."""
    import dataclasses

    from oracle import make_pbd_params
    from opencloth_amd.cloth import PBDCloth

    sc = scenes.pbd_large_scene(64, seed=7)
    sc.params = dataclasses.replace(sc.params, list_order=True, solver_iterations=5)
    rng = np.random.default_rng(8)
    V = (0.05 * rng.standard_normal(sc.X.shape)).astype(np.float32)
    V[sc.pinned.astype(bool)] = 0
    say("pbd list order 64x64 x10")
    cl = PBDCloth(sc.X, sc.distance, sc.bending, sc.pinned, sc.params, V)
    s = oracle_system(o32, sc)
    cl.step(10)
    X, Vr = s.step(sc.X, V, make_pbd_params(sc.params), 0, 10)
    Xg, Vg = cl.get_state()
    assert np.array_equal(bits(Xg), bits(X)) and np.array_equal(bits(Vg), bits(Vr))
    cl.close()
