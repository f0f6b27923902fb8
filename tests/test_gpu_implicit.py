"""GPU parity of the implicit step through the C ABI against the CPU oracle.

Bars: per-spring forces, pair blocks and right-hand sides bit-exact against
the FP32 restatement (same expression order, no FMA on either side); CG
iterates bit-exact when the oracle follows the kernels' reduction tree;
trajectories within 1e-4 max relative position error per step (plain
21x21 scene, 300 steps) against the FP32 restatement, with the FP64 oracle
error reported.
"""
import numpy as np
import pytest

from opencloth_amd import scenes
from tests.helpers import oracle_params, rel_pos_err, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def o32():
    from oracle import Oracle

    return Oracle(32)


@pytest.fixture(scope="module")
def o64():
    from oracle import Oracle

    return Oracle(64)


def make_cloth(sc):
    from opencloth_amd.cloth import ImplicitCloth

    return ImplicitCloth.from_scene(sc)


def random_state(sc, seed, vscale=0.5, xscale=0.05):
    rng = np.random.default_rng(seed)
    X = (sc.X + xscale * rng.standard_normal(sc.X.shape)).astype(np.float32)
    V = (vscale * rng.standard_normal(sc.X.shape)).astype(np.float32)
    V[sc.pinned.astype(bool)] = 0
    return X, V


def test_library_loads_and_sees_device():
    from opencloth_amd._lib import lib

    say("device count", lib().oc_device_count())
    assert lib().oc_device_count() >= 1


@pytest.mark.parametrize("seed", [0, 1])
def test_forces_bitwise(o32, seed):
    sc = scenes.plain_scene()
    X, V = random_state(sc, seed)
    say("forces 21x21 seed", seed)
    cl = make_cloth(sc)
    cl.set_state(X, V)
    F = cl.compute_forces()
    sys_ = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    Fref = sys_.forces_scatter(X, V, sc.params.mass, sc.params.gravity)
    assert np.array_equal(F.view(np.uint32), Fref.view(np.uint32)), np.abs(F - Fref).max()


def test_pairs_match_oracle(o32):
    sc = scenes.plain_scene()
    cl = make_cloth(sc)
    a, b = cl.pairs()
    sys_ = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    ra, rb, _ = sys_.pairs()
    assert np.array_equal(a, ra) and np.array_equal(b, rb)


@pytest.mark.parametrize("seed", [0, 3])
def test_assembly_bitwise(o32, seed):
    sc = scenes.plain_scene()
    X, V = random_state(sc, seed)
    say("assembly 21x21 seed", seed)
    cl = make_cloth(sc)
    cl.set_state(X, V)
    cl.set_dv(None)
    cl.assemble()
    Cb, b = cl.system()
    sys_ = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(sc.params, sc.nx, sc.ny)
    bref, Cref = sys_.assemble(X, V, prm)
    assert np.array_equal(Cb.view(np.uint32), Cref.view(np.uint32)), np.abs(Cb - Cref).max()
    assert np.array_equal(b.view(np.uint32), bref.view(np.uint32)), np.abs(b - bref).max()


@pytest.mark.parametrize("warm", [False, True])
def test_cg_solution_bitwise(o32, warm):
    sc = scenes.plain_scene()
    X, V = random_state(sc, 5)
    say("cg 21x21 warm", warm)
    p = scenes.StepParams(dt=1 / 60, cg_max_iters=25, cg_tol=1e-6, cg_tol_mode=scenes.TOL_ABSOLUTE,
                          warm_start=warm)
    rng = np.random.default_rng(9)
    dv0 = (1e-2 * rng.standard_normal(X.shape)).astype(np.float32)
    dv0[sc.pinned.astype(bool)] = 0
    from opencloth_amd.cloth import ImplicitCloth

    cl = ImplicitCloth(sc.nx, sc.ny, X, sc.springs, sc.pinned, p, V)
    cl.set_dv(dv0)
    cl.assemble()
    it, st = cl.solve_conjugate_gradient()
    dv = cl.get_dv()
    sys_ = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(p, sc.nx, sc.ny)
    B, Cp = sys_.assemble(X, V, prm)
    dvref, itref, stref, _ = sys_.cg(Cp, prm, B, dv0 if warm else None)
    say("iters gpu", it, "oracle", itref, "status", st, stref)
    assert it == itref and st == stref
    assert np.array_equal(dv.view(np.uint32), dvref.view(np.uint32)), np.abs(dv - dvref).max()


def test_plain_rollout_300_steps(o32, o64):
    """21x21, 300 steps: per-step max relative position error <= 1e-4 against
    the FP32 restatement (it is 0: the run is bit-exact) and <= 1e-5 against
    the FP64 restatement (about 1.2e-6 measured)."""
    sc = scenes.plain_scene()
    say("rollout 21x21 x300")
    cl = make_cloth(sc)
    sys32 = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    sys64 = o64.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(sc.params, sc.nx, sc.ny)
    _, _, _, h32, it32, st32 = sys32.rollout(sc.X, sc.V, prm, sc.steps, history=True)
    _, _, _, h64, it64, _ = sys64.rollout(sc.X, sc.V, oracle_params(sc.params, sc.nx, sc.ny, False),
                                          sc.steps, history=True)
    err32 = np.zeros(sc.steps); err64 = np.zeros(sc.steps)
    for k in range(sc.steps):
        cl.step(1)
        X = cl.X
        err32[k] = rel_pos_err(X, h32[k])
        err64[k] = rel_pos_err(X, h64[k])
    say("max rel err vs fp32 %.3e, vs fp64 %.3e" % (err32.max(), err64.max()))
    assert err32.max() <= 1e-4
    assert err64.max() <= 1e-5


def test_large_scene_two_steps(o32):
    """512x512, top edge pinned, 30 fixed CG iterations: two steps against the
    FP32 restatement with the kernels' reduction order."""
    sc = scenes.large_scene()
    say("large 512x512 x2")
    cl = make_cloth(sc)
    sys32 = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    prm = oracle_params(sc.params, sc.nx, sc.ny)
    X, V, DV, _, iters, st = sys32.rollout(sc.X, sc.V, prm, 2)
    cl.step(2)
    Xg, Vg = cl.get_state()
    it, stg = cl.last_status()
    say("err", rel_pos_err(Xg, X), "iters", it, iters, "status", stg, st)
    assert it == 30 and stg == 0
    assert rel_pos_err(Xg, X) <= 1e-6
    assert np.array_equal(Xg.view(np.uint32), X.view(np.uint32))


@pytest.mark.parametrize("scene", ["plain", "large"])
def test_persistent_cg_matches_multi_kernel(o32, scene):
    """OC_FLAG_CG_PERSISTENT (all CG iterations in one cooperative launch)
    runs the same per-tile arithmetic: two steps bit-identical to the FP32
    restatement and to the per-iteration launches."""
    import dataclasses

    from opencloth_amd.cloth import ImplicitCloth

    sc = scenes.plain_scene() if scene == "plain" else scenes.large_scene()
    say("persistent cg", scene)
    p = dataclasses.replace(sc.params, cg_persistent=True)
    cl = ImplicitCloth(sc.nx, sc.ny, sc.X, sc.springs, sc.pinned, p, sc.V)
    ref = make_cloth(sc)
    cl.step(2)
    ref.step(2)
    Xg, Vg = cl.get_state()
    Xm, Vm = ref.get_state()
    assert cl.last_status() == ref.last_status()
    assert np.array_equal(Xg.view(np.uint32), Xm.view(np.uint32))
    assert np.array_equal(Vg.view(np.uint32), Vm.view(np.uint32))
    sys32 = o32.system(sc.n_vertices, sc.springs, sc.pinned)
    X, V, _, _, iters, st = sys32.rollout(sc.X, sc.V, oracle_params(sc.params, sc.nx, sc.ny), 2)
    assert cl.last_status() == (int(np.ravel(iters)[-1]), int(np.ravel(st)[-1]))
    assert np.array_equal(Xg.view(np.uint32), X.view(np.uint32))
